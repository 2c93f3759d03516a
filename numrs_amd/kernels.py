"""Host-side mirror of the reference's `src/kernels` interface over the C ABI.

Names, argument meaning and error behaviour follow the reference's crate-private kernels module
(paths under the reference's src/), so that parity tests read like the reference's own tests:

    reduce.rs:207-421     sum_f64 / mean_f64 / sum_sq_dev_f64 / var_f64 / min_f64 / max_f64 /
                          prod_f64 and their _f32 twins -> a Python float
    gemm.rs:246-268       gemm_2d(m, k, n, a, b, c): overwrites c in place
    array/linalg.rs:56-310 batched_gemm(batch, m, k, n, a, b, c) and Array::matmul's batching
    elementwise.rs:51-258 binary_f64 / binary_f32 / unary_dispatch with op codes instead of closures
    operations.rs:303-318 sum_i64 (Array::<i64>::sum)

Arrays are numpy arrays standing in for Rust slices (host memory, logical order). Every call
goes to the GPU through libnumrs_amd.so; there is no CPU path here.
"""
from __future__ import annotations

import builtins
import ctypes as C

import numpy as np

from ._lib import ScalarOp, call

# NRK_SOP_* / NRK_BOP_* / NRK_AXIS_* codes of include/numrs_amd.h
SOP = {"add": 0, "sub": 1, "mul": 2, "div": 3, "rsub": 4, "rdiv": 5, "neg": 6, "abs": 7, "square": 8, "sqrt": 9,
       "floor": 10, "ceil": 11, "trunc": 12, "round": 13, "fract": 14, "sign": 15, "clamp_lo": 16, "clamp_hi": 17,
       "exp": 18, "exp2": 19, "expm1": 20, "log": 21, "log2": 22, "log10": 23, "log1p": 24, "tanh": 25,
       "sinh": 26, "cosh": 27, "asinh": 28, "acosh": 29, "atanh": 30, "atan": 31, "asin": 32, "acos": 33,
       "sin": 34, "cos": 35, "tan": 36, "relu": 37, "leaky_relu": 38, "elu": 39, "selu": 40,
       "hardsigmoid": 41, "hardswish": 42, "softplus": 43, "mish": 44, "silu": 45, "gelu": 46,
       "logit": 47, "sinc": 48, "smoothstep": 49, "smootherstep": 50, "erf": 51, "erfc": 52,
       "digamma": 53, "pow": 54, "gamma": 55, "lgamma": 56, "cbrt": 57}
BOP = {"add": 0, "sub": 1, "mul": 2, "div": 3, "rem": 4, "maximum": 5, "minimum": 6, "copysign": 7, "hypot": 8,
       "atan2": 9, "logaddexp": 10, "prelu": 11, "power": 12,
       "bitand": 13, "bitor": 14, "bitxor": 15, "shl": 16, "shr": 17}  # the last five: integers only
DTYPE = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int64): 2, np.dtype(np.int32): 3,
         np.dtype(np.uint64): 4, np.dtype(np.uint32): 5}
_INT_SUF = {np.dtype(np.int64): "i64", np.dtype(np.int32): "i32", np.dtype(np.uint64): "u64", np.dtype(np.uint32): "u32"}
AXIS = {"sum": 0, "mean": 1, "var": 2, "min": 3, "max": 4, "argmin": 5, "argmax": 6}

PARALLEL_MIN_LEN = 10_000  # kernels/mod.rs:74 (kept for callers; the GPU has no CPU tiers)
PARALLEL_CHUNK = 8_192  # kernels/mod.rs:92


def chain(*steps) -> tuple[C.Array, int]:
    """Build an nrk_scalar_op chain from ("mul", 1.5), ("add", -0.25), ("neg",) ... steps."""
    arr = (ScalarOp * max(1, len(steps)))()
    for i, st in enumerate(steps):
        name = st[0]
        arr[i].op = SOP[name]
        arr[i].reserved = 0
        arr[i].scalar = float(st[1]) if len(st) > 1 else 0.0
    return arr, len(steps)


def _vec(a, dtype) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype != np.dtype(dtype):
        raise TypeError(f"expected {np.dtype(dtype)}, got {a.dtype}")
    return a.reshape(-1)


def _ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def _scalar_reduce(fn: str, a, dtype, out_ctype, *extra):
    a = _vec(a, dtype)
    out = out_ctype()
    call(fn, _ptr(a), C.c_size_t(a.size), *extra, C.byref(out))
    return out.value


# ---- kernels/reduce.rs ----
def sum_f64(a) -> float: return _scalar_reduce("nrk_sum_f64", a, np.float64, C.c_double)
def sum_f32(a) -> float: return _scalar_reduce("nrk_sum_f32", a, np.float32, C.c_float)
def mean_f64(a) -> float: return _scalar_reduce("nrk_mean_f64", a, np.float64, C.c_double)
def mean_f32(a) -> float: return _scalar_reduce("nrk_mean_f32", a, np.float32, C.c_float)
def min_f64(a) -> float: return _scalar_reduce("nrk_min_f64", a, np.float64, C.c_double)
def min_f32(a) -> float: return _scalar_reduce("nrk_min_f32", a, np.float32, C.c_float)
def max_f64(a) -> float: return _scalar_reduce("nrk_max_f64", a, np.float64, C.c_double)
def max_f32(a) -> float: return _scalar_reduce("nrk_max_f32", a, np.float32, C.c_float)
def prod_f64(a) -> float: return _scalar_reduce("nrk_prod_f64", a, np.float64, C.c_double)
def prod_f32(a) -> float: return _scalar_reduce("nrk_prod_f32", a, np.float32, C.c_float)


def sum_sq_dev_f64(a, mean: float) -> float:
    return _scalar_reduce("nrk_sum_sq_dev_f64", a, np.float64, C.c_double, C.c_double(mean))


def sum_sq_dev_f32(a, mean: float) -> float:
    return _scalar_reduce("nrk_sum_sq_dev_f32", a, np.float32, C.c_float, C.c_float(mean))


def var_f64(a, ddof: int) -> float:
    return _scalar_reduce("nrk_var_f64", a, np.float64, C.c_double, C.c_size_t(ddof))


def var_f32(a, ddof: int) -> float:
    return _scalar_reduce("nrk_var_f32", a, np.float32, C.c_float, C.c_size_t(ddof))


# ---- ufuncs.rs:946-979, 1060 ----
def dot(a, b) -> float:
    """ufuncs.rs:946 dot: the sum of the rounded products; lengths must match (ShapeMismatch)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.size != b.size:
        raise ValueError(f"shape mismatch: {a.size} vs {b.size}")
    f32 = a.dtype == np.float32 and b.dtype == np.float32
    dt, ct = (np.float32, C.c_float) if f32 else (np.float64, C.c_double)
    av, bv = _vec(a, dt), _vec(b, dt)
    out = ct()
    call("nrk_dot_f32" if f32 else "nrk_dot_f64", _ptr(av), _ptr(bv), C.c_size_t(av.size), C.byref(out))
    return out.value


def _float_reduce(name, a, *extra):
    a = np.asarray(a)
    f32 = a.dtype == np.float32
    return _scalar_reduce(f"nrk_{name}_{'f32' if f32 else 'f64'}", a, np.float32 if f32 else np.float64,
                          C.c_float if f32 else C.c_double, *extra)


def norm_l1(a) -> float: return _float_reduce("norm_l1", a)   # :975 sum |x|
def norm_l2(a) -> float: return _float_reduce("norm_l2", a)   # :966 sqrt(sum x * x)
def std(a, ddof: int = 0) -> float: return _float_reduce("std", a, C.c_size_t(ddof))  # :1060 var.sqrt()


def sum_i64(a) -> int:
    """Array::<i64>::sum (operations.rs:303-318): wrapping integer fold."""
    return _scalar_reduce("nrk_sum_i64", a, np.int64, C.c_int64)


def argmax_f64(a) -> int: return _scalar_reduce("nrk_argmax_f64", a, np.float64, C.c_uint64)
def argmax_f32(a) -> int: return _scalar_reduce("nrk_argmax_f32", a, np.float32, C.c_uint64)
def argmin_f64(a) -> int: return _scalar_reduce("nrk_argmin_f64", a, np.float64, C.c_uint64)
def argmin_f32(a) -> int: return _scalar_reduce("nrk_argmin_f32", a, np.float32, C.c_uint64)


def reduce_axis(op: str, x: np.ndarray, outer: int, length: int, inner: int, ddof: int = 0) -> np.ndarray:
    """Reduce the (outer, length, inner) view of x along `length`; returns outer*inner values."""
    x = np.ascontiguousarray(x)
    if x.dtype not in (np.float64, np.float32):
        raise TypeError("reduce_axis takes f64 or f32")
    out_dtype = np.uint64 if op in ("argmax", "argmin") else x.dtype
    out = np.empty(outer * inner, out_dtype)
    fn = "nrk_reduce_axis_f64" if x.dtype == np.float64 else "nrk_reduce_axis_f32"
    call(fn, C.c_int32(AXIS[op]), _ptr(x), C.c_size_t(outer), C.c_size_t(length), C.c_size_t(inner),
         C.c_size_t(ddof), _ptr(out))
    return out


# ---- kernels/elementwise.rs ----
def unary_dispatch(a, steps) -> np.ndarray:
    """unary_dispatch (elementwise.rs:246-258) with an op-coded closure: chain(("mul", 1.5), ...)."""
    a = np.ascontiguousarray(a)
    ops, n_ops = chain(*steps)
    out = np.empty_like(a)
    fn = {np.dtype(np.float64): "nrk_map_f64", np.dtype(np.float32): "nrk_map_f32"}[a.dtype]
    call(fn, C.cast(ops, C.c_void_p), C.c_int32(n_ops), _ptr(a), _ptr(out), C.c_size_t(a.size))
    return out


# ---- ufuncs.rs unary functions (f64 in the reference; f32 arrays are accepted too) as op chains.
# The Rust expression each mirrors is quoted; chains of IEEE steps round exactly like it.

def _ufunc(a, *steps) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    return unary_dispatch(a, list(steps)).reshape(a.shape)


def exp(a): return _ufunc(a, ("exp",))                       # ufuncs.rs:398 x.exp()
def exp2(a): return _ufunc(a, ("exp2",))                     # :408 x.exp2()
def expm1(a): return _ufunc(a, ("expm1",))                   # :418 x.exp_m1()
def log(a): return _ufunc(a, ("log",))                       # :428 x.ln()
def ln(a): return log(a)                                     # :438
def log2(a): return _ufunc(a, ("log2",))                     # :443 x.log2()
def log10(a): return _ufunc(a, ("log10",))                   # :453 x.log10()
def log1p(a): return _ufunc(a, ("log1p",))                   # :463 x.ln_1p()
def tanh(a): return _ufunc(a, ("tanh",))                     # :639 x.tanh()
def sinh(a): return _ufunc(a, ("sinh",))                     # :619 x.sinh()
def cosh(a): return _ufunc(a, ("cosh",))                     # :629 x.cosh()
def arcsinh(a): return _ufunc(a, ("asinh",))                 # :653 x.asinh()
def arccosh(a): return _ufunc(a, ("acosh",))                 # :668 x.acosh()
def arctanh(a): return _ufunc(a, ("atanh",))                 # :683 x.atanh()
asinh, acosh, atanh = arcsinh, arccosh, arctanh
def arctan(a): return _ufunc(a, ("atan",))                   # :572 x.atan()
def arcsin(a): return _ufunc(a, ("asin",))                   # :542 x.asin()
def arccos(a): return _ufunc(a, ("acos",))                   # :557 x.acos()
atan, asin, acos = arctan, arcsin, arccos
def sin(a): return _ufunc(a, ("sin",))                       # :477 x.sin()
def cos(a): return _ufunc(a, ("cos",))                       # :487 x.cos()
def tan(a): return _ufunc(a, ("tan",))                       # :497 x.tan()
def sincos(a): return sin(a), cos(a)                         # :507 (x.sin(), x.cos())
def sigmoid(a): return _ufunc(a, ("neg",), ("exp",), ("add", 1.0), ("rdiv", 1.0))  # :1235 1 / (1 + (-x).exp())
def reciprocal(a): return _ufunc(a, ("rdiv", 1.0))           # :374 1.0 / x
def rsqrt(a): return _ufunc(a, ("sqrt",), ("rdiv", 1.0))     # :384 1.0 / x.sqrt()
def sqrt(a): return _ufunc(a, ("sqrt",))                     # :354 x.sqrt()
def square(a): return _ufunc(a, ("square",))                 # :344 x * x
def cbrt(a): return _ufunc(a, ("cbrt",))                     # :367 x.cbrt()
def negative(a): return _ufunc(a, ("neg",))                  # :319 -x
def absolute(a): return _ufunc(a, ("abs",))                  # :329 x.abs()
def floor(a): return _ufunc(a, ("floor",))                   # :702 x.floor()
def ceil(a): return _ufunc(a, ("ceil",))                     # :712 x.ceil()
def round(a): return _ufunc(a, ("round",))                   # :722 x.round(), half away from zero  # noqa: A001
def trunc(a): return _ufunc(a, ("trunc",))                   # :732 x.trunc()
def fract(a): return _ufunc(a, ("fract",))                   # :742 x - x.trunc()
def sign(a): return _ufunc(a, ("sign",))                     # :756 0 / 1 / -1 (NaN -> -1)


def _bufunc(a, b, op: str) -> np.ndarray:
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch: {a.shape} vs {b.shape}")  # the reference's ShapeMismatch
    dt = np.float32 if a.dtype == np.float32 and b.dtype == np.float32 else np.float64
    fn = "nrk_zip_f32" if dt == np.float32 else "nrk_zip_f64"
    return _binary(fn, a.reshape(-1), b.reshape(-1), op, dt).reshape(a.shape)


def copysign(mag, sgn): return _bufunc(mag, sgn, "copysign")   # ufuncs.rs:784 m.copysign(s)
def hypot(a, b): return _bufunc(a, b, "hypot")                  # :841 a.hypot(b)
def arctan2(y, x): return _bufunc(y, x, "atan2")                # :587 y.atan2(x)
def logaddexp(a, b): return _bufunc(a, b, "logaddexp")          # :916
atan2 = arctan2


def fma(a, b, c):
    """ufuncs.rs:864 a.mul_add(b, c): one rounding; shapes must match (ShapeMismatch)."""
    a, b, c = np.asarray(a), np.asarray(b), np.asarray(c)
    if not (a.shape == b.shape == c.shape):
        raise ValueError(f"shape mismatch: {a.shape}, {b.shape}, {c.shape}")
    f32 = a.dtype == b.dtype == c.dtype == np.float32
    dt = np.float32 if f32 else np.float64
    av, bv, cv = _vec(a, dt), _vec(b, dt), _vec(c, dt)
    y = np.empty_like(av)
    call("nrk_fma_f32" if f32 else "nrk_fma_f64", _ptr(av), _ptr(bv), _ptr(cv), _ptr(y), C.c_size_t(av.size))
    return y.reshape(a.shape)


def lerp(a, b, t: float):
    """ufuncs.rs:893 a + t * (b - a): three roundings; shapes must match (ShapeMismatch)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch: {a.shape} vs {b.shape}")
    f32 = a.dtype == b.dtype == np.float32
    dt = np.float32 if f32 else np.float64
    av, bv = _vec(a, dt), _vec(b, dt)
    y = np.empty_like(av)
    call("nrk_lerp_f32" if f32 else "nrk_lerp_f64", _ptr(av), _ptr(bv), (C.c_float if f32 else C.c_double)(t), _ptr(y),
         C.c_size_t(av.size))
    return y.reshape(a.shape)


_ACC_OPS = {"add": "add", "subtract": "sub", "multiply": "mul", "divide": "div", "maximum": "maximum",
            "minimum": "minimum"}


def accumulate(op: str, a, axis=None):
    """ufunc_accumulate (ufunc_ops.rs:535): op in add / subtract / multiply / divide / maximum /
    minimum; axis None scans the flattened array, otherwise that axis (moved last for the call)."""
    a = np.asarray(a)
    f32 = a.dtype == np.float32
    dt = np.float32 if f32 else np.float64
    if axis is None:
        x = _vec(a, dt)
        rows, length = 1, x.size
        back = lambda y: y  # noqa: E731 (flat, as the reference returns it)
    else:
        ax = axis % a.ndim
        moved = np.ascontiguousarray(np.moveaxis(a.astype(dt, copy=False), ax, -1))
        x = moved.reshape(-1)
        length = moved.shape[-1] if moved.ndim else 1
        rows = x.size // length if length else 0
        back = lambda y: np.moveaxis(y.reshape(moved.shape), -1, ax)  # noqa: E731
    y = np.empty_like(x)
    call("nrk_accumulate_f32" if f32 else "nrk_accumulate_f64", C.c_int32(BOP[_ACC_OPS[op]]), _ptr(x),
         C.c_size_t(rows), C.c_size_t(length), _ptr(y))
    return back(y)


def _ufop_dtype(*arrs):
    f32 = all(np.asarray(x).dtype == np.float32 for x in arrs)
    return (np.float32, "f32") if f32 else (np.float64, "f64")


def outer(op: str, a, b):
    """ufunc_outer (ufunc_ops.rs:583): y[i, j] = a[i] op b[j], shape a.shape + b.shape."""
    a, b = np.asarray(a), np.asarray(b)
    dt, suf = _ufop_dtype(a, b)
    av, bv = _vec(a, dt), _vec(b, dt)
    y = np.empty(av.size * bv.size, dt)
    call(f"nrk_outer_{suf}", C.c_int32(BOP[_ACC_OPS[op]]), _ptr(av), C.c_size_t(av.size), _ptr(bv),
         C.c_size_t(bv.size), _ptr(y))
    return y.reshape(a.shape + b.shape)


def where(op: str, a, b, mask):
    """ufunc_where (ufunc_ops.rs:841): a op b where mask, else a; shapes must match."""
    a, b, m = np.asarray(a), np.asarray(b), np.asarray(mask)
    if a.shape != b.shape or a.shape != m.shape:
        raise ValueError(f"shape mismatch: {a.shape}, {b.shape}, {m.shape}")
    dt, suf = _ufop_dtype(a, b)
    av, bv = _vec(a, dt), _vec(b, dt)
    mv = np.ascontiguousarray(m.reshape(-1), dtype=np.uint8)
    y = np.empty_like(av)
    call(f"nrk_where_{suf}", C.c_int32(BOP[_ACC_OPS[op]]), _ptr(av), _ptr(bv), _ptr(mv), _ptr(y), C.c_size_t(av.size))
    return y.reshape(a.shape)


def add_where(a, b, mask): return where("add", a, b, mask)            # ufunc_ops.rs:893
def subtract_where(a, b, mask): return where("subtract", a, b, mask)  # :902
def multiply_where(a, b, mask): return where("multiply", a, b, mask)  # :911
def divide_where(a, b, mask): return where("divide", a, b, mask)      # :920


def at(op: str, a: np.ndarray, indices, b) -> None:
    """ufunc_at (ufunc_ops.rs:772), in place on the caller's contiguous float array a: for each k in
    order, a[indices[k]] = a[indices[k]] op b[k]; b has shape (len(indices),) + a.shape[1:]."""
    if not isinstance(a, np.ndarray) or a.ndim == 0 or not a.flags.c_contiguous or a.dtype not in (np.float32, np.float64):
        raise ValueError("ufunc_at requires a contiguous float array with at least 1 dimension")
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64).reshape(-1))
    b = np.ascontiguousarray(np.asarray(b, dtype=a.dtype))
    if b.shape != (idx.size,) + a.shape[1:]:
        raise ValueError(f"shape mismatch: b is {b.shape}, expected {(idx.size,) + a.shape[1:]}")
    row_len = int(np.prod(a.shape[1:], dtype=np.int64))
    call("nrk_at_f32" if a.dtype == np.float32 else "nrk_at_f64", C.c_int32(BOP[_ACC_OPS[op]]), _ptr(a),
         C.c_size_t(a.shape[0]), C.c_size_t(row_len), _ptr(idx), C.c_size_t(idx.size), _ptr(b))


def reduceat(op: str, a, indices, axis=None):
    """ufunc_reduceat (ufunc_ops.rs:713): numpy's op.reduceat segment rule along axis (None: the
    flattened array); an index out of bounds raises."""
    a = np.asarray(a)
    f32 = a.dtype == np.float32
    dt = np.float32 if f32 else np.float64
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64).reshape(-1))
    if axis is None:
        x, rows, length = _vec(a, dt), 1, a.size
        back = lambda y: y  # noqa: E731
    else:
        ax = axis % a.ndim
        moved = np.ascontiguousarray(np.moveaxis(a.astype(dt, copy=False), ax, -1))
        x, length = moved.reshape(-1), moved.shape[-1]
        rows = x.size // length if length else 0
        back = lambda y: np.moveaxis(y.reshape(moved.shape[:-1] + (idx.size,)), -1, ax)  # noqa: E731
    y = np.empty(rows * idx.size, dt)
    call("nrk_reduceat_f32" if f32 else "nrk_reduceat_f64", C.c_int32(BOP[_ACC_OPS[op]]), _ptr(x), C.c_size_t(rows),
         C.c_size_t(length), _ptr(idx), C.c_size_t(idx.size), _ptr(y))
    return back(y)


def cumsum(a, axis=None): return accumulate("add", a, axis)          # statistics/mod.rs:584
def cumprod(a, axis=None): return accumulate("multiply", a, axis)    # statistics/mod.rs:795


def _softmax(fn, a):
    a = np.asarray(a)
    f32 = a.dtype == np.float32
    x = _vec(a, np.float32 if f32 else np.float64)
    y = np.empty_like(x)
    call(f"nrk_{fn}_{'f32' if f32 else 'f64'}", _ptr(x), _ptr(y), C.c_size_t(x.size))
    return y.reshape(a.shape)


def softmax(a): return _softmax("softmax", a)                  # ufuncs.rs:1434 over all elements
def log_softmax(a): return _softmax("log_softmax", a)          # ufuncs.rs:1411


def prelu(a, alpha): return _bufunc(a, alpha, "prelu")  # :1394 x > 0 ? x : alpha * x


# ufuncs.rs:182-303 arithmetic by name (the ufunc closures :150-173, one rounding each), their
# scalar forms (Array::*_scalar, operations.rs:448-474), and the flat f64 reductions (:990-1060)
def add(a, b): return _bufunc(a, b, "add")
def subtract(a, b): return _bufunc(a, b, "sub")
def multiply(a, b): return _bufunc(a, b, "mul")
def divide(a, b): return _bufunc(a, b, "div")
def maximum(a, b): return _bufunc(a, b, "maximum")            # :170 if a > b {a} else {b}
def minimum(a, b): return _bufunc(a, b, "minimum")            # :174
def add_scalar(a, s: float): return _ufunc(a, ("add", s))
def subtract_scalar(a, s: float): return _ufunc(a, ("sub", s))
def multiply_scalar(a, s: float): return _ufunc(a, ("mul", s))
def divide_scalar(a, s: float): return _ufunc(a, ("div", s))
abs = absolute  # noqa: A001
def sum(a): return sum_f64(np.asarray(a, np.float64).reshape(-1))    # noqa: A001
def mean(a): return mean_f64(np.asarray(a, np.float64).reshape(-1))
def var(a): return var_f64(np.asarray(a, np.float64).reshape(-1), 0)


def power(a, b):
    """ufuncs.rs:226 power through the power ufunc (:166 a.powf(b)), numpy-style broadcasting."""
    a, b = np.broadcast_arrays(np.asarray(a), np.asarray(b))
    return _bufunc(np.ascontiguousarray(a), np.ascontiguousarray(b), "power")


def power_scalar(a, b: float): return _ufunc(a, ("pow", b))  # :305 x.powf(b)
def gamma(a): return _ufunc(a, ("gamma",))                    # :1122 Lanczos, reflection below 0.5
def lgamma(a): return _ufunc(a, ("lgamma",))                  # :1183 gamma_single(x).abs().ln()
gammaln = lgamma                                              # :1193          # :1394 x > 0 ? x : alpha * x


# activations, special functions and steps: the reference's f64 expressions, one rounding per
# operation (ufuncs.rs:520-1480)
def relu(a): return _ufunc(a, ("relu",))                     # :1380 x.max(0.0)
def leaky_relu(a, alpha: float = 0.01): return _ufunc(a, ("leaky_relu", alpha))  # :1390
def elu(a, alpha: float = 1.0): return _ufunc(a, ("elu", alpha))  # :1322
def selu(a): return _ufunc(a, ("selu",))                     # :1334
def hardsigmoid(a): return _ufunc(a, ("hardsigmoid",))       # :1344
def hardswish(a): return _ufunc(a, ("hardswish",))           # :1362
def softplus(a): return _ufunc(a, ("softplus",))             # :1291
def mish(a): return _ufunc(a, ("mish",))                     # :1309
def swish(a): return _ufunc(a, ("silu",))                    # :1275 x / (1 + (-x).exp())
silu = swish                                                 # :1279
def gelu(a): return _ufunc(a, ("gelu",))                     # :1262 (tanh form)
def logit(a): return _ufunc(a, ("logit",))                   # :1251 (x / (1 - x)).ln()
def sinc(a): return _ufunc(a, ("sinc",))                     # :527 sin(pi x) / (pi x)
def erf(a): return _ufunc(a, ("erf",))                       # :1081 erf_scalar
def erfc(a): return _ufunc(a, ("erfc",))                     # :1118 1 - erf_scalar(x)
def digamma(a): return _ufunc(a, ("digamma",))               # :1198


def _step(op, edge0: float, edge1: float, x):
    """t = ((v - edge0) / (edge1 - edge0)).clamp(0.0, 1.0) (ufuncs.rs:1467, 1480): edge1 - edge0
    is rounded once, as the reference computes it; clamp's NaN passes through."""
    return _ufunc(x, ("sub", edge0), ("div", float(np.float64(edge1) - np.float64(edge0))),
                  ("clamp_lo", 0.0), ("clamp_hi", 1.0), (op,))


def smoothstep(edge0, edge1, x): return _step("smoothstep", edge0, edge1, x)       # :1460
def smootherstep(edge0, edge1, x): return _step("smootherstep", edge0, edge1, x)   # :1473


def deg2rad(a):
    """ufuncs.rs:807 x.to_radians() = x * (pi / 180): one multiply by the rounded constant."""
    return _ufunc(a, ("mul", float(np.pi) / 180.0))


def rad2deg(a):
    """ufuncs.rs:822 x.to_degrees() = x * (180 / pi): one multiply by the rounded constant."""
    return _ufunc(a, ("mul", 180.0 / float(np.pi)))


radians, degrees = deg2rad, rad2deg


def clip(a, lo: float, hi: float):
    """ufuncs.rs:774 x.clamp(min, max): NaN stays NaN; Rust panics when min > max or either bound
    is NaN, which raises ValueError here."""
    if not (lo <= hi):
        raise ValueError(f"clip bounds must satisfy min <= max, got {lo}, {hi}")
    return _ufunc(a, ("clamp_lo", lo), ("clamp_hi", hi))


def binary_f64(a, b, op: str) -> np.ndarray:
    """binary_f64 (elementwise.rs:70-89) with an op-coded closure."""
    return _binary("nrk_zip_f64", a, b, op, np.float64)


def binary_f32(a, b, op: str) -> np.ndarray:
    """binary_f32 (elementwise.rs:90-109) with an op-coded closure."""
    return _binary("nrk_zip_f32", a, b, op, np.float32)


class IntOp(C.Structure):
    """nrk_int_op"""
    _fields_ = [("op", C.c_int32), ("reserved", C.c_int32), ("scalar", C.c_int64)]


def binary_dispatch(a, b, op: str) -> np.ndarray:
    """binary_dispatch (elementwise.rs:217-235) for any T the library runs: f64 / f32 through the
    float zip, i64 / i32 / u64 / u32 through the integer zip (wrapping, truncating division,
    masked shifts; a zero divisor raises like the reference's panic). Equal lengths."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype in (np.float64, np.float32):
        return _binary("nrk_zip_f64" if a.dtype == np.float64 else "nrk_zip_f32", a, b, op, a.dtype)
    suf = _INT_SUF.get(a.dtype)
    if suf is None:
        raise TypeError(f"no integer kernel for {a.dtype}")
    return _binary(f"nrk_zip_{suf}", a, b, op, a.dtype)


def unary_dispatch_int(a, steps) -> np.ndarray:
    """unary_dispatch (elementwise.rs:237-258) for the integer T: steps like ("mul", 3), ("neg",)
    with integer scalars; + - * and negation wrap as Rust's release arithmetic does."""
    a = np.ascontiguousarray(a)
    suf = _INT_SUF.get(a.dtype)
    if suf is None:
        raise TypeError(f"no integer kernel for {a.dtype}")
    arr = (IntOp * max(1, len(steps)))()
    for i, st in enumerate(steps):
        arr[i].op = SOP[st[0]]
        arr[i].reserved = 0
        arr[i].scalar = int(st[1]) if len(st) > 1 else 0
    out = np.empty_like(a)
    call(f"nrk_map_{suf}", C.cast(arr, C.c_void_p), C.c_int32(len(steps)), _ptr(a), _ptr(out), C.c_size_t(a.size))
    return out


def _broadcast_op(a, b, op: str) -> np.ndarray:
    """Array::broadcast_op + binary_serial (operations.rs:260-284, 370-440): y = a op b over the
    numpy broadcast of the two shapes, one GPU pass (no broadcast_to copies)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype != b.dtype:
        raise TypeError("operands of a broadcast op share one dtype")
    if a.dtype not in DTYPE:
        raise TypeError(f"no kernel for {a.dtype}")
    try:
        shape = np.broadcast_shapes(a.shape, b.shape)
    except ValueError as e:
        raise ValueError(f"shape mismatch: {a.shape} vs {b.shape}") from e
    out = np.empty(shape, a.dtype)
    sa = (C.c_size_t * max(1, a.ndim))(*a.shape)
    sb = (C.c_size_t * max(1, b.ndim))(*b.shape)
    call("nrk_zip_broadcast", C.c_int32(DTYPE[a.dtype]), C.c_int32(BOP[op]), _ptr(a), C.c_int32(a.ndim), sa, _ptr(b),
         C.c_int32(b.ndim), sb, _ptr(out))
    return out


def add_broadcast(a, b): return _broadcast_op(a, b, "add")          # operations.rs:370
def subtract_broadcast(a, b): return _broadcast_op(a, b, "sub")     # operations.rs:392
def multiply_broadcast(a, b): return _broadcast_op(a, b, "mul")     # operations.rs:414
def divide_broadcast(a, b): return _broadcast_op(a, b, "div")       # operations.rs:436


def zip_strided(op: str, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """nrk_dev_zip_strided on numpy views as they are (any strides, 0 for broadcast axes), so a
    transposed or sliced operand is read in place; b may be another integer dtype for shifts
    (bitwise_ops.rs:190-243 left_shift<T, U> / right_shift). The result has a's dtype."""
    from . import device

    a, b = np.asarray(a), np.asarray(b)
    shape = np.broadcast_shapes(a.shape, b.shape)
    av, bv = np.broadcast_to(a, shape), np.broadcast_to(b, shape)
    out = np.empty(shape, a.dtype)
    if out.size == 0:
        return out

    def upload(v):  # the smallest buffer that holds every element the view can address
        lo = builtins.sum(builtins.min(0, st * (n - 1)) for st, n in zip(v.strides, v.shape))
        hi = builtins.sum(builtins.max(0, st * (n - 1)) for st, n in zip(v.strides, v.shape)) + v.itemsize
        raw = np.ctypeslib.as_array((C.c_uint8 * (hi - lo)).from_address(v.ctypes.data + lo)).copy()
        d = device.DeviceArray((raw.size,), np.uint8)
        d.fill_from(raw)
        return d, -lo

    da, offa = upload(av)
    db, offb = upload(bv)
    dy = device.DeviceArray(shape, a.dtype)
    nd = len(shape)
    shp = (C.c_size_t * max(1, nd))(*shape)
    stra = (C.c_int64 * max(1, nd))(*[st // av.itemsize for st in av.strides])
    strb = (C.c_int64 * max(1, nd))(*[st // bv.itemsize for st in bv.strides])
    call("nrk_dev_zip_strided", C.c_int32(DTYPE[a.dtype]), C.c_int32(BOP[op]), C.c_int32(nd), shp,
         C.c_void_p(da.ptr + offa), stra, C.c_void_p(db.ptr + offb), C.c_int32(DTYPE[b.dtype]), strb, C.c_void_p(dy.ptr), None)
    device.synchronize()
    return dy.download()


def left_shift(a, b): return zip_strided("shl", a, b)     # bitwise_ops.rs:190
def right_shift(a, b): return zip_strided("shr", a, b)    # bitwise_ops.rs:228


def _binary(fn, a, b, op, dtype):
    a = _vec(a, dtype)
    b = _vec(b, dtype)
    if a.size != b.size:
        raise ValueError("binary kernels require equal-length slices")
    out = np.empty_like(a)
    call(fn, C.c_int32(BOP[op]), _ptr(a), _ptr(b), _ptr(out), C.c_size_t(a.size))
    return out


# ---- kernels/gemm.rs ----
def gemm_2d(m: int, k: int, n: int, a, b, c: np.ndarray) -> None:
    """gemm_2d (gemm.rs:246-268): C := A*B into the caller's c (overwrite). Slice lengths that
    do not match m*k, k*n, m*n zero c and read nothing, like gemm_generic's defensive branch."""
    if not isinstance(c, np.ndarray) or not c.flags.c_contiguous:
        raise TypeError("c must be a C-contiguous numpy array")
    dtype = c.dtype
    a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    b = np.ascontiguousarray(b, dtype=dtype).reshape(-1)
    fn = {np.dtype(np.float64): "nrk_gemm_2d_f64", np.dtype(np.float32): "nrk_gemm_2d_f32"}[dtype]
    call(fn, C.c_size_t(m), C.c_size_t(k), C.c_size_t(n), _ptr(a), C.c_size_t(a.size), _ptr(b), C.c_size_t(b.size),
         _ptr(c), C.c_size_t(c.size))


def batched_gemm(batch: int, m: int, k: int, n: int, a, b, c: np.ndarray) -> None:
    """batched_gemm (array/linalg.rs:56-136): C[t] := A[t] B[t] over contiguous panels, into the
    caller's c. batch, m or n = 0, or slice lengths other than batch*m*k / batch*k*n /
    batch*m*n, leave c untouched; otherwise panel t is gemm_2d's result on that panel."""
    if not isinstance(c, np.ndarray) or not c.flags.c_contiguous:
        raise TypeError("c must be a C-contiguous numpy array")
    dtype = c.dtype
    a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    b = np.ascontiguousarray(b, dtype=dtype).reshape(-1)
    fn = {np.dtype(np.float64): "nrk_batched_gemm_f64", np.dtype(np.float32): "nrk_batched_gemm_f32"}[dtype]
    call(fn, C.c_size_t(batch), C.c_size_t(m), C.c_size_t(k), C.c_size_t(n), _ptr(a), C.c_size_t(a.size), _ptr(b),
         C.c_size_t(b.size), _ptr(c), C.c_size_t(c.size))


def matmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Array::matmul (array/linalg.rs:150-250) for f64 / f32 (MFMA) and the integer types
    (gemm_generic's wrapping kernel, gemm.rs:408-445): 2-D x 2-D is gemm_2d; a 1-D operand
    is promoted to [1, k] (left) or [k, 1] (right) and the promoted axis stays in the result;
    batch axes (all but the last two) broadcast numpy-style. Instead of broadcast_to copies, an
    operand whose batch axes are all broadcast is passed once with panel stride 0; batch shapes
    that broadcast on both sides go panel by panel, each operand indexed in place. Errors (the
    reference's DimensionMismatch / ShapeMismatch) raise ValueError."""
    from . import device
    from ._lib import call as _call

    a = np.asarray(a)
    b = np.asarray(b)
    dtype = np.result_type(a.dtype, b.dtype)
    cdt = _compute_dtype(dtype)  # integers: gemm_generic in the 32- or 64-bit wrapping kernel
    if cdt != dtype:
        return matmul(a.astype(cdt), b.astype(cdt)).astype(dtype)
    if a.ndim == 2 and b.ndim == 2 and dtype in (np.float64, np.float32):
        if a.shape[1] != b.shape[0]:
            raise ValueError(f"shape mismatch: {a.shape} x {b.shape}")
        c = np.empty((a.shape[0], b.shape[1]), dtype)
        gemm_2d(a.shape[0], a.shape[1], b.shape[1], a, b, c)
        return c
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if b.ndim == 1:
        b = b.reshape(-1, 1)
    if a.ndim < 2 or b.ndim < 2:
        raise ValueError("matmul requires operands with at least 1 dimension")
    m, k = a.shape[-2:]
    k2, n = b.shape[-2:]
    if k != k2:
        raise ValueError(f"shape mismatch: {a.shape} x {b.shape}")
    ba, bb = a.shape[:-2], b.shape[:-2]
    try:
        bshape = tuple(np.broadcast_shapes(ba, bb))
    except ValueError as e:
        raise ValueError(f"batch shapes {ba} and {bb} do not broadcast") from e
    batch = int(np.prod(bshape, dtype=np.int64))
    out = np.empty(bshape + (m, n), dtype)
    if batch == 0 or m == 0 or n == 0:
        return out
    a = np.ascontiguousarray(a, dtype)
    b = np.ascontiguousarray(b, dtype)
    fn = {np.dtype(np.float64): "nrk_dev_batched_gemm_f64", np.dtype(np.float32): "nrk_dev_batched_gemm_f32",
          np.dtype(np.int64): "nrk_dev_gemm_generic_i64", np.dtype(np.int32): "nrk_dev_gemm_generic_i32"}[np.dtype(dtype)]
    da, db = device.DeviceArray.upload(a.reshape(-1)), device.DeviceArray.upload(b.reshape(-1))
    dc = device.DeviceArray((out.size,), dtype)
    isz = np.dtype(dtype).itemsize

    def run(nb, pa, sa, pb, sb, pc):
        _call(fn, C.c_size_t(nb), C.c_size_t(m), C.c_size_t(k), C.c_size_t(n), C.c_void_p(pa), C.c_size_t(sa),
              C.c_void_p(pb), C.c_size_t(sb), C.c_void_p(pc), None)

    def stride(sh, panel):  # one call over the whole batch: full batch shape, or a single panel
        if sh == bshape:
            return panel
        if int(np.prod(sh, dtype=np.int64)) == 1:
            return 0
        return None

    sa, sb = stride(ba, m * k), stride(bb, k * n)
    if sa is not None and sb is not None:
        run(batch, da.ptr, sa, db.ptr, sb, dc.ptr)
    else:  # both sides broadcast along different axes: panel by panel, operands in place
        ia = np.broadcast_to(np.arange(int(np.prod(ba, dtype=np.int64))).reshape(ba), bshape).reshape(-1)
        ib = np.broadcast_to(np.arange(int(np.prod(bb, dtype=np.int64))).reshape(bb), bshape).reshape(-1)
        for t in range(batch):
            run(1, da.ptr + int(ia[t]) * m * k * isz, 0, db.ptr + int(ib[t]) * k * n * isz, 0, dc.ptr + t * m * n * isz)
    device.synchronize()
    res = dc.download().reshape(out.shape)
    del da, db, dc
    return res


def _compute_dtype(dtype) -> np.dtype:
    """The kernel a matmul of `dtype` runs in: f64 / f32 natively; integers in the wrapping 64-bit
    (64-bit types) or 32-bit (narrower types) gemm_generic kernel -- the product modulo 2^32 taken
    modulo 2^w is the product modulo 2^w, so casting back gives the reference's wrapped w-bit
    result. Anything else is refused."""
    dtype = np.dtype(dtype)
    if dtype in (np.float64, np.float32):
        return dtype
    if np.issubdtype(dtype, np.integer):
        return np.dtype(np.int64) if dtype.itemsize == 8 else np.dtype(np.int32)
    raise TypeError(f"matmul has no GPU kernel for {dtype}")


def gemm_generic(m: int, k: int, n: int, a, b, c: np.ndarray) -> None:
    """gemm_generic (gemm.rs:408-445) for integer c (i64 / u64 in the 64-bit kernel, i32 / u32
    in the 32-bit one): c is zeroed over its own length; if a, b or c is shorter than m*k, k*n,
    m*n it stays zero, otherwise the leading parts are multiplied, wrapping."""
    if not isinstance(c, np.ndarray) or not c.flags.c_contiguous:
        raise TypeError("c must be a C-contiguous numpy array")
    if c.dtype.itemsize not in (4, 8) or not np.issubdtype(c.dtype, np.integer):
        raise TypeError("gemm_generic takes 32- or 64-bit integers")
    w = np.int64 if c.dtype.itemsize == 8 else np.int32
    a = np.ascontiguousarray(np.asarray(a).astype(c.dtype, copy=False)).reshape(-1).view(w)
    b = np.ascontiguousarray(np.asarray(b).astype(c.dtype, copy=False)).reshape(-1).view(w)
    call("nrk_gemm_generic_i64" if w == np.int64 else "nrk_gemm_generic_i32", C.c_size_t(m), C.c_size_t(k), C.c_size_t(n),
         _ptr(a), C.c_size_t(a.size), _ptr(b), C.c_size_t(b.size), _ptr(c), C.c_size_t(c.size))


def gemm_f32_bf16x(a: np.ndarray, b: np.ndarray, products: int = 3) -> np.ndarray:
    """FP32 GEMM as BF16 split products (3 or 6 MFMA products; 1 = inputs rounded to bf16)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    m, k = a.shape
    n = b.shape[1]
    c = np.empty((m, n), np.float32)
    call("nrk_gemm_f32_bf16x", C.c_size_t(m), C.c_size_t(k), C.c_size_t(n), _ptr(a), _ptr(b), _ptr(c), C.c_int32(products))
    return c


def gemm_f64_ozaki(a: np.ndarray, b: np.ndarray, moduli: int = 0) -> np.ndarray:
    """FP64 GEMM by the Ozaki scheme II on INT8 matrix cores (moduli 0 = default 16)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    m, k = a.shape
    n = b.shape[1]
    c = np.empty((m, n), np.float64)
    call("nrk_gemm_f64_ozaki", C.c_size_t(m), C.c_size_t(k), C.c_size_t(n), _ptr(a), _ptr(b), _ptr(c), C.c_int32(moduli))
    return c


# ---- expr/owned.rs ExprNode trees and expr/fused_eval.rs eval ----
class ExprNodeC(C.Structure):
    """nrk_expr_node"""
    _fields_ = [("kind", C.c_int32), ("op", C.c_int32), ("scalar", C.c_double)]


_EXPR_BIN = {"add": 0, "sub": 1, "mul": 2, "div": 3}
_EXPR_RHS = {"add": 0, "sub": 1, "mul": 2, "div": 3}       # x op s
_EXPR_LHS = {"add": 0, "sub": 4, "mul": 2, "div": 5}       # s op x (s + x = x + s, s * x = x * s)
_EXPR_UNARY = {"neg": 6, "abs": 7, "sqrt": 9, "exp": 18, "ln": 21}


class Expr:
    """An ExprNode tree (expr/owned.rs:204-223): Leaf, Binary, ScalarRhs, ScalarLhs, Unary, Fma,
    built with the operators (Expr.leaf(a) * Expr.leaf(b) + Expr.leaf(c), e * 2.0, 1.0 - e, -e,
    e.sqrt(), e.exp(), e.ln(), Expr.fma(a, b, c)) and evaluated in one GPU pass by eval()."""

    def __init__(self, kind, op=None, kids=(), scalar=0.0, array=None):
        self.kind, self.op, self.kids, self.scalar, self.array = kind, op, tuple(kids), scalar, array

    @staticmethod
    def leaf(a) -> "Expr":
        return Expr("leaf", array=np.ascontiguousarray(a))

    @staticmethod
    def fma(a, b, c) -> "Expr":
        return Expr("fma", kids=(_as_expr(a), _as_expr(b), _as_expr(c)))

    def _bin(self, other, op, reverse=False):
        if isinstance(other, Expr):
            return Expr("bin", op, (other, self) if reverse else (self, other))
        return Expr("lhs" if reverse else "rhs", op, (self,), float(other))

    def __add__(self, o): return self._bin(o, "add")
    def __radd__(self, o): return self._bin(o, "add", True)
    def __sub__(self, o): return self._bin(o, "sub")
    def __rsub__(self, o): return self._bin(o, "sub", True)
    def __mul__(self, o): return self._bin(o, "mul")
    def __rmul__(self, o): return self._bin(o, "mul", True)
    def __truediv__(self, o): return self._bin(o, "div")
    def __rtruediv__(self, o): return self._bin(o, "div", True)
    def __neg__(self): return Expr("unary", "neg", (self,))
    def abs(self): return Expr("unary", "abs", (self,))
    def sqrt(self): return Expr("unary", "sqrt", (self,))
    def exp(self): return Expr("unary", "exp", (self,))
    def ln(self): return Expr("unary", "ln", (self,))

    def program(self):
        """Postfix nodes [(kind, op, scalar)] and the distinct operand arrays, in first-use order."""
        leaves, nodes = [], []

        def walk(e):
            if e.kind == "leaf":
                for j, a in enumerate(leaves):
                    if a is e.array:
                        break
                else:
                    j = len(leaves)
                    leaves.append(e.array)
                nodes.append((0, j, 0.0))
                return
            for k in e.kids:
                walk(k)
            if e.kind == "bin":
                nodes.append((2, _EXPR_BIN[e.op], 0.0))
            elif e.kind == "rhs":
                nodes.append((1, _EXPR_RHS[e.op], e.scalar))
            elif e.kind == "lhs":
                nodes.append((1, _EXPR_LHS[e.op], e.scalar))
            elif e.kind == "unary":
                nodes.append((1, _EXPR_UNARY[e.op], 0.0))
            else:
                nodes.append((3, 0, 0.0))

        walk(self)
        return nodes, leaves

    def eval(self) -> np.ndarray:
        """fused_eval.rs:635-697 eval: the whole tree in one pass, each node rounded once."""
        nodes, leaves = self.program()
        dt = leaves[0].dtype
        if dt not in (np.float64, np.float32) or any(a.dtype != dt or a.shape != leaves[0].shape for a in leaves):
            raise ValueError("expression operands must share one float dtype and one shape")
        arr = (ExprNodeC * len(nodes))(*[ExprNodeC(k, o, s) for k, o, s in nodes])
        ptrs = (C.c_void_p * len(leaves))(*[a.ctypes.data for a in leaves])
        out = np.empty_like(leaves[0])
        call("nrk_expr_eval_f64" if dt == np.float64 else "nrk_expr_eval_f32", C.cast(arr, C.c_void_p), C.c_int32(len(nodes)),
             C.cast(ptrs, C.c_void_p), C.c_int32(len(leaves)), _ptr(out), C.c_size_t(out.size))
        return out


def _as_expr(v):
    return v if isinstance(v, Expr) else Expr.leaf(v)
