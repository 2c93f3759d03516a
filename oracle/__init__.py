"""TEST INFRASTRUCTURE ONLY: ctypes binding of the C oracle (oracle/oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker or the timed CPU baseline -- never as a product code path.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

vp = C.c_void_p
sz = C.c_size_t


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `python -m oracle.build`")
        L = C.CDLL(str(LIB_PATH))
        for suf, ct in (("f64", C.c_double), ("f32", C.c_float)):
            for name in ("ref_sum", "ref_mean", "ref_max", "ref_min", "ref_prod"):
                f = getattr(L, f"{name}_{suf}")
                f.argtypes = [vp, sz]
                f.restype = ct
            f = getattr(L, f"ref_var_{suf}")
            f.argtypes = [vp, sz, sz]
            f.restype = ct
            f = getattr(L, f"ref_sum_sq_dev_{suf}")
            f.argtypes = [vp, sz, ct]
            f.restype = ct
            for name in ("ref_argmax", "ref_argmin"):
                f = getattr(L, f"{name}_{suf}")
                f.argtypes = [vp, sz]
                f.restype = C.c_uint64
            f = getattr(L, f"ref_axis_{suf}")
            f.argtypes = [C.c_int, vp, sz, sz, sz, sz, vp]
            f.restype = None
            f = getattr(L, f"ref_map_{suf}")
            f.argtypes = [vp, vp, C.c_int, vp, vp, sz]
            f.restype = None
            f = getattr(L, f"ref_zip_{suf}")
            f.argtypes = [C.c_int, vp, vp, vp, sz]
            f.restype = None
            f = getattr(L, f"dd_stats_{suf}")
            f.argtypes = [vp, sz, sz, vp]
            f.restype = None
            f = getattr(L, f"dd_row_stats_{suf}")
            f.argtypes = [vp, sz, sz, sz, vp]
            f.restype = None
            f = getattr(L, f"dd_gemm_entries_{suf}")
            f.argtypes = [sz, sz, vp, vp, sz, vp, vp, vp]
            f.restype = None
            f = getattr(L, f"ref_gemm_generic_{suf}")
            f.argtypes = [sz, sz, sz, vp, vp, vp]
            f.restype = None
        for f in (L.ref_gemm_generic_i64, L.ref_gemm_generic_i32):
            f.argtypes = [sz, sz, sz, vp, vp, vp]
            f.restype = None
        L.ref_gemm_rowsplit_f64.argtypes = [sz, sz, sz, vp, vp, vp]
        L.ref_gemm_rowsplit_f64.restype = None
        L.dd_gemm_f64.argtypes = [sz, sz, sz, vp, vp, vp]
        L.dd_gemm_f64.restype = None
        L.abs_gemm_entries_f64.argtypes = [sz, sz, vp, vp, sz, vp, vp, vp]
        L.abs_gemm_entries_f64.restype = None
        L.ref_sum_i64.argtypes = [vp, sz]
        L.ref_sum_i64.restype = C.c_int64
        L.exact_sum_i64.argtypes = [vp, sz, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]
        L.exact_sum_i64.restype = None
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(vp)


def _suf(a: np.ndarray) -> str:
    return {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}[a.dtype]


def set_threads(n: int) -> None:
    lib().oracle_set_threads(n)


def ref(name: str, a: np.ndarray, *extra):
    """ref_<name>_<dtype>(a, n, *extra) -- reduction restatements returning a scalar."""
    a = np.ascontiguousarray(a).reshape(-1)
    f = getattr(lib(), f"ref_{name}_{_suf(a)}")
    return f(_p(a), a.size, *extra)


def ref_axis(op: int, x: np.ndarray, outer: int, length: int, inner: int, ddof: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x)
    out = np.empty(outer * inner, np.uint64 if op >= 5 else x.dtype)
    getattr(lib(), f"ref_axis_{_suf(x)}")(op, _p(x), outer, length, inner, ddof, _p(out))
    return out


def ref_map(ops: list[int], scalars: list[float], x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    o = np.asarray(ops, np.int32)
    s = np.asarray(scalars, np.float64)
    y = np.empty_like(x)
    getattr(lib(), f"ref_map_{_suf(x)}")(_p(o), _p(s), len(ops), _p(x), _p(y), x.size)
    return y


def ref_zip(op: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    y = np.empty_like(a)
    getattr(lib(), f"ref_zip_{_suf(a)}")(op, _p(a), _p(b), _p(y), a.size)
    return y


_ISUF = {np.dtype(np.int64): "i64", np.dtype(np.int32): "i32", np.dtype(np.uint64): "u64", np.dtype(np.uint32): "u32"}


def ref_zip_int(op: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Integer binary_serial (elementwise.rs:51-66) under Rust's release arithmetic (oracle.c)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, a.dtype)
    y = np.empty_like(a)
    getattr(lib(), f"ref_zip_int_{_ISUF[a.dtype]}")(op, _p(a), _p(b), _p(y), C.c_size_t(a.size))
    return y


def ref_map_int(ops: list[int], scalars: list[int], x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    o = np.asarray(ops, np.int32)
    s = np.asarray(scalars, np.int64)
    y = np.empty_like(x)
    getattr(lib(), f"ref_map_int_{_ISUF[x.dtype]}")(_p(o), _p(s), len(ops), _p(x), _p(y), C.c_size_t(x.size))
    return y


def ref_expr(nodes, leaves) -> np.ndarray:
    """eval_eager (fused_eval.rs:605-658) of a postfix program [(kind, op, scalar), ...] over equally
    long leaves, one rounded operation per node (oracle.c ref_expr)."""
    leaves = [np.ascontiguousarray(v) for v in leaves]
    dt = leaves[0].dtype
    kind = np.array([nd[0] for nd in nodes], np.int32)
    op = np.array([nd[1] for nd in nodes], np.int32)
    sc = np.array([nd[2] if len(nd) > 2 else 0.0 for nd in nodes], np.float64)
    ptrs = (C.c_void_p * len(leaves))(*[v.ctypes.data for v in leaves])
    y = np.empty_like(leaves[0])
    getattr(lib(), f"ref_expr_{_suf(leaves[0])}")(_p(kind), _p(op), _p(sc), len(nodes), ptrs, _p(y), C.c_size_t(y.size))
    return y


def ref_fma(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    a, b, c = (np.ascontiguousarray(v) for v in (a, b, c))
    y = np.empty_like(a)
    f = getattr(lib(), f"ref_fma_{_suf(a)}")
    f.argtypes = [C.c_void_p] * 4 + [C.c_size_t]
    f(_p(a), _p(b), _p(c), _p(y), a.size)
    return y


def ref_lerp(a: np.ndarray, b: np.ndarray, t: float) -> np.ndarray:
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    y = np.empty_like(a)
    f = getattr(lib(), f"ref_lerp_{_suf(a)}")
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_double if a.dtype == np.float64 else C.c_float, C.c_void_p, C.c_size_t]
    f(_p(a), _p(b), t, _p(y), a.size)
    return y


def sum_i64(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a, np.int64)
    return int(lib().ref_sum_i64(_p(a), a.size))


def exact_sum_i64(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a, np.int64)
    hi, lo = C.c_int64(), C.c_uint64()
    lib().exact_sum_i64(_p(a), a.size, C.byref(hi), C.byref(lo))
    return (hi.value << 64) | lo.value


def dd_stats(x: np.ndarray, ddof: int = 0) -> dict:
    """Double-double sum (hi+lo), mean and var(ddof) of a flat array."""
    x = np.ascontiguousarray(x).reshape(-1)
    out = np.empty(4)
    getattr(lib(), f"dd_stats_{_suf(x)}")(_p(x), x.size, ddof, _p(out))
    return {"sum_hi": out[0], "sum_lo": out[1], "sum": out[0] + out[1], "mean": out[2], "var": out[3]}


def dd_row_stats(x: np.ndarray, ddof: int = 0) -> np.ndarray:
    """rows x 4: [sum_hi, sum_lo, mean, var] per row of a 2-D array."""
    x = np.ascontiguousarray(x)
    out = np.empty((x.shape[0], 4))
    getattr(lib(), f"dd_row_stats_{_suf(x)}")(_p(x), x.shape[0], x.shape[1], ddof, _p(out))
    return out


def gemm_generic(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    m, k = a.shape
    n = b.shape[1]
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    c = np.empty((m, n), a.dtype)
    suf = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32", np.dtype(np.int64): "i64",
           np.dtype(np.int32): "i32"}[a.dtype]
    getattr(lib(), f"ref_gemm_generic_{suf}")(m, k, n, _p(a), _p(b), _p(c))
    return c


def gemm_rowsplit_f64(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    m, k = a.shape
    n = b.shape[1]
    c = np.empty((m, n), np.float64)
    lib().ref_gemm_rowsplit_f64(m, k, n, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), _p(c))
    return c


def dd_gemm(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    m, k = a.shape
    n = b.shape[1]
    c = np.empty((m, n), np.float64)
    lib().dd_gemm_f64(m, k, n, _p(np.ascontiguousarray(a, np.float64)), _p(np.ascontiguousarray(b, np.float64)), _p(c))
    return c


def dd_gemm_entries(a: np.ndarray, b: np.ndarray, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    k = a.shape[1]
    n = b.shape[1]
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    out = np.empty(rows.size)
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    getattr(lib(), f"dd_gemm_entries_{_suf(a)}")(k, n, _p(a), _p(b), rows.size, _p(rows), _p(cols), _p(out))
    return out


def abs_gemm_entries(a: np.ndarray, b: np.ndarray, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    k = a.shape[1]
    n = b.shape[1]
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    out = np.empty(rows.size)
    lib().abs_gemm_entries_f64(k, n, _p(np.ascontiguousarray(a, np.float64)), _p(np.ascontiguousarray(b, np.float64)),
                               rows.size, _p(rows), _p(cols), _p(out))
    return out


def ref_ufunc_fold(op: str, seq) -> float:
    """ufunc_ops.rs:240-257 reduce_lane without an initial value: fold from the first element with
    UfuncOp::apply (ufunc_ops.rs:156-180). maximum / minimum: NaN in either operand gives NaN,
    otherwise `a > b ? a : b` (`a < b ? a : b`), so an equal later operand wins, which fixes the
    sign of a zero result. Pure Python: for short segments only."""
    it = iter(seq)
    acc = next(it)
    for x in it:
        if op in ("maximum", "minimum"):
            if math.isnan(acc) or math.isnan(x):
                acc = math.nan
            elif (acc > x) if op == "maximum" else (acc < x):
                pass
            else:
                acc = x
        elif op == "add":
            acc = acc + x
        elif op == "multiply":
            acc = acc * x
        else:
            raise ValueError(op)
    return acc
