"""Expression fusion (expr/fused_eval.rs:421-697) on the GPU: trees of up to 4 operands in one pass,
and a two-operand tree feeding the fused reduction.

Bars: bit-exact against the oracle's eval_eager restatement (one rounding per node, no FMA) for
every shape fused_specialized handles (:421-497) and for general trees; trees with exp / ln are
bit-exact against the eager evaluation through this library's own map kernels (fused == eager, as
the reference's own tests demand). The reduction of a fused two-operand tree is bit-identical to
the fused reduction of the materialised tree, and within the reduction bars of the double-double
statistics."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng
from numrs_amd._lib import ReduceOutputs, call
from numrs_amd.device import DeviceArray
from test_gpu_fused import run_fused

pytestmark = pytest.mark.gpu
E = K.Expr
OPS = ("add", "sub", "mul", "div")


def _bits(x):
    return x.view(np.uint64 if x.dtype == np.float64 else np.uint32)


def _leaves(dt, n, count=4):
    r = np.random.default_rng(n + count)
    return [(r.standard_normal(n) * (j + 1) + 0.5 * j).astype(dt) for j in range(count)]


def _check(expr, oracle):
    nodes, leaves = expr.program()
    got = expr.eval()
    want = oracle.ref_expr(nodes, leaves)
    np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=str(nodes))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 7, 4096 + 3, 300_001])
def test_specialised_shapes_bit_exact(gpu, oracle, dt, n):
    a, b, c, d = (E.leaf(v) for v in _leaves(dt, n))
    binop = {"add": lambda x, y: x + y, "sub": lambda x, y: x - y, "mul": lambda x, y: x * y, "div": lambda x, y: x / y}
    for o in OPS:
        f = binop[o]
        for g in OPS:
            h = binop[g]
            _check(f(h(a, b), c), oracle)          # (Leaf op Leaf) op Leaf   (a * b + c)
            _check(f(a, h(b, c)), oracle)          # Leaf op (Leaf op Leaf)
            _check(f(h(a, b), h(c, d)), oracle)    # (Leaf op Leaf) op (Leaf op Leaf)
        _check(f(a, b), oracle)                    # Leaf op Leaf
        _check(f(a * 2.0, b), oracle)              # (Leaf sop k) op Leaf  -- axpy
        _check(f(a, b * 2.0), oracle)              # Leaf op (Leaf sop k)
        _check(f(a, 3.0), oracle)                  # ScalarRhs
        _check(f(3.0, a) if o != "add" else 3.0 + a, oracle)  # ScalarLhs
    _check(E.fma(a, b, c), oracle)                 # Fma: (a * b) + c, two roundings
    _check(-a, oracle)
    _check(a.abs().sqrt(), oracle)
    _check(a, oracle)                              # a lone leaf is a copy


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_general_trees_bit_exact(gpu, oracle, dt):
    a, b, c, d = (E.leaf(v) for v in _leaves(dt, 100_003))
    for expr in (((a + b) * c - d) / a, (a - 1.0) * (b + 2.0) - (c * d).abs(), ((a * b) * (c * d)) + (a - b),
                 E.fma(a + b, c - d, a * 0.5), 1.0 / ((a * a + b * b).sqrt() + 1.0), (a / b) - (c / d) + 7.0):
        _check(expr, oracle)


def test_fma_is_two_roundings(gpu):
    """Fma(a, b, c) rounds the product (fused_eval.rs:493): 1 + 2^-30 squared minus 1 loses the
    2^-60 term that a fused multiply-add would keep."""
    x = np.array([1.0 + 2.0 ** -30])
    got = E.fma(x, x, np.array([-1.0])).eval()
    assert got[0] == 2.0 ** -29  # (1 + 2^-29 + 2^-60) rounds to 1 + 2^-29 before the add; fma gives 2^-29 + 2^-60


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_exp_ln_trees_equal_eager(gpu, dt):
    """Trees with exp / ln: fused equals the node-by-node evaluation through the library's own
    kernels (map for exp / ln, zip for the binary ops) to the bit."""
    av, bv = (np.abs(v) + 0.1 for v in _leaves(dt, 50_001, 2))
    a, b = E.leaf(av), E.leaf(bv)
    got = ((a * 0.5).exp() - b.ln() * 2.0).eval()
    t1 = K.unary_dispatch(av, [("mul", 0.5), ("exp",)])
    t2 = K.unary_dispatch(bv, [("log",), ("mul", 2.0)])
    want = K.binary_dispatch(t1, t2, "sub")
    np.testing.assert_array_equal(_bits(got), _bits(want))


def run_expr_reduce(nodes, leaves, rows, cols, ddof=0, want_rows=True):
    dt = leaves[0].dtype
    dls = [DeviceArray.upload(v.reshape(-1)) for v in leaves]
    outs, ro, fo = {}, ReduceOutputs(), ReduceOutputs()
    for f in ("sum", "mean", "var", "min", "max", "argmin", "argmax"):
        odt = np.uint64 if f.startswith("arg") else dt
        if want_rows:
            outs["row_" + f] = DeviceArray((rows,), odt)
            setattr(ro, f, outs["row_" + f].ptr)
        outs["full_" + f] = DeviceArray((1,), odt)
        setattr(fo, f, outs["full_" + f].ptr)
    arr = (K.ExprNodeC * len(nodes))(*[K.ExprNodeC(k, o, s) for k, o, s in nodes])
    ptrs = (C.c_void_p * len(dls))(*[d.ptr for d in dls])
    call("nrk_dev_expr_reduce_f64" if dt == np.float64 else "nrk_dev_expr_reduce_f32", C.cast(arr, C.c_void_p),
         C.c_int32(len(nodes)), C.cast(ptrs, C.c_void_p), C.c_int32(len(dls)), C.c_size_t(rows), C.c_size_t(cols),
         C.c_size_t(ddof), C.byref(ro) if want_rows else None, C.byref(fo), None)
    call("nrk_device_synchronize")
    return {k: v.download() for k, v in outs.items()}


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(2048, 1024), (1024, 4096), (37, 1000), (1, 1 << 20), (1, (1 << 20) + 333)])
def test_two_operand_tree_feeds_the_reduction(gpu, oracle, dt, shape):
    """y = (a * 1.5) - (b + 0.25), reduced along rows and over the whole array in one pass over a
    and b: every output bit-identical to the fused reduction of the materialised y (the same
    kernels see the same y), and max / argmax bit-exact against the oracle."""
    rows, cols = shape
    av = rng.normal(rng.BASE_SEED + 5100 + rows, rows * cols, dt).reshape(rows, cols)
    bv = rng.normal(rng.BASE_SEED + 5200 + cols, rows * cols, dt).reshape(rows, cols)
    a, b = E.leaf(av), E.leaf(bv)
    nodes, leaves = ((a * 1.5) - (b + 0.25)).program()
    y = oracle.ref_expr(nodes, [v.reshape(-1) for v in leaves]).reshape(rows, cols)
    want_rows = rows > 1
    got = run_expr_reduce(nodes, leaves, rows, cols, want_rows=want_rows)
    ref = run_fused(y, [], want_rows=want_rows)
    for k in got:
        np.testing.assert_array_equal(_bits(got[k]) if got[k].dtype != np.uint64 else got[k],
                                      _bits(ref[k]) if ref[k].dtype != np.uint64 else ref[k], err_msg=k)
    assert got["full_argmax"][0] == oracle.ref("argmax", y) and got["full_max"][0] == oracle.ref("max", y)
    ex = oracle.dd_stats(y, 0)
    e = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    s_abs = float(np.abs(y.astype(np.float64)).sum())
    assert abs(got["full_sum"][0] - ex["sum"]) <= 8 * e * s_abs + e * abs(ex["sum"])
    assert abs(got["full_var"][0] - ex["var"]) <= 64 * e * ex["var"]
    # a tree outside the one-pass form (three operands) takes the workspace path, same bounds
    c = E.leaf(av * dt(0.5))
    nodes3, leaves3 = ((a - b) * c).program()
    y3 = oracle.ref_expr(nodes3, [v.reshape(-1) for v in leaves3]).reshape(rows, cols)
    got3 = run_expr_reduce(nodes3, leaves3, rows, cols, want_rows=want_rows)
    ref3 = run_fused(y3, [], want_rows=want_rows)
    for k in got3:
        np.testing.assert_array_equal(got3[k], ref3[k], err_msg=k)
