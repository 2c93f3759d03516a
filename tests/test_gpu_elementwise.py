"""GPU elementwise map/zip vs the restatement: bit-exact (one rounding per op, no FMA)."""
import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 1023, 100_003, 2_000_000])
def test_scalar_chain_bit_exact(gpu, oracle, dtype, n):
    x = rng.normal(rng.BASE_SEED + 900 + n, n, dtype)
    y = K.unary_dispatch(x, [("mul", 1.5), ("add", -0.25)])
    np.testing.assert_array_equal(y, oracle.ref_map([2, 0], [1.5, -0.25], x))
    steps = [("sub", 0.125), ("div", 3.0), ("rsub", 2.0), ("rdiv", 1.0), ("neg",), ("abs",), ("square",), ("sqrt",)]
    y = K.unary_dispatch(x, steps)
    np.testing.assert_array_equal(y, oracle.ref_map([K.SOP[s[0]] for s in steps], [s[1] if len(s) > 1 else 0 for s in steps], x))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("step", [("add", 0.1), ("sub", 0.3), ("mul", 1.1), ("div", 3.0), ("rsub", 2.0), ("rdiv", 1.0),
                                  ("neg",), ("abs",), ("square",), ("sqrt",)])
def test_each_scalar_op_correctly_rounded(gpu, oracle, dtype, step):
    """Every op code alone, bit-exact against the CPU's IEEE result (division and square root
    included: the CPU rounds them correctly, so the GPU must too)."""
    x = np.abs(rng.normal(rng.BASE_SEED + 950, 200_003, dtype)) + dtype(0.01)
    y = K.unary_dispatch(x, [step])
    np.testing.assert_array_equal(y, oracle.ref_map([K.SOP[step[0]]], [step[1] if len(step) > 1 else 0], x))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
# hypot, atan2, logaddexp and power go through elementary functions: their ulp bounds are tested in
# test_gpu_ufuncs.py; every other binary op is one IEEE operation and must match bit for bit
@pytest.mark.parametrize("op", [o for o in K.BOP if o not in ("hypot", "atan2", "logaddexp", "power") and K.BOP[o] < 13])
def test_zip_bit_exact(gpu, oracle, dtype, op):
    n = 300_007
    a = rng.normal(rng.BASE_SEED + 1000, n, dtype)
    b = rng.normal(rng.BASE_SEED + 1001, n, dtype)
    b[::1000] = np.nan
    fn = K.binary_f64 if dtype == np.float64 else K.binary_f32
    np.testing.assert_array_equal(fn(a, b, op), oracle.ref_zip(K.BOP[op], a, b))


def test_float_zip_refuses_integer_only_ops(gpu):
    for op in ("bitand", "shl"):
        with pytest.raises(gpu.NrkError):
            K.binary_f64(np.ones(4), np.ones(4), op)


def test_numpy_arithmetic_fixtures(gpu, numpy_ref):
    bops = {"add": "add", "subtract": "sub", "multiply": "mul", "divide": "div"}
    sops = {"scalar_add": "add", "scalar_subtract": "sub", "scalar_multiply": "mul", "scalar_divide": "div"}
    for name, case in numpy_ref["arithmetic_operations"].items():
        want = np.array(case["result"]["data"], np.float64)
        if case["operation"] in bops:
            got = K.binary_f64(np.array(case["input1"]["data"]), np.array(case["input2"]["data"]), bops[case["operation"]])
        else:
            got = K.unary_dispatch(np.array(case["input"]["data"], np.float64), [(sops[case["operation"]], case["scalar"])])
        np.testing.assert_array_equal(got.reshape(-1), want.reshape(-1), err_msg=name)


def test_rem_known(gpu, known):
    k = known["rem_known"]
    np.testing.assert_array_equal(K.binary_f64(np.array(k["a"]), np.array(k["b"]), "rem"), np.array(k["out"]))


def test_misaligned_views(gpu, oracle):
    """A base pointer off 16-byte alignment takes the scalar path and gives the same bits."""
    x = rng.normal(5, 10_001, np.float32)
    v = x[1:]
    np.testing.assert_array_equal(K.unary_dispatch(v, [("mul", 1.5), ("add", -0.25)]), oracle.ref_map([2, 0], [1.5, -0.25], v))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 3, 1001, 300_007, 1 << 22])
def test_fma_and_lerp_bit_exact(gpu, oracle, dtype, n):
    """fma (ufuncs.rs:864, one rounding) and lerp (:893, a + t * (b - a)) against the oracle."""
    a = rng.normal(rng.BASE_SEED + 2000 + n % 97, n, dtype)
    b = rng.normal(rng.BASE_SEED + 2001, n, dtype)
    c = (-(a * b) + rng.normal(rng.BASE_SEED + 2002, n, dtype) * dtype(1e-3)).astype(dtype)
    if n > 10:
        a[::1009] = np.nan
        b[5::1013] = np.inf
    np.testing.assert_array_equal(K.fma(a, b, c), oracle.ref_fma(a, b, c))
    for t in (0.0, 0.37, 1.0, -1.5):
        np.testing.assert_array_equal(K.lerp(a, b, t), oracle.ref_lerp(a, b, t))
    # an input off 16-byte alignment takes the scalar path and gives the same bits
    if n > 2:
        np.testing.assert_array_equal(K.fma(a[1:], b[1:], c[1:]), oracle.ref_fma(a[1:], b[1:], c[1:]))


def test_fma_lerp_shape_mismatch(gpu):
    with pytest.raises(ValueError):
        K.fma(np.ones(3), np.ones(3), np.ones(4))
    with pytest.raises(ValueError):
        K.lerp(np.ones(3), np.ones(2), 0.5)
