"""Plain (reference-order) reductions through nrk_dev_map_reduce_plain_{f32,f64}.

Bar: bit-exact. Every sum, mean and variance is the sequence of rounded additions of the
oracle's restatement of the reference's kernels (kernels/reduce.rs:207-314 for the whole array,
the left-to-right axis folds for rows), applied to y = (x * 1.5) + (-0.25) materialised by the
oracle's map. The reference's own pinned bits (tests/test_reduction_determinism.rs:126-197) hold
too. Extremum fields come from the fused kernel and are checked against the oracle as well.
"""
import numpy as np
import pytest

from conftest import det_data
from test_gpu_fused import run_fused
from numrs_amd import kernels as K
from numrs_amd import rng

pytestmark = pytest.mark.gpu
SUMS = ("sum", "mean", "var")
STEPS = [("mul", 1.5), ("add", -0.25)]


def bits(v):
    v = np.asarray(v)
    return v.view(np.uint64 if v.dtype == np.float64 else np.uint32)


def mapped(oracle, x):
    return oracle.ref_map([2, 0], [1.5, -0.25], x).reshape(x.shape)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
# across every tier boundary: simd_sum below 10000, chunks of 8192 from 10000, the variance's one
# decision at 40000; ragged last chunks and 0..3-element tails
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 1023, 8192, 9999, 10_000, 10_001, 16_385, 39_999, 40_000,
                               40_001, 100_003, 1_000_000, 2_000_003])
@pytest.mark.parametrize("ddof", [0, 1])
def test_full_bit_exact(gpu, oracle, dtype, n, ddof):
    x = rng.normal(rng.BASE_SEED + 3100 + n % 977, n, dtype)
    x[::7] *= dtype(1000.0)  # spread the magnitudes so that the order of additions shows in the bits
    got = run_fused(x.reshape(1, n), STEPS, ddof=ddof, want_rows=False, plain=True, fields=SUMS)
    y = mapped(oracle, x)
    assert bits(got["full_sum"][0]) == bits(dtype(oracle.ref("sum", y)))
    assert bits(got["full_mean"][0]) == bits(dtype(oracle.ref("mean", y)))
    assert bits(got["full_var"][0]) == bits(dtype(oracle.ref("var", y, ddof)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (4, 7), (257, 100), (16, 8193), (3, 40_000), (1024, 1024), (600, 2050),
                                   (5000, 3)])
def test_rows_bit_exact(gpu, oracle, dtype, shape):
    rows, cols = shape
    x = rng.normal(rng.BASE_SEED + 3200 + rows + cols, rows * cols, dtype).reshape(rows, cols)
    x[:, ::5] *= dtype(300.0)
    ddof = 1 if cols > 1 else 0
    got = run_fused(x, STEPS, ddof=ddof, plain=True)
    y = mapped(oracle, x)
    for op, name in ((0, "sum"), (1, "mean"), (2, "var")):
        np.testing.assert_array_equal(bits(got["row_" + name]), bits(oracle.ref_axis(op, y, rows, cols, 1, ddof)), err_msg=name)
    # whole-array fields of the same call
    assert bits(got["full_sum"][0]) == bits(dtype(oracle.ref("sum", y)))
    assert bits(got["full_var"][0]) == bits(dtype(oracle.ref("var", y, ddof)))
    # the extrema come from the fused kernel
    np.testing.assert_array_equal(got["row_argmax"], oracle.ref_axis(6, y, rows, cols, 1))
    np.testing.assert_array_equal(got["row_min"], oracle.ref_axis(3, y, rows, cols, 1))
    assert got["full_argmin"][0] == oracle.ref("argmin", y) and got["full_max"][0] == oracle.ref("max", y)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_misaligned_and_unmapped(gpu, oracle, dtype):
    """A pointer off 16-byte alignment (element-wise staging) and an empty chain give the same bits."""
    rows, cols = 300, 4097
    x = rng.normal(rng.BASE_SEED + 3300, rows * cols, dtype).reshape(rows, cols)
    a = run_fused(x, [], plain=True, fields=SUMS)
    b = run_fused(x, [], plain=True, fields=SUMS, offset_elems=1)
    for k in a:
        np.testing.assert_array_equal(bits(a[k]), bits(b[k]), err_msg=k)
    np.testing.assert_array_equal(bits(a["row_sum"]), bits(oracle.ref_axis(0, x, rows, cols, 1)))
    assert bits(a["full_sum"][0]) == bits(dtype(oracle.ref("sum", x)))
    assert bits(a["full_var"][0]) == bits(dtype(oracle.ref("var", x, 0)))


def test_determinism_pins(gpu, known):
    """tests/test_reduction_determinism.rs:126-197: x_i = i * 0.5 - 12345, n = 50000."""
    pins = known["determinism_pins"]
    x = det_data(np.float64)
    got = run_fused(x.reshape(1, -1), [], want_rows=False, plain=True, fields=SUMS)
    assert int(bits(got["full_sum"])[0]) == pins["sum_f64_bits"]
    assert int(bits(got["full_mean"])[0]) == pins["mean_f64_bits"]
    assert int(bits(got["full_var"])[0]) == pins["var_ddof0_f64_bits"]
    assert int(bits(np.sqrt(got["full_var"]))[0]) == pins["std_ddof0_f64_bits"]
    x32 = det_data(np.float32)
    got = run_fused(x32.reshape(1, -1), [], want_rows=False, plain=True, fields=("sum",))
    assert int(bits(got["full_sum"])[0]) == pins["sum_f32_bits"]


def test_plain_order_differs_from_compensated(gpu, oracle):
    """A sum where the order of additions matters: 2^24 + 1 + 1 + ... in binary32 (exact sum
    2^24 + 65535). The plain order loses the ones that meet 2^24 in the first partial sum (the
    reference's result, 2047 short); the compensated path is within one ulp (2) of the exact sum."""
    n = 1 << 16
    x = np.ones(n, np.float32)
    x[0] = np.float32(2.0 ** 24)
    exact = 2 ** 24 + n - 1
    plain = run_fused(x.reshape(1, n), [], want_rows=False, plain=True, fields=("sum",))["full_sum"][0]
    comp = run_fused(x.reshape(1, n), [], want_rows=False, fields=("sum",))["full_sum"][0]
    assert bits(plain) == bits(np.float32(oracle.ref("sum", x)))
    assert exact - int(plain) == 2047
    assert abs(int(comp) - exact) <= 2


def test_empty_and_ddof_errors(gpu):
    """Empty: every output 0 (kernels/reduce.rs:207-421); var along an axis no longer than ddof is
    refused as on the compensated path."""
    got = run_fused(np.zeros((3, 0), np.float64), [], plain=True, fields=("sum", "mean"))
    assert np.all(got["row_sum"] == 0) and np.all(got["row_mean"] == 0) and got["full_sum"][0] == 0
    got = run_fused(np.zeros((3, 0), np.float64), [], want_rows=False, plain=True, fields=("var",))
    assert got["full_var"][0] == 0
    with pytest.raises(gpu.NrkError):
        run_fused(np.ones((4, 2), np.float64), [], ddof=2, plain=True, fields=("var",))
