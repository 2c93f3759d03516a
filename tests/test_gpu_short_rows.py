"""Row reductions of short rows (at most 1024 binary32 / 512 binary64 elements): a few lanes per
row, the row held in registers, a two-pass variance, and the whole-array results from the same
launch. Checked against the oracle: extrema and indices bit-exact, sums and variances within the
reduction bars of the double-double row statistics; a pointer off 16-byte alignment (scalar loads)
gives the same bits as the aligned run; repeated calls give the same bits."""
import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng
from test_gpu_fused import EPS, run_fused

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (3, 2), (64, 1), (777, 3), (1000, 16), (100_003, 7), (50_000, 8), (4097, 12), (4099, 17), (5000, 64),
          (2048, 100), (1000, 255), (333, 256),
          (20_000, 64), (3000, 511), (2500, 512), (1500, 513), (1200, 671), (1100, 1000), (1030, 1024),
          (999, 1025), (1500, 1500), (700, 2048)]


def _check_rows(got, y, oracle, ddof=0):
    rows, cols = y.shape
    for op in ("max", "min", "argmax", "argmin"):
        want = oracle.ref_axis(K.AXIS[op], y, rows, cols, 1)
        g = got["row_" + op]
        if op in ("max", "min"):
            np.testing.assert_array_equal(np.isnan(g), np.isnan(want), err_msg=op)
            ok = ~np.isnan(want)
            np.testing.assert_array_equal(g[ok], want[ok], err_msg=op)
        else:
            np.testing.assert_array_equal(g, want, err_msg=op)
    ex = oracle.dd_row_stats(y, ddof)
    y64 = y.astype(np.float64)
    fin = np.isfinite(y64).all(axis=1)  # rows with NaN or inf: NaN where binary64 sum / var is NaN
    with np.errstate(invalid="ignore"):
        for k, want in (("row_sum", y64.sum(axis=1)), ("row_var", y64.var(axis=1))):
            np.testing.assert_array_equal(np.isnan(got[k][~fin]), np.isnan(want[~fin]), err_msg=k)
    ex = ex[fin]
    s_abs = np.abs(y[fin].astype(np.float64)).sum(axis=1)
    e = EPS[y.dtype.type]
    assert np.all(np.abs(got["row_sum"][fin] - (ex[:, 0] + ex[:, 1])) <= 8 * e * s_abs + e * np.abs(ex[:, 0]))
    assert np.all(np.abs(got["row_mean"][fin] - ex[:, 2]) <= 8 * e * s_abs / cols + 2 * e * np.abs(ex[:, 2]))
    if cols > 1 + ddof:
        assert np.all(np.abs(got["row_var"][fin] - ex[:, 3]) <= 64 * e * ex[:, 3])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_short_rows_vs_oracle(gpu, oracle, dtype, shape):
    rows, cols = shape
    x = rng.normal(rng.BASE_SEED + 1700 + rows + 7 * cols, rows * cols, dtype).reshape(rows, cols)
    steps = [("mul", 1.5), ("add", -0.25)]
    got = run_fused(x, steps)
    y = oracle.ref_map([2, 0], [1.5, -0.25], x).reshape(rows, cols)
    _check_rows(got, y, oracle)
    for op in ("max", "min", "argmax", "argmin"):
        assert got["full_" + op][0] == oracle.ref(op, y), op
    ex = oracle.dd_stats(y, 0)
    e = EPS[dtype]
    s_abs = float(np.abs(y.astype(np.float64)).sum())
    assert abs(got["full_sum"][0] - ex["sum"]) <= 8 * e * s_abs + e * abs(ex["sum"])
    if rows * cols > 1:
        assert abs(got["full_var"][0] - ex["var"]) <= 64 * e * ex["var"]
    # the same bits: off 16-byte alignment, a second call, and rows only
    b = run_fused(x, steps, offset_elems=1)
    c = run_fused(x, steps)
    d = run_fused(x, steps, want_full=False)
    for k in got:
        np.testing.assert_array_equal(got[k], b[k], err_msg=k)
        np.testing.assert_array_equal(got[k], c[k], err_msg=k)
        if k.startswith("row_"):
            np.testing.assert_array_equal(got[k], d[k], err_msg=k)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_short_rows_nan_inf_ties_and_ddof(gpu, oracle, dtype):
    rows, cols = 4096, 200
    x = rng.normal(rng.BASE_SEED + 1800, rows * cols, dtype).reshape(rows, cols)
    x[1, 0] = np.nan                  # first element NaN: argmax / argmin 0, max NaN
    x[2, 150] = np.nan                # interior NaN: skipped by argmax, poisons max
    x[3, :] = -np.inf                 # no winner: index 0
    x[4, :] = -0.0
    x[4, 120] = 0.0                   # zero maximum: the first zero's sign
    x[5, 7] = x[5, 190] = 99.0        # tie: the first index
    x[6, 33] = x[6, 34] = -99.0
    x[7, :] = 3.25                    # a constant row: variance 0
    x[8, :] = 1e4                     # a large mean over a small spread
    x[8, ::2] += 1.0
    got = run_fused(x, [], want_full=True)
    _check_rows(got, x, oracle)
    assert got["row_argmax"][1] == 0 and np.isnan(got["row_max"][1]) and np.isnan(got["row_max"][2])
    assert got["row_argmax"][3] == 0 and got["row_argmax"][5] == 7 and got["row_argmin"][6] == 33
    assert got["row_max"][4] == 0.0 and np.signbit(got["row_max"][4]) == np.signbit(oracle.ref_axis(4, x, rows, cols, 1)[4])
    assert got["row_argmax"][4] == 0
    assert got["row_var"][7] == 0.0
    assert abs(got["row_var"][8] - 0.25) <= 64 * EPS[dtype] * 0.25
    assert got["full_argmax"][0] == oracle.ref("argmax", x)
    got1 = run_fused(x, [], want_full=False)
    for k in got1:
        np.testing.assert_array_equal(got[k], got1[k], err_msg=k)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_short_rows_general_chain_and_ddof(gpu, oracle, dtype):
    """A chain outside y = a x + b (abs then sqrt then mul), ddof 1, against the oracle's map."""
    rows, cols = 3001, 97
    x = rng.normal(rng.BASE_SEED + 1900, rows * cols, dtype).reshape(rows, cols)
    got = run_fused(x, [("abs",), ("sqrt",), ("mul", 3.0)], ddof=1)
    y = K.unary_dispatch(x, [("abs",), ("sqrt",), ("mul", 3.0)])
    _check_rows(got, y, oracle, ddof=1)
