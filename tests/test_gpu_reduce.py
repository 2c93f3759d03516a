"""GPU parity for the reductions, through the C ABI, against the oracle.

Bars: integer sums, argmax/argmin indices, min/max values and every strided (axis-0) fold are
bit-exact. Floating-point sums/means/variances of the segment kernel are measured against the
double-double oracle: |gpu - exact| <= TOL * sum(|x|) (sums) and relative error <= TOL_VAR
(variance), with the tolerances written below.
"""
import math

import numpy as np
import pytest

from conftest import det_data, fval
from numrs_amd import kernels as K
from numrs_amd import rng

pytestmark = pytest.mark.gpu

EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}
SUM_C = 8.0        # |err| <= 8 eps sum|x|: per-thread pairwise (<= 5 levels) + double-word combine
VAR_C = 64.0       # relative: <= 64 eps (d = x - center, d*d, pairwise, then FP64 Chan combine)
SUF = {np.float64: "f64", np.float32: "f32"}


def kfn(name, dtype):
    return getattr(K, f"{name}_{SUF[dtype]}")


def test_determinism_pins(gpu, known):
    """The reference's pinned bits (test_reduction_determinism.rs:126-197). Every partial sum of
    this data is exact, so any sufficiently exact summation order must reproduce them."""
    pins = known["determinism_pins"]
    x = det_data(np.float64)
    bits = lambda v: int(np.float64(v).view(np.uint64))  # noqa: E731
    assert bits(K.sum_f64(x)) == pins["sum_f64_bits"]
    assert bits(K.mean_f64(x)) == pins["mean_f64_bits"]
    var = K.var_f64(x, 0)
    assert bits(var) == pins["var_ddof0_f64_bits"]
    assert bits(math.sqrt(var)) == pins["std_ddof0_f64_bits"]
    x32 = det_data(np.float32)
    assert int(np.float32(K.sum_f32(x32)).view(np.uint32)) == pins["sum_f32_bits"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_known_answers(gpu, known, dtype):
    v = np.array(known["var_small"]["x"], dtype)
    assert kfn("var", dtype)(v, 0) == 2.0 and kfn("var", dtype)(v, 1) == 2.5
    assert kfn("sum_sq_dev", dtype)(v, 3.0) == 10.0
    mm = np.array(known["min_max_known"]["x"], dtype)
    assert kfn("min", dtype)(mm) == -10.0 and kfn("max", dtype)(mm) == 8.0
    for case in known["min_max_nan"]["cases"]:
        x = np.array([fval(c) for c in case], dtype)
        assert math.isnan(kfn("min", dtype)(x)) and math.isnan(kfn("max", dtype)(x))
    for case in known["min_max_inf"]["cases"]:
        x = np.array([fval(c) for c in case["x"]], dtype)
        assert kfn("min", dtype)(x) == fval(case["min"]) and kfn("max", dtype)(x) == fval(case["max"])
    la = known["lane_aligned_nan"]
    for pos in la["nan_positions"]:
        d = np.full(la["len"], la["fill"], dtype)
        d[0] = la["first"]
        d[pos] = np.nan
        assert math.isnan(kfn("max", dtype)(d)) and math.isnan(kfn("min", dtype)(d))
    for key, fn in (("argmax_doc", "argmax"), ("argmin_doc", "argmin")):
        k = known[key]
        x = np.array(k["x"], dtype)
        assert kfn(fn, dtype)(x) == k["flat"]
        assert list(K.reduce_axis(fn, x, 2, 3, 1)) == k["axis1"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_empty_inputs(gpu, dtype):
    e = np.zeros(0, dtype)
    for name in ("sum", "mean", "min", "max"):
        assert kfn(name, dtype)(e) == 0.0
    assert kfn("var", dtype)(e, 0) == 0.0
    assert kfn("prod", dtype)(e) == 1.0
    with pytest.raises(gpu.NrkError):
        kfn("argmax", dtype)(e)
    assert K.sum_i64(np.zeros(0, np.int64)) == 0


SIZES = [1, 3, 63, 64, 255, 256, 1023, 1024, 4095, 4096, 4097, 8191, 8192, 8193, 9999, 10000, 10001,
         40000, 80001, 262144, 1_000_003]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sum_mean_var_vs_double_double(gpu, oracle, dtype):
    worst = {"sum": 0.0, "var": 0.0}
    for i, n in enumerate(SIZES):
        x = rng.normal(rng.BASE_SEED + 100 + i, n, dtype)
        ex = oracle.dd_stats(x, 0)
        s_abs = float(np.abs(x.astype(np.float64)).sum())
        s = kfn("sum", dtype)(x)
        assert abs(s - ex["sum"]) <= SUM_C * EPS[dtype] * s_abs + abs(ex["sum"]) * EPS[dtype], (n, s, ex["sum"])
        m = kfn("mean", dtype)(x)
        assert abs(m - ex["mean"]) <= (SUM_C * EPS[dtype] * s_abs) / n + abs(ex["mean"]) * EPS[dtype] * 2, n
        if n > 1:
            for ddof in (0, 1):
                ev = oracle.dd_stats(x, ddof)["var"]
                v = kfn("var", dtype)(x, ddof)
                rel = abs(v - ev) / ev
                worst["var"] = max(worst["var"], rel)
                assert rel <= VAR_C * EPS[dtype], (n, ddof, v, ev)
        worst["sum"] = max(worst["sum"], abs(s - ex["sum"]) / max(abs(ex["sum"]), 1e-300))
    print(f"{SUF[dtype]}: max rel err sum={worst['sum']:.3e} var={worst['var']:.3e}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sum_sq_dev_vs_exact_sum_of_terms(gpu, dtype):
    """sum_sq_dev (reduce.rs:239-260): the reference sums the terms d * d with d = x - mean, each
    operation rounded in T. numpy forms the same rounded terms (IEEE arithmetic in T); the kernel
    may only differ in the order of the sum: |got - fsum(terms)| <= SUM_C eps sum(terms) over the
    whole size grid, with the mean both exact-ish (the FP64 mean) and deliberately off."""
    worst = 0.0
    for i, n in enumerate(SIZES):
        x = rng.normal(rng.BASE_SEED + 150 + i, n, dtype) + dtype(3.0)
        for mean in (dtype(float(np.mean(x.astype(np.float64)))), dtype(2.5)):
            d = (x - mean).astype(dtype)
            terms = (d * d).astype(dtype)
            exact = math.fsum(terms.astype(np.float64).tolist())
            got = kfn("sum_sq_dev", dtype)(x, float(mean))
            err = abs(got - exact)
            assert err <= SUM_C * EPS[dtype] * exact + EPS[dtype] * exact, (n, float(mean), got, exact)
            worst = max(worst, err / exact if exact else err)
    print(f"{SUF[dtype]}: sum_sq_dev max rel err {worst:.3e}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_min_max_argmax_bit_exact(gpu, oracle, dtype):
    for i, n in enumerate(SIZES):
        x = rng.normal(rng.BASE_SEED + 200 + i, n, dtype)
        assert kfn("max", dtype)(x) == oracle.ref("max", x)
        assert kfn("min", dtype)(x) == oracle.ref("min", x)
        assert kfn("argmax", dtype)(x) == oracle.ref("argmax", x)
        assert kfn("argmin", dtype)(x) == oracle.ref("argmin", x)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_nan_and_tie_placement(gpu, oracle, dtype):
    """NaN in the first/middle/last segment and at lane-aligned spots; ties between segments;
    the NaN-at-0 argmax pin and the interior-NaN skip (statistics/mod.rs:92-101)."""
    n = 3 * 8192 + 77
    base = rng.normal(rng.BASE_SEED + 300, n, dtype)
    for pos in [0, 1, 63, 64, 4095, 4096, 8191, 8192, 16384 + 5, n - 1]:
        x = base.copy()
        x[pos] = np.nan
        assert math.isnan(kfn("max", dtype)(x)) and math.isnan(kfn("min", dtype)(x))
        assert kfn("argmax", dtype)(x) == oracle.ref("argmax", x)
        assert kfn("argmin", dtype)(x) == oracle.ref("argmin", x)
    x = np.full(n, -1.0, dtype)
    for p in (n - 1, 20000, 9000, 8191, 3):  # equal maxima: the first index must win
        x[p] = 7.0
        assert kfn("argmax", dtype)(x) == p == oracle.ref("argmax", x)
    allnan = np.full(100, np.nan, dtype)
    assert kfn("argmax", dtype)(allnan) == 0
    z = np.array([-0.0, 0.0, -0.0], dtype)  # max value keeps the sign of the first maximal zero
    assert np.signbit(kfn("max", dtype)(z)) == np.signbit(oracle.ref("max", z))
    inf = np.array([1.0, np.inf, -np.inf, np.inf], dtype)
    assert kfn("argmax", dtype)(inf) == 1 and kfn("argmin", dtype)(inf) == 2


def test_i64_sum_bit_exact(gpu, oracle):
    for i, n in enumerate([1, 5, 4095, 4096, 4097, 100_000, 1 << 22, (1 << 22) + 3]):
        x = rng.int_uniform(rng.BASE_SEED + 400 + i, n, -(1 << 20), 1 << 20)
        assert K.sum_i64(x) == oracle.sum_i64(x) == oracle.exact_sum_i64(x)
    wrap = np.array([np.iinfo(np.int64).max, 1, 5], np.int64)
    assert K.sum_i64(wrap) == oracle.sum_i64(wrap)
    assert K.sum_i64(wrap[1:]) == 6
    assert K.sum_i64(np.arange(1, 10001, dtype=np.int64)) == 10000 * 10001 // 2  # closed form


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_axis1_rows_vs_oracle(gpu, oracle, dtype):
    """Axis-1 (row) reductions through the segment kernel: max/min/argmax/argmin bit-exact;
    sum/mean/var within the stated tolerance of the double-double row statistics."""
    for rows, cols in [(1, 1), (3, 5), (7, 8192), (5, 8193), (2, 100_000), (33, 300), (1000, 16)]:
        x = rng.normal(rng.BASE_SEED + 500 + rows + cols, rows * cols, dtype).reshape(rows, cols)
        for op in ("max", "min", "argmax", "argmin"):
            got = K.reduce_axis(op, x, rows, cols, 1)
            np.testing.assert_array_equal(got, oracle.ref_axis(K.AXIS[op], x, rows, cols, 1), err_msg=f"{op} {rows}x{cols}")
        ex = oracle.dd_row_stats(x, 0)
        s_abs = np.abs(x.astype(np.float64)).sum(axis=1)
        s = K.reduce_axis("sum", x, rows, cols, 1).astype(np.float64)
        assert np.all(np.abs(s - (ex[:, 0] + ex[:, 1])) <= SUM_C * EPS[dtype] * s_abs + np.abs(ex[:, 0]) * EPS[dtype])
        if cols > 1:
            v = K.reduce_axis("var", x, rows, cols, 1, ddof=1).astype(np.float64)
            ev = oracle.dd_row_stats(x, 1)[:, 3]
            assert np.all(np.abs(v - ev) <= VAR_C * EPS[dtype] * ev)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_axis0_strided_many_outputs_bit_exact(gpu, oracle, dtype):
    """Strided (inner > 1) reductions with at least 2^18 outputs fold each output in the
    reference's own left-to-right order: every op is bit-exact against the restatement
    (math/aggregation.rs:548-561 etc.)."""
    for outer, length, inner in [(1, 5, 300_000), (3, 4, 90_000), (1, 37, 1 << 18)]:
        x = rng.normal(rng.BASE_SEED + 600 + length, outer * length * inner, dtype)
        x[::97] = x[::97] * 3
        if length == 37:  # loads run 8 rows ahead of the fold: NaNs first, inside a batch, last; ties
            v = x.reshape(length, inner)
            v[0, 10] = v[9, 11] = v[36, 12] = np.nan
            v[:, 13] = 2.0
            v[[3, 20], 14] = 50.0
        for op in ("sum", "mean", "var", "min", "max", "argmin", "argmax"):
            got = K.reduce_axis(op, x, outer, length, inner, ddof=1 if op == "var" else 0)
            want = oracle.ref_axis(K.AXIS[op], x, outer, length, inner, 1 if op == "var" else 0)
            np.testing.assert_array_equal(got, want, err_msg=f"{op} {outer}x{length}x{inner}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_axis0_strided_few_outputs(gpu, oracle, dtype):
    """Strided reductions with fewer outputs: sums, means and extrema by column tiles over ranges
    of the axis, variances with the axis transposed into rows for the row kernels. Extrema and
    their indices are bit-exact against the restatement (they do not depend on the order); sums,
    means and variances are within the row reductions' bars of the double-double statistics of
    each output's elements."""
    for outer, length, inner in [(1, 7, 13), (1, 1000, 300), (4, 50, 70), (2, 3, 1025), (1, 70_000, 3), (5, 999, 65),
                                 (1, 4097, 129)]:
        x = rng.normal(rng.BASE_SEED + 600 + length, outer * length * inner, dtype)
        x[::97] = x[::97] * 3
        for op in ("min", "max", "argmin", "argmax"):
            got = K.reduce_axis(op, x, outer, length, inner)
            want = oracle.ref_axis(K.AXIS[op], x, outer, length, inner, 0)
            np.testing.assert_array_equal(got, want, err_msg=f"{op} {outer}x{length}x{inner}")
        cols = x.reshape(outer, length, inner).transpose(0, 2, 1).reshape(outer * inner, length)
        ex = oracle.dd_row_stats(np.ascontiguousarray(cols), 1)
        s_abs = np.abs(cols.astype(np.float64)).sum(axis=1)
        e = EPS[dtype]
        s = K.reduce_axis("sum", x, outer, length, inner).astype(np.float64)
        assert np.all(np.abs(s - (ex[:, 0] + ex[:, 1])) <= 8 * e * s_abs + e * np.abs(ex[:, 0])), f"sum {outer}x{length}x{inner}"
        m = K.reduce_axis("mean", x, outer, length, inner).astype(np.float64)
        assert np.all(np.abs(m - ex[:, 2]) <= 8 * e * s_abs / length + 2 * e * np.abs(ex[:, 2])), f"mean {outer}x{length}x{inner}"
        if length > 1:
            v = K.reduce_axis("var", x, outer, length, inner, ddof=1).astype(np.float64)
            assert np.all(np.abs(v - ex[:, 3]) <= VAR_C * e * ex[:, 3]), f"var {outer}x{length}x{inner}"
    y = rng.normal(1, 4 * 9, dtype).reshape(4, 9)
    y[2, 3] = np.nan
    mx = K.reduce_axis("max", y, 1, 4, 9)
    assert math.isnan(mx[3]) and not np.isnan(np.delete(mx, 3)).any()
    assert K.reduce_axis("argmax", y, 1, 4, 9)[3] == oracle.ref_axis(6, y.reshape(-1), 1, 4, 9, 0)[3]


def test_axis_errors(gpu):
    x = np.ones((3, 4))
    with pytest.raises(gpu.NrkError):
        K.reduce_axis("var", x, 3, 4, 1, ddof=4)
    with pytest.raises(gpu.NrkError):
        K.reduce_axis("argmax", np.zeros(0), 3, 0, 1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_deterministic_across_launches(gpu, dtype):
    x = rng.normal(rng.BASE_SEED + 700, 3_000_017, dtype)
    a = [kfn("sum", dtype)(x) for _ in range(3)] + [kfn("var", dtype)(x, 0) for _ in range(3)]
    assert a[0] == a[1] == a[2] and a[3] == a[4] == a[5]


def test_prod_close_to_exact(gpu):
    x = 1.0 + rng.uniform(rng.BASE_SEED + 800, 100_000, -1e-4, 1e-4)
    exact = math.exp(math.fsum(np.log(x)))
    assert abs(K.prod_f64(x) - exact) / exact < 1e-12
    assert abs(K.prod_f32(x.astype(np.float32)) - exact) / exact < 1e-3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_axis0_extrema_special_values(gpu, oracle, dtype):
    """Extrema along a strided axis with few outputs (column tiles over ranges of the axis), bit-exact
    against the restatement on the cases where the order of the fold shows: a NaN first element
    (argmax / argmin 0, max / min NaN), a NaN after the winner, columns of all -inf / +inf (no
    strict winner: index 0), ties across tiles and ranges (the lowest index wins), +-0."""
    rows, cols = 9000, 70
    x = rng.normal(rng.BASE_SEED + 640, rows * cols, dtype).reshape(rows, cols)
    x[0, 1] = np.nan
    x[5000, 2] = np.nan
    x[:, 3] = -np.inf
    x[:, 4] = np.inf
    x[:, 5] = 7.0
    x[:, 6] = 0.0
    x[3000::1000, 6] = -0.0
    x[[130, 4100, 8999], 7] = 50.0
    x[[64, 4096, 8191], 8] = -50.0
    x[:, 9] = np.nan
    x[8999, 10] = 60.0
    x[[0, 8999], 11] = np.nan
    for op in ("min", "max", "argmin", "argmax"):
        got = K.reduce_axis(op, x.reshape(-1), 1, rows, cols)
        want = oracle.ref_axis(K.AXIS[op], x.reshape(-1), 1, rows, cols, 0)
        np.testing.assert_array_equal(np.asarray(got).view(np.uint8), np.asarray(want).view(np.uint8), err_msg=op)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_axis0_sums_special_values(gpu, dtype):
    """Sums, means and variances along a strided axis with few outputs (column tiles over ranges of
    the axis): an infinity gives an infinite sum (not the NaN of a TwoSum error term), +inf with
    -inf and NaN give NaN, columns of small integers are exact; variances as stated below."""
    rows, cols = 5000, 12
    x = np.ones((rows, cols), dtype)
    x[17, 1] = np.inf
    x[100, 2] = np.inf
    x[4000, 2] = -np.inf
    x[2500, 3] = np.nan
    x[:, 4] = np.arange(rows, dtype=dtype)
    s = K.reduce_axis("sum", x.reshape(-1), 1, rows, cols)
    m = K.reduce_axis("mean", x.reshape(-1), 1, rows, cols)
    assert s[0] == rows and s[1] == np.inf and np.isnan(s[2]) and np.isnan(s[3])
    assert s[4] == rows * (rows - 1) // 2 and m[4] == (rows - 1) / 2
    assert m[0] == 1.0 and m[1] == np.inf
    # variances (Chan records over 16-row blocks): a constant column is exactly 0; an infinite or
    # NaN element gives NaN, as the reference's mean-then-squares does (inf - inf); 0 .. n-1 has
    # the closed form (n^2 - 1) / 12 (ddof 0)
    v = K.reduce_axis("var", x.reshape(-1), 1, rows, cols)
    assert v[0] == 0.0 and np.isnan(v[1]) and np.isnan(v[2]) and np.isnan(v[3])
    want = (rows * rows - 1) / 12
    assert abs(float(v[4]) - want) <= VAR_C * EPS[dtype] * want
