"""The oracle pinned against the reference's own vectors, closed forms and brute force (CPU)."""
import math

import numpy as np
import pytest

from conftest import det_data, fval

F64_BITS = lambda x: int(np.float64(x).view(np.uint64))  # noqa: E731
F32_BITS = lambda x: int(np.float32(x).view(np.uint32))  # noqa: E731


@pytest.mark.parametrize("threads", [1, 8])
def test_determinism_pins_f64(oracle, known, threads):
    """tests/test_reduction_determinism.rs:126-176 bit patterns, under 1 and 8 threads."""
    pins = known["determinism_pins"]
    oracle.set_threads(threads)
    x = det_data(np.float64)
    assert F64_BITS(oracle.ref("sum", x)) == pins["sum_f64_bits"]
    assert F64_BITS(oracle.ref("mean", x)) == pins["mean_f64_bits"]
    var = oracle.ref("var", x, 0)
    assert F64_BITS(var) == pins["var_ddof0_f64_bits"]
    assert F64_BITS(math.sqrt(var)) == pins["std_ddof0_f64_bits"]
    oracle.set_threads(1)


def test_determinism_pin_f32(oracle, known):
    x = det_data(np.float32)
    assert F32_BITS(oracle.ref("sum", x)) == known["determinism_pins"]["sum_f32_bits"]


def test_dd_stats_match_pins_and_closed_form(oracle, known):
    """The double-double reference hits the pinned values (exact here) and Faulhaber's closed form."""
    pins = known["determinism_pins"]
    x = det_data(np.float64)
    st = oracle.dd_stats(x, 0)
    assert F64_BITS(st["sum"]) == pins["sum_f64_bits"]
    assert F64_BITS(st["mean"]) == pins["mean_f64_bits"]
    assert F64_BITS(st["var"]) == pins["var_ddof0_f64_bits"]
    # closed form: sum_{i<n} (i/2 - c) = n(n-1)/4 - n c ; var = (1/4)(n^2 - 1)/12
    n = 50000
    assert st["sum"] == n * (n - 1) / 4 - n * 12345.0
    assert st["var"] == pytest.approx(0.25 * (n * n - 1) / 12.0, rel=1e-15)


def test_dd_sum_is_exact_beyond_double(oracle):
    """A case plain double summation gets wrong: 1e16 + 1 - 1e16 + ... (catastrophic cancellation)."""
    x = np.array([1e16, 1.0, -1e16, 1.0, 1e-3] * 100)
    st = oracle.dd_stats(x)
    assert st["sum"] == pytest.approx(200.1, rel=1e-14)
    naive = 0.0
    for v in x:
        naive = naive + v
    assert naive != pytest.approx(200.1, rel=1e-3)  # a left-to-right double fold loses the 1s


def test_var_small_known(oracle, known):
    k = known["var_small"]
    x = np.array(k["x"])
    assert oracle.ref("var", x, 0) == k["ddof0"]
    assert oracle.ref("var", x, 1) == k["ddof1"]
    assert oracle.ref("sum_sq_dev", x, 3.0) == 10.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_min_max_known_nan_inf(oracle, known, dtype):
    k = known["min_max_known"]
    x = np.array(k["x"], dtype)
    assert oracle.ref("min", x) == k["min"] and oracle.ref("max", x) == k["max"]
    for case in known["min_max_nan"]["cases"]:
        x = np.array([fval(v) for v in case], dtype)
        assert math.isnan(oracle.ref("min", x)) and math.isnan(oracle.ref("max", x))
    for case in known["min_max_inf"]["cases"]:
        x = np.array([fval(v) for v in case["x"]], dtype)
        assert oracle.ref("min", x) == fval(case["min"]) and oracle.ref("max", x) == fval(case["max"])
    la = known["lane_aligned_nan"]
    for pos in la["nan_positions"]:
        d = np.full(la["len"], la["fill"], dtype)
        d[0] = la["first"]
        d[pos] = np.nan
        assert math.isnan(oracle.ref("max", d)) and math.isnan(oracle.ref("min", d))
    assert oracle.ref("min", np.zeros(0, dtype)) == 0.0 and oracle.ref("max", np.zeros(0, dtype)) == 0.0


def test_min_max_tiers_match_numpy(oracle):
    """No-NaN min/max are order independent: every tier equals numpy (reduce.rs:929-965 grid)."""
    rng = np.random.default_rng(3)
    for n in [1, 63, 64, 9999, 10000, 10001, 79999, 80000, 80001, 200000]:
        x = rng.standard_normal(n)
        assert oracle.ref("min", x) == x.min() and oracle.ref("max", x) == x.max()
        nan_at = [0, n // 2, n - 1]
        for p in nan_at:
            y = x.copy()
            y[p] = np.nan
            assert math.isnan(oracle.ref("max", y))


def test_argmax_semantics(oracle, known):
    for key, fn in (("argmax_doc", "argmax"), ("argmin_doc", "argmin")):
        k = known[key]
        x = np.array(k["x"])
        assert oracle.ref(fn, x) == k["flat"]
        r, c = k["shape"]
        assert list(oracle.ref_axis(6 if fn == "argmax" else 5, x, r, c, 1)) == k["axis1"]
    # first strict maximum; interior NaN skipped; NaN at 0 pins 0 (statistics/mod.rs:92-101)
    assert oracle.ref("argmax", np.array([1.0, 5.0, 5.0, 2.0])) == 1
    assert oracle.ref("argmax", np.array([1.0, np.nan, 3.0])) == 2
    assert oracle.ref("argmax", np.array([np.nan, 9.0, 3.0])) == 0
    assert oracle.ref("argmax", np.array([-0.0, 0.0])) == 0
    # brute force against numpy on NaN-free data (numpy's argmax is also first-occurrence)
    rng = np.random.default_rng(5)
    for n in [1, 2, 100, 4097]:
        x = rng.integers(-5, 5, n).astype(np.float64)
        assert oracle.ref("argmax", x) == int(np.argmax(x))
        assert oracle.ref("argmin", x) == int(np.argmin(x))


def test_numpy_statistics_fixtures(oracle, numpy_ref):
    """The reference's NumPy-generated statistics (1e-12, as test_array_operations_reference.rs:486-551)."""
    ops = {"sum": 0, "mean": 1, "var": 2, "min": 3, "max": 4}
    checked = 0
    for name, case in numpy_ref["statistical_operations"].items():
        if case["operation"] not in ops and case["operation"] != "std":
            continue  # percentile is outside the kernels path
        checked += 1
        x = np.array(case["input"]["data"], np.float64)
        shape = case["input"]["shape"]
        res = case["result"]
        want = np.atleast_1d(np.array(res["data"] if isinstance(res, dict) else res, np.float64)).reshape(-1)
        op, axis = case["operation"], case.get("axis")
        if axis is None:
            if op == "std":
                got = np.array([math.sqrt(oracle.ref("var", x, 0))])
            elif op == "var":
                got = np.array([oracle.ref("var", x, 0)])
            else:
                got = np.array([oracle.ref(op, x)])
        else:
            rows, cols = (shape + [1])[:2] if len(shape) == 2 else (1, shape[0])
            outer, length, inner = (rows, cols, 1) if axis == 1 else (1, rows, cols)
            if op == "std":
                got = np.sqrt(oracle.ref_axis(2, x, outer, length, inner, 0))
            else:
                got = oracle.ref_axis(ops[op], x, outer, length, inner)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12, err_msg=name)
    assert checked >= 30


def test_numpy_arithmetic_fixtures(oracle, numpy_ref):
    bops = {"add": 0, "subtract": 1, "multiply": 2, "divide": 3}
    sops = {"scalar_add": 0, "scalar_subtract": 1, "scalar_multiply": 2, "scalar_divide": 3}
    for name, case in numpy_ref["arithmetic_operations"].items():
        want = np.array(case["result"]["data"], np.float64)
        if case["operation"] in bops:
            a = np.array(case["input1"]["data"], np.float64)
            b = np.array(case["input2"]["data"], np.float64)
            got = oracle.ref_zip(bops[case["operation"]], a, b)
        else:
            x = np.array(case["input"]["data"], np.float64)
            got = oracle.ref_map([sops[case["operation"]]], [case["scalar"]], x)
        np.testing.assert_array_equal(got, want, err_msg=name)


def test_copysign_matches_ieee(oracle):
    """Rust's f64::copysign is the IEEE copySign: magnitude of a, sign bit of b, NaN signs too."""
    a = np.array([0.0, -0.0, 1.5, -2.0, np.inf, -np.inf, np.nan, 5e-324, 3.0, -4.0])
    b = np.array([-1.0, 1.0, -0.0, 0.0, -np.nan, np.nan, -1.0, -np.inf, np.nan, -np.nan])
    for t in (np.float64, np.float32):
        want, got = np.copysign(a.astype(t), b.astype(t)), oracle.ref_zip(7, a.astype(t), b.astype(t))
        np.testing.assert_array_equal(got.view(np.uint64 if t == np.float64 else np.uint32),
                                      want.view(np.uint64 if t == np.float64 else np.uint32))


def test_prelu_matches_numpy(oracle):
    """prelu (ufuncs.rs:1405): x where x > 0, else alpha * x; NaN x gives NaN."""
    x = np.array([1.5, -2.0, 0.0, -0.0, np.nan, -np.inf, np.inf, -3.0])
    al = np.array([0.1, 0.25, 0.5, 0.5, 0.5, 0.5, 0.5, -1.0])
    np.testing.assert_array_equal(oracle.ref_zip(11, x, al), [1.5, -0.5, 0.0, -0.0, np.nan, -np.inf, np.inf, 3.0])


def test_fma_is_one_rounding(oracle):
    """ufuncs.rs:864 mul_add: the exact a*b + c rounded once, checked with rational arithmetic."""
    from fractions import Fraction
    r = np.random.default_rng(77)
    a, b = r.normal(size=3000), r.normal(size=3000)
    c = -(a * b) + r.normal(size=3000) * 1e-17  # cancellation: the low half of a*b decides
    for t in (np.float64, np.float32):
        A, B, Cc = a.astype(t), b.astype(t), c.astype(t)
        got = oracle.ref_fma(A, B, Cc)
        exact = [Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z)) for x, y, z in zip(A, B, Cc)]
        if t == np.float64:  # float(Fraction) is the correctly rounded double
            np.testing.assert_array_equal(got, np.array([float(e) for e in exact]))
        else:  # the nearest binary32: no neighbour is closer to the exact value
            for g, e in zip(got, exact):
                lo, hi = np.nextafter(g, np.float32(-np.inf)), np.nextafter(g, np.float32(np.inf))
                d = abs(Fraction(float(g)) - e)
                assert d <= abs(Fraction(float(lo)) - e) and d <= abs(Fraction(float(hi)) - e)


def test_lerp_is_three_roundings(oracle):
    r = np.random.default_rng(78)
    a, b = r.normal(size=5000), r.normal(size=5000)
    for t in (0.0, 0.3, 1.0, -2.5):
        np.testing.assert_array_equal(oracle.ref_lerp(a, b, t), a + t * (b - a))
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    np.testing.assert_array_equal(oracle.ref_lerp(a32, b32, 0.3), a32 + np.float32(0.3) * (b32 - a32))


def test_two_roundings_not_fma(oracle):
    """y = x*1.5 + (-0.25) is two roundings: bit-equal to numpy's separate ops, and for some x
    different from the fused (single-rounding) value."""
    x = np.random.default_rng(7).standard_normal(100000)
    y = oracle.ref_map([2, 0], [1.5, -0.25], x)
    np.testing.assert_array_equal(y, (x * 1.5) + (-0.25))
    x32 = x.astype(np.float32)
    y32 = oracle.ref_map([2, 0], [1.5, -0.25], x32)
    np.testing.assert_array_equal(y32, (x32 * np.float32(1.5)) + np.float32(-0.25))
    fused = (x32.astype(np.float64) * 1.5 - 0.25).astype(np.float32)
    assert np.any(fused != y32)


def test_axis_restatement_brute_force(oracle):
    rng = np.random.default_rng(11)
    x = rng.standard_normal((7, 13))
    # row folds left to right == python loop in the same order
    s = oracle.ref_axis(0, x, 7, 13, 1)
    for r in range(7):
        acc = 0.0
        for v in x[r]:
            acc = acc + v
        assert s[r] == acc
    # column view (axis 0)
    c = oracle.ref_axis(0, x, 1, 7, 13)
    for j in range(13):
        acc = 0.0
        for v in x[:, j]:
            acc = acc + v
        assert c[j] == acc
    np.testing.assert_array_equal(oracle.ref_axis(6, x, 7, 13, 1), np.argmax(x, axis=1))
    np.testing.assert_array_equal(oracle.ref_axis(5, x, 1, 7, 13), np.argmin(x, axis=0))
    np.testing.assert_allclose(oracle.ref_axis(2, x, 7, 13, 1, 1), np.var(x, axis=1, ddof=1), rtol=1e-13)
    y = x.copy()
    y[2, 5] = np.nan
    mx = oracle.ref_axis(4, y, 7, 13, 1)
    assert math.isnan(mx[2]) and not np.isnan(np.delete(mx, 2)).any()


def test_i64_sum_exact_and_wrapping(oracle):
    rng = np.random.default_rng(13)
    x = rng.integers(-(1 << 20), (1 << 20) + 1, 1 << 16, dtype=np.int64)
    assert oracle.sum_i64(x) == int(x.sum(dtype=np.int64)) == oracle.exact_sum_i64(x)
    big = np.array([np.iinfo(np.int64).max, 1], np.int64)
    assert oracle.sum_i64(big) == np.iinfo(np.int64).min  # release-build wrap
    assert oracle.exact_sum_i64(big) == (1 << 63)


def test_gemm_generic_known_and_brute_force(oracle, known):
    k = known["matmul_3x3"]
    a = np.array(k["a"], np.float64).reshape(3, 3)
    np.testing.assert_array_equal(oracle.gemm_generic(a, a).reshape(-1), np.array(k["c"], np.float64))
    np.testing.assert_array_equal(oracle.gemm_generic(a, np.array(k["v"], np.float64).reshape(3, 1)).reshape(-1),
                                  np.array(k["av"], np.float64))
    rng = np.random.default_rng(17)
    ai = rng.integers(-50, 50, (70, 130)).astype(np.int64)
    bi = rng.integers(-50, 50, (130, 65)).astype(np.int64)
    np.testing.assert_array_equal(oracle.gemm_generic(ai, bi), ai @ bi)
    # wrapping (Rust release arithmetic): numpy's integer matmul wraps the same way
    r2 = np.random.default_rng(18)
    for dt in (np.int64, np.int32):
        info = np.iinfo(dt)
        aw = r2.integers(info.min, info.max, (9, 70), dtype=dt, endpoint=True)
        bw = r2.integers(info.min, info.max, (70, 11), dtype=dt, endpoint=True)
        with np.errstate(over="ignore"):
            np.testing.assert_array_equal(oracle.gemm_generic(aw, bw), aw @ bw)
    # f64: ordered k-sum == explicit loop
    af = rng.standard_normal((5, 70))
    bf = rng.standard_normal((70, 4))
    c = oracle.gemm_generic(af, bf)
    for i in range(5):
        for j in range(4):
            acc = 0.0
            for p in range(70):
                acc = acc + af[i, p] * bf[p, j]
            assert c[i, j] == acc


def test_dd_gemm_exact_on_integers_and_close_to_long_sum(oracle):
    rng = np.random.default_rng(19)
    a = rng.integers(-1000, 1000, (9, 33)).astype(np.float64)
    b = rng.integers(-1000, 1000, (33, 7)).astype(np.float64)
    np.testing.assert_array_equal(oracle.dd_gemm(a, b), a @ b)
    a = rng.standard_normal((6, 257))
    b = rng.standard_normal((257, 5))
    import fractions
    exact = [[float(sum(fractions.Fraction(a[i, p]) * fractions.Fraction(b[p, j]) for p in range(257)))
              for j in range(5)] for i in range(6)]
    np.testing.assert_array_equal(oracle.dd_gemm(a, b), np.array(exact))
    rows = np.array([0, 3, 5])
    cols = np.array([4, 0, 2])
    np.testing.assert_array_equal(oracle.dd_gemm_entries(a, b, rows, cols), np.array(exact)[rows, cols])


def test_rowsplit_stand_in_close_to_exact(oracle):
    rng = np.random.default_rng(23)
    a = rng.uniform(-1, 1, (300, 200))
    b = rng.uniform(-1, 1, (200, 170))
    for t in (1, 8):
        oracle.set_threads(t)
        c = oracle.gemm_rowsplit_f64(a, b)
        np.testing.assert_allclose(c, oracle.dd_gemm(a, b), rtol=0, atol=1e-12)
    oracle.set_threads(1)


def test_planted_mistake_is_caught(oracle, known):
    """Sensitivity check: a restatement that drops the last element, or sums in f32 instead of f64,
    misses the pinned bits -- so the pin test above would catch such a mistake."""
    x = det_data(np.float64)
    pins = known["determinism_pins"]
    assert F64_BITS(oracle.ref("sum", x[:-1])) != pins["sum_f64_bits"]
    assert F64_BITS(oracle.ref("var", x, 1)) != pins["var_ddof0_f64_bits"]


def _mlp_case(K=3, N=9, seed=31):
    from oracle import mlp

    r = np.random.default_rng(seed)
    p = r.normal(0, 0.4, mlp.param_count(K))
    x = r.normal(0, 1, (N, 6))
    t = r.normal(0, 1, (N, K))
    mask = (r.uniform(size=(N, K)) > 0.25).astype(np.float64)
    return mlp, p, x, t, mask


def test_mlp_oracle_gradient_matches_finite_differences():
    """Central differences (h = 1e-6, FP64) pin the analytic backward of the MLP oracle."""
    mlp, p, x, t, mask = _mlp_case()
    _, _, g = mlp.loss_and_grad(p, 3, x, t, mask)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp.loss_and_grad(p + e, 3, x, t, mask)[1] - mlp.loss_and_grad(p - e, 3, x, t, mask)[1]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_mlp_oracle_planted_mistake_is_caught():
    """A backward that forgets the ReLU mask of the second layer no longer matches differences."""
    mlp, p, x, t, mask = _mlp_case(seed=32)
    K = 3
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p, K)
    y, (z1, h1, z2, h2) = mlp.forward(p, K, x)
    dy = 2.0 * (y - t) * mask / mask.sum()
    wrong_gb2 = (dy @ W3).sum(axis=0)  # missing * (z2 > 0)
    _, _, g = mlp.loss_and_grad(p, K, x, t, mask)
    o = 6 * 64 + 64 + 64 * 64
    assert (z2 <= 0).any() and not np.allclose(wrong_gb2, g[o:o + 64])


def test_mlp_oracle_adam_closed_form():
    """First Adam step from m = v = 0 moves every parameter by lr * sign(g) (up to eps)."""
    from oracle import mlp

    g = np.array([0.5, -2.0, 1e-3, -7.0])
    p1, m, v = mlp.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    np.testing.assert_allclose(p1, -0.01 * np.sign(g), rtol=1e-4)


def test_ufunc_fold_oracle_pins():
    """ref_ufunc_fold (reduce_lane + UfuncOp::apply): maximum / minimum equal numpy's reductions on
    tie-free data; on an all-zero segment the result is the last element (a tie goes to the later
    operand), so its sign is the last zero's; NaN anywhere gives NaN; add of small integers is the
    closed-form n(n-1)/2."""
    import math as _m

    import oracle

    r = np.random.default_rng(41)
    x = r.normal(size=300)
    assert oracle.ref_ufunc_fold("maximum", x) == x.max()
    assert oracle.ref_ufunc_fold("minimum", x) == x.min()
    for last in (0.0, -0.0):
        z = [(-0.0 if r.integers(2) else 0.0) for _ in range(50)] + [last]
        assert _m.copysign(1.0, oracle.ref_ufunc_fold("maximum", z)) == _m.copysign(1.0, last)
        assert _m.copysign(1.0, oracle.ref_ufunc_fold("minimum", z)) == _m.copysign(1.0, last)
    assert _m.isnan(oracle.ref_ufunc_fold("maximum", [1.0, _m.nan, 3.0]))
    assert _m.isnan(oracle.ref_ufunc_fold("minimum", [_m.nan, 1.0]))
    assert oracle.ref_ufunc_fold("add", [float(i) for i in range(1000)]) == 1000 * 999 / 2


def test_integer_elementwise_oracle_pins():
    """ref_zip_int / ref_map_int (Rust release arithmetic): the reference's doc examples
    (bitwise_ops.rs:25-28 and, or, xor; :186-189 left_shift; :224-227 right_shift), numpy's
    wrapping + - * & | ^ on full-range values, Python's exact integers for truncated / and % and
    for masked shifts."""
    import fractions

    import oracle

    a = np.array([13, 17, 21], np.int64)
    b = np.array([9, 7, 15], np.int64)
    assert list(oracle.ref_zip_int(13, a, b)) == [9, 1, 5]
    assert list(oracle.ref_zip_int(14, a, b)) == [13, 23, 31]
    assert list(oracle.ref_zip_int(15, a, b)) == [4, 22, 26]
    assert list(oracle.ref_zip_int(16, np.array([5, 10, 15], np.int32), np.array([1, 2, 3], np.int32))) == [10, 40, 120]
    assert list(oracle.ref_zip_int(17, np.array([40, 80, 120], np.int32), np.array([1, 2, 3], np.int32))) == [20, 20, 15]
    r = np.random.default_rng(61)
    for dt, bits in ((np.int64, 64), (np.int32, 32), (np.uint64, 64), (np.uint32, 32)):
        info = np.iinfo(dt)
        x = r.integers(info.min, info.max, 2000, dtype=dt, endpoint=True)
        y = r.integers(info.min, info.max, 2000, dtype=dt, endpoint=True)
        with np.errstate(over="ignore"):
            for op, f in ((0, np.add), (1, np.subtract), (2, np.multiply), (13, np.bitwise_and), (14, np.bitwise_or),
                          (15, np.bitwise_xor), (5, np.maximum), (6, np.minimum)):
                np.testing.assert_array_equal(oracle.ref_zip_int(op, x, y), f(x, y))
        y[y == 0] = 3
        if info.min < 0:
            y[(y == -1) & (x == info.min)] = 5
        q, rm = oracle.ref_zip_int(3, x, y), oracle.ref_zip_int(4, x, y)
        for i in range(200):
            tq = int(fractions.Fraction(int(x[i]), int(y[i])))  # truncates toward zero
            assert int(q[i]) == tq and int(rm[i]) == int(x[i]) - int(y[i]) * tq
        s = r.integers(-200, 200, 2000).astype(dt)
        sh = oracle.ref_zip_int(16, x, s)
        for i in range(200):
            amt = int(s[i]) & (bits - 1)
            want = (int(x[i]) << amt) & ((1 << bits) - 1)
            assert int(sh[i].astype(np.uint64 if bits == 64 else np.uint32)) == want
    v = np.array([-5, 0, 7, np.iinfo(np.int64).min], np.int64)
    got = oracle.ref_map_int([2, 0, 6, 7], [3, 1, 0, 0], v)  # |-(3v + 1)|, wrapping
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(got, np.abs(-(v * 3 + 1)))
