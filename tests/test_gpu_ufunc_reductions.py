"""dot, norm_l1, norm_l2 and std (ufuncs.rs:946-979, 1060) through the C ABI on the GPU.

The reference sums the rounded products (or |x|, x * x) one after another in f64; the kernels
sum the same rounded terms in a fixed tree. Parity is therefore a bound, not bits: against the
exact sum of the same rounded terms (math.fsum), a tree of depth d adds at most d roundings to
each term, |err| <= d * 2^-53 * sum |term| (2^-24 for binary32), with d = 64 covering the
in-thread, wave, workgroup and final trees for every n tested here.
"""
import math

import numpy as np
import pytest

from numrs_amd import kernels as K

pytestmark = pytest.mark.gpu
GAMMA = {np.float64: 64 * 2.0**-53, np.float32: 64 * 2.0**-24}
SIZES = [1, 7, 255, 4096, 4097, 65_536 + 3, 1_000_003, 1 << 22]


def _bound(terms, dt):
    return GAMMA[dt] * math.fsum(np.abs(terms.astype(np.float64)))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", SIZES)
def test_dot_within_tree_bound(gpu, dt, n):
    r = np.random.default_rng(n)
    a = r.normal(size=n).astype(dt)
    b = r.normal(size=n).astype(dt)
    prods = (a * b).astype(dt)  # the reference's rounded products
    got = K.dot(a, b)
    exact = math.fsum(prods.astype(np.float64))
    assert abs(got - exact) <= _bound(prods, dt) + 1e-300
    assert K.dot(b, a) == got  # commutative to the bit (test_proptest_array_properties.rs:477)
    assert K.dot(a, b) == got  # the tree depends only on n


def test_dot_known_values_and_edges(gpu):
    assert K.dot(np.array([1.0, 2.0, 3.0]), np.array([4.0, 5.0, 6.0])) == 32.0
    assert K.dot(np.array([], np.float64), np.array([], np.float64)) == 0.0
    with pytest.raises(ValueError):
        K.dot(np.ones(3), np.ones(4))
    assert math.isnan(K.dot(np.array([1.0, np.nan]), np.array([1.0, 1.0])))
    assert K.dot(np.array([np.inf, 1.0]), np.array([2.0, 1.0])) == np.inf


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", SIZES)
def test_norms_within_tree_bound(gpu, dt, n):
    r = np.random.default_rng(n + 1)
    x = r.normal(size=n).astype(dt)
    l1 = K.norm_l1(x)
    assert abs(l1 - math.fsum(np.abs(x.astype(np.float64)))) <= _bound(x, dt)
    sq = (x * x).astype(dt)
    s = math.fsum(sq.astype(np.float64))
    l2 = K.norm_l2(x)
    # sqrt halves the relative error of the sum and adds one rounding
    eps = 2.0**-53 if dt == np.float64 else 2.0**-24
    assert abs(l2 - math.sqrt(s)) <= (0.5 * _bound(sq, dt) / s + eps) * math.sqrt(s)


def test_norm_and_std_known_values(gpu):
    assert K.norm_l2(np.array([3.0, 4.0])) == 5.0
    assert K.norm_l1(np.array([-1.0, 2.0, -3.0])) == 6.0
    assert K.norm_l2(np.array([], np.float64)) == 0.0
    assert K.std(np.array([2.0, 4.0, 4.0, 4.0, 5.0, 5.0, 7.0, 9.0])) == 2.0  # the textbook population std
    assert K.std(np.array([1.0, 2.0, 3.0, 4.0]), ddof=1) == pytest.approx(math.sqrt(5.0 / 3.0), rel=1e-15)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_std_is_sqrt_of_var(gpu, dt):
    x = np.random.default_rng(9).normal(3.0, 2.0, 1_000_003).astype(dt)
    for ddof in (0, 1):
        v = K.var_f64(x, ddof) if dt == np.float64 else K.var_f32(x, ddof)
        got = K.std(x, ddof)
        want = float(np.sqrt(dt(v)))  # var's bits, one correctly rounded sqrt
        assert got == want
        assert got == pytest.approx(float(np.std(x.astype(np.float64), ddof=ddof)), rel=1e-12 if dt == np.float64 else 1e-5)


LD = np.longdouble


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 5, 4097, 1_000_003])
def test_softmax_and_log_softmax(gpu, dt, n):
    """softmax / log_softmax (ufuncs.rs:1411-1457) against the exact values in long double. Every
    term of the sum is positive, so the tree's relative error is at most 64 eps; exp adds 1 ulp
    and the division one rounding."""
    r = np.random.default_rng(n + 7)
    x = (r.normal(size=n) * 8).astype(dt)
    xl = x.astype(LD)
    e = np.exp(xl - xl.max())
    s = e.sum()
    want = (e / s).astype(np.float64)
    eps = 2.0**-53 if dt == np.float64 else 2.0**-24
    got = K.softmax(x)
    assert got.dtype == dt and got.shape == x.shape
    assert np.all(np.abs(got - want) <= 70 * eps * want + 1e-300)
    lwant = ((xl - xl.max()) - np.log(s)).astype(np.float64)
    lgot = K.log_softmax(x)
    assert np.all(np.abs(lgot - lwant) <= 70 * eps * (1 + np.abs(np.log(np.float64(s)))) + 2 * eps * np.abs(lwant))


def test_softmax_special_inputs(gpu):
    assert K.softmax(np.array([], np.float64)).size == 0
    np.testing.assert_array_equal(K.softmax(np.array([3.0])), [1.0])
    np.testing.assert_array_equal(K.softmax(np.array([1.0, 1.0, 1.0, 1.0])), [0.25] * 4)
    np.testing.assert_array_equal(K.log_softmax(np.array([2.0, 2.0])), [-np.log(2.0)] * 2)
    # shift by the max: no overflow for large inputs
    big = np.array([1000.0, 1000.0 + np.log(3.0)])
    e = np.exp(LD(big[1] - big[0]))  # the difference is exact (Sterbenz); ~3 as rounded
    np.testing.assert_allclose(K.softmax(big), [float(1 / (1 + e)), float(e / (1 + e))], rtol=1e-15)
    # a NaN anywhere makes every output NaN, as in the reference (NaN reaches the sum)
    assert np.isnan(K.softmax(np.array([1.0, np.nan, 2.0]))).all()
    assert np.isnan(K.log_softmax(np.array([1.0, np.nan, 2.0]))).all()
    # 2-D input: the reference flattens, and so does this
    m = np.arange(6.0).reshape(2, 3)
    np.testing.assert_allclose(K.softmax(m).sum(), 1.0, rtol=1e-15)


# ---- ufunc_accumulate (ufunc_ops.rs:535-580): scans along contiguous rows

ACC_LENS = [1, 7, 2047, 2048, 2049, 70_001, 5_000_003]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", ACC_LENS)
def test_cummax_cummin_bit_exact(gpu, dt, n):
    """Maximum / minimum are exact and associative: the scan gives the sequential fold's bits,
    NaN propagating from its position on (UfuncOp::apply, ufunc_ops.rs:156)."""
    r = np.random.default_rng(n)
    x = r.normal(size=n).astype(dt)
    if n > 100:
        x[n // 2] = np.nan
    np.testing.assert_array_equal(K.accumulate("maximum", x), np.maximum.accumulate(x))
    np.testing.assert_array_equal(K.accumulate("minimum", x), np.minimum.accumulate(x))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", ACC_LENS)
def test_cumsum_and_subtract_within_tree_bound(gpu, dt, n):
    """The scan's sums run in a fixed tree (in-thread, wavefront, workgroup, carry levels): each
    prefix is within 128 eps * (prefix of |x|) of the exact prefix sum."""
    r = np.random.default_rng(n + 1)
    x = r.normal(size=n).astype(dt)
    eps = 2.0**-53 if dt == np.float64 else 2.0**-24
    exact = np.cumsum(x.astype(LD))
    bound = 128 * eps * np.cumsum(np.abs(x).astype(LD)) + 1e-300
    got = K.cumsum(x)
    assert got.dtype == dt
    assert np.all(np.abs(got.astype(LD) - exact) <= bound)
    sub = np.concatenate([[x[0]], x[0] - np.cumsum(x[1:].astype(LD))]) if n > 1 else x.astype(LD)
    assert np.all(np.abs(K.accumulate("subtract", x).astype(LD) - sub) <= bound)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cumprod_and_divide(gpu, dt):
    """Any bracketing of k products has k - 1 roundings: relative error <= k eps (2k for divide,
    which multiplies by rounded reciprocals)."""
    r = np.random.default_rng(5)
    n = 100_003
    x = r.uniform(0.999, 1.001, n).astype(dt)
    eps = 2.0**-53 if dt == np.float64 else 2.0**-24
    k = np.arange(1, n + 1, dtype=np.float64)
    exact = np.cumprod(x.astype(LD))
    assert np.all(np.abs(K.cumprod(x).astype(LD) / exact - 1) <= (k + 1) * eps)
    quot = x[0] / np.concatenate([[LD(1)], np.cumprod(x[1:].astype(LD))])
    assert np.all(np.abs(K.accumulate("divide", x).astype(LD) / quot - 1) <= (2 * k + 1) * eps)


def test_accumulate_axes_and_edges(gpu):
    r = np.random.default_rng(6)
    a = r.normal(size=(3, 4100, 5))
    for ax in (0, 1, 2, -1):
        got = K.cumsum(a, axis=ax)
        assert got.shape == a.shape
        np.testing.assert_allclose(got, np.cumsum(a, axis=ax), rtol=0, atol=1e-12)
        np.testing.assert_array_equal(K.accumulate("maximum", a, axis=ax), np.maximum.accumulate(a, axis=ax))
    np.testing.assert_allclose(K.cumsum(a), np.cumsum(a), rtol=0, atol=1e-11)  # axis None: flattened
    # the running sum starts at +0, so -0 inputs give +0 (statistics/mod.rs:584)
    z = K.cumsum(np.array([-0.0, -0.0, 1.0]))
    assert not np.signbit(z[0]) and not np.signbit(z[1])
    assert K.cumsum(np.array([], np.float64)).size == 0
    np.testing.assert_array_equal(K.cumsum(np.array([1.0, 2.0, 3.0, 4.0])), [1.0, 3.0, 6.0, 10.0])
    np.testing.assert_array_equal(K.accumulate("subtract", np.array([10.0, 1.0, 2.0])), [10.0, 9.0, 7.0])
    from numrs_amd._lib import call
    import ctypes as C
    x = np.ones(4)
    with pytest.raises(Exception):
        call("nrk_accumulate_f64", C.c_int32(K.BOP["rem"]), x.ctypes.data_as(C.c_void_p), C.c_size_t(1), C.c_size_t(4),
             x.ctypes.data_as(C.c_void_p))


# ---- ufunc_outer / ufunc_where (ufunc_ops.rs:583, 841)

NP_UFUNC = {"add": np.add, "subtract": np.subtract, "multiply": np.multiply, "divide": np.divide,
            "maximum": np.maximum, "minimum": np.minimum}  # numpy's maximum / minimum propagate NaN too


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("op", list(NP_UFUNC))
def test_outer_and_where_bit_exact(gpu, dt, op):
    r = np.random.default_rng(11)
    a = r.normal(size=(37, 3)).astype(dt)
    b = r.normal(size=3001).astype(dt)
    a[5, 1] = np.nan
    b[17] = np.nan
    b[18] = 0.0
    with np.errstate(all="ignore"):
        want = NP_UFUNC[op].outer(a, b)
    got = K.outer(op, a, b)
    assert got.shape == a.shape + b.shape and got.dtype == dt
    np.testing.assert_array_equal(got, want)
    x, y = r.normal(size=100_003).astype(dt), r.normal(size=100_003).astype(dt)
    x[::101] = np.nan
    m = r.random(100_003) < 0.5
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(K.where(op, x, y, m), np.where(m, NP_UFUNC[op](x, y), x))


def test_outer_where_edges(gpu):
    assert K.outer("add", np.array([], np.float64), np.ones(3)).shape == (0, 3)
    np.testing.assert_array_equal(K.add_where(np.array([1.0, 2.0]), np.array([10.0, 20.0]), np.array([True, False])),
                                  [11.0, 2.0])
    with pytest.raises(ValueError):
        K.where("add", np.ones(3), np.ones(3), np.ones(2, bool))


# ---- ufunc_reduceat (ufunc_ops.rs:619-733)

@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_reduceat(gpu, dt):
    r = np.random.default_rng(12)
    x = r.normal(size=200_003).astype(dt)
    x[777] = np.nan
    idx = np.array([0, 5, 5, 3, 1000, 70_000, 199_999, 10])  # repeats, decreasing, last segment to the end
    # maximum / minimum: exact, numpy's reduceat rule and NaN propagation, to the bit
    np.testing.assert_array_equal(K.reduceat("maximum", x, idx), np.maximum.reduceat(x, idx))
    np.testing.assert_array_equal(K.reduceat("minimum", x, idx), np.minimum.reduceat(x, idx))
    # add / subtract / multiply: the fixed tree against the exact segment values
    eps = 2.0**-53 if dt == np.float64 else 2.0**-24
    y = x.copy()
    y[777] = 0.5
    want = np.add.reduceat(y.astype(LD), idx)
    absum = np.add.reduceat(np.abs(y).astype(LD), idx)
    assert np.all(np.abs(K.reduceat("add", y, idx).astype(LD) - want) <= 64 * eps * absum)
    sub = np.array([y[j] - (np.sum(y[j + 1:e].astype(LD)) if e > j else 0)
                    for j, e in zip(idx, list(idx[1:]) + [y.size])], LD)
    assert np.all(np.abs(K.reduceat("subtract", y, idx).astype(LD) - sub) <= 64 * eps * absum)
    p = r.uniform(0.999, 1.001, 50_000).astype(dt)
    pidx = np.array([0, 100, 40_000])
    np.testing.assert_allclose(K.reduceat("multiply", p, pidx).astype(np.float64),
                               np.multiply.reduceat(p.astype(LD), pidx).astype(np.float64), rtol=50_000 * eps)
    # along an axis; signed zeros of a one-element and an all-zero segment are kept
    m = r.normal(size=(6, 9000)).astype(dt)
    np.testing.assert_array_equal(K.reduceat("maximum", m, [0, 10, 4000], axis=1), np.maximum.reduceat(m, [0, 10, 4000], axis=1))
    np.testing.assert_array_equal(K.reduceat("maximum", m, [0, 2], axis=0), np.maximum.reduceat(m, [0, 2], axis=0))
    z = np.array([-0.0, -0.0, 1.0], dt)
    got = K.reduceat("add", z, [0, 2])
    assert got[0] == 0 and np.signbit(got[0]) and got[1] == 1.0
    with pytest.raises(Exception):
        K.reduceat("add", x, [0, x.size])


# ---- ufunc_at (ufunc_ops.rs:772-838)

@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("op", list(NP_UFUNC))
def test_ufunc_at_bit_exact(gpu, dt, op):
    """Repeated indices fold in k order, as numpy's unbuffered op.at does: the same bits."""
    r = np.random.default_rng(13)
    a = r.uniform(0.5, 2.0, size=(300, 7)).astype(dt)
    idx = r.integers(0, 300, 5000)
    idx[:50] = 17  # a long run of one row
    b = r.uniform(0.5, 2.0, size=(5000, 7)).astype(dt)
    b[3, 2] = np.nan
    want = a.copy()
    with np.errstate(all="ignore"):
        NP_UFUNC[op].at(want, idx, b)
    got = a.copy()
    K.at(op, got, idx, b)
    np.testing.assert_array_equal(got, want)
    v = r.normal(size=1000).astype(dt)  # 1-D
    vi = r.integers(0, 1000, 100_000)
    vb = r.normal(size=100_000).astype(dt)
    wv = v.copy()
    with np.errstate(all="ignore"):
        NP_UFUNC[op].at(wv, vi, vb)
    K.at(op, v, vi, vb)
    np.testing.assert_array_equal(v, wv)


def test_ufunc_at_errors(gpu):
    a = np.zeros((4, 2))
    with pytest.raises(Exception):
        K.at("add", a, [0, 4], np.ones((2, 2)))
    assert not a.any()  # untouched after the rejected call
    with pytest.raises(ValueError):
        K.at("add", a, [0], np.ones((1, 3)))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_reduceat_max_min_zero_sign_follows_reference_fold(gpu, oracle, dt):
    """Segments longer than a workgroup (256 threads) of zeros of mixed signs: the reference's fold
    (reduce_lane, ufunc_ops.rs:240-257) lets an equal later operand win, so maximum / minimum return
    the segment's last zero. Also NaN results are the canonical NaN (UfuncOp::apply's T::nan())."""
    r = np.random.default_rng(51)
    x = np.where(r.integers(0, 2, 5000) == 1, -0.0, 0.0).astype(dt)
    idx = np.array([0, 300, 1299, 1300, 4000])
    x[1299], x[3999], x[4999] = -0.0, 0.0, -0.0
    for op in ("maximum", "minimum"):
        got = K.reduceat(op, x, idx)
        ends = list(idx[1:]) + [x.size]
        want = np.array([oracle.ref_ufunc_fold(op, x[j:e].tolist()) for j, e in zip(idx, ends)], dt)
        np.testing.assert_array_equal(np.signbit(got), np.signbit(want), err_msg=op)
        np.testing.assert_array_equal(got, want)
    y = np.ones(1000, dt)
    y[500] = -np.nan
    got = K.reduceat("maximum", y, [0, 600])
    nan_bits = np.array([np.nan], dt).view(np.uint64 if dt == np.float64 else np.uint32)[0]
    assert got.view(nan_bits.dtype)[0] == nan_bits and got[1] == 1.0


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_dot_bits_do_not_depend_on_alignment(gpu, dt):
    """nrk_dev_dot on operands offset by one element (8-byte f64, 4-byte f32: the scalar path)
    gives the same bits as on 16-byte aligned copies (the vector path), for full and tail
    segments."""
    import ctypes as C

    from numrs_amd import device
    from numrs_amd._lib import call

    fn = "nrk_dev_dot_f64" if dt == np.float64 else "nrk_dev_dot_f32"
    isz = np.dtype(dt).itemsize
    st = device.Stream()
    for n in (4096, 3 * 4096 + 100, 1_000_003):
        r = np.random.default_rng(n)
        a = r.normal(size=n).astype(dt)
        b = r.normal(size=n).astype(dt)
        da, db = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
        pa = device.DeviceArray.upload(np.concatenate([np.zeros(1, dt), a]))
        pb = device.DeviceArray.upload(np.concatenate([np.zeros(1, dt), b]))
        out = device.DeviceArray((2,), dt)
        print(f"dot alignment n={n} {np.dtype(dt).name}", flush=True)
        call(fn, C.c_void_p(da.ptr), C.c_void_p(db.ptr), C.c_size_t(n), C.c_void_p(out.ptr), C.c_void_p(st.handle))
        call(fn, C.c_void_p(pa.ptr + isz), C.c_void_p(pb.ptr + isz), C.c_size_t(n), C.c_void_p(out.ptr + isz),
             C.c_void_p(st.handle))
        st.synchronize()
        got = out.download()
        assert got[0].tobytes() == got[1].tobytes(), (n, got)
        assert got[0] == K.dot(a, b)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_at_device_form_skips_out_of_range_indices(gpu, dt):
    """Device form of ufunc_at: indices >= n0 are dropped, including ones whose low bits equal a
    valid row (10 and 2^40 + 2 against n0 = 5 would share the sort key of row 2 if only the low
    bits were sorted); the valid updates apply in k order, as numpy's op.at over the valid ones."""
    import ctypes as C

    from numrs_amd import device
    from numrs_amd._lib import call

    n0, row = 5, 3
    a = np.arange(n0 * row, dtype=dt).reshape(n0, row) + 1
    idx = np.array([2, 10, 2, (1 << 40) + 2, 4, 5, 7, 2], np.uint64)
    b = (np.arange(idx.size * row, dtype=dt).reshape(idx.size, row) + 1) * dt(0.5)
    valid = idx < n0
    want = a.copy()
    np.multiply.at(want, idx[valid].astype(np.int64), b[valid])
    st = device.Stream()
    da, di, dbb = device.DeviceArray.upload(a), device.DeviceArray.upload(idx), device.DeviceArray.upload(b)
    fn = "nrk_dev_at_f64" if dt == np.float64 else "nrk_dev_at_f32"
    call(fn, C.c_int32(2), C.c_void_p(da.ptr), C.c_size_t(n0), C.c_size_t(row), C.c_void_p(di.ptr), C.c_size_t(idx.size),
         C.c_void_p(dbb.ptr), C.c_void_p(st.handle))
    st.synchronize()
    np.testing.assert_array_equal(da.download(), want)


def test_softmax_f32_special_inputs(gpu):
    """binary32: a NaN anywhere makes every output NaN, +inf likewise (inf - inf); -inf entries give
    exactly 0; equal entries give exactly 1/n."""
    x = np.array([1.0, np.nan, 2.0] * 5000, np.float32)
    assert np.isnan(K.softmax(x)).all()
    assert np.isnan(K.softmax(np.array([1.0, np.inf, 2.0] * 5000, np.float32))).all()
    y = K.softmax(np.array([0.0, -np.inf, 0.0, 0.0] * 4096, np.float32))
    assert np.all(y[1::4] == 0) and np.all(y[0::4] == np.float32(1 / 12288))
