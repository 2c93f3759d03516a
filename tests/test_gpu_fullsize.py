"""The measured workloads at their full sizes, checked through size-independent properties.

- Fused y = 1.5 x - 0.25 -> sum/max/argmax/var on x (8192, 8192) f32 and f64 (the headline step):
  row max/argmax bit-exact against the restatement (numpy over the oracle's two-rounding map);
  whole-array max/argmax equal to the max of the row maxima and its flat index; whole-array
  sum and var within the stated bounds of the double-double statistics of y; the whole-array sum
  within the same bound of the double-double sum of the row sums (a checksum of checksums).
- i64 sum of 2^28 values uniform in [-2^20, 2^20]: exact.
- Ozaki-II FP64 GEMM (14 and 16 moduli) at 4096^3 and 8192^3, and the FP32 split GEMMs at
  8192^3: on 512 sampled entries against the double-double dot products of the oracle (error scale
  sum_k |A_ik||B_kj|), never less accurate than the native FP64 GEMM, resp. within the split bars."""
import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng
from test_gpu_fused import run_fused

pytestmark = pytest.mark.gpu
EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_headline_fused_step(gpu, oracle, dtype):
    rows = cols = 8192
    x = rng.normal(rng.BASE_SEED + (2 if dtype == np.float32 else 3), rows * cols, dtype).reshape(rows, cols)
    out = run_fused(x, [("mul", 1.5), ("add", -0.25)])
    y = oracle.ref_map([2, 0], [1.5, -0.25], x)
    rmax = y.max(axis=1)
    rarg = y.argmax(axis=1)
    np.testing.assert_array_equal(out["row_max"], rmax)
    np.testing.assert_array_equal(out["row_argmax"], rarg.astype(np.uint64))
    # whole array: the max of the row maxima, at its first flat index
    r = int(np.argmax(rmax))
    assert out["full_max"][0] == rmax[r]
    assert int(out["full_argmax"][0]) == r * cols + int(rarg[r]) == int(np.argmax(y))
    # sums and variance against the double-double statistics of y
    ex = oracle.dd_stats(y)
    abs_sum = float(np.abs(y).astype(np.float64).sum())
    assert abs(float(out["full_sum"][0]) - ex["sum"]) <= 8 * EPS[dtype] * abs_sum + EPS[dtype] * abs(ex["sum"])
    assert abs(float(out["full_var"][0]) - ex["var"]) <= 64 * EPS[dtype] * ex["var"]
    # checksum of checksums: the whole-array sum agrees with the exact sum of the row sums
    row_sums = out["row_sum"].astype(np.float64)
    s_rows = oracle.dd_stats(row_sums)["sum"]
    assert abs(float(out["full_sum"][0]) - s_rows) <= 8 * EPS[dtype] * abs_sum + rows * EPS[dtype] * np.abs(row_sums).max()


def test_i64_sum_2_28_exact(gpu):
    n = 1 << 28
    x = rng.int_uniform(rng.BASE_SEED + 4, n, -(1 << 20), 1 << 20)
    assert K.sum_i64(x) == int(x.sum(dtype=np.int64))


def _sampled(oracle, a, b, cs, count, seed):
    """max over `count` sampled entries of |C - C*| / sum_k |A_ik||B_kj| for each C in cs, C* the
    double-double dot product of the oracle."""
    n, m = b.shape[1], a.shape[0]
    pick = np.random.default_rng(seed)
    rows, cols = pick.integers(0, m, count), pick.integers(0, n, count)
    exact = oracle.dd_gemm_entries(a, b, rows, cols)
    scale = oracle.abs_gemm_entries(a, b, rows, cols)
    return [float((np.abs(c[rows, cols].astype(np.float64) - exact) / scale).max()) for c in cs]


@pytest.mark.parametrize("n", [4096, 8192])
def test_ozaki14_not_less_accurate_than_native(gpu, oracle, n):
    """The benchmark's choice at native-or-better error, on the benchmark's own inputs (same seeds):
    Ozaki-II with 14 moduli against the native FP64 MFMA GEMM on 512 sampled entries; 16 moduli
    (the default) likewise."""
    a = rng.scaled_pow2(rng.BASE_SEED + 40 + n, n * n).reshape(n, n)
    b = rng.scaled_pow2(rng.BASE_SEED + 41 + n, n * n).reshape(n, n)
    c_nat = np.empty((n, n))
    K.gemm_2d(n, n, n, a, b, c_nat)
    c14 = K.gemm_f64_ozaki(a, b, 14)
    c16 = K.gemm_f64_ozaki(a, b, 16)
    e_nat, e14, e16 = _sampled(oracle, a, b, [c_nat, c14, c16], 512, 5 + n)
    print(f"{n}^3 max error over 512 entries: native {e_nat:.3e}, ozaki14 {e14:.3e}, ozaki16 {e16:.3e}")
    assert e14 <= e_nat and e16 <= e_nat


def test_split_gemm_8192_f32(gpu, oracle):
    """FP32 at 8192^3 on the benchmark's inputs: native FP32 (128 x 128 tile), BF16x3 (the
    128 x 256 split tile) and BF16x6 (128 x 128) on 512 sampled entries against the double-double
    oracle. Bars: native within gamma_k = k u / (1 - k u) of sum |A||B| (the classical dot-product
    bound, u = 2^-24); BF16x6 (A0*B0 accumulated apart from the five smaller pairs) no less
    accurate than native; BF16x3 <= 2^-13 (its dropped pairs are of order 2^-16 per product)."""
    n = 8192
    a = rng.scaled_pow2(rng.BASE_SEED + 40 + n, n * n, dtype=np.float32).reshape(n, n)
    b = rng.scaled_pow2(rng.BASE_SEED + 41 + n, n * n, dtype=np.float32).reshape(n, n)
    c_nat = np.empty((n, n), np.float32)
    K.gemm_2d(n, n, n, a, b, c_nat)
    c3 = K.gemm_f32_bf16x(a, b, 3)
    c6 = K.gemm_f32_bf16x(a, b, 6)
    e_nat, e3, e6 = _sampled(oracle, a, b, [c_nat, c3, c6], 512, 17)
    print(f"8192^3 f32 max error over 512 entries: native {e_nat:.3e}, bf16x3 {e3:.3e}, bf16x6 {e6:.3e}")
    u = 2.0 ** -24
    assert e_nat <= n * u / (1 - n * u)
    assert e6 <= e_nat
    assert e3 <= 2.0 ** -13


@pytest.mark.parametrize("n", [(1 << 32) + 4099, 5 << 30])
def test_whole_array_beyond_2_32_elements(gpu, n):
    """Whole-array reductions over more than 2^32 binary32 elements (16 and 20 GiB, filled on the
    device): ones, with 5 and -3 planted past index 2^32 -- in the tail after the last whole chunk
    for 2^32 + 4099, inside whole chunks for 5 * 2^30. The sum is exact in binary64 (every partial
    is an integer below 2^53), so it is the binary32 rounding of n; max / argmax and min / argmin
    give the planted values at their 64-bit indices; the variance is 32 / n."""
    import ctypes as C
    from numrs_amd._lib import ReduceOutputs, call
    from numrs_amd.device import DeviceArray

    print(f"whole array of {n} f32", flush=True)
    d = DeviceArray((n,), np.float32)
    call("nrk_memset_zero", C.c_void_p(d.ptr), C.c_size_t(n * 4), None)
    ops, n_ops = K.chain(("add", 1.0))
    call("nrk_dev_map_f32", C.cast(ops, C.c_void_p), n_ops, C.c_void_p(d.ptr), C.c_void_p(d.ptr), C.c_size_t(n), None)
    i_max, i_min = (1 << 32) + 17, (1 << 32) + 4000
    for i, v in ((i_max, 5.0), (i_min, -3.0)):
        val = np.array([v], np.float32)
        call("nrk_memcpy_htod", C.c_void_p(d.ptr + 4 * i), val.ctypes.data_as(C.c_void_p), C.c_size_t(4), None)
    outs, fo = {}, ReduceOutputs()
    for f in ("sum", "mean", "var", "min", "max", "argmin", "argmax"):
        outs[f] = DeviceArray((1,), np.uint64 if f.startswith("arg") else np.float32)
        setattr(fo, f, outs[f].ptr)
    empty, _ = K.chain()
    call("nrk_dev_map_reduce_f32", C.cast(empty, C.c_void_p), 0, C.c_void_p(d.ptr), C.c_size_t(1), C.c_size_t(n), C.c_size_t(0),
         None, C.byref(fo), None)
    call("nrk_device_synchronize")
    got = {k: v.download()[0] for k, v in outs.items()}
    assert got["sum"] == np.float32(float(n))
    assert got["mean"] == 1.0
    assert got["max"] == 5.0 and got["argmax"] == i_max
    assert got["min"] == -3.0 and got["argmin"] == i_min
    assert abs(float(got["var"]) - 32.0 / n) <= 1e-6 * 32.0 / n
    del d


@pytest.mark.parametrize("shape", [((1 << 22) + 3, 1025), ((1 << 21) + 1, 2049)])
def test_rows_beyond_2_32_elements(gpu, shape):
    """Row and whole-array results over more than 2^32 binary32 elements, rows off the 16-byte grid
    ((2^22 + 3) x 1025: the short-row kernel; (2^21 + 1) x 2049: the general kernel, its
    whole-array result through the record tree): ones, with 5 planted in a late row. Every row sum
    is cols (cols + 4 in the planted row) and every row mean 1; the planted row's argmax is its
    column; the whole-array argmax is the planted element's 64-bit flat index and the whole-array
    sum is exact before its rounding."""
    import ctypes as C
    from numrs_amd._lib import ReduceOutputs, call
    from numrs_amd.device import DeviceArray

    rows, cols = shape
    n = rows * cols
    print(f"rows {rows} x {cols} f32", flush=True)
    d = DeviceArray((n,), np.float32)
    call("nrk_memset_zero", C.c_void_p(d.ptr), C.c_size_t(n * 4), None)
    ops, n_ops = K.chain(("add", 1.0))
    call("nrk_dev_map_f32", C.cast(ops, C.c_void_p), n_ops, C.c_void_p(d.ptr), C.c_void_p(d.ptr), C.c_size_t(n), None)
    r5, c5 = rows - 2, 1000
    val = np.array([5.0], np.float32)
    call("nrk_memcpy_htod", C.c_void_p(d.ptr + 4 * (r5 * cols + c5)), val.ctypes.data_as(C.c_void_p), C.c_size_t(4), None)
    outs, ro, fo = {}, ReduceOutputs(), ReduceOutputs()
    for f in ("sum", "mean", "max", "argmax"):
        odt = np.uint64 if f == "argmax" else np.float32
        outs["row_" + f] = DeviceArray((rows,), odt)
        outs["full_" + f] = DeviceArray((1,), odt)
        setattr(ro, f, outs["row_" + f].ptr)
        setattr(fo, f, outs["full_" + f].ptr)
    empty, _ = K.chain()
    call("nrk_dev_map_reduce_f32", C.cast(empty, C.c_void_p), 0, C.c_void_p(d.ptr), C.c_size_t(rows), C.c_size_t(cols),
         C.c_size_t(0), C.byref(ro), C.byref(fo), None)
    call("nrk_device_synchronize")
    got = {k: v.download() for k, v in outs.items()}
    want_sum = np.full(rows, float(cols), np.float32)
    want_sum[r5] = cols + 4.0
    np.testing.assert_array_equal(got["row_sum"], want_sum)
    assert got["row_mean"][0] == 1.0 and got["row_mean"][-1] == 1.0
    assert got["row_argmax"][r5] == c5 and got["row_max"][r5] == 5.0 and got["row_argmax"][0] == 0
    assert got["full_argmax"][0] == r5 * cols + c5 and got["full_max"][0] == 5.0
    assert got["full_sum"][0] == np.float32(float(n + 4))
    del d
