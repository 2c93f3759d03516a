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
