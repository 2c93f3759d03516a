"""The measured workloads at their full sizes, checked through size-independent properties.

- Fused y = 1.5 x - 0.25 -> sum/max/argmax/var on x (8192, 8192) f32 and f64 (the headline step):
  row max/argmax bit-exact against the restatement (numpy over the oracle's two-rounding map);
  whole-array max/argmax equal to the max of the row maxima and its flat index; whole-array
  sum and var within the stated bounds of the double-double statistics of y; the whole-array sum
  within the same bound of the double-double sum of the row sums (a checksum of checksums).
- i64 sum of 2^28 values uniform in [-2^20, 2^20]: exact.
- Ozaki-II FP64 GEMM at 4096^3: on sampled entries, never less accurate than the native FP64
  GEMM against correctly rounded dot products (error scale sum_k |A_ik||B_kj|)."""
import math

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


def _split(x):
    c = x * 134217729.0
    hi = c - (c - x)
    return hi, x - hi


def _exact_dot(a, b):
    ah, al = _split(a)
    bh, bl = _split(b)
    p = a * b
    e = al * bl - (((p - ah * bh) - al * bh) - ah * bl)  # Dekker: p + e = a*b exactly
    return math.fsum(np.concatenate([p, e]).tolist())


def test_ozaki_4096_not_less_accurate_than_native(gpu):
    n = 4096
    a = rng.scaled_pow2(rng.BASE_SEED + 2400, n * n).reshape(n, n)
    b = rng.scaled_pow2(rng.BASE_SEED + 2401, n * n).reshape(n, n)
    c_nat = np.empty((n, n))
    K.gemm_2d(n, n, n, a, b, c_nat)
    c_oz = K.gemm_f64_ozaki(a, b, 15)
    pick = np.random.default_rng(5)
    worst_nat = worst_oz = 0.0
    for _ in range(32):
        i, j = int(pick.integers(n)), int(pick.integers(n))
        ex = _exact_dot(a[i], b[:, j])
        scale = float(np.abs(a[i]) @ np.abs(b[:, j]))
        worst_nat = max(worst_nat, abs(c_nat[i, j] - ex) / scale)
        worst_oz = max(worst_oz, abs(c_oz[i, j] - ex) / scale)
    print(f"4096^3 sampled max error: native {worst_nat:.2e}, ozaki15 {worst_oz:.2e}")
    assert worst_oz <= worst_nat
