"""Split-precision GEMMs vs the double-double oracle (no reference counterpart: FP64 oracle).

Error measure: e_ij = |C_ij - C*_ij| / sum_k |A_ik||B_kj|, the scale GEMM error bounds use
(C* = double-double product). Bars, written here:
  BF16x6 (f32):  max e <= 4 * max e(native FP32 MFMA GEMM) + 2^-22   ("same error" as native)
  BF16x3 (f32):  max e <= 2^-13                                      (drops ~2^-16-order terms)
  bf16 inputs:   exact products of the rounded inputs, FP32 accumulation only
  Ozaki-II (f64, 16 moduli): max e <= 2 * max e(native FP64 MFMA GEMM) + 2^-52
Inputs follow the split-precision workload: s * 2^e * u, e uniform in [-10, 10], u in [1, 2)."""
import math

import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng

pytestmark = pytest.mark.gpu
SHAPES = [(17, 31, 13), (64, 64, 64), (200, 300, 100), (512, 512, 512), (257, 1000, 129)]


def rel_err(c, exact, scale):
    return np.abs(c.astype(np.float64) - exact) / np.maximum(scale, 1e-300)


def native(a, b):
    m, k = a.shape
    n = b.shape[1]
    c = np.empty((m, n), a.dtype)
    K.gemm_2d(m, k, n, a, b, c)
    return c


def full_refs(oracle, a, b):
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    return oracle.dd_gemm(a64, b64), np.abs(a64) @ np.abs(b64)


@pytest.mark.parametrize("shape", SHAPES)
def test_bf16x_vs_native_fp32(gpu, oracle, shape):
    m, k, n = shape
    a = rng.scaled_pow2(rng.BASE_SEED + 2000 + m, m * k, dtype=np.float32).reshape(m, k)
    b = rng.scaled_pow2(rng.BASE_SEED + 2100 + n, k * n, dtype=np.float32).reshape(k, n)
    exact, scale = full_refs(oracle, a, b)
    e_nat = rel_err(native(a, b), exact, scale).max()
    e6 = rel_err(K.gemm_f32_bf16x(a, b, 6), exact, scale).max()
    e3 = rel_err(K.gemm_f32_bf16x(a, b, 3), exact, scale).max()
    print(f"{shape}: native {e_nat:.2e}  bf16x6 {e6:.2e}  bf16x3 {e3:.2e}")
    assert e6 <= 4 * e_nat + 2.0 ** -22
    assert e3 <= 2.0 ** -13


def test_bf16_plain_products_exact(gpu, oracle):
    m, k, n = 96, 160, 80
    a = rng.normal(1, m * k, np.float32).reshape(m, k)
    b = rng.normal(2, k * n, np.float32).reshape(k, n)
    # round inputs to bf16 (RNE) on the host; then products are exact, only FP32 accumulation rounds
    def to_bf16(x):
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32)
    ab, bb = to_bf16(a), to_bf16(b)
    exact, scale = full_refs(oracle, ab, bb)
    e = rel_err(K.gemm_f32_bf16x(a, b, 1), exact, scale).max()
    assert e <= k * 2.0 ** -24


def test_bf16x_degenerate(gpu):
    c = K.gemm_f32_bf16x(np.zeros((3, 0), np.float32), np.zeros((0, 4), np.float32), 3)
    assert c.shape == (3, 4) and np.all(c == 0)


@pytest.mark.parametrize("shape", [s for s in SHAPES if s[1] >= 100] + [(256, 4096, 64)])
def test_ozaki_vs_native_fp64(gpu, oracle, shape):
    """The 'error <= native DGEMM' claim is made for inner dimensions k >= 100 (the workload is
    4096^3): there the row/column scaling error of the Ozaki split is below the k-growing
    accumulation error of FP64. For tiny k the derived truncation bound (next test) is the bar."""
    m, k, n = shape
    a = rng.scaled_pow2(rng.BASE_SEED + 2200 + m, m * k).reshape(m, k)
    b = rng.scaled_pow2(rng.BASE_SEED + 2300 + n, k * n).reshape(k, n)
    exact, scale = full_refs(oracle, a, b)
    e_nat = rel_err(native(a, b), exact, scale).max()
    e_oz = rel_err(K.gemm_f64_ozaki(a, b), exact, scale).max()
    print(f"{shape}: native fp64 {e_nat:.2e}  ozaki16 {e_oz:.2e}")
    assert e_oz <= 2 * e_nat + 2.0 ** -52


def test_ozaki_integer_inputs_exact(gpu):
    r = np.random.default_rng(9)
    a = r.integers(-1000, 1000, (100, 333)).astype(np.float64)
    b = r.integers(-1000, 1000, (333, 77)).astype(np.float64)
    np.testing.assert_array_equal(K.gemm_f64_ozaki(a, b), a @ b)


MODULI = [255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173, 167]


def ozaki_row_bound(nm, k):
    """The 2-norm every scaled row of A / column of B is held to with N moduli: X^2 <= (M/2)(1 - 2^-20)
    (X <= 2^62), less sqrt(k)/2 for the rounding of k entries and a unit. Cauchy-Schwarz then keeps
    |C'_ij| <= |A'_i| |B'_j| inside the CRT range (-M/2, M/2) for every k."""
    M = math.prod(MODULI[:nm])
    return min(math.sqrt(M / 2 * (1 - 2.0 ** -20)), 2.0 ** 62) - 0.5 * math.sqrt(k) - 1


def test_ozaki_error_within_truncation_bound(gpu, oracle):
    """The only error is rounding A's rows / B's columns to integers after scaling by the largest
    power of two 2^s with 2^s |A_i|_2 <= X: |dA_ik| <= 2^-s / 2 <= |A_i|_2 / X, likewise for B, so
    |C - C*|_ij <= (|A_i|_2 sum_k |B_kj| + |B_j|_2 sum_k |A_ik|) / X + k |A_i|_2 |B_j|_2 / X^2 (+ the
    final rounding). That bound is never above the per-entry rule's (every scaled entry below 2^P,
    P the largest with k 2^(2P) < M/4), which the test checks too."""
    m = k = n = 256
    a = rng.scaled_pow2(5, m * k).reshape(m, k)
    b = rng.scaled_pow2(6, k * n).reshape(k, n)
    exact, _ = full_refs(oracle, a, b)
    ra = np.sqrt((a * a).sum(axis=1))[:, None]
    cb = np.sqrt((b * b).sum(axis=0))[None, :]
    sa = np.abs(a).sum(axis=1)[:, None]
    sb = np.abs(b).sum(axis=0)[None, :]
    rowmax = np.abs(a).max(axis=1)[:, None]
    colmax = np.abs(b).max(axis=0)[None, :]
    errs = {}
    for nm in (8, 14, 16):
        X = ozaki_row_bound(nm, k)
        err = np.abs(K.gemm_f64_ozaki(a, b, nm) - exact)
        ulp = np.abs(exact) * 2.0 ** -52
        bound = (ra * sb + cb * sa) / X + k * ra * cb / X ** 2
        assert np.all(err <= bound * 1.001 + ulp), nm
        P = min(53, math.floor((math.log2(math.prod(MODULI[:nm])) - 2 - math.log2(k)) / 2))
        assert np.all(err <= 2.0 ** -P * (rowmax * sb + colmax * sa) * 1.001 + ulp), nm
        errs[nm] = err.max()
    assert errs[16] <= errs[14] <= errs[8]


@pytest.mark.parametrize("nm", [3, 8, 10, 16, 20])
def test_ozaki_products_at_the_edge_of_the_crt_range(gpu, nm):
    """A row of A parallel to a column of B makes Cauchy-Schwarz tight: C'_ii = |A'_i|_2^2 for
    B = A^T, anywhere in (X^2 / 4, X^2] by where |A_i|_2 falls between powers of two. With
    integer entries the scaling is exact, so the product must come back exact. One row is picked
    to land within 2 % of the top of the CRT range; with 64 rows spread over the octave, a range
    rule too loose by even sqrt(2) would push some products past M/2 and wrap them. With 16 and
    20 moduli (20: the 2^62 cap on X) the scaled entries reach ~2^60, past the 2^53 of an FP64
    significand, which exercises the residue split of large integers. The moduli counts take the
    CRT through 1, 2, 2, 3 and 4 limbs (3 moduli: entries kept below 39 so that X, ~2^11.5, still
    scales them by a power of two >= 1)."""
    k = 64
    X = ozaki_row_bound(nm, k)
    hi = 38 if nm < 6 else 1000
    r = np.random.default_rng(21 + nm)
    rows = [r.integers(-hi, hi + 1, k).astype(np.float64) for _ in range(63)]
    for _ in range(20000):
        v = r.integers(-hi, hi + 1, k).astype(np.float64)
        nrm = math.sqrt(float(v @ v))
        s = math.floor(math.log2(X / nrm))
        if 0.99 * X < nrm * 2.0 ** s < (1 - 1e-6) * X:
            break
    else:
        pytest.fail("no integer vector with a scaled norm within 1% of the bound")
    if nm <= 16:  # below the 2^62 cap the bound is the CRT range itself
        assert (nrm * 2.0 ** s) ** 2 > 0.98 * math.prod(MODULI[:nm]) / 2
    a = np.stack(rows + [v])
    b = a.T.copy()
    np.testing.assert_array_equal(K.gemm_f64_ozaki(a, b, nm), a @ b)


def test_ozaki_nonfinite_rows_are_nan(gpu):
    a = np.ones((4, 8))
    b = np.ones((8, 5))
    a[2, 3] = np.inf
    c = K.gemm_f64_ozaki(a, b)
    assert np.all(np.isnan(c[2])) and np.all(c[[0, 1, 3]] == 8.0)


def test_ozaki_zero_k_and_zero_rows(gpu):
    c = K.gemm_f64_ozaki(np.zeros((3, 0)), np.zeros((0, 4)))
    assert np.all(c == 0)
    a = np.zeros((5, 7))
    b = rng.normal(3, 7 * 6).reshape(7, 6)
    assert np.all(K.gemm_f64_ozaki(a, b) == 0)


@pytest.mark.parametrize("shape", [(300, 200, 520), (513, 257, 129), (1024, 640, 768)])
def test_ozaki_every_int8_tile_config_bit_identical(gpu, shape):
    """All INT8 GEMM tile configurations compute the same exact integer residues, so the whole
    Ozaki product (16 moduli) is bit-identical across them, and equal to the default path's."""
    import ctypes as C
    from numrs_amd import device
    from numrs_amd import selector as S
    from numrs_amd._lib import call
    m, k, n = shape
    a = rng.scaled_pow2(70, m * k).reshape(m, k)
    b = rng.scaled_pow2(71, k * n).reshape(k, n)
    want = K.gemm_f64_ozaki(a, b, 16)
    st = device.Stream()
    dA, dB = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
    dC = device.DeviceArray((m, n), np.float64)
    for cfg in range(S.config_count(S.F64_OZAKI)):
        print(f"ozaki tile config {cfg}, shape {shape}", flush=True)
        call("nrk_dev_run_config", S.F64_OZAKI, cfg, m, n, k, C.c_void_p(dA.ptr), C.c_void_p(dB.ptr),
             C.c_void_p(dC.ptr), C.c_void_p(st.handle))
        st.synchronize()
        got = dC.download()
        np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64), err_msg=f"config {cfg}")
