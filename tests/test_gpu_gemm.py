"""Plain GEMM (f64 / f32 on MFMA) vs the oracle, through the C ABI.

Tolerance (tests/test_matmul_dispatch.rs:91-107 style, magnitude-relative):
|C - C_exact| <= tol * max(|C_exact|, 1), tol = 1e-12 (f64) and 1e-4 (f32), with C_exact the
double-double product. Also: the 3x3 known answer, degenerate dims, k = 0 overwrite of NaN
garbage, the length-mismatch rule and bit-identity across launches."""
import numpy as np
import pytest

from numrs_amd import kernels as K
from numrs_amd import rng

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-4}
SHAPES = [(1, 1, 1), (3, 3, 3), (17, 31, 13), (64, 64, 64), (63, 65, 127), (128, 16, 128), (129, 257, 130),
          (200, 1, 300), (1, 500, 1), (512, 512, 512)]


def gemm(a, b):
    m, k = a.shape
    n = b.shape[1]
    c = np.full((m, n), np.nan, a.dtype)
    K.gemm_2d(m, k, n, a, b, c)
    return c


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemm_vs_double_double(gpu, oracle, dtype):
    worst = 0.0
    for i, (m, k, n) in enumerate(SHAPES):
        a = rng.uniform(rng.BASE_SEED + 1500 + i, m * k).reshape(m, k).astype(dtype)
        b = rng.uniform(rng.BASE_SEED + 1600 + i, k * n).reshape(k, n).astype(dtype)
        c = gemm(a, b).astype(np.float64)
        ex = oracle.dd_gemm(a.astype(np.float64), b.astype(np.float64))
        err = np.abs(c - ex) / np.maximum(np.abs(ex), 1.0)
        worst = max(worst, float(err.max()))
        assert err.max() <= TOL[dtype], (m, k, n, err.max())
    print(f"max scaled err {worst:.3e}")


def test_known_3x3(gpu, known):
    k = known["matmul_3x3"]
    a = np.array(k["a"], np.float64).reshape(3, 3)
    np.testing.assert_array_equal(gemm(a, a).reshape(-1), np.array(k["c"], np.float64))
    np.testing.assert_array_equal(gemm(a, np.array(k["v"], np.float64).reshape(3, 1)).reshape(-1), np.array(k["av"], np.float64))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_degenerate_dims(gpu, dtype):
    c = np.full((4, 5), np.nan, dtype)
    K.gemm_2d(4, 0, 5, np.zeros(0, dtype), np.zeros(0, dtype), c)  # k == 0 -> zeros (gemm.rs:295-303)
    assert np.all(c == 0)
    c = np.full((0, 5), np.nan, dtype)
    K.gemm_2d(0, 3, 5, np.zeros(0, dtype), np.ones(15, dtype), c)
    c = np.full((3, 5), 7.0, dtype)
    K.gemm_2d(3, 2, 5, np.ones(5, dtype), np.ones(10, dtype), c)  # a_len != m*k -> zero C, read nothing
    assert np.all(c == 0)


def test_integer_valued_exact_and_identity(gpu):
    r = np.random.default_rng(3)
    a = r.integers(-100, 100, (130, 70)).astype(np.float64)
    b = r.integers(-100, 100, (70, 90)).astype(np.float64)
    np.testing.assert_array_equal(gemm(a, b), a @ b)
    e = np.eye(70)
    np.testing.assert_array_equal(gemm(a, e), a)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bit_identical_across_launches(gpu, dtype):
    a = rng.uniform(7, 300 * 400).reshape(300, 400).astype(dtype)
    b = rng.uniform(8, 400 * 500).reshape(400, 500).astype(dtype)
    c1, c2 = gemm(a, b), gemm(a, b)
    np.testing.assert_array_equal(c1, c2)


def test_large_f64_sampled(gpu, oracle):
    """2048^3 f64 (big-tile path) checked on 512 sampled entries against double-double."""
    n = 2048
    a = rng.uniform(rng.BASE_SEED + 1700, n * n).reshape(n, n)
    b = rng.uniform(rng.BASE_SEED + 1701, n * n).reshape(n, n)
    c = gemm(a, b)
    r = np.random.default_rng(1)
    rows, cols = r.integers(0, n, 512), r.integers(0, n, 512)
    ex = oracle.dd_gemm_entries(a, b, rows, cols)
    err = np.abs(c[rows, cols] - ex) / np.maximum(np.abs(ex), 1.0)
    assert err.max() <= 1e-12


# ---- batched_gemm (array/linalg.rs:56-136) and Array::matmul's batch axes (linalg.rs:150-250)

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(5, 37, 53, 29), (3, 128, 96, 64), (2, 300, 17, 260)])
def test_batched_panels_bit_identical_to_gemm_2d(gpu, dtype, shape):
    """Panel t of a batch is gemm_2d on that panel, bit for bit (same kernel, same k order), and
    gemm_2d is pinned against the double-double oracle above."""
    batch, m, k, n = shape
    a = rng.uniform(rng.BASE_SEED + 1700 + m, batch * m * k).astype(dtype)
    b = rng.uniform(rng.BASE_SEED + 1800 + n, batch * k * n).astype(dtype)
    c = np.full(batch * m * n, np.nan, dtype)
    K.batched_gemm(batch, m, k, n, a, b, c)
    for t in range(batch):
        want = gemm(a[t * m * k:(t + 1) * m * k].reshape(m, k), b[t * k * n:(t + 1) * k * n].reshape(k, n))
        np.testing.assert_array_equal(c[t * m * n:(t + 1) * m * n].reshape(m, n).view(np.uint64 if dtype == np.float64 else np.uint32),
                                      want.view(np.uint64 if dtype == np.float64 else np.uint32))


def test_batched_degenerate_and_length_rules(gpu):
    """linalg.rs:60-73: batch, m or n = 0, or a slice length mismatch, return without writing C
    (unlike gemm_2d's zero fill); k = 0 panels are gemm_2d's, i.e. zero-filled."""
    sentinel = np.full(12, 7.0)
    for batch, m, k, n, la, lb in [(0, 2, 3, 2, 0, 0), (2, 0, 3, 2, 0, 12), (2, 2, 3, 0, 12, 0),
                                   (2, 2, 3, 3, 11, 18), (2, 2, 3, 3, 12, 17)]:
        c = sentinel.copy()
        K.batched_gemm(batch, m, k, n, np.ones(la), np.ones(lb), c[:batch * m * n] if batch * m * n else c)
        np.testing.assert_array_equal(c, sentinel)
    c = np.full(12, np.nan)
    K.batched_gemm(2, 2, 0, 3, np.ones(0), np.ones(0), c)
    assert np.all(c == 0.0)


def test_batched_integer_panels_exact_and_many_launches(gpu):
    """Integer-valued panels multiply exactly, so the result equals numpy's exactly. 70 000
    one-element panels take two launches (at most 65 535 panels per launch)."""
    r = np.random.default_rng(31)
    batch, m, k, n = 6, 33, 70, 45
    a = r.integers(-50, 51, (batch, m, k)).astype(np.float64)
    b = r.integers(-50, 51, (batch, k, n)).astype(np.float64)
    c = np.empty(batch * m * n)
    K.batched_gemm(batch, m, k, n, a, b, c)
    np.testing.assert_array_equal(c.reshape(batch, m, n), a @ b)
    big = 70_000
    x = r.integers(-1000, 1001, big).astype(np.float32)
    y = r.integers(-1000, 1001, big).astype(np.float32)
    z = np.empty(big, np.float32)
    K.batched_gemm(big, 1, 1, 1, x, y, z)
    np.testing.assert_array_equal(z, x * y)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matmul_batch_broadcast(gpu, dtype):
    """matmul's batch broadcast: one side unbatched (panel stride 0), equal batch axes, and axes
    broadcast on both sides (panel by panel), with 1-D promotion. Integer-valued entries make
    every product exact, so each case must equal numpy's matmul exactly (same broadcast rules)."""
    r = np.random.default_rng(41)
    ints = lambda *s: r.integers(-20, 21, s).astype(dtype)  # noqa: E731
    cases = [(ints(4, 9, 7), ints(7, 5)), (ints(9, 7), ints(3, 7, 5)), (ints(2, 3, 9, 7), ints(2, 3, 7, 5)),
             (ints(4, 1, 9, 7), ints(1, 3, 7, 5)), (ints(7), ints(2, 7, 5)), (ints(2, 9, 7), ints(7)),
             (ints(1, 9, 7), ints(6, 7, 5))]
    for a, b in cases:
        got = K.matmul(a, b)
        want = np.matmul(a.reshape(1, -1) if a.ndim == 1 else a, b.reshape(-1, 1) if b.ndim == 1 else b)
        assert got.shape == want.shape, (a.shape, b.shape)
        np.testing.assert_array_equal(got, want)
    with pytest.raises(ValueError):
        K.matmul(ints(2, 3, 4), ints(5, 6))
    with pytest.raises(ValueError):
        K.matmul(ints(2, 3, 4), ints(3, 4, 6))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matmul_non_contiguous_operands(gpu, oracle, dtype):
    """tests/test_matmul_dispatch.rs:283-325: operands that are transposed views (not C-contiguous)
    are read in logical order (borrow::operand, borrow.rs:59-64, copies them), so the product is
    the contiguous product: bit-identical to the call on contiguous copies, and within the
    reference's 1e-9 (f64) / 1e-4 (f32) of the naive product. Also for a batch of transposed
    panels (matmul's batch axes) and a column-sliced (strided) left operand."""
    m, k, n = 33, 65, 31
    a = rng.uniform(rng.BASE_SEED + 1900, k * m).reshape(k, m).astype(dtype).T
    b = rng.uniform(rng.BASE_SEED + 1901, n * k).reshape(n, k).astype(dtype).T
    assert not a.flags.c_contiguous and not b.flags.c_contiguous
    ex = oracle.dd_gemm(a.astype(np.float64), b.astype(np.float64))
    ac, bc = np.ascontiguousarray(a), np.ascontiguousarray(b)
    want = K.matmul(ac, bc)
    for x, y in ((a, b), (a, bc), (ac, b)):
        got = K.matmul(x, y)
        np.testing.assert_array_equal(got, want)
        assert np.all(np.abs(got - ex) <= (1e-9 if dtype == np.float64 else 1e-4) * np.maximum(np.abs(ex), 1.0))
    wide = rng.uniform(rng.BASE_SEED + 1902, m * 2 * k).reshape(m, 2 * k).astype(dtype)[:, ::2]
    np.testing.assert_array_equal(K.matmul(wide, b), K.matmul(np.ascontiguousarray(wide), bc))
    pa = rng.uniform(rng.BASE_SEED + 1903, 4 * k * m).reshape(4, k, m).astype(dtype).transpose(0, 2, 1)
    got = K.matmul(pa, b)
    for t in range(4):
        np.testing.assert_array_equal(got[t], K.matmul(np.ascontiguousarray(pa[t]), bc))
