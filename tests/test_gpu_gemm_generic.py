"""gemm_generic (kernels/gemm.rs:408-445) on the GPU: the integer GEMM behind Array<i32/i64>::matmul
and the f32/f64 path for slice lengths that do not match exactly (gemm.rs:250-253).

Bars: integers bit-exact against the oracle's restatement (wrapping, like Rust's release
arithmetic) and against numpy's integer matmul (which wraps the same way); floats of the generic
path bit-exact against the restatement's ascending-k, two-rounding order. The reference's own cases:
the boundary size grid with seq_i32 data (tests/test_matmul_dispatch.rs:205-219), the batched i32
case (:435-458) and the transposed-view generic case (:330-345)."""
import numpy as np
import pytest

from numrs_amd import kernels as K

pytestmark = pytest.mark.gpu

SIZE_GRID = [(1, 1, 1), (1, 32, 1), (2, 31, 33), (31, 32, 33), (32, 32, 32), (33, 64, 31), (63, 64, 65), (64, 64, 64),
             (65, 63, 127), (127, 128, 2)]  # tests/test_matmul_dispatch.rs:133-143


def seq_i32(n, modulus, offset):  # tests/test_matmul_dispatch.rs:121-123
    return (np.arange(n) % modulus + offset).astype(np.int32)


def test_i32_boundary_grid_matches_naive(gpu, oracle):
    for m, k, n in SIZE_GRID:
        a = seq_i32(m * k, 7, -3).reshape(m, k)
        b = seq_i32(k * n, 5, -2).reshape(k, n)
        got = K.matmul(a, b)
        assert got.dtype == np.int32 and got.shape == (m, n)
        np.testing.assert_array_equal(got, oracle.gemm_generic(a, b))
        np.testing.assert_array_equal(got, a @ b)


def test_i32_batched_and_transposed_generic(gpu):
    batch, m, k, n = 4, 5, 3, 6  # tests/test_matmul_dispatch.rs:435-458
    a = seq_i32(batch * m * k, 7, -3).reshape(batch, m, k)
    b = seq_i32(batch * k * n, 5, -2).reshape(batch, k, n)
    np.testing.assert_array_equal(K.matmul(a, b), np.matmul(a, b))
    t = np.array([1, 2, 3, 4, 5, 6], np.int32).reshape(2, 3).T  # :330-345, a transposed view
    assert not t.flags.c_contiguous
    np.testing.assert_array_equal(K.matmul(t, np.eye(2, dtype=np.int32)), np.array([[1, 4], [2, 5], [3, 6]], np.int32))


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.uint64, np.uint32, np.int16, np.uint8])
def test_integer_matmul_wraps_like_release_rust(gpu, oracle, dtype):
    """Full-range values: products and sums overflow and wrap modulo 2^w; 64-bit and 32-bit types
    run their own kernels, narrower ones the 32-bit kernel cast back (still exact modulo 2^w)."""
    r = np.random.default_rng(np.dtype(dtype).itemsize * 7 + 1)
    info = np.iinfo(dtype)
    m, k, n = 67, 130, 75
    a = r.integers(info.min, info.max, (m, k), dtype=dtype, endpoint=True)
    b = r.integers(info.min, info.max, (k, n), dtype=dtype, endpoint=True)
    got = K.matmul(a, b)
    assert got.dtype == dtype
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(got, np.matmul(a, b))
    if dtype in (np.int64, np.int32):
        np.testing.assert_array_equal(got, oracle.gemm_generic(a, b))


def test_integer_matmul_large_and_broadcast(gpu, oracle):
    """A shape with many tiles and partial edges, and a batch whose B is broadcast (panel stride
    0), against the restatement."""
    r = np.random.default_rng(5)
    a = r.integers(-(1 << 40), 1 << 40, (300, 517), dtype=np.int64)
    b = r.integers(-(1 << 40), 1 << 40, (517, 129), dtype=np.int64)
    np.testing.assert_array_equal(K.matmul(a, b), oracle.gemm_generic(a, b))
    a3 = r.integers(-1000, 1000, (3, 70, 40), dtype=np.int32)
    b2 = r.integers(-1000, 1000, (40, 90), dtype=np.int32)
    got = K.matmul(a3, b2)
    for t in range(3):
        np.testing.assert_array_equal(got[t], oracle.gemm_generic(a3[t], b2))


def test_host_slice_contract(gpu):
    """gemm.rs:408-423: C zeroed over its own length; a short slice leaves it zero; longer slices
    are fine (their leading parts are used); k = 0 gives zeros."""
    c = np.full(12, 7, np.int64)
    K.gemm_generic(3, 2, 4, np.ones(5, np.int64), np.ones(8, np.int64), c)  # a_len 5 < 6
    assert not c.any()
    c = np.full(15, 7, np.int64)  # c longer than m*n: the tail is zeroed too
    a = np.arange(6 + 3, dtype=np.int64)  # longer than m*k
    b = np.arange(8, dtype=np.int64)
    K.gemm_generic(3, 2, 4, a, b, c)
    np.testing.assert_array_equal(c[:12], (a[:6].reshape(3, 2) @ b.reshape(2, 4)).reshape(-1))
    assert not c[12:].any()
    c32 = np.full(6, 3, np.int32)
    K.gemm_generic(2, 0, 3, np.zeros(0, np.int32), np.zeros(0, np.int32), c32)
    assert not c32.any()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_float_generic_path_for_long_slices(gpu, oracle, dtype):
    """gemm_2d with a slice longer than its matrix goes through gemm_generic (gemm.rs:250-253):
    ascending k, product then sum each rounded -- bit-exact against the restatement."""
    r = np.random.default_rng(9)
    m, k, n = 70, 129, 66
    a = r.standard_normal(m * k + 5).astype(dtype)
    b = r.standard_normal(k * n).astype(dtype)
    c = np.full(m * n, np.nan, dtype)
    K.gemm_2d(m, k, n, a, b, c)
    want = oracle.gemm_generic(a[:m * k].reshape(m, k), b.reshape(k, n)).reshape(-1)
    np.testing.assert_array_equal(c.view(np.uint64 if dtype == np.float64 else np.uint32),
                                  want.view(np.uint64 if dtype == np.float64 else np.uint32))
