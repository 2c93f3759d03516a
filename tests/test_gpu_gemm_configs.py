"""Every GEMM tile configuration gives the same bits as the default dispatch.

The reference keeps K unsplit so that results do not depend on the thread count
(kernels/gemm.rs:364-381, tested at gemm.rs:784-807). Here every tile configuration of a family
runs the same k order per output, so the tile (picked by the built-in rule, the tuner or the
learned selector) never changes a result bit. Each configuration is forced through
nrk_dev_run_config on shapes with partial tiles at both edges, large enough (>= 1024) that the
default rule itself picks the bigger tiles, and compared bit for bit with the default entry point.
The split products are also checked against the double-double oracle on 512 sampled entries, with
the bars of test_gpu_gemm_split.py (BF16x6 <= 4x native FP32 + 2^-22, BF16x3 <= 2^-13)."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device, rng
from numrs_amd import selector as S
from numrs_amd._lib import call

pytestmark = pytest.mark.gpu

SHAPES = [(1100, 1029, 1300), (2048, 2048, 512), (1024, 1536, 777)]  # (m, n, k)


def _run_default(fn, m, n, k, dA, dB, dC, st, *extra):
    call(fn, m, k, n, C.c_void_p(dA.ptr), C.c_void_p(dB.ptr), C.c_void_p(dC.ptr), *extra, C.c_void_p(st.handle))
    st.synchronize()
    return dC.download()


def _run_config(family, cfg, m, n, k, dA, dB, dC, st):
    print(f"family {family} config {cfg} shape m={m} n={n} k={k}", flush=True)
    call("nrk_dev_run_config", family, cfg, m, n, k, C.c_void_p(dA.ptr), C.c_void_p(dB.ptr), C.c_void_p(dC.ptr),
         C.c_void_p(st.handle))
    st.synchronize()
    return dC.download()


def _bits(x):
    return x.view(np.uint64 if x.dtype == np.float64 else np.uint32)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_native_gemm_every_tile_config_bit_identical(gpu, dtype, shape):
    m, n, k = shape
    a = rng.uniform(rng.BASE_SEED + 3000 + m, m * k).reshape(m, k).astype(dtype)
    b = rng.uniform(rng.BASE_SEED + 3100 + n, k * n).reshape(k, n).astype(dtype)
    st = device.Stream()
    dA, dB = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
    dC = device.DeviceArray((m, n), dtype)
    fam = S.F64 if dtype == np.float64 else S.F32
    want = _run_default("nrk_dev_gemm_f64" if dtype == np.float64 else "nrk_dev_gemm_f32", m, n, k, dA, dB, dC, st)
    for cfg in range(S.config_count(fam)):
        got = _run_config(fam, cfg, m, n, k, dA, dB, dC, st)
        np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=f"config {cfg}")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("products", [3, 6])
def test_split_gemm_every_tile_config_bit_identical(gpu, oracle, products, shape):
    m, n, k = shape
    a = rng.scaled_pow2(rng.BASE_SEED + 3200 + m, m * k, dtype=np.float32).reshape(m, k)
    b = rng.scaled_pow2(rng.BASE_SEED + 3300 + n, k * n, dtype=np.float32).reshape(k, n)
    st = device.Stream()
    dA, dB = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
    dC = device.DeviceArray((m, n), np.float32)
    want = _run_default("nrk_dev_gemm_f32_bf16x", m, n, k, dA, dB, dC, st, C.c_int32(products))
    fam = S.F32_BF16X3 if products == 3 else S.F32_BF16X6
    for cfg in range(S.config_count(fam)):
        got = _run_config(fam, cfg, m, n, k, dA, dB, dC, st)
        np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=f"products {products} config {cfg}")
    # the default result against the double-double oracle on 512 sampled entries
    pick = np.random.default_rng(m + n + k + products)
    rows, cols = pick.integers(0, m, 512), pick.integers(0, n, 512)
    exact = oracle.dd_gemm_entries(a, b, rows, cols)
    scale = oracle.abs_gemm_entries(a, b, rows, cols)
    nat = _run_default("nrk_dev_gemm_f32", m, n, k, dA, dB, dC, st)
    e_nat = float((np.abs(nat[rows, cols].astype(np.float64) - exact) / scale).max())
    e = float((np.abs(want[rows, cols].astype(np.float64) - exact) / scale).max())
    print(f"{shape} bf16x{products}: {e:.3e} (native fp32 {e_nat:.3e})")
    assert e <= (4 * e_nat + 2.0 ** -22 if products == 6 else 2.0 ** -13)
