"""The host-pointer entry points move data through pinned buffers of 8 MiB (four per host thread)
filled and drained by a small host-copy pool. Sizes around the buffer boundaries (one element
either side of 1, 4 and 5 buffers, so the ring wraps) must give the same results as the device
entry points on the same data: bit for bit, since the kernels are the same."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device
from numrs_amd import kernels as K
from numrs_amd import rng
from numrs_amd._lib import call

pytestmark = pytest.mark.gpu

CHUNK_F32 = (8 << 20) // 4


@pytest.mark.parametrize("n", [CHUNK_F32 - 1, CHUNK_F32 + 1, 4 * CHUNK_F32, 5 * CHUNK_F32 + 3])
def test_host_sum_and_map_match_device(gpu, n):
    x = rng.normal(rng.BASE_SEED + 7100 + n, n, np.float32)
    print(f"host staging n={n}", flush=True)
    got = K.sum_f32(x)
    dx = device.DeviceArray.upload(x)
    out = device.DeviceArray((1,), np.float32)
    call("nrk_dev_sum_f32", C.c_void_p(dx.ptr), C.c_size_t(n), C.c_void_p(out.ptr), None)
    device.synchronize()
    assert np.float32(got).tobytes() == out.download()[0].tobytes()
    # both directions: host in, host out (every element crosses the ring twice)
    y = K.unary_dispatch(x, [("mul", 1.5), ("add", -0.25)])
    want = ((x * np.float32(1.5)).astype(np.float32) + np.float32(-0.25)).astype(np.float32)
    np.testing.assert_array_equal(y, want)
    # the upload / download helpers take the same path
    np.testing.assert_array_equal(dx.download(), x)


def test_host_gemm_through_staging(gpu):
    """A host-pointer GEMM whose operands span several buffers (1536^2 f64 = 18 MiB each)."""
    m = k = n = 1536
    a = rng.uniform(rng.BASE_SEED + 7200, m * k, -1.0, 1.0)
    b = rng.uniform(rng.BASE_SEED + 7201, k * n, -1.0, 1.0)
    c = np.full(m * n, np.nan)
    K.gemm_2d(m, k, n, a, b, c)
    da, db = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
    dc = device.DeviceArray((m * n,), np.float64)
    call("nrk_dev_gemm_f64", C.c_size_t(m), C.c_size_t(k), C.c_size_t(n), C.c_void_p(da.ptr), C.c_void_p(db.ptr),
         C.c_void_p(dc.ptr), None)
    device.synchronize()
    np.testing.assert_array_equal(c, dc.download())
