"""Out-of-memory paths (an allocator hook injects the failures; no reference counterpart).

nrk_debug_fail_allocations makes one of the library's own device allocations fail as a real
out-of-memory does. The call that needed it returns NRK_ERR_OUT_OF_MEMORY with what it was
allocating; nothing is left half-allocated, and the next call (the hook disarmed) allocates and
computes the right result."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device, rng
from numrs_amd._lib import NrkError, ReduceOutputs, call

pytestmark = pytest.mark.gpu
OOM = 3  # NRK_ERR_OUT_OF_MEMORY


def _fused_sum(dx, rows, cols, st):
    out = device.DeviceArray((rows,), np.float64)
    ro = ReduceOutputs()
    ro.sum = out.ptr
    call("nrk_dev_map_reduce_f64", None, 0, C.c_void_p(dx.ptr), rows, cols, 0, C.byref(ro), None, C.c_void_p(st.handle))
    st.synchronize()
    return out.download()


def test_injected_failure_reported_then_recovered(gpu, oracle):
    rows, cols = 3000, 777
    x = rng.normal(rng.BASE_SEED + 4100, rows * cols, np.float64).reshape(rows, cols)
    dx = device.DeviceArray.upload(x)  # before arming: the operand itself is allocated
    st = device.Stream()  # a fresh stream: its reduction scratch is allocated on first use
    call("nrk_debug_fail_allocations", 0)
    with pytest.raises(NrkError) as e:
        _fused_sum(dx, rows, cols, st)
    assert e.value.status == OOM and "injected" in str(e.value)
    # disarmed after the one failure: the same call allocates and computes
    got = _fused_sum(dx, rows, cols, st)
    ex = oracle.dd_row_stats(x, 0)
    assert np.all(np.abs(got - (ex[:, 0] + ex[:, 1])) <= 8 * 2.0 ** -53 * np.abs(x).sum(axis=1))
    st.close()


def test_injected_failure_of_a_later_allocation(gpu):
    """The second allocation from now fails: the first succeeds and nothing leaks into the next
    calls; nrk_device_malloc reports the failure and returns no pointer."""
    call("nrk_debug_fail_allocations", 1)
    a = device.DeviceBuffer(1 << 20)  # allocation 0 succeeds
    with pytest.raises(NrkError) as e:
        device.DeviceBuffer(1 << 20)  # allocation 1 fails
    assert e.value.status == OOM
    b = device.DeviceBuffer(1 << 20)  # disarmed
    assert a.ptr and b.ptr and a.ptr != b.ptr
    call("nrk_debug_fail_allocations", 5)
    call("nrk_debug_fail_allocations", -1)  # disarming before it fires
    c = [device.DeviceBuffer(4096) for _ in range(8)]
    assert all(x.ptr for x in c)


def test_real_out_of_memory_is_reported(gpu):
    """An allocation beyond the device's memory: NRK_ERR_OUT_OF_MEMORY, and the device still works."""
    with pytest.raises(NrkError) as e:
        device.DeviceBuffer(1 << 50)  # 1 PiB
    assert e.value.status == OOM
    x = device.DeviceArray.upload(np.arange(10.0))
    np.testing.assert_array_equal(x.download(), np.arange(10.0))
