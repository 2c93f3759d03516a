"""nrk_dev_map_reduce_f64_noalloc: the fused map -> reduce on the device's fixed scratch (no
hipMalloc), against nrk_dev_map_reduce_f64 on the same inputs -- the same kernels, so the same
bits -- on consecutive calls with shrinking inputs (the fixed scratch keeps the previous call's
records, which must not leak into the next), and the size limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device, kernels, rng
from numrs_amd._lib import NrkError, ReduceOutputs, call

pytestmark = pytest.mark.gpu
FIELDS = ("sum", "mean", "var", "min", "max", "argmin", "argmax")


def _outs(rows):
    ro, fo, keep = ReduceOutputs(), ReduceOutputs(), {}
    for f in FIELDS:
        dt = np.uint64 if f.startswith("arg") else np.float64
        keep["row_" + f] = device.DeviceArray((rows,), dt)
        keep["full_" + f] = device.DeviceArray((1,), dt)
        setattr(ro, f, keep["row_" + f].ptr)
        setattr(fo, f, keep["full_" + f].ptr)
    return ro, fo, keep


def _run(fn, dx, rows, cols, ddof, st, with_rows=True):
    ops, n = kernels.chain(("mul", 1.5), ("add", -0.25))
    ro, fo, keep = _outs(rows)
    print(f"{fn} rows={rows} cols={cols} rows_out={with_rows}", flush=True)
    call(fn, C.cast(ops, C.c_void_p), n, C.c_void_p(dx.ptr), C.c_size_t(rows), C.c_size_t(cols), C.c_size_t(ddof),
         C.byref(ro) if with_rows else None, C.byref(fo), C.c_void_p(st.handle))
    st.synchronize()
    return {k: v.download() for k, v in keep.items() if with_rows or k.startswith("full_")}


@pytest.mark.parametrize("shapes", [[(300, 5000), (40, 777)], [(8192, 1024), (3, 100_003)], [(20_000, 300), (50, 7)]])
def test_noalloc_matches_allocating_entry_point(gpu, shapes):
    st = device.Stream()
    for i, (rows, cols) in enumerate(shapes):  # the second call is smaller than the first
        x = rng.normal(rng.BASE_SEED + 4000 + i, rows * cols).reshape(rows, cols)
        x[rows // 2, cols // 3] = np.nan if i == 1 else x[rows // 2, cols // 3]
        dx = device.DeviceArray.upload(x)
        want = _run("nrk_dev_map_reduce_f64", dx, rows, cols, 1, st)
        got = _run("nrk_dev_map_reduce_f64_noalloc", dx, rows, cols, 1, st)
        for k in want:
            np.testing.assert_array_equal(got[k].view(np.uint64), want[k].view(np.uint64), err_msg=f"{k} {rows}x{cols}")


def test_noalloc_whole_array_only_and_size_limit(gpu):
    st = device.Stream()
    n = 3_000_017
    x = rng.normal(rng.BASE_SEED + 4010, n)
    dx = device.DeviceArray.upload(x)
    want = _run("nrk_dev_map_reduce_f64", dx, 1, n, 0, st, with_rows=False)
    got = _run("nrk_dev_map_reduce_f64_noalloc", dx, 1, n, 0, st, with_rows=False)
    for k in want:
        np.testing.assert_array_equal(got[k].view(np.uint64), want[k].view(np.uint64), err_msg=k)
    # row results for 100 000 rows of 1025 (the record tree's path) need ~27 MB of records:
    # refused, nothing launched
    big = device.DeviceArray((100_000 * 1025,), np.float64)
    with pytest.raises(NrkError) as e:
        _run("nrk_dev_map_reduce_f64_noalloc", big, 100_000, 1025, 0, st)
    assert "fixed scratch" in str(e.value)
    # and the entry point still works after the refusal
    got2 = _run("nrk_dev_map_reduce_f64_noalloc", dx, 1, n, 0, st, with_rows=False)
    for k in want:
        np.testing.assert_array_equal(got2[k].view(np.uint64), want[k].view(np.uint64), err_msg=k)
