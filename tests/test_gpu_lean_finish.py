"""The fused map -> reduce finishes the whole-array result inside its one launch: the last
workgroup to arrive (counted on two levels of agent-scope counters, 32 workgroups per set)
merges every workgroup's record in a fixed order; row results wait in an LDS buffer of 16 rows
per wavefront. These shapes reach the cases the headline shape does not: unequal chunk counts
per wavefront (the merges with division), a last set of fewer than 32 workgroups, and more than
16 rows per wavefront (the row buffer flushed inside the segment loop). The whole-array result
must not depend on which workgroup arrives last (repeated calls, bit for bit), and a captured
graph's replays must equal the eager call (the counters return to zero inside the kernel)."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device
from numrs_amd import kernels as K
from numrs_amd import rng
from numrs_amd._lib import ReduceOutputs, call
from numrs_amd.grouped import Graph

pytestmark = pytest.mark.gpu
EPS = {np.float64: 2.0 ** -53, np.float32: 2.0 ** -24}
FIELDS = ("sum", "var", "max", "argmax", "min", "argmin")


def _outs(rows, dt):
    ro, fo, keep = ReduceOutputs(), ReduceOutputs(), {}
    for f in FIELDS:
        odt = np.uint64 if f.startswith("arg") else dt
        keep["row_" + f] = device.DeviceArray((rows,), odt)
        keep["full_" + f] = device.DeviceArray((1,), odt)
        setattr(ro, f, keep["row_" + f].ptr)
        setattr(fo, f, keep["full_" + f].ptr)
    return ro, fo, keep


def _call(dx, rows, cols, dt, ro, fo, stream):
    ops, n = K.chain(("mul", 1.5), ("add", -0.25))
    fn = "nrk_dev_map_reduce_f64" if dt == np.float64 else "nrk_dev_map_reduce_f32"
    call(fn, C.cast(ops, C.c_void_p), n, C.c_void_p(dx.ptr), C.c_size_t(rows), C.c_size_t(cols), C.c_size_t(0),
         C.byref(ro), C.byref(fo), C.c_void_p(stream))


@pytest.mark.parametrize("dtype,shape", [
    (np.float32, (1500, 1024)),    # 94 workgroups: unequal chunk counts, a last set of 30
    (np.float64, (1500, 1024)),    # 188 wavefronts of 8 chunks: the last ones short
    (np.float32, (40000, 1024)),   # 19-20 rows per wavefront: the row buffer flushes in the loop
    (np.float64, (70000, 1024)),   # 18 rows per wavefront
])
def test_lean_finish_cases_vs_oracle(gpu, oracle, dtype, shape):
    rows, cols = shape
    print(f"lean finish {np.dtype(dtype).name} {rows}x{cols}", flush=True)
    x = rng.normal(rng.BASE_SEED + 7300 + rows, rows * cols, dtype).reshape(rows, cols)
    dx = device.DeviceArray.upload(x)
    st = device.Stream()
    ro, fo, keep = _outs(rows, dtype)
    _call(dx, rows, cols, dtype, ro, fo, st.handle)
    st.synchronize()
    got = {k: v.download() for k, v in keep.items()}
    y = oracle.ref_map([2, 0], [1.5, -0.25], x).reshape(rows, cols)
    for op in ("max", "min", "argmax", "argmin"):
        np.testing.assert_array_equal(got["row_" + op], oracle.ref_axis(K.AXIS[op], y, rows, cols, 1), err_msg=op)
    assert got["full_argmax"][0] == oracle.ref("argmax", y) and got["full_argmin"][0] == oracle.ref("argmin", y)
    assert got["full_max"][0] == oracle.ref("max", y) and got["full_min"][0] == oracle.ref("min", y)
    e = EPS[dtype]
    ex_rows = oracle.dd_row_stats(y, 0)
    s_abs = np.abs(y.astype(np.float64)).sum(axis=1)
    assert np.all(np.abs(got["row_sum"] - (ex_rows[:, 0] + ex_rows[:, 1])) <= 8 * e * s_abs + e * np.abs(ex_rows[:, 0]))
    assert np.all(np.abs(got["row_var"] - ex_rows[:, 3]) <= 64 * e * ex_rows[:, 3])
    ex = oracle.dd_stats(y, 0)
    assert abs(got["full_sum"][0] - ex["sum"]) <= 8 * e * float(s_abs.sum()) + e * abs(ex["sum"])
    assert abs(got["full_var"][0] - ex["var"]) <= 64 * e * ex["var"]
    # repeated calls: whichever workgroup arrives last, the same bits
    for _ in range(3):
        _call(dx, rows, cols, dtype, ro, fo, st.handle)
        st.synchronize()
        for k in keep:
            np.testing.assert_array_equal(keep[k].download().view(np.uint8), got[k].view(np.uint8), err_msg=k)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lean_finish_graph_replays(gpu, dtype):
    """Captured after one eager call (the scratch exists then): every replay, with eager calls on
    other data in between, reproduces the eager results bit for bit."""
    rows, cols = 8192, 1024
    x = rng.normal(rng.BASE_SEED + 7400, rows * cols, dtype).reshape(rows, cols)
    z = rng.normal(rng.BASE_SEED + 7401, rows * cols, dtype).reshape(rows, cols)
    dx, dz = device.DeviceArray.upload(x), device.DeviceArray.upload(z)
    st = device.Stream()
    ro, fo, keep = _outs(rows, dtype)
    _call(dx, rows, cols, dtype, ro, fo, st.handle)
    st.synchronize()
    want = {k: v.download() for k, v in keep.items()}
    ro2, fo2, keep2 = _outs(rows, dtype)
    graph = Graph(st.handle, lambda: _call(dx, rows, cols, dtype, ro, fo, st.handle))
    try:
        for rep in range(4):
            print(f"graph replay {rep}", flush=True)
            for v in keep.values():
                call("nrk_memset_zero", C.c_void_p(v.ptr), C.c_size_t(v.nbytes), C.c_void_p(st.handle))
            graph.launch()
            _call(dz, rows, cols, dtype, ro2, fo2, st.handle)  # eager work between replays
            st.synchronize()
            for k in keep:
                np.testing.assert_array_equal(keep[k].download().view(np.uint8), want[k].view(np.uint8), err_msg=f"{k} replay {rep}")
    finally:
        graph.close()
