"""Ragged batch in one persistent launch vs the per-call loop and the oracle; hipGraph replays.

GEMMs: bf16 inputs (exact products), fp32 accumulation. Bar: on sampled entries,
|C - C*| <= k * 2^-24 * sum_k |a||b| with C* the double-double product of the bf16 inputs, and the
grouped result bit-identical to the unbatched per-call loop (same tile code, same k order).
Row sums: |s - s*| <= 8 * 2^-24 * sum|x| + 2^-24 |s*| against the double-double row sums."""
import numpy as np
import pytest

from numrs_amd import device, rng
from numrs_amd.grouped import Graph, GroupedPlan, bf16_bits_to_f32, gemm_bf16, to_bf16_bits

pytestmark = pytest.mark.gpu


def log_uniform_int(seed, n, lo, hi):
    u = rng.uniform(seed, n, 0.0, 1.0)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi + 1) - np.log(lo)))).astype(np.int64).clip(lo, hi)


def make_batch(n_gemm, n_rows, seed, max_len=65536):
    shapes = np.stack([log_uniform_int(seed + i, n_gemm, 16, 512) for i in range(3)], axis=1)
    probs = []
    for g, (m, n, k) in enumerate(shapes):
        a = to_bf16_bits(rng.normal(seed + 10 + g, int(m * k), np.float32))
        b = to_bf16_bits(rng.normal(seed + 20000 + g, int(k * n), np.float32))
        probs.append((int(m), int(n), int(k), a, b))
    lens = log_uniform_int(seed + 7, n_rows, 32, max_len)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = rng.normal(seed + 8, int(offs[-1]), np.float32)
    return probs, offs, x


def upload(probs):
    dev = []
    for (m, n, k, a, b) in probs:
        da, db = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
        dc = device.DeviceArray((m, n), np.float32)
        dev.append((m, n, k, da, db, dc))
    return dev


def descs_of(dev):
    return [(m, n, k, da.ptr, db.ptr, dc.ptr) for (m, n, k, da, db, dc) in dev]


@pytest.fixture(scope="module")
def batch(gpu):
    probs, offs, x = make_batch(160, 1500, rng.BASE_SEED + 3000)
    return probs, offs, x


def test_grouped_matches_loop_and_oracle(gpu, oracle, batch):
    probs, offs, x = batch
    dev = upload(probs)
    dx = device.DeviceArray.upload(x)
    dout = device.DeviceArray((len(offs) - 1,), np.float32)
    plan = GroupedPlan(descs_of(dev), dx.ptr, offs, dout.ptr)
    plan.run()
    device.synchronize()
    grouped = [dc.download() for (*_, dc) in dev]
    sums = dout.download()
    # per-call loop into fresh outputs
    loop_out = []
    for (m, n, k, da, db, dc) in dev:
        d2 = device.DeviceArray((m, n), np.float32)
        gemm_bf16(m, n, k, da.ptr, db.ptr, d2.ptr)
        loop_out.append(d2)
    device.synchronize()
    r = np.random.default_rng(0)
    for i, ((m, n, k, a, b), cg, d2) in enumerate(zip(probs, grouped, loop_out)):
        np.testing.assert_array_equal(cg, d2.download(), err_msg=f"problem {i}")
        A = bf16_bits_to_f32(a).reshape(m, k)
        B = bf16_bits_to_f32(b).reshape(k, n)
        rows, cols = r.integers(0, m, 24), r.integers(0, n, 24)
        ex = oracle.dd_gemm_entries(A.astype(np.float64), B.astype(np.float64), rows, cols)
        scale = oracle.abs_gemm_entries(A, B, rows, cols)
        assert np.all(np.abs(cg[rows, cols] - ex) <= k * 2.0 ** -24 * scale), i
    for rr in range(len(offs) - 1):
        seg = x[offs[rr]:offs[rr + 1]]
        ex = oracle.dd_stats(seg)["sum"]
        assert abs(sums[rr] - ex) <= 8 * 2.0 ** -24 * float(np.abs(seg).astype(np.float64).sum()) + 2.0 ** -24 * abs(ex), rr
    plan.close()


def test_grouped_replay_and_graph(gpu, batch):
    probs, offs, x = batch
    probs = probs[:40]
    dev = upload(probs)
    dx = device.DeviceArray.upload(x)
    dout = device.DeviceArray((len(offs) - 1,), np.float32)
    plan = GroupedPlan(descs_of(dev), dx.ptr, offs, dout.ptr)
    st = device.Stream()
    plan.run(st.handle)
    st.synchronize()
    want = [dc.download() for (*_, dc) in dev] + [dout.download()]
    # eager replays: the kernel resets its own queue
    for _ in range(2):
        plan.run(st.handle)
    st.synchronize()
    assert all(np.array_equal(a, b) for a, b in zip(want, [dc.download() for (*_, dc) in dev] + [dout.download()]))
    graph = Graph(st.handle, lambda: plan.run(st.handle))
    for rep in range(3):
        for (*_, dc) in dev:  # eager work between replays: clear the outputs
            device.call_memset(dc)
        device.call_memset(dout)
        st.synchronize()
        graph.launch()
        st.synchronize()
        got = [dc.download() for (*_, dc) in dev] + [dout.download()]
        assert all(np.array_equal(a, b) for a, b in zip(want, got)), f"replay {rep}"
    graph.close()
    plan.close()
    st.close()


def test_empty_and_rows_only(gpu, oracle):
    plan = GroupedPlan([])
    plan.run()
    plan.close()
    x = rng.normal(5, 100_000, np.float32)
    offs = np.array([0, 10, 10, 5000, 100_000], np.int64)  # includes an empty row and a long row
    dx = device.DeviceArray.upload(x)
    dout = device.DeviceArray((4,), np.float32)
    plan = GroupedPlan([], dx.ptr, offs, dout.ptr)
    plan.run()
    device.synchronize()
    s = dout.download()
    assert s[1] == 0.0
    for rr in range(4):
        seg = x[offs[rr]:offs[rr + 1]]
        assert abs(s[rr] - oracle.dd_stats(seg)["sum"]) <= 8 * 2.0 ** -24 * float(np.abs(seg).sum()) + 1e-30
    plan.close()


def test_packed_unaligned_operands_and_per_call_loop(gpu, oracle, batch):
    """Operands packed back to back in one arena each (bases and rows on 2-byte boundaries only:
    the aligned-block loads and their shifts): the grouped result is bit-identical to each GEMM
    run alone on its own 16-byte-aligned copy, and nrk_dev_ragged_per_call (one launch per GEMM
    and per row from C) gives the same GEMM bits and row sums within the bar."""
    import ctypes as C

    from numrs_amd._lib import GemmDesc, call

    probs, offs, x = batch
    probs = probs[:60]
    a_off = np.concatenate([[1], 1 + np.cumsum([p[0] * p[2] for p in probs])])  # start one element in
    b_off = np.concatenate([[3], 3 + np.cumsum([p[2] * p[1] for p in probs])])
    arena_a = np.zeros(int(a_off[-1]) + 8, np.uint16)
    arena_b = np.zeros(int(b_off[-1]) + 8, np.uint16)
    for g, (m, n, k, a, b) in enumerate(probs):
        arena_a[a_off[g]:a_off[g] + m * k] = a
        arena_b[b_off[g]:b_off[g] + k * n] = b
    dA, dB = device.DeviceArray.upload(arena_a), device.DeviceArray.upload(arena_b)
    outs = [device.DeviceArray((m, n), np.float32) for (m, n, *_r) in probs]
    descs = [(m, n, k, dA.ptr + 2 * int(a_off[g]), dB.ptr + 2 * int(b_off[g]), outs[g].ptr)
             for g, (m, n, k, *_r) in enumerate(probs)]
    dx = device.DeviceArray.upload(x)
    dout = device.DeviceArray((len(offs) - 1,), np.float32)
    plan = GroupedPlan(descs, dx.ptr, offs, dout.ptr)
    plan.run()
    device.synchronize()
    grouped = [o.download() for o in outs]
    sums = dout.download()
    plan.close()
    aligned = upload(probs)
    for (m, n, k, da, db, dc) in aligned:
        gemm_bf16(m, n, k, da.ptr, db.ptr, dc.ptr)
    device.synchronize()
    for g, (cg, (*_r, dc)) in enumerate(zip(grouped, aligned)):
        np.testing.assert_array_equal(cg, dc.download(), err_msg=f"problem {g}")
    # the per-call loop over the same packed operands
    for o in outs:
        device.call_memset(o)
    dout2 = device.DeviceArray((len(offs) - 1,), np.float32)
    arr = (GemmDesc * len(descs))(*[GemmDesc(m, n, k, 0, a, b, c) for (m, n, k, a, b, c) in descs])
    roff = np.ascontiguousarray(offs, np.int64)
    call("nrk_dev_ragged_per_call", len(descs), C.cast(arr, C.c_void_p), len(offs) - 1, C.c_void_p(dx.ptr),
         roff.ctypes.data_as(C.c_void_p), C.c_void_p(dout2.ptr), None)
    device.synchronize()
    for g, (o, cg) in enumerate(zip(outs, grouped)):
        np.testing.assert_array_equal(o.download(), cg, err_msg=f"per-call problem {g}")
    s2 = dout2.download()
    for rr in range(len(offs) - 1):
        seg = x[offs[rr]:offs[rr + 1]]
        ex = oracle.dd_stats(seg)["sum"]
        bar = 8 * 2.0 ** -24 * float(np.abs(seg).astype(np.float64).sum()) + 2.0 ** -24 * abs(ex)
        assert abs(s2[rr] - ex) <= bar and abs(sums[rr] - ex) <= bar, rr
