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
