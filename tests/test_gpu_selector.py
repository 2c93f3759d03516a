"""The MLP cost model's HIP kernels against the FP64 oracle (oracle/mlp.py), and what the trained
selector gains over the built-in heuristics it replaces (no reference counterpart).

Bars, written here: forward |y - y*| <= 1e-4 (1 + |y*|); gradient of the mean loss within
2e-3 * max|g*| per parameter block (fp32 sums over 512 samples); Adam equal to the oracle's update
applied to the GPU's own gradient within 1e-5 relative (moments: 1e-6 of their largest value); the
learned selector's regret on held-out shapes below the heuristic's."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device
from numrs_amd import selector as S
from numrs_amd._lib import call

pytestmark = pytest.mark.gpu


def _case(K, N, seed):
    from oracle import mlp

    r = np.random.default_rng(seed)
    p = r.normal(0, 0.3, mlp.param_count(K)).astype(np.float32)
    x = r.normal(0, 1, (N, 6)).astype(np.float32)
    t = r.normal(0, 1, (N, K)).astype(np.float32)
    mask = (r.uniform(size=(N, K)) > 0.2).astype(np.float32)
    return mlp, p, x, t, mask


def _train_step(p, x, t, mask, K, lr, step=1, m=None, v=None):
    dp, dx, dt, dm = (device.DeviceArray.upload(a) for a in (p, x, t, mask))
    dmm = device.DeviceArray.upload(np.zeros_like(p) if m is None else m)
    dvv = device.DeviceArray.upload(np.zeros_like(p) if v is None else v)
    g = device.DeviceArray((len(p),), np.float32)
    loss = device.DeviceArray((1,), np.float32)
    call("nrk_mlp_train_step", C.c_void_p(dp.ptr), C.c_void_p(dmm.ptr), C.c_void_p(dvv.ptr), step, K, C.c_void_p(dx.ptr),
         C.c_void_p(dt.ptr), C.c_void_p(dm.ptr), len(x), C.c_float(lr), C.c_void_p(loss.ptr), C.c_void_p(g.ptr), None)
    device.synchronize()
    return dp.download(), g.download(), float(loss.download()[0]), dmm.download(), dvv.download()


@pytest.mark.parametrize("K", [1, 5, 8])
def test_forward_matches_oracle(gpu, K):
    mlp, p, x, _, _ = _case(K, 777, 40 + K)
    dp, dx = device.DeviceArray.upload(p), device.DeviceArray.upload(x)
    dy = device.DeviceArray((len(x), K), np.float32)
    call("nrk_mlp_forward", C.c_void_p(dp.ptr), K, C.c_void_p(dx.ptr), len(x), C.c_void_p(dy.ptr), None)
    y = dy.download()
    ref, _ = mlp.forward(p.astype(np.float64), K, x.astype(np.float64))
    assert np.all(np.abs(y - ref) <= 1e-4 * (1 + np.abs(ref)))


@pytest.mark.parametrize("K", [3, 5])
def test_gradients_match_oracle(gpu, K):
    mlp, p, x, t, mask = _case(K, 512, 50 + K)
    p1, g, loss, _, _ = _train_step(p, x, t, mask, K, lr=0.0)
    np.testing.assert_array_equal(p1, p)  # lr = 0: no update
    total, _, gref = mlp.loss_and_grad(p.astype(np.float64), K, x.astype(np.float64), t.astype(np.float64),
                                       mask.astype(np.float64))
    assert abs(loss - total) <= 1e-4 * total
    o = 0
    for size in (64 * 6, 64, 64 * 64, 64, K * 64, K):
        blk, ref = g[o:o + size], gref[o:o + size]
        assert np.all(np.abs(blk - ref) <= 2e-3 * np.abs(ref).max() + 1e-7), (o, size)
        o += size


def test_adam_step_matches_oracle(gpu):
    mlp, p, x, t, mask = _case(4, 256, 61)
    r = np.random.default_rng(62)
    m0 = r.normal(0, 1e-3, len(p)).astype(np.float32)
    v0 = np.abs(r.normal(0, 1e-5, len(p))).astype(np.float32)
    p1, g, _, m1, v1 = _train_step(p, x, t, mask, 4, lr=1e-2, step=3, m=m0, v=v0)
    # the kernel's constants are fp32 (0.999f is 0.99900001287...): use the same values
    b1, b2 = float(np.float32(0.9)), float(np.float32(0.999))
    pr, mr, vr = mlp.adam(p.astype(np.float64), g.astype(np.float64), m0.astype(np.float64), v0.astype(np.float64), 3,
                          float(np.float32(1e-2)), b1=b1, b2=b2)
    # fp32 moment updates: b1 m + (1 - b1) g can cancel, so the bound is relative to the largest term
    np.testing.assert_allclose(m1, mr, rtol=1e-5, atol=1e-6 * np.abs(mr).max())
    np.testing.assert_allclose(v1, vr, rtol=1e-5, atol=1e-6 * np.abs(vr).max())
    np.testing.assert_allclose(p1, pr, rtol=1e-5, atol=1e-7)


def test_training_fits_a_cost_surface(gpu):
    """Full-batch Adam on the GPU fits a smooth function of the features (loss drops > 20x)."""
    r = np.random.default_rng(70)
    f = S.features(*(S.log_uniform_int(r, 2000, 16, 512) for _ in range(3)), r.integers(0, 3, 2000))
    T = np.stack([f[:, 0] + f[:, 1] + 0.5 * f[:, 2] + 0.3 * j * f[:, 3] - 0.1 * j * f[:, 0] for j in range(5)], axis=1)
    model = S.CostModel(5, seed=3)
    hist = model.fit(f, T.astype(np.float32), steps=600)
    assert hist[-1] < hist[0] / 20, hist
    err = np.abs(model.predict(f) - T)
    assert np.median(err) < 0.1


def test_selector_gains_over_heuristic(gpu):
    """Tune GEMM shapes drawn as in the ragged batch, train on 1200, choose on 400 held-out: the
    learned choice has lower regret against exhaustive search than the built-in tile rule."""
    r = np.random.default_rng(80)
    train = S.tune_gemm_set(*S.gemm_shapes(r, 1200), reps=3)
    held = S.tune_gemm_set(*S.gemm_shapes(r, 400), reps=3)
    model = S.CostModel(S.config_count(S.F64), seed=4)
    model.fit(train.X, np.log2(train.times), steps=800)
    res = S.evaluate(model, held, margin=0.0)
    print(res)
    assert res["regret_model"] < res["regret_heuristic"], res
    assert res["regret_model_with_fallback"] <= res["regret_model"] + 1e-12
