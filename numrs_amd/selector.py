"""Autotuner + learned kernel selector (no reference counterpart).

The reference fixes its dispatch tiers at compile time (kernels/mod.rs:63-106 thresholds,
kernels/gemm.rs:290-312). Here every operation family has a table of launch configurations
(`nrk_config_count`); the tuner times all of them per shape (`nrk_tune`, exhaustive search), and
a 6 -> 64 -> 64 -> K MLP cost model (`nrk_mlp_forward` / `nrk_mlp_train_step`, trained and run by
the library's own HIP kernels) predicts log2(time) of every configuration so that the fastest one
is picked without timing. When the model's two best predictions are closer than a margin, the
selector falls back to timing the shape (exhaustive), and caches the answer.

Inputs of the model (one row per call): log2 d0, log2 d1, log2 d2, dtype id (0 f64, 1 f32,
2 bf16), transA, transB -- for GEMMs d0, d1, d2 = m, n, k; for row reductions rows, cols, 1.

    python -m numrs_amd.selector [--train 20000] [--eval 100000] [--steps 1500] [--seed S]

samples shapes as the ragged-batch workload does (log-uniform integers in [16, 512] for m, n, k
across f64 / f32 / bf16; row sums with lengths log-uniform in [32, 65536]), tunes them, trains the
models on the GPU and prints one JSON line: regret (sum of selected times / sum of exhaustive-best
times - 1) on the held-out stream for the model, for the built-in heuristics it replaces, and the
fallback rate and inference throughput.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import device
from ._lib import call, load

# operation families (include/numrs_amd.h)
F64, F32, BF16, ROWSUM_F32, F32_BF16X3, F64_OZAKI, F32_BF16X6 = 0, 1, 2, 3, 4, 5, 6
GEMM_FAMILY_DTYPE = {F64: 0, F32: 1, BF16: 2}
IN, H = 6, 64
# the trained GEMM cost model, persisted (numpy .npz, no pickles) and registered with the library's
# dispatch by install()
MODEL_PATH = Path(__file__).resolve().parent / "selector_model.npz"


def config_count(family: int) -> int:
    k = C.c_int32()
    call("nrk_config_count", family, C.byref(k))
    return k.value


def features(d0, d1, d2, dtype_id, ta=0, tb=0) -> np.ndarray:
    d0, d1, d2 = (np.asarray(v, np.float64) for v in (d0, d1, d2))
    n = len(d0)
    f = np.empty((n, IN), np.float32)
    f[:, 0] = np.log2(d0)
    f[:, 1] = np.log2(d1)
    f[:, 2] = np.log2(np.maximum(d2, 1))
    f[:, 3] = np.broadcast_to(np.asarray(dtype_id, np.float32), (n,))
    f[:, 4] = ta
    f[:, 5] = tb
    return f


def tune(family: int, d0, d1, d2=None, reps: int = 3, batch: int = 2048) -> np.ndarray:
    """Exhaustive timings (ns per launch, mean of `reps` after one warm-up): (N, K) float32."""
    d0 = np.ascontiguousarray(d0, np.int32)
    d1 = np.ascontiguousarray(d1, np.int32)
    d2 = np.ascontiguousarray(np.zeros_like(d0) if d2 is None else d2, np.int32)
    K = config_count(family)
    out = np.empty((len(d0), K), np.float32)
    for s in range(0, len(d0), batch):
        e = min(len(d0), s + batch)
        call("nrk_tune", family, e - s, d0[s:e].ctypes.data_as(C.c_void_p), d1[s:e].ctypes.data_as(C.c_void_p),
             d2[s:e].ctypes.data_as(C.c_void_p), reps, out[s:e].ctypes.data_as(C.c_void_p))
    return out


class CostModel:
    """The MLP, parameters resident on the device; trained by full-batch Adam steps on the GPU."""

    def __init__(self, K: int, seed: int = 0):
        load()
        self.K = K
        self.P = int(call_size("nrk_mlp_param_count", K))
        r = np.random.default_rng(seed)
        p = np.concatenate([
            r.normal(0, math.sqrt(2.0 / IN), H * IN), np.zeros(H),
            r.normal(0, math.sqrt(2.0 / H), H * H), np.zeros(H),
            r.normal(0, math.sqrt(1.0 / H), K * H), np.zeros(K)]).astype(np.float32)
        assert len(p) == self.P
        self.params = device.DeviceArray.upload(p)
        self.m = device.DeviceArray.upload(np.zeros(self.P, np.float32))
        self.v = device.DeviceArray.upload(np.zeros(self.P, np.float32))
        self.step = 0
        self.x_mean = np.zeros(IN, np.float32)
        self.x_scale = np.ones(IN, np.float32)
        self.y_offset = 0.0

    def _norm(self, X: np.ndarray) -> np.ndarray:
        return ((X - self.x_mean) / self.x_scale).astype(np.float32)

    def fit(self, X: np.ndarray, T: np.ndarray, mask: np.ndarray | None = None, steps: int = 1500, lr: float = 3e-3,
            stream=None) -> list[float]:
        """Full-batch training on log2 targets T (N, K); returns the loss every 100 steps."""
        X = np.asarray(X, np.float32)
        self.x_mean = X.mean(axis=0)
        sc = X.std(axis=0)
        self.x_scale = np.where(sc > 0, sc, 1.0).astype(np.float32)
        mask = np.ones_like(T, np.float32) if mask is None else np.asarray(mask, np.float32)
        self.y_offset = float(np.sum(T * mask) / max(1.0, float(mask.sum())))
        dx = device.DeviceArray.upload(self._norm(X))
        dt = device.DeviceArray.upload(np.ascontiguousarray(T - self.y_offset, np.float32))
        dm = device.DeviceArray.upload(mask)
        loss = device.DeviceArray((1,), np.float32)
        n = len(X)
        hist = []
        s = C.c_void_p(stream)
        for i in range(steps):
            self.step += 1
            want = (i % 100 == 0) or i == steps - 1
            call("nrk_mlp_train_step", C.c_void_p(self.params.ptr), C.c_void_p(self.m.ptr), C.c_void_p(self.v.ptr),
                 self.step, self.K, C.c_void_p(dx.ptr), C.c_void_p(dt.ptr), C.c_void_p(dm.ptr), n, C.c_float(lr),
                 C.c_void_p(loss.ptr) if want else None, None, s)
            if want:
                hist.append(float(loss.download()[0]) / float(mask.sum()))
        return hist

    def predict(self, X: np.ndarray, stream=None) -> np.ndarray:
        """Predicted log2(time ns), (N, K)."""
        X = np.asarray(X, np.float32)
        dx = device.DeviceArray.upload(self._norm(X))
        dy = device.DeviceArray((len(X), self.K), np.float32)
        call("nrk_mlp_forward", C.c_void_p(self.params.ptr), self.K, C.c_void_p(dx.ptr), len(X), C.c_void_p(dy.ptr),
             C.c_void_p(stream))
        return dy.download() + self.y_offset


def save_model(path, model: "CostModel", X: np.ndarray, margin: float) -> None:
    """Persist a trained GEMM cost model: parameters, input normalisation, target offset, the
    log2 range of the training shapes (the library consults the model only inside it) and the
    fallback margin."""
    X = np.asarray(X, np.float32)
    np.savez(path, params=model.params.download(), K=np.int32(model.K), x_mean=model.x_mean, x_scale=model.x_scale,
             y_offset=np.float32(model.y_offset), domain_lo=X[:, :3].min(axis=0), domain_hi=X[:, :3].max(axis=0),
             margin=np.float32(margin))


def install(path=None, families=(F64, F32, BF16)) -> list[int]:
    """Register the persisted GEMM cost model with the library's dispatch for `families` (one model
    for all three GEMM dtypes, told apart by the dtype feature). Returns the families installed:
    none when the file is absent (the built-in rules then stay)."""
    path = MODEL_PATH if path is None else Path(path)
    if not path.exists():
        return []
    z = np.load(path, allow_pickle=False)
    K = int(z["K"])
    arrs = {k: np.ascontiguousarray(z[k], np.float32) for k in ("params", "x_mean", "x_scale", "domain_lo", "domain_hi")}
    ptr = {k: v.ctypes.data_as(C.c_void_p) for k, v in arrs.items()}
    done = []
    for fam in families:
        call("nrk_selector_set_model", fam, C.c_float(GEMM_FAMILY_DTYPE[fam]), ptr["params"], K, ptr["x_mean"], ptr["x_scale"],
             ptr["domain_lo"], ptr["domain_hi"], C.c_float(float(z["y_offset"])), C.c_float(float(z["margin"])))
        done.append(fam)
    return done


def choose(family: int, d0: int, d1: int, d2: int) -> int:
    """The tile configuration a dispatch of this shape takes now (-1: the built-in rule)."""
    c = C.c_int32()
    call("nrk_selector_choose", family, d0, d1, d2, C.byref(c))
    return c.value


def stats(family: int) -> dict:
    v = [C.c_int64() for _ in range(4)]
    call("nrk_selector_stats", family, *[C.byref(x) for x in v])
    return dict(zip(("rule", "model", "timing", "table"), (x.value for x in v)))


def table() -> list[tuple[int, int, int, int, int]]:
    """The timed decisions (family, d0, d1, d2, config), e.g. to persist them."""
    n = C.c_int32()
    call("nrk_selector_table_size", C.byref(n))
    out = []
    for i in range(n.value):
        v = [C.c_int32() for _ in range(5)]
        call("nrk_selector_table_get", i, *[C.byref(x) for x in v])
        out.append(tuple(x.value for x in v))
    return out


def call_size(name: str, *args) -> int:
    """Entry points that return a size_t instead of a status."""
    return int(getattr(load(), name)(*args))


def heuristic(family: int, d0, d1, d2=None) -> np.ndarray:
    """The built-in choice the selector replaces: native GEMM takes 128x128 tiles (config 4) when
    there are at least as many 128x128 tiles as CUs, else 64x64 (config 1) when there are more
    64x64 tiles than CUs, else 32x32 (config 0); the bf16 GEMM uses 64x64 (config 1); row sums
    use 8 workgroups per CU (config 3)."""
    d0 = np.asarray(d0)
    if family in (F64, F32):
        cus = device_cus()
        d1 = np.asarray(d1)
        t128 = np.ceil(d0 / 128) * np.ceil(d1 / 128)
        t64 = np.ceil(d0 / 64) * np.ceil(d1 / 64)
        return np.where(t128 >= cus, 4, np.where(t64 > cus, 1, 0))
    if family == BF16:
        return np.full(len(d0), 1)
    if family == ROWSUM_F32:
        return np.full(len(d0), 3)
    return np.full(len(d0), 0)


def device_cus() -> int:
    n = C.c_int32()
    call("nrk_device_cu_count", C.byref(n))
    return n.value


class Selector:
    """Model-based choice with exhaustive fallback when the top-2 predictions are within `margin`
    (log2 units, i.e. a ratio of 2^margin)."""

    def __init__(self, model: CostModel, margin: float = 0.0):
        self.model = model
        self.margin = margin

    def choose(self, X: np.ndarray):
        pred = self.model.predict(X)
        order = np.argsort(pred, axis=1)
        best = order[:, 0]
        gap = np.take_along_axis(pred, order[:, 1:2], axis=1)[:, 0] - np.take_along_axis(pred, order[:, :1], axis=1)[:, 0]
        fallback = gap < self.margin
        return best, fallback


def regret(times: np.ndarray, chosen: np.ndarray) -> float:
    """sum of chosen times / sum of the per-shape best - 1."""
    sel = times[np.arange(len(times)), chosen]
    return float(sel.sum() / times.min(axis=1).sum() - 1.0)


# ------------------------------------------------------------------------------------------------
# shapes as the ragged-batch workload draws them

def log_uniform_int(r: np.random.Generator, n: int, lo: int, hi: int) -> np.ndarray:
    u = r.uniform(0.0, 1.0, n)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi + 1) - np.log(lo)))).astype(np.int64).clip(lo, hi)


def gemm_shapes(r: np.random.Generator, n: int):
    """(family, m, n, k): m, n, k log-uniform integers in [16, 512], dtype uniform over f64/f32/bf16."""
    fam = r.choice([F64, F32, BF16], n)
    return fam, log_uniform_int(r, n, 16, 512), log_uniform_int(r, n, 16, 512), log_uniform_int(r, n, 16, 512)


def row_shapes(r: np.random.Generator, n: int, max_elems: int = 1 << 24):
    """(rows, cols): cols log-uniform in [32, 65536], rows log-uniform in [1, 4096], rows*cols capped."""
    cols = log_uniform_int(r, n, 32, 65536)
    rows = log_uniform_int(r, n, 1, 4096)
    rows = np.minimum(rows, np.maximum(1, max_elems // cols))
    return rows, cols


class TunedSet:
    """Shapes of one model with their exhaustive timings (log2 ns targets)."""

    def __init__(self, feats: np.ndarray, times: np.ndarray, heur: np.ndarray, kind: np.ndarray):
        self.X, self.times, self.heur, self.kind = feats, times, heur, kind


def tune_gemm_set(fam, m, n, k, reps) -> TunedSet:
    K = config_count(F64)
    times = np.empty((len(m), K), np.float32)
    heur = np.empty(len(m), np.int64)
    for f in (F64, F32, BF16):
        sel = fam == f
        if sel.any():
            times[sel] = tune(f, m[sel], n[sel], k[sel], reps)
            heur[sel] = heuristic(f, m[sel], n[sel])
    return TunedSet(features(m, n, k, np.vectorize(GEMM_FAMILY_DTYPE.get)(fam)), times, heur, fam)


def tune_row_set(rows, cols, reps) -> TunedSet:
    times = tune(ROWSUM_F32, rows, cols, None, reps)
    return TunedSet(features(rows, cols, np.ones_like(rows), 1), times, heuristic(ROWSUM_F32, rows, cols),
                    np.full(len(rows), ROWSUM_F32))


def evaluate(model: CostModel, ts: TunedSet, margin: float) -> dict:
    sel = Selector(model, margin)
    t0 = time.perf_counter()
    chosen, fb = sel.choose(ts.X)
    infer_s = time.perf_counter() - t0
    best = ts.times.argmin(axis=1)
    with_fb = np.where(fb, best, chosen)
    return {"n": len(ts.X), "regret_model": regret(ts.times, chosen), "regret_model_with_fallback": regret(ts.times, with_fb),
            "fallback_rate": float(fb.mean()), "regret_heuristic": regret(ts.times, ts.heur),
            "top1_agreement": float((chosen == best).mean()), "inference_s": infer_s,
            "stream_ms": {"exhaustive": float(ts.times.min(axis=1).sum() / 1e6),
                          "model_with_fallback": float(ts.times[np.arange(len(best)), with_fb].sum() / 1e6),
                          "heuristic": float(ts.times[np.arange(len(best)), ts.heur].sum() / 1e6)}}


def run_evaluation(n_train: int = 20000, n_train_rows: int = 4000, n_eval: int = 100000, steps: int = 1500, reps: int = 3,
                   margin: float = 0.05, seed: int = 0x4E554D5253, save=None, log=None) -> dict:
    """Tune n_train GEMM and n_train_rows row-sum shapes, train the two cost models on the GPU, and
    evaluate model-based selection (with timing fallback) against exhaustive search and the
    built-in rules on a held-out stream of n_eval calls (80% GEMMs, 20% row sums). When the
    persisted GEMM model exists, it is scored on the same held-out GEMM calls too."""
    log = log or (lambda s: print(s, file=sys.stderr, flush=True))
    r = np.random.default_rng(seed)
    t0 = time.time()
    tr_g = tune_gemm_set(*gemm_shapes(r, n_train), reps)
    tr_r = tune_row_set(*row_shapes(r, n_train_rows), reps)
    log(f"tuned training shapes in {time.time() - t0:.1f}s")
    n_eval_g = int(round(n_eval * 0.8))
    t1 = time.time()
    ev_g = tune_gemm_set(*gemm_shapes(r, n_eval_g), reps)
    ev_r = tune_row_set(*row_shapes(r, n_eval - n_eval_g), reps)
    log(f"tuned held-out stream in {time.time() - t1:.1f}s")
    t2 = time.time()
    mg = CostModel(config_count(F64), seed=1)
    hist_g = mg.fit(tr_g.X, np.log2(tr_g.times), steps=steps)
    mr = CostModel(config_count(ROWSUM_F32), seed=2)
    hist_r = mr.fit(tr_r.X, np.log2(tr_r.times), steps=steps)
    log(f"trained in {time.time() - t2:.1f}s; loss gemm {hist_g[0]:.3f} -> {hist_g[-1]:.4f}, rows {hist_r[0]:.3f} -> {hist_r[-1]:.4f}")
    if save:
        save_model(save, mg, tr_g.X, margin)
        log(f"saved the GEMM model to {save}")
    eg, er = evaluate(mg, ev_g, margin), evaluate(mr, ev_r, margin)
    times = np.concatenate([ev_g.times.min(axis=1), ev_r.times.min(axis=1)])
    # the mixed stream: every held-out call with the model's (or the fallback's) choice
    cg, fg = Selector(mg, margin).choose(ev_g.X)
    cr, fr = Selector(mr, margin).choose(ev_r.X)
    sel_g = np.where(fg, ev_g.times.argmin(1), cg)
    sel_r = np.where(fr, ev_r.times.argmin(1), cr)
    t_sel = ev_g.times[np.arange(len(sel_g)), sel_g].sum() + ev_r.times[np.arange(len(sel_r)), sel_r].sum()
    t_heur = ev_g.times[np.arange(len(sel_g)), ev_g.heur].sum() + ev_r.times[np.arange(len(sel_r)), ev_r.heur].sum()
    big = np.tile(ev_g.X, (max(1, 100000 // max(1, len(ev_g.X))), 1))[:100000]
    mg.predict(big)
    ti = time.perf_counter()
    mg.predict(big)
    infer = time.perf_counter() - ti
    out = {"train_shapes": {"gemm": n_train, "rows": n_train_rows}, "eval_calls": n_eval, "margin_log2": margin,
           "stream": {"regret_model_with_fallback": float(t_sel / times.sum() - 1.0),
                      "regret_heuristic": float(t_heur / times.sum() - 1.0),
                      "fallback_rate": float((fg.sum() + fr.sum()) / n_eval),
                      "exhaustive_ms": float(times.sum() / 1e6), "selected_ms": float(t_sel / 1e6),
                      "heuristic_ms": float(t_heur / 1e6)},
           "gemm": eg, "rows": er, "mlp_inference_per_s": len(big) / infer,
           "loss": {"gemm": [hist_g[0], hist_g[-1]], "rows": [hist_r[0], hist_r[-1]]},
           "seconds": {"tune": round(t2 - t0, 1), "train": round(time.time() - t2, 1)}}
    if MODEL_PATH.exists():
        z = np.load(MODEL_PATH, allow_pickle=False)
        shipped = CostModel(int(z["K"]), seed=0)
        shipped.params = device.DeviceArray.upload(np.ascontiguousarray(z["params"], np.float32))
        shipped.x_mean, shipped.x_scale = z["x_mean"].astype(np.float32), z["x_scale"].astype(np.float32)
        shipped.y_offset = float(z["y_offset"])
        out["shipped_model_gemm"] = evaluate(shipped, ev_g, float(z["margin"]))
    return out


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train", type=int, default=20000, help="GEMM training shapes")
    ap.add_argument("--train-rows", type=int, default=4000, help="row-reduction training shapes")
    ap.add_argument("--eval", type=int, default=100000, help="held-out calls (80%% GEMM, 20%% row sums)")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--margin", type=float, default=0.05, help="fallback when the top-2 predictions are closer (log2)")
    ap.add_argument("--seed", type=int, default=0x4E554D5253)
    ap.add_argument("--save", default=None, help="write the trained GEMM model here (.npz)")
    a = ap.parse_args(argv)
    device.init(0)
    out = run_evaluation(a.train, a.train_rows, a.eval, a.steps, a.reps, a.margin, a.seed, a.save)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
