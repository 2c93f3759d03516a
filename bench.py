#!/usr/bin/env python3
"""Benchmark of the kernels dispatch path on one MI355X.

Headline workload (one "step"): for x of shape (8192, 8192) with x ~ N(0,1), in f32 and in f64,
y = 1.5 * x - 0.25 (two roundings) followed by sum, max, argmax and var along axis 1 AND over the
whole array -- all eight results from ONE fused pass over x per dtype -- plus the exact sum of an
i64 array of 2^28 values uniform in [-2^20, 2^20]. Inputs are resident in HBM before timing.

value = algorithmic bytes of one step (each input read once: 2^26*4 + 2^26*8 + 2^28*8 B) divided
by the measured step time, in GB/s; the timed region holds the K steps and nothing else. A second
pass times each call on its own (HIP events on the launching stream) for the per-kernel figures
and the roofline of the dominant kernel against MI355X's 8 TB/s.

Also reported (same JSON line): the split-precision GEMMs against the native MFMA GEMMs (f32
8192^3 BF16x3/x6, f64 4096^3 and 8192^3 Ozaki-II with 13-16 moduli; error of sampled entries
against correctly rounded dot products) and the ragged batch (4096 bf16 GEMMs + 65536 f32 row
sums, one persistent launch replayed from a hipGraph vs the per-call loop). cpu_baseline times the
oracle's restatement of the reference's CPU path (unfused) on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-gemm] [--no-ragged]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SAMPLES = 512  # sampled GEMM entries for the error of each engine
ROWS = COLS = 8192
N_I64 = 1 << 28
# HBM bytes per launch measured with `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` over
# `python3 bench.py --pmc-pass`, each counter scaled by its reading for a plain 2 GiB copy
# (FETCH factor 2097152 / 1048730.438 KiB, WRITE factor 2097152 / 2097212.25 KiB). The run and its
# figures are kept in profiles/hbm_traffic_reduce_step.md; the i64 sum fetches 0.9999x its
# algorithmic bytes (no re-reads).
_FETCH, _WRITE = 2097152 / 1048730.438, 2097152 / 2097212.25
MEASURED_TRAFFIC = {"sum_i64": (1048619.062 * _FETCH + 32.062 * _WRITE) * 1024,
                    "fused_map_reduce_f32": (132561.5 * _FETCH + 385.156 * _WRITE) * 1024,
                    "fused_map_reduce_f64": (263728.375 * _FETCH + 320.656 * _WRITE) * 1024}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-gemm", action="store_true")
    p.add_argument("--no-ragged", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-ufuncs", action="store_true")
    p.add_argument("--no-selector", action="store_true")
    p.add_argument("--no-shapes", action="store_true")
    p.add_argument("--pmc-pass", action="store_true",
                   help="run each hot call once after a plain copy of known size (2 GiB, the counter "
                        "calibration), for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE; prints no JSON line")
    return p.parse_args()


def pmc_pass() -> int:
    """Dispatches, in order: a plain copy of 2^28 f64 (2 GiB read + 2 GiB written; the elementwise
    map with an empty chain), the fused f32 and f64 calls and the i64 sum of the headline step."""
    from numrs_amd import device, kernels, rng
    from numrs_amd._lib import ReduceOutputs, call

    device.init(0)
    stream = device.Stream()
    s = stream.handle
    src = device.DeviceArray.upload(rng.normal(rng.BASE_SEED + 6, N_I64, np.float64))
    dst = device.DeviceArray((N_I64,), np.float64)
    call("nrk_dev_map_f64", None, 0, C.c_void_p(src.ptr), C.c_void_p(dst.ptr), C.c_size_t(N_I64), C.c_void_p(s))
    stream.synchronize()
    del src, dst
    ops, n_ops = kernels.chain(("mul", 1.5), ("add", -0.25))
    for dt, fn, seed in ((np.float32, "nrk_dev_map_reduce_f32", 2), (np.float64, "nrk_dev_map_reduce_f64", 3)):
        x = device.DeviceArray.upload(rng.normal(rng.BASE_SEED + seed, ROWS * COLS, dt))
        ro, fo, keep = ReduceOutputs(), ReduceOutputs(), []
        for f in ("sum", "max", "argmax", "var"):
            r, g = device.DeviceArray((ROWS,), np.float64), device.DeviceArray((1,), np.float64)
            keep += [r, g]
            setattr(ro, f, r.ptr)
            setattr(fo, f, g.ptr)
        call(fn, C.cast(ops, C.c_void_p), n_ops, C.c_void_p(x.ptr), ROWS, COLS, 0, C.byref(ro), C.byref(fo), C.c_void_p(s))
        stream.synchronize()
    xi = device.DeviceArray.upload(rng.int_uniform(rng.BASE_SEED + 4, N_I64, -(1 << 20), 1 << 20))
    out = device.DeviceArray((1,), np.int64)
    call("nrk_dev_sum_i64", C.c_void_p(xi.ptr), C.c_size_t(N_I64), C.c_void_p(out.ptr), C.c_void_p(s))
    stream.synchronize()
    log("pmc pass done")
    return 0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "nrk runs on one GPU only; --gpus must be 1 "
                          "(how the path would shard is a design note, not a launched configuration)"}))
        return 2
    if args.pmc_pass:
        return pmc_pass()

    from numrs_amd import device, kernels, rng
    from numrs_amd._lib import ReduceOutputs, call

    device.init(0)
    t0 = time.time()
    x32 = rng.normal(rng.BASE_SEED + 2, ROWS * COLS, np.float32).reshape(ROWS, COLS)
    x64 = rng.normal(rng.BASE_SEED + 3, ROWS * COLS, np.float64).reshape(ROWS, COLS)
    xi = rng.int_uniform(rng.BASE_SEED + 4, N_I64, -(1 << 20), 1 << 20)
    log(f"inputs generated in {time.time() - t0:.1f}s")
    d32 = device.DeviceArray.upload(x32)
    d64 = device.DeviceArray.upload(x64)
    dxi = device.DeviceArray.upload(xi)
    stream = device.Stream()
    s = stream.handle

    fields = ("sum", "max", "argmax", "var")

    def outputs(dt):
        ro, fo = ReduceOutputs(), ReduceOutputs()
        keep = []
        for f in fields:
            odt = np.uint64 if f == "argmax" else dt
            r = device.DeviceArray((ROWS,), odt)
            g = device.DeviceArray((1,), odt)
            keep += [r, g]
            setattr(ro, f, r.ptr)
            setattr(fo, f, g.ptr)
        return ro, fo, keep

    ro32, fo32, k32 = outputs(np.float32)
    ro64, fo64, k64 = outputs(np.float64)
    isum = device.DeviceArray((1,), np.int64)
    ops, n_ops = kernels.chain(("mul", 1.5), ("add", -0.25))
    ops_p = C.cast(ops, C.c_void_p)

    def k_f32():
        call("nrk_dev_map_reduce_f32", ops_p, n_ops, C.c_void_p(d32.ptr), ROWS, COLS, 0, C.byref(ro32), C.byref(fo32), C.c_void_p(s))

    def k_f64():
        call("nrk_dev_map_reduce_f64", ops_p, n_ops, C.c_void_p(d64.ptr), ROWS, COLS, 0, C.byref(ro64), C.byref(fo64), C.c_void_p(s))

    def k_i64():
        call("nrk_dev_sum_i64", C.c_void_p(dxi.ptr), C.c_size_t(N_I64), C.c_void_p(isum.ptr), C.c_void_p(s))

    kern = [("fused_map_reduce_f32", k_f32, ROWS * COLS * 4),
            ("fused_map_reduce_f64", k_f64, ROWS * COLS * 8),
            ("sum_i64", k_i64, N_I64 * 8)]
    step_bytes = sum(b for _, _, b in kern)

    for _ in range(args.warmup):
        for _, f, _ in kern:
            f()
    k_i64()
    stream.synchronize()
    # correctness guard on the measured data: the i64 sum must be exact
    got = int(isum.download()[0])
    want = int(xi.sum(dtype=np.int64))
    if got != want:
        raise SystemExit(f"i64 sum mismatch: {got} != {want}")

    # ---- the timed region: K steps, nothing else on the stream
    start, stop = device.Event(), device.Event()
    device.synchronize()
    start.record(s)
    for _ in range(args.steps):
        for _, f, _ in kern:
            f()
    stop.record(s)
    stream.synchronize()
    device.synchronize()
    ms_per_step = start.elapsed_ms(stop) / args.steps
    value = step_bytes / (ms_per_step * 1e-3) / 1e9
    log(f"step {ms_per_step * 1e3:.1f} us, {value:.1f} GB/s")

    # ---- per-call durations: each call launched K times back to back between two events
    per = {}
    for name, f, _ in kern:
        a, b = device.Event(), device.Event()
        a.record(s)
        for _ in range(args.steps):
            f()
        b.record(s)
        stream.synchronize()
        per[name] = a.elapsed_ms(b) / args.steps
    kernels_out = {}
    for name, _, nbytes in kern:
        gbs = nbytes / (per[name] * 1e-3) / 1e9
        kernels_out[name] = {"ms": round(per[name], 5), "bytes": nbytes, "GB/s": round(gbs, 1),
                             "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(kern, key=lambda k: per[k[0]])
    dom_gbs = dom[2] / (per[dom[0]] * 1e-3) / 1e9
    # errors of the measured results (outside every timed region) against an extended-precision
    # evaluation of the same y = 1.5 x - 0.25 (two roundings in T)
    kernels_out["fused_map_reduce_f32"]["errors"] = fused_errors(x32, k32)
    kernels_out["fused_map_reduce_f64"]["errors"] = fused_errors(x64, k64)
    log("fused errors: f32 " + json.dumps(kernels_out["fused_map_reduce_f32"]["errors"]) + " f64 "
        + json.dumps(kernels_out["fused_map_reduce_f64"]["errors"]))
    kernels_out["unfused_reference_style"] = unfused_reference_style(device, call, ReduceOutputs, ops_p, n_ops, s, stream,
                                                                     ((np.float32, d32), (np.float64, d64)), args.steps)
    log("unfused reference-style sequence: " + json.dumps(kernels_out["unfused_reference_style"]))
    kernels_out["plain_vs_compensated"] = plain_vs_compensated(device, call, ReduceOutputs, ops_p, n_ops, s, stream,
                                                               ((x32, d32), (x64, d64)), args.steps)
    log("plain vs compensated: " + json.dumps(kernels_out["plain_vs_compensated"]))
    del x64

    # host-pointer boundary (the reference's &[T] slices): sum of x (f32, 256 MiB) through
    # nrk_sum_f32, which stages the slice to the device; PCIe-inclusive, not the headline value
    host_ms = []
    for _ in range(3):
        t = time.perf_counter()
        kernels.sum_f32(x32.reshape(-1))
        host_ms.append((time.perf_counter() - t) * 1e3)
    host_path = {"call": "nrk_sum_f32 (host slice, staged to HBM)", "bytes": x32.nbytes, "ms": round(min(host_ms), 3),
                 "GB/s": round(x32.nbytes / min(host_ms) / 1e6, 1)}
    log(f"host-pointer sum_f32: {host_path['ms']} ms, {host_path['GB/s']} GB/s")

    gemm = None if args.no_gemm else bench_gemm(device, rng, call, s, stream)
    ragged = None if args.no_ragged else bench_ragged(device, rng, call, s, stream)
    batched = None if args.no_batched else bench_batched(device, rng, call, s, stream)
    ufuncs = None if args.no_ufuncs else bench_ufuncs(device, kernels, call, s, stream, d32, d64)
    shapes = None if args.no_shapes else bench_shapes(device, kernels, call, ReduceOutputs, s, stream, d32, d64)
    sel = None if args.no_selector else bench_selector()
    cpu = None if args.no_cpu_baseline else cpu_baseline(x32, d64, xi, step_bytes)

    line = {
        "metric": "GEMM TFLOP/s (split-precision vs native, same err) + reduction GB/s % of HBM",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64+i64",
        "data": "synthetic (counter-based splitmix64: N(0,1) x, uniform-int i64)",
        "config": {"workload": "fused y=1.5x-0.25 -> sum/max/argmax/var along axis 1 and full, x (8192,8192) f32 and f64; "
                               "sum of 2^28 i64 in [-2^20,2^20]",
                   "rows": ROWS, "cols": COLS, "n_i64": N_I64, "step_bytes": step_bytes,
                   "frac_hbm_step": round(value / HBM_PEAK_GBS, 4)},
        "kernels": kernels_out,
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get(dom[0]), "algorithmic_bytes": dom[2]},
        "host_pointer_path": host_path,
        "gemm": gemm,
        "ragged": ragged,
        "batched": batched,
        "ufuncs": ufuncs,
        "general_shapes": shapes,
        "selector": sel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def _exact_stats(y):
    """Row and whole-array sums and population variances of y in 80-bit long double (pairwise)."""
    LD = np.longdouble
    yl = y.astype(LD)
    n = y.shape[1]
    row_sum = yl.sum(axis=1)
    row_var = ((yl - (row_sum / n)[:, None]) ** 2).sum(axis=1) / n
    full_sum = row_sum.sum()
    full_var = ((yl - full_sum / y.size) ** 2).sum() / y.size
    return row_sum, row_var, full_sum, full_var


def unfused_reference_style(device, call, ReduceOutputs, ops_p, n_ops, s, stream, data, steps) -> dict:
    """The reference's unfused sequence on this GPU, for comparison with the fused pass: y = 1.5 x - 0.25
    written to HBM (one map launch; the reference's two maps round the same way), then each of the
    eight reductions (sum, max, argmax, var along axis 1 and over the whole array) as a pass of its
    own over y. Bytes counted: write y once, read x once and y eight times."""
    out = {}
    for dt, dx in data:
        suf = "f32" if dt == np.float32 else "f64"
        isz = np.dtype(dt).itemsize
        n = ROWS * COLS
        dy = device.DeviceArray((n,), dt)
        bufs = []
        outs = []
        for full in (False, True):
            for f in ("sum", "max", "argmax", "var"):
                o = ReduceOutputs()
                b = device.DeviceArray((1 if full else ROWS,), np.uint64 if f == "argmax" else dt)
                bufs.append(b)
                setattr(o, f, b.ptr)
                outs.append((full, o))

        def once():
            call(f"nrk_dev_map_{suf}", ops_p, n_ops, C.c_void_p(dx.ptr), C.c_void_p(dy.ptr), C.c_size_t(n), C.c_void_p(s))
            for full, o in outs:
                call(f"nrk_dev_map_reduce_{suf}", None, 0, C.c_void_p(dy.ptr), ROWS, COLS, 0,
                     None if full else C.byref(o), C.byref(o) if full else None, C.c_void_p(s))
        once()
        stream.synchronize()
        a, b = device.Event(), device.Event()
        a.record(s)
        for _ in range(steps):
            once()
        b.record(s)
        stream.synchronize()
        ms = a.elapsed_ms(b) / steps
        moved = (1 + 1 + 8) * n * isz
        out[suf] = {"ms": round(ms, 5), "launches": 9, "bytes_moved": moved, "GB/s_moved": round(moved / (ms * 1e-3) / 1e9, 1)}
        del dy, bufs
    return out


def plain_vs_compensated(device, call, ReduceOutputs, ops_p, n_ops, s, stream, data, steps) -> dict:
    """The same y = 1.5 x - 0.25 -> row and whole-array sum and var (ddof 0) through the plain call
    (the reference's own order of rounded additions: chunked simd_sum + left-to-right folds, several
    passes) and the compensated one (one pass, double-word merges): time per call and max relative
    error of each against long double."""
    out = {}
    for x, dx in data:
        dt = x.dtype.type
        suf = "f32" if dt == np.float32 else "f64"
        bufs = {k: device.DeviceArray((ROWS if k.startswith("row") else 1,), dt)
                for k in ("row_sum", "row_var", "full_sum", "full_var")}
        ro, fo = ReduceOutputs(), ReduceOutputs()
        ro.sum, ro.var = bufs["row_sum"].ptr, bufs["row_var"].ptr
        fo.sum, fo.var = bufs["full_sum"].ptr, bufs["full_var"].ptr
        y = ((x * dt(1.5)).astype(dt) + dt(-0.25)).astype(dt)
        ex = _exact_stats(y)
        del y
        entry = {}
        for mode, fn in (("plain", f"nrk_dev_map_reduce_plain_{suf}"), ("compensated", f"nrk_dev_map_reduce_{suf}")):
            def once():
                call(fn, ops_p, n_ops, C.c_void_p(dx.ptr), ROWS, COLS, 0, C.byref(ro), C.byref(fo), C.c_void_p(s))
            once()
            stream.synchronize()
            a, b = device.Event(), device.Event()
            a.record(s)
            for _ in range(steps):
                once()
            b.record(s)
            stream.synchronize()
            ms = a.elapsed_ms(b) / steps
            got = [bufs[k].download() for k in ("row_sum", "row_var", "full_sum", "full_var")]
            LD = np.longdouble

            def rel(g, e):
                return float(f"{float(np.max(np.abs(np.asarray(g, LD) - e) / np.abs(e))):.3e}")
            entry[mode] = {"ms": round(ms, 5), "GB/s_one_read": round(x.nbytes / (ms * 1e-3) / 1e9, 1),
                           "row_sum_max_rel_err": rel(got[0], ex[0]), "row_var_max_rel_err": rel(got[1], ex[1]),
                           "full_sum_rel_err": rel(got[2][0], ex[2]), "full_var_rel_err": rel(got[3][0], ex[3])}
        entry["fields"] = "row and whole-array sum, var (ddof 0)"
        out[suf] = entry
    return out


def fused_errors(x, keep) -> dict:
    """Max relative error of the fused call's sums and variances (rows and whole array) against
    y = (x * 1.5) + (-0.25) rounded in T, summed in 80-bit long double (pairwise, so its own error
    is ~2^-60 relative, far below the kernels' bounds); max / argmax must be exact."""
    dt = x.dtype.type
    y = (x * dt(1.5)).astype(dt) + dt(-0.25)
    y = y.astype(dt)
    got = [a.download() for a in keep]  # row/full for sum, max, argmax, var
    LD = np.longdouble
    yl = y.astype(LD)
    n = y.shape[1]
    row_sum = yl.sum(axis=1)
    row_mean = row_sum / n
    row_var = ((yl - row_mean[:, None]) ** 2).sum(axis=1) / n
    full_sum = row_sum.sum()
    full_mean = full_sum / y.size
    full_var = ((yl - full_mean) ** 2).sum() / y.size
    del yl

    def rel(a, b):
        return float(np.max(np.abs(np.asarray(a, LD) - b) / np.abs(b)))

    rmax, rarg = y.max(axis=1), y.argmax(axis=1)
    return {"row_sum_max_rel_err": float(f"{rel(got[0], row_sum):.3e}"),
            "full_sum_rel_err": float(f"{rel(got[1][0], full_sum):.3e}"),
            "row_var_max_rel_err": float(f"{rel(got[6], row_var):.3e}"),
            "full_var_rel_err": float(f"{rel(got[7][0], full_var):.3e}"),
            "row_max_exact": bool(np.array_equal(got[2], rmax)),
            "row_argmax_exact": bool(np.array_equal(got[4], rarg.astype(np.uint64))),
            "full_argmax_exact": bool(int(got[5][0]) == int(np.argmax(y))),
            "unit_roundoff": float(np.finfo(dt).eps / 2)}


# ---------------------------------------------------------------------------------------------
# split-precision GEMMs

def _split(x):
    c = x * 134217729.0  # Veltkamp split for FP64 (2^27 + 1)
    hi = c - (c - x)
    return hi, x - hi


def _exact_dot(a: np.ndarray, b: np.ndarray) -> float:
    """Correctly rounded sum_k a_k b_k of FP64 vectors: Dekker's exact products + math.fsum."""
    ah, al = _split(a)
    bh, bl = _split(b)
    p = a * b
    e = al * bl - (((p - ah * bh) - al * bh) - ah * bl)
    return math.fsum(np.concatenate([p, e]).tolist())


def _sample_refs(A, B, idx):
    """C*_ij (correctly rounded) and sum_k |A_ik||B_kj| for the sampled entries, computed once."""
    exact, scale = np.empty(len(idx)), np.empty(len(idx))
    for t, (i, j) in enumerate(idx):
        a = A[i].astype(np.float64)
        b = B[:, j].astype(np.float64)
        exact[t] = _exact_dot(a, b)
        scale[t] = float(np.abs(a) @ np.abs(b))
    return exact, scale


def _sampled_error(Cm, idx, refs) -> float:
    """max over sampled (i, j) of |C_ij - C*_ij| / sum_k |A_ik||B_kj| (the GEMM error-bound scale)."""
    exact, scale = refs
    rows = np.array([i for i, _ in idx])
    cols = np.array([j for _, j in idx])
    return float((np.abs(Cm[rows, cols].astype(np.float64) - exact) / scale).max())


_ROCBLAS = None


def _rocblas():
    """rocBLAS from the ROCm install (the same HIP runtime as the library, so it runs on the
    library's device buffers and stream): a handle, or None when it cannot be loaded."""
    global _ROCBLAS
    if _ROCBLAS is None:
        try:
            lib = C.CDLL("/opt/rocm/lib/librocblas.so")
            h = C.c_void_p()
            if lib.rocblas_create_handle(C.byref(h)) != 0:
                raise OSError("rocblas_create_handle failed")
            for fn, ct in (("rocblas_sgemm", C.c_float), ("rocblas_dgemm", C.c_double)):
                getattr(lib, fn).argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(ct),
                                             C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(ct), C.c_void_p, C.c_int]
            lib.rocblas_set_stream.argtypes = [C.c_void_p, C.c_void_p]
            _ROCBLAS = (lib, h)
        except OSError as e:
            log(f"library GEMM skipped: {e}")
            _ROCBLAS = False
    return _ROCBLAS or None


def _library_gemm(n, dt, pa, pb, pc, s):
    """C = A B (row-major n x n) by rocBLAS's sgemm/dgemm, as the column-major C^T = B^T A^T:
    a callable for the timing loop, or None when rocBLAS is unavailable."""
    rb = _rocblas()
    if rb is None:
        return None
    lib, h = rb
    if lib.rocblas_set_stream(h, C.c_void_p(s)) != 0:
        log("library GEMM skipped: rocblas_set_stream failed")
        return None
    ct = C.c_float if dt == np.float32 else C.c_double
    fn = lib.rocblas_sgemm if dt == np.float32 else lib.rocblas_dgemm
    one, zero = ct(1.0), ct(0.0)
    none = 111  # rocblas_operation_none

    def run():
        st = fn(h, none, none, n, n, n, C.byref(one), pb, n, pa, n, C.byref(zero), pc, n)
        if st != 0:
            raise RuntimeError(f"rocBLAS gemm returned status {st}")
    return run


def bench_gemm(device, rng, call, s, stream):
    def timeit(f, reps):
        f()
        stream.synchronize()
        a, b = device.Event(), device.Event()
        a.record(s)
        for _ in range(reps):
            f()
        b.record(s)
        stream.synchronize()
        return a.elapsed_ms(b) / reps

    out = {}
    pick = np.random.default_rng(11)
    cases = [("f32_8192", np.float32, 8192), ("f64_4096", np.float64, 4096), ("f64_8192", np.float64, 8192)]
    for tag, dt, n in cases:
        t0 = time.time()
        A = rng.scaled_pow2(rng.BASE_SEED + 40 + n, n * n, dtype=dt).reshape(n, n)
        B = rng.scaled_pow2(rng.BASE_SEED + 41 + n, n * n, dtype=dt).reshape(n, n)
        dA, dB = device.DeviceArray.upload(A), device.DeviceArray.upload(B)
        dC = device.DeviceArray((n, n), dt)
        idx = [(int(pick.integers(n)), int(pick.integers(n))) for _ in range(SAMPLES)]
        refs = _sample_refs(A, B, idx)
        flop = 2.0 * n ** 3
        pa, pb, pc = C.c_void_p(dA.ptr), C.c_void_p(dB.ptr), C.c_void_p(dC.ptr)
        if dt == np.float32:
            runs = [("native_fp32", lambda: call("nrk_dev_gemm_f32", n, n, n, pa, pb, pc, C.c_void_p(s)), 1)]
            for p in (3, 6):
                runs.append((f"bf16x{p}", (lambda p=p: call("nrk_dev_gemm_f32_bf16x", n, n, n, pa, pb, pc, p, C.c_void_p(s))), p))
        else:
            runs = [("native_fp64", lambda: call("nrk_dev_gemm_f64", n, n, n, pa, pb, pc, C.c_void_p(s)), 1)]
            for mo in (13, 14, 15, 16):
                runs.append((f"ozaki{mo}", (lambda mo=mo: call("nrk_dev_gemm_f64_ozaki", n, n, n, pa, pb, pc, mo, C.c_void_p(s))), mo))
        res = {}
        for name, f, products in runs:
            ms = timeit(f, 3)
            err = _sampled_error(dC.download(), idx, refs)
            res[name] = {"ms": round(ms, 3), "TFLOP/s": round(flop / ms / 1e9, 2), "max_rel_err": float(f"{err:.3e}"),
                         "mfma_products": products}
        lib = _library_gemm(n, dt, pa, pb, pc, s)
        libname = "library_fp32" if dt == np.float32 else "library_fp64"
        if lib is not None:
            ms = timeit(lib, 3)
            err = _sampled_error(dC.download(), idx, refs)
            res[libname] = {"ms": round(ms, 3), "TFLOP/s": round(flop / ms / 1e9, 2), "max_rel_err": float(f"{err:.3e}"),
                            "what": "rocBLAS sgemm/dgemm (ROCm install) on the same device operands"}
        natives = [runs[0][0]] + ([libname] if lib is not None else [])
        native = res[runs[0][0]]
        # a split engine qualifies at the error of the own native kernel; the speed-up is taken
        # against the faster of the own native kernel and the installed library
        same_err = [k for k in res if k not in natives and res[k]["max_rel_err"] <= native["max_rel_err"]]
        best = min(same_err, key=lambda k: res[k]["ms"]) if same_err else None
        fastest_native = min(natives, key=lambda k: res[k]["ms"])
        res["best_split_at_native_err"] = best
        res["native_reference"] = fastest_native
        res["speedup_vs_native"] = round(res[fastest_native]["ms"] / res[best]["ms"], 3) if best else None
        out[tag] = res
        log(f"gemm {tag}: " + ", ".join(f"{k} {v['ms']}ms err {v['max_rel_err']}" for k, v in res.items() if isinstance(v, dict))
            + f"  ({time.time() - t0:.1f}s)")
        del dA, dB, dC, A, B
    out["error_measure"] = (f"max over {SAMPLES} sampled entries of |C - C*| / sum_k |A_ik||B_kj|, C* = correctly rounded "
                            "dot product (Dekker products + fsum); inputs sign*2^e*u, e in [-10,10], u in [1,2)")
    return out


# ---------------------------------------------------------------------------------------------
# elementwise ufuncs (ufuncs.rs exp / log / tanh ...) over the resident 8192x8192 inputs

def bench_ufuncs(device, kernels, call, s, stream, d32, d64):
    t0 = time.time()
    n = ROWS * COLS
    out = {"workload": f"y = f(x) over x ({ROWS},{COLS}) f32 and f64, N(0,1); bytes = read x + write y",
           "hbm_peak_GB/s": HBM_PEAK_GBS}
    for dt, dx, tag in ((np.float64, d64, "f64"), (np.float32, d32, "f32")):
        dy = device.DeviceArray((n,), dt)
        fn = f"nrk_dev_map_{tag}"
        row = {}
        isz = np.dtype(dt).itemsize

        def timed(name, f, nbytes):
            f()
            stream.synchronize()
            a, b = device.Event(), device.Event()
            a.record(s)
            for _ in range(10):
                f()
            b.record(s)
            stream.synchronize()
            ms = a.elapsed_ms(b) / 10
            gbs = nbytes / ms / 1e6
            row[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 3)}

        for name, steps in (("copy", ()), ("exp", (("exp",),)), ("log", (("abs",), ("log",))), ("tanh", (("tanh",),)),
                            ("sigmoid", (("neg",), ("exp",), ("add", 1.0), ("rdiv", 1.0))), ("floor", (("floor",),)),
                            ("sin", (("sin",),)), ("pow2.5", (("abs",), ("pow", 2.5))), ("gelu", (("gelu",),)),
                            ("erf", (("erf",),))):
            ops, k = kernels.chain(*steps)
            timed(name, lambda: call(fn, C.cast(ops, C.c_void_p), k, C.c_void_p(dx.ptr), C.c_void_p(dy.ptr), C.c_size_t(n),
                                     C.c_void_p(s)), 2 * n * isz)
        # two- and three-operand streams, a dot product and softmax (bytes: each input read once,
        # the output written once; softmax's own passes re-read x and y)
        dout = device.DeviceArray((16,), dt)
        timed("dot", lambda: call(f"nrk_dev_dot_{tag}", C.c_void_p(dx.ptr), C.c_void_p(dy.ptr), C.c_size_t(n),
                                  C.c_void_p(dout.ptr), C.c_void_p(s)), 2 * n * isz)
        half = n // 2
        timed("fma", lambda: call(f"nrk_dev_fma_{tag}", C.c_void_p(dx.ptr), C.c_void_p(dx.ptr + half * isz),
                                  C.c_void_p(dy.ptr), C.c_void_p(dy.ptr), C.c_size_t(half), C.c_void_p(s)), 4 * half * isz)
        timed("softmax", lambda: call(f"nrk_dev_softmax_{tag}", C.c_void_p(dx.ptr), C.c_void_p(dy.ptr), C.c_size_t(n),
                                      C.c_void_p(s)), 2 * n * isz)
        timed("cumsum", lambda: call(f"nrk_dev_accumulate_{tag}", C.c_int32(0), C.c_void_p(dx.ptr), C.c_size_t(1),
                                     C.c_size_t(n), C.c_void_p(dy.ptr), C.c_void_p(s)), 2 * n * isz)
        del dout
        out[tag] = row
        log(f"ufuncs {tag}: " + ", ".join(f"{k} {v['ms']:.3f} ms ({v['GB/s']:.0f} GB/s)" for k, v in row.items()))
        del dy
    out["seconds"] = round(time.time() - t0, 1)
    return out


# ---------------------------------------------------------------------------------------------
# the fused reduction on shapes other than the headline's: the first rows * cols elements of the
# headline's x (2^26 elements), the same y = 1.5 x - 0.25 and sum / max / argmax / var

SHAPES = (  # (rows, cols, row results, whole-array results)
    (8191, 8191, True, True),      # long rows off the 16-byte grid (general kernel)
    (8192, 8188, True, False),     # long aligned rows, not a multiple of the segment (general kernel)
    (512, 131072, True, False),    # few long rows (general kernel, rows split)
    (1 << 20, 64, True, False),    # short rows (short-row kernel, 4 lanes per row)
    (1 << 20, 64, True, True),
    (262144, 256, True, False),    # 16 lanes per row
    (100000, 671, True, False),    # 32 lanes per row, rows off the grid
    (1 << 23, 8, True, False),     # one lane per row
    (1, (1 << 26) - 1, False, True),  # whole array, not a whole number of chunks (lean + tail)
    (1, 1000003, False, True),     # whole array, under 1024 chunks
    (1, 1000, False, True),        # whole array, one workgroup
)


# (rows, cols) reduced along axis 0: 8192 / 64 outputs by column tiles, 2^20 outputs one thread each
AXIS0_SHAPES = ((8192, 8192), (1 << 20, 64), (64, 1 << 20))
AXIS0_OPS = (("sum", 0), ("max", 4), ("var", 2))


def bench_shapes(device, kernels, call, ReduceOutputs, s, stream, d32, d64):
    t0 = time.time()
    ops, n_ops = kernels.chain(("mul", 1.5), ("add", -0.25))
    ops_p = C.cast(ops, C.c_void_p)
    out = {"workload": "fused y = 1.5x - 0.25 -> sum/max/argmax/var, row and/or whole-array results, on the first "
                       "rows*cols elements of the headline x; and sum, max and var along axis 0 (nrk_dev_reduce_axis) of "
                       "the same data; GB/s counts the input bytes once",
           "hbm_peak_GB/s": HBM_PEAK_GBS}
    for dt, dx, tag in ((np.float32, d32, "f32"), (np.float64, d64, "f64")):
        fn = f"nrk_dev_map_reduce_{tag}"
        row = {}
        for rows, cols, want_rows, want_full in SHAPES:
            ro, fo, keep = ReduceOutputs(), ReduceOutputs(), []
            for f in ("sum", "max", "argmax", "var"):
                odt = np.uint64 if f == "argmax" else dt
                if want_rows:
                    keep.append(device.DeviceArray((rows,), odt))
                    setattr(ro, f, keep[-1].ptr)
                if want_full:
                    keep.append(device.DeviceArray((1,), odt))
                    setattr(fo, f, keep[-1].ptr)
            args_ = (ops_p, n_ops, C.c_void_p(dx.ptr), rows, cols, 0, C.byref(ro) if want_rows else None,
                     C.byref(fo) if want_full else None, C.c_void_p(s))
            f = lambda: call(fn, *args_)
            f()
            stream.synchronize()
            a, b = device.Event(), device.Event()
            a.record(s)
            for _ in range(10):
                f()
            b.record(s)
            stream.synchronize()
            ms = a.elapsed_ms(b) / 10
            nbytes = rows * cols * np.dtype(dt).itemsize
            key = f"{rows}x{cols}" + (":rows" if want_rows else "") + (":full" if want_full else "")
            row[key] = {"us": round(ms * 1e3, 2), "GB/s": round(nbytes / ms / 1e6, 1),
                        "frac_hbm": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3)}
            del keep
        # reductions along axis 0 (a strided axis: outer = 1, len = rows, inner = cols) of the same data
        dsum = device.DeviceArray((1 << 20,), dt)
        for rows, cols in AXIS0_SHAPES:
            for name, op in AXIS0_OPS:
                f = lambda: call(f"nrk_dev_reduce_axis_{tag}", C.c_int32(op), C.c_void_p(dx.ptr), C.c_size_t(1),
                                 C.c_size_t(rows), C.c_size_t(cols), C.c_size_t(0), C.c_void_p(dsum.ptr), C.c_void_p(s))
                f()
                stream.synchronize()
                a, b = device.Event(), device.Event()
                a.record(s)
                for _ in range(10):
                    f()
                b.record(s)
                stream.synchronize()
                ms = a.elapsed_ms(b) / 10
                nbytes = rows * cols * np.dtype(dt).itemsize
                row[f"{rows}x{cols}:{name}_axis0"] = {"us": round(ms * 1e3, 2), "GB/s": round(nbytes / ms / 1e6, 1),
                                                      "frac_hbm": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3)}
        del dsum
        out[tag] = row
        log(f"general shapes {tag}: " + ", ".join(f"{k} {v['us']:.1f} us ({v['frac_hbm']:.2f})" for k, v in row.items()))
    out["seconds"] = round(time.time() - t0, 1)
    return out


# ---------------------------------------------------------------------------------------------
# uniform batches (batched_gemm / matmul batch axes): 64 panels of the 512^3 plain-GEMM shape

def bench_batched(device, rng, call, s, stream):
    t0 = time.time()
    batch, n = 64, 512
    flop = 2.0 * batch * n ** 3
    out = {"workload": f"{batch} panels of {n}x{n}x{n}, U[-1,1], contiguous [batch, m, k] x [batch, k, n]"}
    for dt, peak, tag in ((np.float64, 78.6, "f64"), (np.float32, 157.3, "f32")):
        A = rng.uniform(rng.BASE_SEED + 60, batch * n * n).astype(dt)
        B = rng.uniform(rng.BASE_SEED + 61, batch * n * n).astype(dt)
        dA, dB = device.DeviceArray.upload(A), device.DeviceArray.upload(B)
        dC = device.DeviceArray((batch * n * n,), dt)
        isz = np.dtype(dt).itemsize
        fb, f1 = f"nrk_dev_batched_gemm_{tag}", f"nrk_dev_gemm_{tag}"
        pa, pb, pc = C.c_void_p(dA.ptr), C.c_void_p(dB.ptr), C.c_void_p(dC.ptr)

        def timeit(f, reps=5):
            f()
            stream.synchronize()
            a, b = device.Event(), device.Event()
            a.record(s)
            for _ in range(reps):
                f()
            b.record(s)
            stream.synchronize()
            return a.elapsed_ms(b) / reps

        one = timeit(lambda: call(fb, batch, n, n, n, pa, n * n, pb, n * n, pc, C.c_void_p(s)))
        bcast = timeit(lambda: call(fb, batch, n, n, n, pa, n * n, pb, 0, pc, C.c_void_p(s)))

        def loop():
            for t in range(batch):
                call(f1, n, n, n, C.c_void_p(dA.ptr + t * n * n * isz), C.c_void_p(dB.ptr + t * n * n * isz),
                     C.c_void_p(dC.ptr + t * n * n * isz), C.c_void_p(s))
        per_call = timeit(loop)
        single = timeit(lambda: call(f1, n, n, n, pa, pb, pc, C.c_void_p(s)), 20)
        out[tag] = {"batched_ms": round(one, 4), "TFLOP/s": round(flop / one / 1e9, 2),
                    "frac_of_mfma_peak": round(flop / one / 1e9 / peak, 3), "peak_TFLOP/s": peak,
                    "broadcast_b_ms": round(bcast, 4), "per_panel_loop_ms": round(per_call, 4),
                    "speedup_vs_loop": round(per_call / one, 2),
                    "single_512_ms": round(single, 4), "single_512_TFLOP/s": round(2.0 * n ** 3 / single / 1e9, 2)}
        log(f"batched {tag}: one launch {one:.3f} ms ({flop / one / 1e9:.1f} TFLOP/s), B broadcast {bcast:.3f} ms, "
            f"per-panel loop {per_call:.3f} ms, single 512^3 {single * 1e3:.1f} us")
        del dA, dB, dC
    out["seconds"] = round(time.time() - t0, 1)
    return out


# ---------------------------------------------------------------------------------------------
# ragged batch: 4096 bf16 GEMMs + 65536 f32 row sums in one persistent launch

def bench_selector() -> dict:
    """The learned tile selector on a held-out stream: 20000 GEMM + 4000 row-sum training shapes
    tuned exhaustively, the two cost models trained on the GPU, then 100000 held-out calls (80%
    GEMMs of m, n, k log-uniform in [16, 512] over f64 / f32 / bf16, 20% row sums) chosen by the
    models with timing fallback, against exhaustive search and the built-in rules; plus the shipped
    model (numrs_amd/selector_model.npz, the one the GEMM dispatch consults) on the same calls."""
    from numrs_amd import selector as S

    t0 = time.time()
    res = S.run_evaluation(20000, 4000, 100000, log=log)
    st = res["stream"]
    out = {"train_shapes": res["train_shapes"], "eval_calls": res["eval_calls"], "margin_log2": res["margin_log2"],
           "stream_regret": round(st["regret_model_with_fallback"], 5), "stream_regret_builtin_rules": round(st["regret_heuristic"], 5),
           "fallback_rate": round(st["fallback_rate"], 5), "stream_ms": {k: round(st[k], 3) for k in
                                                                          ("exhaustive_ms", "selected_ms", "heuristic_ms")},
           "gemm": {k: round(res["gemm"][k], 5) for k in ("regret_model", "regret_model_with_fallback", "fallback_rate",
                                                           "regret_heuristic", "top1_agreement")},
           "rows": {k: round(res["rows"][k], 5) for k in ("regret_model", "regret_model_with_fallback", "fallback_rate",
                                                           "regret_heuristic")},
           "mlp_inferences_per_s": round(res["mlp_inference_per_s"]), "seconds": round(time.time() - t0, 1)}
    if "shipped_model_gemm" in res:
        out["shipped_model_gemm"] = {k: round(res["shipped_model_gemm"][k], 5)
                                     for k in ("regret_model", "regret_model_with_fallback", "fallback_rate", "top1_agreement")}
    log("selector: " + json.dumps(out))
    return out


def bench_ragged(device, rng, call, s, stream):
    from numrs_amd.grouped import Graph, GroupedPlan, to_bf16_bits

    t0 = time.time()
    seed = rng.BASE_SEED + 5000

    def log_uniform_int(sd, n, lo, hi):
        u = rng.uniform(sd, n, 0.0, 1.0)
        return np.floor(np.exp(np.log(lo) + u * (np.log(hi + 1) - np.log(lo)))).astype(np.int64).clip(lo, hi)

    n_gemm, n_rows = 4096, 65536
    m, n, k = (log_uniform_int(seed + i, n_gemm, 16, 512) for i in range(3))
    a_off = np.concatenate([[0], np.cumsum(m * k)])
    b_off = np.concatenate([[0], np.cumsum(k * n)])
    c_off = np.concatenate([[0], np.cumsum(m * n)])
    lens = log_uniform_int(seed + 7, n_rows, 32, 65536)
    r_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    flops = float(np.sum(2.0 * m * n * k))
    row_bytes = int(r_off[-1]) * 4

    # operands: a 16 Mi-element block of N(0,1) values repeated over each arena (timing only)
    block = rng.normal(seed + 9, 1 << 24, np.float32)
    bblock = to_bf16_bits(block)

    def arena(total, src):
        buf = device.DeviceArray((int(total),), src.dtype)
        blk = device.DeviceArray.upload(src)
        done = 0
        while done < total:
            cnt = min(len(src), int(total) - done)
            call("nrk_memcpy_dtod", C.c_void_p(buf.ptr + done * src.itemsize), C.c_void_p(blk.ptr), cnt * src.itemsize, C.c_void_p(s))
            done += cnt
        stream.synchronize()
        return buf

    dA, dB = arena(a_off[-1], bblock), arena(b_off[-1], bblock)
    dC = device.DeviceArray((int(c_off[-1]),), np.float32)
    dX = arena(r_off[-1], block)
    dR = device.DeviceArray((n_rows,), np.float32)
    descs = [(int(m[g]), int(n[g]), int(k[g]), dA.ptr + 2 * int(a_off[g]), dB.ptr + 2 * int(b_off[g]), dC.ptr + 4 * int(c_off[g]))
             for g in range(n_gemm)]
    plan = GroupedPlan(descs, dX.ptr, r_off, dR.ptr)
    info = plan.info()
    graph = Graph(s, lambda: plan.run(s))

    def timed(f, reps):
        f()
        stream.synchronize()
        a, b = device.Event(), device.Event()
        a.record(s)
        for _ in range(reps):
            f()
        b.record(s)
        stream.synchronize()
        return a.elapsed_ms(b) / reps

    ms_graph = timed(graph.launch, 5)
    ms_plan = timed(lambda: plan.run(s), 5)

    from numrs_amd._lib import GemmDesc

    # the per-call baseline, issued by the library's own C loop (one launch per GEMM and per row)
    darr = (GemmDesc * n_gemm)(*[GemmDesc(*d[:3], 0, *d[3:]) for d in descs])
    roff = np.ascontiguousarray(r_off, np.int64)

    def loop():
        call("nrk_dev_ragged_per_call", n_gemm, C.cast(darr, C.c_void_p), n_rows, C.c_void_p(dX.ptr),
             roff.ctypes.data_as(C.c_void_p), C.c_void_p(dR.ptr), C.c_void_p(s))

    ms_loop = timed(loop, 2)
    graph.close()
    plan.close()
    log(f"ragged: graph {ms_graph:.3f} ms, plan {ms_plan:.3f} ms, per-call loop {ms_loop:.3f} ms ({time.time() - t0:.1f}s)")
    return {"gemms": n_gemm, "rows": n_rows, "gemm_flop": flops, "row_bytes": row_bytes, "plan": info,
            "graph_ms": round(ms_graph, 4), "launch_ms": round(ms_plan, 4), "per_call_loop_ms": round(ms_loop, 3),
            "per_call_loop": "nrk_dev_ragged_per_call: one nrk_dev_gemm_bf16 launch per GEMM and one nrk_dev_sum_f32 "
                             "launch per row, issued from C on one stream",
            "row_floor_ms_at_8TBps": round(row_bytes / 8e12 * 1e3, 4),
            "speedup_vs_loop": round(ms_loop / ms_graph, 2),
            "graph_gemm_TFLOP/s_if_alone": round(flops / ms_graph / 1e9, 2),
            "graph_row_GB/s_if_alone": round(row_bytes / ms_graph / 1e6, 1),
            "data": "m,n,k log-uniform-int [16,512]; row lengths log-uniform-int [32,65536]; operand values: one seeded "
                    "N(0,1) block of 2^24 elements repeated (timing only)"}


# ---------------------------------------------------------------------------------------------

def cpu_baseline(x32, d64, xi, step_bytes):
    """The oracle's restatement of the reference's unfused CPU sequence, timed on this host:
    minimum over 3 samples of one full step."""
    import oracle

    x64 = d64.download()
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(threads)
    best = None
    for _ in range(3):
        t = time.perf_counter()
        for x in (x32, x64):
            y = oracle.ref_map([2], [1.5], x)        # &x * 1.5   (operations.rs:462, one map pass)
            y = oracle.ref_map([0], [-0.25], y)      # ... + -0.25 (operations.rs:448, second pass)
            for op in (0, 4, 6):                     # math::sum / max / argmax along axis 1
                oracle.ref_axis(op, y, ROWS, COLS, 1)
            oracle.ref_axis(2, y, ROWS, COLS, 1, 0)  # math::var along axis 1
            oracle.ref("sum", y)                     # kernels::reduce::sum (chunked, threads)
            oracle.ref("max", y)
            oracle.ref("argmax", y)
            oracle.ref("var", y, 0)
        oracle.sum_i64(xi)                           # Array::<i64>::sum (serial fold)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(step_bytes / best / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "seconds": round(best, 3), "nproc": os.cpu_count(), "cpu": model,
            "sample": "one full step of the same workload (both dtypes + the i64 sum) through the C restatement, "
                      "minimum of 3; axis folds single-threaded as in the reference, chunked full reductions on the pool"}


if __name__ == "__main__":
    raise SystemExit(main())
