"""The selector on the dispatch path (no reference counterpart: it replaces the compile-time tiers
of kernels/mod.rs:63-106 and kernels/gemm.rs:290-312).

A dispatch with no forced tile takes its configuration from a recorded decision, else from the
registered cost model inside its trained shapes (timed when the model is unsure), else from the
built-in rule. Every configuration runs the same k order per output, so whatever the selector
picks, the result bits equal the built-in rule's: checked bit for bit for every configuration of
every family that consults it."""
import ctypes as C

import numpy as np
import pytest

from numrs_amd import device, rng
from numrs_amd import selector as S
from numrs_amd._lib import NrkError, call

pytestmark = pytest.mark.gpu

M, N, K = 300, 200, 150


@pytest.fixture
def clean(gpu):
    call("nrk_selector_clear", -1)
    yield
    call("nrk_selector_clear", -1)
    S.install()


def _bits(x):
    return x.view(np.uint64 if x.dtype == np.float64 else np.uint32)


def _operands(dt, seed):
    a = rng.normal(rng.BASE_SEED + seed, M * K, dt).reshape(M, K)
    b = rng.normal(rng.BASE_SEED + seed + 1, K * N, dt).reshape(K, N)
    return device.DeviceArray.upload(a), device.DeviceArray.upload(b), device.DeviceArray((M, N), dt)


def _runner(fam):
    st = device.Stream()
    if fam in (S.F64, S.F32, S.F32_BF16X3, S.F32_BF16X6, S.F64_OZAKI):
        dt = np.float64 if fam in (S.F64, S.F64_OZAKI) else np.float32
        dA, dB, dC = _operands(dt, 900 + fam)
        fn, extra = {S.F64: ("nrk_dev_gemm_f64", ()), S.F32: ("nrk_dev_gemm_f32", ()),
                     S.F32_BF16X3: ("nrk_dev_gemm_f32_bf16x", (C.c_int32(3),)),
                     S.F32_BF16X6: ("nrk_dev_gemm_f32_bf16x", (C.c_int32(6),)),
                     S.F64_OZAKI: ("nrk_dev_gemm_f64_ozaki", (C.c_int32(14),))}[fam]

        def run():
            call(fn, M, K, N, C.c_void_p(dA.ptr), C.c_void_p(dB.ptr), C.c_void_p(dC.ptr), *extra, C.c_void_p(st.handle))
            st.synchronize()
            return dC.download()
        return run, (dA, dB, dC, st)
    # bf16 bits in, f32 out
    r = np.random.default_rng(fam)
    a = (r.integers(0x3c00, 0x4000, M * K) | 0).astype(np.uint16)
    b = (r.integers(0x3c00, 0x4000, K * N) | 0).astype(np.uint16)
    dA, dB = device.DeviceArray.upload(a), device.DeviceArray.upload(b)
    dC = device.DeviceArray((M, N), np.float32)

    class Desc(C.Structure):
        _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32), ("reserved", C.c_int32), ("a", C.c_void_p),
                    ("b", C.c_void_p), ("c", C.c_void_p)]
    d = Desc(M, N, K, 0, dA.ptr, dB.ptr, dC.ptr)

    def run():
        call("nrk_dev_gemm_bf16", C.byref(d), C.c_void_p(st.handle))
        st.synchronize()
        return dC.download()
    return run, (dA, dB, dC, st, d)


@pytest.mark.parametrize("fam", [S.F64, S.F32, S.BF16, S.F32_BF16X3, S.F32_BF16X6, S.F64_OZAKI])
def test_recorded_decision_drives_dispatch_same_bits(clean, fam):
    run, keep = _runner(fam)
    assert S.choose(fam, M, N, K) == -1  # nothing registered: the built-in rule
    want = run()
    for cfg in range(S.config_count(fam)):
        print(f"family {fam} config {cfg}", flush=True)
        call("nrk_selector_table_put", fam, M, N, K, cfg)
        assert S.choose(fam, M, N, K) == cfg
        before = S.stats(fam)["table"]
        got = run()
        assert S.stats(fam)["table"] == before + 1
        np.testing.assert_array_equal(_bits(got), _bits(want), err_msg=f"config {cfg}")
    assert (fam, M, N, K, S.config_count(fam) - 1) in S.table()


def _constant_model(fam, best, margin):
    """A cost model whose prediction ignores the shape: config `best` cheapest by 1 (log2)."""
    k = S.config_count(fam)
    P = S.call_size("nrk_mlp_param_count", k)
    p = np.zeros(P, np.float32)
    p[P - k:] = 10.0
    p[P - k + best] = 9.0
    z = np.zeros(6, np.float32)
    one = np.ones(6, np.float32)
    lo, hi = np.full(3, 4.0, np.float32), np.full(3, 9.0, np.float32)
    call("nrk_selector_set_model", fam, C.c_float(S.GEMM_FAMILY_DTYPE.get(fam, 0)), p.ctypes.data_as(C.c_void_p), k,
         z.ctypes.data_as(C.c_void_p), one.ctypes.data_as(C.c_void_p), lo.ctypes.data_as(C.c_void_p),
         hi.ctypes.data_as(C.c_void_p), C.c_float(0.0), C.c_float(margin))


@pytest.mark.parametrize("fam", [S.F64, S.F32, S.BF16])
def test_model_drives_dispatch_within_its_shapes(clean, fam):
    run, keep = _runner(fam)
    want = run()
    _constant_model(fam, best=2, margin=0.5)
    assert S.choose(fam, M, N, K) == 2
    got = run()
    np.testing.assert_array_equal(_bits(got), _bits(want))
    assert S.stats(fam)["model"] >= 1
    # outside the trained shapes (log2 d in [4, 9] +- 0.5): the built-in rule
    assert S.choose(fam, 4096, 4096, 4096) == -1
    assert S.choose(fam, 8, 300, 300) == -1


@pytest.mark.parametrize("fam", [S.F64, S.BF16])
def test_unsure_model_falls_back_to_timing(clean, fam):
    """Top-2 predictions closer than the margin: the shape is timed, the fastest configuration
    recorded, and the result bits are still the built-in rule's."""
    run, keep = _runner(fam)
    want = run()
    _constant_model(fam, best=1, margin=2.0)
    got = run()
    np.testing.assert_array_equal(_bits(got), _bits(want))
    assert S.stats(fam)["timing"] == 1
    rec = [e for e in S.table() if e[:4] == (fam, M, N, K)]
    assert len(rec) == 1 and 0 <= rec[0][4] < S.config_count(fam)
    run()
    assert S.stats(fam)["table"] >= 1 and S.stats(fam)["timing"] == 1  # recorded, not timed again


def test_model_argument_checks(clean):
    k = S.config_count(S.F64)
    p = np.zeros(S.call_size("nrk_mlp_param_count", k), np.float32)
    z = np.zeros(6, np.float32)
    with pytest.raises(NrkError):  # a model with the wrong number of outputs
        call("nrk_selector_set_model", S.F64, C.c_float(0), p.ctypes.data_as(C.c_void_p), k + 1, z.ctypes.data_as(C.c_void_p),
             z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), C.c_float(0), C.c_float(0))
    with pytest.raises(NrkError):  # x_scale must be positive
        call("nrk_selector_set_model", S.F64, C.c_float(0), p.ctypes.data_as(C.c_void_p), k, z.ctypes.data_as(C.c_void_p),
             z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), C.c_float(0), C.c_float(0))
    with pytest.raises(NrkError):
        call("nrk_selector_table_put", S.F64, M, N, K, k)


def test_persisted_model_installed(gpu):
    """The shipped model (trained on this GPU's timings) is registered at init for the three GEMM
    dtypes and consulted inside its shapes only."""
    if not S.MODEL_PATH.exists():
        pytest.skip("no persisted selector model in this tree")
    call("nrk_selector_clear", -1)
    assert S.install() == [S.F64, S.F32, S.BF16]
    for fam in (S.F64, S.F32, S.BF16):
        assert 0 <= S.choose(fam, 256, 256, 256) < S.config_count(fam)
        assert S.choose(fam, 8192, 8192, 8192) == -1
