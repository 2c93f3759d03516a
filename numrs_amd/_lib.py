"""ctypes binding of libnumrs_amd.so (the C ABI declared in include/numrs_amd.h).

Loading fails loudly: there is no CPU fallback anywhere in this package. If the shared object is
missing, build it with `python -m numrs_amd.build`.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnumrs_amd.so"

c_size = C.c_size_t
c_i32 = C.c_int32
vp = C.c_void_p
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)


class ScalarOp(C.Structure):
    """nrk_scalar_op"""
    _fields_ = [("op", C.c_int32), ("reserved", C.c_int32), ("scalar", C.c_double)]


class ReduceOutputs(C.Structure):
    """nrk_reduce_outputs (device pointers, NULL to skip)"""
    _fields_ = [("sum", vp), ("mean", vp), ("var", vp), ("min", vp), ("max", vp), ("argmin", vp), ("argmax", vp)]


class GemmDesc(C.Structure):
    """nrk_gemm_desc: C (m x n f32) := A (m x k bf16) * B (k x n bf16), device pointers"""
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32), ("reserved", C.c_int32),
                ("a", vp), ("b", vp), ("c", vp)]


# name -> argtypes (all return nrk_status = int32)
_SIGS: dict[str, list] = {
    "nrk_init": [C.c_int],
    "nrk_device_cu_count": [C.POINTER(c_i32)],
    "nrk_device_malloc": [C.POINTER(vp), c_size],
    "nrk_device_free": [vp],
    "nrk_memcpy_htod": [vp, vp, c_size, vp],
    "nrk_memcpy_dtoh": [vp, vp, c_size, vp],
    "nrk_memcpy_dtod": [vp, vp, c_size, vp],
    "nrk_memset_zero": [vp, c_size, vp],
    "nrk_stream_create": [C.POINTER(vp)],
    "nrk_stream_destroy": [vp],
    "nrk_stream_synchronize": [vp],
    "nrk_device_synchronize": [],
    "nrk_event_create": [C.POINTER(vp)],
    "nrk_event_destroy": [vp],
    "nrk_event_record": [vp, vp],
    "nrk_stream_wait_event": [vp, vp],
    "nrk_event_elapsed_ms": [C.POINTER(C.c_float), vp, vp],
    # elementwise
    "nrk_map_f64": [vp, c_i32, vp, vp, c_size],
    "nrk_map_f32": [vp, c_i32, vp, vp, c_size],
    "nrk_dev_map_f64": [vp, c_i32, vp, vp, c_size, vp],
    "nrk_dev_map_f32": [vp, c_i32, vp, vp, c_size, vp],
    "nrk_zip_f64": [c_i32, vp, vp, vp, c_size],
    "nrk_zip_f32": [c_i32, vp, vp, vp, c_size],
    "nrk_dev_zip_f64": [c_i32, vp, vp, vp, c_size, vp],
    "nrk_dev_zip_f32": [c_i32, vp, vp, vp, c_size, vp],
    # flat reductions (host)
    "nrk_sum_f64": [vp, c_size, vp], "nrk_sum_f32": [vp, c_size, vp],
    "nrk_mean_f64": [vp, c_size, vp], "nrk_mean_f32": [vp, c_size, vp],
    "nrk_sum_sq_dev_f64": [vp, c_size, C.c_double, vp], "nrk_sum_sq_dev_f32": [vp, c_size, C.c_float, vp],
    "nrk_var_f64": [vp, c_size, c_size, vp], "nrk_var_f32": [vp, c_size, c_size, vp],
    "nrk_dot_f64": [vp, vp, c_size, vp], "nrk_dot_f32": [vp, vp, c_size, vp],
    "nrk_fma_f64": [vp, vp, vp, vp, c_size], "nrk_fma_f32": [vp, vp, vp, vp, c_size],
    "nrk_accumulate_f64": [c_i32, vp, c_size, c_size, vp], "nrk_accumulate_f32": [c_i32, vp, c_size, c_size, vp],
    "nrk_at_f64": [c_i32, vp, c_size, c_size, vp, c_size, vp], "nrk_at_f32": [c_i32, vp, c_size, c_size, vp, c_size, vp],
    "nrk_dev_at_f64": [c_i32, vp, c_size, c_size, vp, c_size, vp, vp],
    "nrk_dev_at_f32": [c_i32, vp, c_size, c_size, vp, c_size, vp, vp],
    "nrk_reduceat_f64": [c_i32, vp, c_size, c_size, vp, c_size, vp], "nrk_reduceat_f32": [c_i32, vp, c_size, c_size, vp, c_size, vp],
    "nrk_dev_reduceat_f64": [c_i32, vp, c_size, c_size, vp, c_size, vp, vp],
    "nrk_dev_reduceat_f32": [c_i32, vp, c_size, c_size, vp, c_size, vp, vp],
    "nrk_outer_f64": [c_i32, vp, c_size, vp, c_size, vp], "nrk_outer_f32": [c_i32, vp, c_size, vp, c_size, vp],
    "nrk_dev_outer_f64": [c_i32, vp, c_size, vp, c_size, vp, vp], "nrk_dev_outer_f32": [c_i32, vp, c_size, vp, c_size, vp, vp],
    "nrk_where_f64": [c_i32, vp, vp, vp, vp, c_size], "nrk_where_f32": [c_i32, vp, vp, vp, vp, c_size],
    "nrk_dev_where_f64": [c_i32, vp, vp, vp, vp, c_size, vp], "nrk_dev_where_f32": [c_i32, vp, vp, vp, vp, c_size, vp],
    "nrk_dev_accumulate_f64": [c_i32, vp, c_size, c_size, vp, vp], "nrk_dev_accumulate_f32": [c_i32, vp, c_size, c_size, vp, vp],
    "nrk_softmax_f64": [vp, vp, c_size], "nrk_softmax_f32": [vp, vp, c_size],
    "nrk_log_softmax_f64": [vp, vp, c_size], "nrk_log_softmax_f32": [vp, vp, c_size],
    "nrk_dev_softmax_f64": [vp, vp, c_size, vp], "nrk_dev_softmax_f32": [vp, vp, c_size, vp],
    "nrk_dev_log_softmax_f64": [vp, vp, c_size, vp], "nrk_dev_log_softmax_f32": [vp, vp, c_size, vp],
    "nrk_dev_fma_f64": [vp, vp, vp, vp, c_size, vp], "nrk_dev_fma_f32": [vp, vp, vp, vp, c_size, vp],
    "nrk_lerp_f64": [vp, vp, C.c_double, vp, c_size], "nrk_lerp_f32": [vp, vp, C.c_float, vp, c_size],
    "nrk_dev_lerp_f64": [vp, vp, C.c_double, vp, c_size, vp], "nrk_dev_lerp_f32": [vp, vp, C.c_float, vp, c_size, vp],
    "nrk_dev_dot_f64": [vp, vp, c_size, vp, vp], "nrk_dev_dot_f32": [vp, vp, c_size, vp, vp],
    "nrk_norm_l1_f64": [vp, c_size, vp], "nrk_norm_l1_f32": [vp, c_size, vp],
    "nrk_norm_l2_f64": [vp, c_size, vp], "nrk_norm_l2_f32": [vp, c_size, vp],
    "nrk_dev_norm_l1_f64": [vp, c_size, vp, vp], "nrk_dev_norm_l1_f32": [vp, c_size, vp, vp],
    "nrk_dev_norm_l2_f64": [vp, c_size, vp, vp], "nrk_dev_norm_l2_f32": [vp, c_size, vp, vp],
    "nrk_std_f64": [vp, c_size, c_size, vp], "nrk_std_f32": [vp, c_size, c_size, vp],
    "nrk_dev_std_f64": [vp, c_size, c_size, vp, vp], "nrk_dev_std_f32": [vp, c_size, c_size, vp, vp],
    "nrk_min_f64": [vp, c_size, vp], "nrk_min_f32": [vp, c_size, vp],
    "nrk_max_f64": [vp, c_size, vp], "nrk_max_f32": [vp, c_size, vp],
    "nrk_prod_f64": [vp, c_size, vp], "nrk_prod_f32": [vp, c_size, vp],
    "nrk_sum_i64": [vp, c_size, vp],
    "nrk_argmax_f64": [vp, c_size, vp], "nrk_argmax_f32": [vp, c_size, vp],
    "nrk_argmin_f64": [vp, c_size, vp], "nrk_argmin_f32": [vp, c_size, vp],
    # flat reductions (device)
    "nrk_dev_sum_f64": [vp, c_size, vp, vp], "nrk_dev_sum_f32": [vp, c_size, vp, vp],
    "nrk_dev_mean_f64": [vp, c_size, vp, vp], "nrk_dev_mean_f32": [vp, c_size, vp, vp],
    "nrk_dev_var_f64": [vp, c_size, c_size, vp, vp], "nrk_dev_var_f32": [vp, c_size, c_size, vp, vp],
    "nrk_dev_min_f64": [vp, c_size, vp, vp], "nrk_dev_min_f32": [vp, c_size, vp, vp],
    "nrk_dev_max_f64": [vp, c_size, vp, vp], "nrk_dev_max_f32": [vp, c_size, vp, vp],
    "nrk_dev_prod_f64": [vp, c_size, vp, vp], "nrk_dev_prod_f32": [vp, c_size, vp, vp],
    "nrk_dev_sum_i64": [vp, c_size, vp, vp],
    "nrk_dev_argmax_f64": [vp, c_size, vp, vp], "nrk_dev_argmax_f32": [vp, c_size, vp, vp],
    "nrk_dev_argmin_f64": [vp, c_size, vp, vp], "nrk_dev_argmin_f32": [vp, c_size, vp, vp],
    # axis reductions
    "nrk_reduce_axis_f64": [c_i32, vp, c_size, c_size, c_size, c_size, vp],
    "nrk_reduce_axis_f32": [c_i32, vp, c_size, c_size, c_size, c_size, vp],
    "nrk_dev_reduce_axis_f64": [c_i32, vp, c_size, c_size, c_size, c_size, vp, vp],
    "nrk_dev_reduce_axis_f32": [c_i32, vp, c_size, c_size, c_size, c_size, vp, vp],
    # fused
    "nrk_dev_map_reduce_f64": [vp, c_i32, vp, c_size, c_size, c_size, vp, vp, vp],
    "nrk_dev_map_reduce_f64_noalloc": [vp, c_i32, vp, c_size, c_size, c_size, vp, vp, vp],
    "nrk_zip_i64": [c_i32, vp, vp, vp, c_size], "nrk_dev_zip_i64": [c_i32, vp, vp, vp, c_size, vp],
    "nrk_map_i64": [vp, c_i32, vp, vp, c_size], "nrk_dev_map_i64": [vp, c_i32, vp, vp, c_size, vp],
    "nrk_zip_i32": [c_i32, vp, vp, vp, c_size], "nrk_dev_zip_i32": [c_i32, vp, vp, vp, c_size, vp],
    "nrk_map_i32": [vp, c_i32, vp, vp, c_size], "nrk_dev_map_i32": [vp, c_i32, vp, vp, c_size, vp],
    "nrk_zip_u64": [c_i32, vp, vp, vp, c_size], "nrk_dev_zip_u64": [c_i32, vp, vp, vp, c_size, vp],
    "nrk_map_u64": [vp, c_i32, vp, vp, c_size], "nrk_dev_map_u64": [vp, c_i32, vp, vp, c_size, vp],
    "nrk_zip_u32": [c_i32, vp, vp, vp, c_size], "nrk_dev_zip_u32": [c_i32, vp, vp, vp, c_size, vp],
    "nrk_map_u32": [vp, c_i32, vp, vp, c_size], "nrk_dev_map_u32": [vp, c_i32, vp, vp, c_size, vp],
    "nrk_expr_eval_f64": [vp, c_i32, vp, c_i32, vp, c_size], "nrk_expr_eval_f32": [vp, c_i32, vp, c_i32, vp, c_size],
    "nrk_dev_expr_eval_f64": [vp, c_i32, vp, c_i32, vp, c_size, vp],
    "nrk_dev_expr_eval_f32": [vp, c_i32, vp, c_i32, vp, c_size, vp],
    "nrk_dev_expr_reduce_f64": [vp, c_i32, vp, c_i32, c_size, c_size, c_size, vp, vp, vp],
    "nrk_dev_expr_reduce_f32": [vp, c_i32, vp, c_i32, c_size, c_size, c_size, vp, vp, vp],
    "nrk_zip_broadcast": [c_i32, c_i32, vp, c_i32, vp, vp, c_i32, vp, vp],
    "nrk_dev_zip_strided": [c_i32, c_i32, c_i32, vp, vp, vp, vp, c_i32, vp, vp, vp],
    "nrk_gemm_generic_i64": [c_size, c_size, c_size, vp, c_size, vp, c_size, vp, c_size],
    "nrk_gemm_generic_i32": [c_size, c_size, c_size, vp, c_size, vp, c_size, vp, c_size],
    "nrk_dev_gemm_generic_i64": [c_size, c_size, c_size, c_size, vp, c_size, vp, c_size, vp, vp],
    "nrk_dev_gemm_generic_i32": [c_size, c_size, c_size, c_size, vp, c_size, vp, c_size, vp, vp],
    "nrk_dev_map_reduce_f32": [vp, c_i32, vp, c_size, c_size, c_size, vp, vp, vp],
    "nrk_dev_map_reduce_plain_f64": [vp, c_i32, vp, c_size, c_size, c_size, vp, vp, vp],
    "nrk_dev_map_reduce_plain_f32": [vp, c_i32, vp, c_size, c_size, c_size, vp, vp, vp],
    # gemm
    "nrk_gemm_2d_f64": [c_size, c_size, c_size, vp, c_size, vp, c_size, vp, c_size],
    "nrk_gemm_2d_f32": [c_size, c_size, c_size, vp, c_size, vp, c_size, vp, c_size],
    "nrk_dev_gemm_f64": [c_size, c_size, c_size, vp, vp, vp, vp],
    "nrk_dev_gemm_f32": [c_size, c_size, c_size, vp, vp, vp, vp],
    "nrk_batched_gemm_f64": [c_size, c_size, c_size, c_size, vp, c_size, vp, c_size, vp, c_size],
    "nrk_batched_gemm_f32": [c_size, c_size, c_size, c_size, vp, c_size, vp, c_size, vp, c_size],
    "nrk_dev_batched_gemm_f64": [c_size, c_size, c_size, c_size, vp, c_size, vp, c_size, vp, vp],
    "nrk_dev_batched_gemm_f32": [c_size, c_size, c_size, c_size, vp, c_size, vp, c_size, vp, vp],
    "nrk_gemm_f32_bf16x": [c_size, c_size, c_size, vp, vp, vp, c_i32],
    "nrk_dev_gemm_f32_bf16x": [c_size, c_size, c_size, vp, vp, vp, c_i32, vp],
    "nrk_gemm_f64_ozaki": [c_size, c_size, c_size, vp, vp, vp, c_i32],
    "nrk_dev_gemm_f64_ozaki": [c_size, c_size, c_size, vp, vp, vp, c_i32, vp],
    # ragged batches
    "nrk_grouped_plan_create": [c_i32, vp, c_i32, vp, vp, vp, vp, C.POINTER(vp)],
    "nrk_grouped_plan_run": [vp, vp],
    "nrk_grouped_plan_info": [vp, C.POINTER(c_i32), C.POINTER(c_i32), C.POINTER(c_i32)],
    "nrk_grouped_plan_destroy": [vp],
    "nrk_dev_gemm_bf16": [vp, vp],
    "nrk_graph_begin_capture": [vp],
    "nrk_graph_end_capture": [vp, C.POINTER(vp)],
    "nrk_graph_launch": [vp, vp],
    "nrk_graph_destroy": [vp],
    # tuner / selector
    "nrk_config_count": [c_i32, C.POINTER(c_i32)],
    "nrk_dev_run_config": [c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp],
    "nrk_tune": [c_i32, c_i32, vp, vp, vp, c_i32, vp],
    "nrk_mlp_forward": [vp, c_i32, vp, c_i32, vp, vp],
    "nrk_mlp_train_step": [vp, vp, vp, c_i32, c_i32, vp, vp, vp, c_i32, C.c_float, vp, vp, vp],
    "nrk_selector_set_model": [c_i32, C.c_float, vp, c_i32, vp, vp, vp, vp, C.c_float, C.c_float],
    "nrk_selector_clear": [c_i32],
    "nrk_debug_fail_allocations": [c_i32],
    "nrk_dev_ragged_per_call": [c_i32, vp, c_i32, vp, vp, vp, vp],
    "nrk_selector_table_put": [c_i32, c_i32, c_i32, c_i32, c_i32],
    "nrk_selector_table_size": [vp],
    "nrk_selector_table_get": [c_i32, vp, vp, vp, vp, vp],
    "nrk_selector_choose": [c_i32, c_i32, c_i32, c_i32, vp],
    "nrk_selector_stats": [c_i32, vp, vp, vp, vp],
}
_SIZE_FNS = {"nrk_mlp_param_count": [c_i32]}

_lock = threading.Lock()
_lib: C.CDLL | None = None


class NrkError(RuntimeError):
    """A non-zero nrk_status."""

    def __init__(self, fn: str, status: int, message: str):
        super().__init__(f"{fn} -> status {status}: {message}")
        self.status = status


def load() -> C.CDLL:
    """Load libnumrs_amd.so once; raise if it is missing (no fallback)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m numrs_amd.build`")
        lib = C.CDLL(os.fspath(LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int32
        for name, args in _SIZE_FNS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = c_size
        for name in ("nrk_status_string", "nrk_last_error", "nrk_version"):
            getattr(lib, name).restype = C.c_char_p
        lib.nrk_status_string.argtypes = [C.c_int32]
        lib.nrk_last_error.argtypes = []
        lib.nrk_version.argtypes = []
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Call an nrk_* entry point and raise NrkError on a non-zero status."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        msg = lib.nrk_last_error().decode(errors="replace")
        raise NrkError(name, status, msg)


def declared_symbols() -> list[str]:
    """Every nrk_* function declared in include/numrs_amd.h (parsed from the header)."""
    import re

    header = Path(__file__).resolve().parent.parent / "include" / "numrs_amd.h"
    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:nrk_status|const char \*|size_t)\s*(nrk_\w+)\s*\(", text, re.M)))
