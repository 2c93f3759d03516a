"""Ragged batches (bf16 GEMMs + f32 row sums) in one persistent launch, and hipGraph capture."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import GemmDesc, call


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even f32 -> bf16, returned as uint16 bits."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_bits_to_f32(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


class GroupedPlan:
    """nrk_grouped_plan: descs are (m, n, k, a_ptr, b_ptr, c_ptr) with device pointers."""

    def __init__(self, descs, rows_x_ptr=None, row_offsets=None, rows_out_ptr=None, item_cost_ns=None):
        n = len(descs)
        arr = (GemmDesc * max(1, n))()
        for i, (m, nn, k, a, b, c) in enumerate(descs):
            arr[i] = GemmDesc(m, nn, k, 0, a, b, c)
        self._descs = arr
        offs = None if row_offsets is None else np.ascontiguousarray(row_offsets, np.int64)
        rows = 0 if offs is None else len(offs) - 1
        cost = None if item_cost_ns is None else np.ascontiguousarray(item_cost_ns, np.float32)
        p = C.c_void_p()
        call("nrk_grouped_plan_create", C.c_int32(n), C.cast(arr, C.c_void_p), C.c_int32(rows), C.c_void_p(rows_x_ptr),
             None if offs is None else offs.ctypes.data_as(C.c_void_p), C.c_void_p(rows_out_ptr),
             None if cost is None else cost.ctypes.data_as(C.c_void_p), C.byref(p))
        self.handle = p.value

    def run(self, stream=None) -> None:
        call("nrk_grouped_plan_run", C.c_void_p(self.handle), C.c_void_p(stream))

    def info(self) -> dict:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        call("nrk_grouped_plan_info", C.c_void_p(self.handle), C.byref(a), C.byref(b), C.byref(c))
        return {"items": a.value, "gemm_tiles": b.value, "grid": c.value}

    def close(self) -> None:
        if self.handle:
            call("nrk_grouped_plan_destroy", C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm_bf16(m, n, k, a_ptr, b_ptr, c_ptr, stream=None) -> None:
    d = GemmDesc(m, n, k, 0, a_ptr, b_ptr, c_ptr)
    call("nrk_dev_gemm_bf16", C.byref(d), C.c_void_p(stream))


class Graph:
    """A hipGraph captured from the work a callable enqueues on `stream`."""

    def __init__(self, stream, enqueue):
        call("nrk_graph_begin_capture", C.c_void_p(stream))
        try:
            enqueue()
        finally:
            g = C.c_void_p()
            call("nrk_graph_end_capture", C.c_void_p(stream), C.byref(g))
        self.handle = g.value
        self.stream = stream

    def launch(self) -> None:
        call("nrk_graph_launch", C.c_void_p(self.handle), C.c_void_p(self.stream))

    def close(self) -> None:
        if self.handle:
            call("nrk_graph_destroy", C.c_void_p(self.handle))
            self.handle = None
