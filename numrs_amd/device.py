"""Device buffers, streams and events over the C ABI (no torch: the library owns device memory)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, load


class DeviceBuffer:
    """A device allocation made by nrk_device_malloc. Frees itself on close()/GC."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("nrk_device_malloc", C.byref(p), C.c_size_t(max(self.nbytes, 16)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        return DeviceArray.upload(a, stream)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            try:
                call("nrk_device_free", C.c_void_p(self.ptr))
            finally:
                self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray(DeviceBuffer):
    """A typed, shaped device buffer."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1
        super().__init__(self.size * self.dtype.itemsize)

    @classmethod
    def upload(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            call("nrk_memcpy_htod", C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                 C.c_void_p(stream))
        return d

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            call("nrk_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(out.nbytes),
                 C.c_void_p(stream))
        return out

    def fill_from(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        call("nrk_memcpy_htod", C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
             C.c_void_p(stream))


class Stream:
    def __init__(self):
        s = C.c_void_p()
        call("nrk_stream_create", C.byref(s))
        self.handle = s.value

    def wait(self, event: "Event") -> None:
        """Later work on this stream waits for `event` (recorded on any stream)."""
        call("nrk_stream_wait_event", C.c_void_p(self.handle), C.c_void_p(event.handle))

    def synchronize(self) -> None:
        call("nrk_stream_synchronize", C.c_void_p(self.handle))

    def close(self) -> None:
        if self.handle:
            call("nrk_stream_destroy", C.c_void_p(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        call("nrk_event_create", C.byref(e))
        self.handle = e.value

    def record(self, stream=None) -> None:
        call("nrk_event_record", C.c_void_p(self.handle), C.c_void_p(stream))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        call("nrk_event_elapsed_ms", C.byref(ms), C.c_void_p(self.handle), C.c_void_p(end.handle))
        return float(ms.value)

    def close(self) -> None:
        if self.handle:
            call("nrk_event_destroy", C.c_void_p(self.handle))
            self.handle = None


def synchronize() -> None:
    call("nrk_device_synchronize")


def init(device: int = 0) -> None:
    """Select the device, and register the persisted tile-selection model (if present) with the
    library's GEMM dispatch."""
    load()
    call("nrk_init", C.c_int(device))
    from . import selector

    selector.install()


def call_memset(buf: DeviceBuffer, stream=None) -> None:
    """Zero a device buffer (eager; never recorded into a graph by this library)."""
    call("nrk_memset_zero", C.c_void_p(buf.ptr), C.c_size_t(buf.nbytes), C.c_void_p(stream))
