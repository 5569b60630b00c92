"""Device memory, streams and events through the C ABI (no torch needed)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi


def device_count() -> int:
    n = C.c_int(0)
    st = _ffi.hip().cublox_device_count(C.byref(n))
    return n.value if st == 0 else 0


def device_name() -> str:
    buf = C.create_string_buffer(256)
    _ffi.check(_ffi.hip().cublox_device_name(buf, 256))
    return buf.value.decode()


class DeviceBuffer:
    """An owned device allocation of `nbytes`."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _ffi.check(_ffi.hip().cublox_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ffi.check(_ffi.hip().cublox_memcpy_h2d_async(
            self.ptr, a.ctypes.data, a.nbytes, stream))
        _ffi.check(_ffi.hip().cublox_stream_synchronize(stream))

    def download(self, dtype, count: int, stream=None) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _ffi.check(_ffi.hip().cublox_memcpy_d2h_async(
            out.ctypes.data, self.ptr, out.nbytes, stream))
        _ffi.check(_ffi.hip().cublox_stream_synchronize(stream))
        return out

    def zero(self, stream=None) -> None:
        _ffi.check(_ffi.hip().cublox_memset_async(self.ptr, 0, self.nbytes, stream))
        _ffi.check(_ffi.hip().cublox_stream_synchronize(stream))

    def free(self) -> None:
        if self.ptr:
            _ffi.hip().cublox_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ffi.check(_ffi.hip().cublox_stream_create(C.byref(s)))
        self.handle = s.value

    def synchronize(self) -> None:
        _ffi.check(_ffi.hip().cublox_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                _ffi.hip().cublox_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ffi.check(_ffi.hip().cublox_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None) -> None:
        _ffi.check(_ffi.hip().cublox_event_record(self.handle, stream))

    def synchronize(self) -> None:
        _ffi.check(_ffi.hip().cublox_event_synchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        _ffi.check(_ffi.hip().cublox_event_elapsed_ms(self.handle, end.handle,
                                                      C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _ffi.hip().cublox_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def synchronize() -> None:
    _ffi.check(_ffi.hip().cublox_device_synchronize())


# cublox_index_case / cublox_index_chain of include/cublox/capi.h.
INDEX_CASE = np.dtype([("local", "<i4", 3), ("map", "<i4", 3), ("half", "<i4", 3),
                       ("origin", "<i4", 3), ("res", "<f4"), ("inv", "<f4"),
                       ("center", "<i4")])
INDEX_CHAIN = np.dtype([("g", "<i4", 3), ("l_back", "<i4", 3), ("hash", "<i4"),
                        ("l_hash", "<i4", 3), ("g_hash", "<i4", 3), ("pos", "<f4", 3),
                        ("g_pos", "<i4", 3)])


def index_chain(cases: np.ndarray) -> np.ndarray:
    """Runs the device index chain (cublox_launch_index_chain) over an array
    of INDEX_CASE records; returns INDEX_CHAIN records."""
    cases = np.ascontiguousarray(cases, INDEX_CASE)
    n = len(cases)
    if n == 0:
        return np.zeros(0, INDEX_CHAIN)
    d_in = DeviceBuffer.from_array(cases.view(np.uint8))
    d_out = DeviceBuffer(n * INDEX_CHAIN.itemsize)
    _ffi.check(_ffi.hip().cublox_launch_index_chain(d_in.ptr, n, d_out.ptr, None))
    raw = d_out.download(np.uint8, n * INDEX_CHAIN.itemsize)
    return raw.view(INDEX_CHAIN).copy()

