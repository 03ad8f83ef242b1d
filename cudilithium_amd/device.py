"""Device memory, streams and events through the library's own C-ABI helpers,
so the host side needs no other GPU runtime."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class DeviceBuffer:
    """A device allocation of `nbytes` bytes (hipMalloc through the library)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.check(_lib.load().dil_malloc_device(ctypes.byref(p), max(1, self.nbytes)), "malloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _lib.check(_lib.load().dil_memcpy_h2d(self.ptr + offset, _ptr(a), a.nbytes, stream), "h2d")
        if stream is None:
            synchronize()

    def download(self, shape, dtype, stream=None, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _lib.check(_lib.load().dil_memcpy_d2h(_ptr(out), self.ptr + offset, out.nbytes, stream), "d2h")
        stream_sync(stream)
        return out

    def memset(self, value: int, stream=None) -> None:
        _lib.check(_lib.load().dil_memset_device(self.ptr, value, self.nbytes, stream), "memset")

    def free(self) -> None:
        if self.ptr:
            _lib.load().dil_free_device(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PitchedBuffer(DeviceBuffer):
    """hipMallocPitch rows, as the reference's callers allocate pools
    (tests/test_vectors.cu:70,100,128,166)."""

    def __init__(self, width: int, height: int):
        p = ctypes.c_void_p()
        pitch = ctypes.c_size_t()
        _lib.check(_lib.load().dil_malloc_pitch(ctypes.byref(p), ctypes.byref(pitch), width, height),
                   "malloc_pitch")
        self.ptr = p.value
        self.pitch = pitch.value
        self.nbytes = self.pitch * max(1, height)


class HostBuffer:
    """Pinned host memory viewed as a numpy array."""

    def __init__(self, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = ctypes.c_void_p()
        _lib.check(_lib.load().dil_malloc_host(ctypes.byref(p), max(1, n)), "malloc_host")
        self.ptr = p.value
        buf = (ctypes.c_uint8 * max(1, n)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=n).view(dt).reshape(shape)

    def free(self) -> None:
        if self.ptr:
            self.array = None
            _lib.load().dil_free_host(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _lib.check(_lib.load().dil_stream_create(ctypes.byref(p)), "stream_create")
        self.handle = p.value

    def synchronize(self) -> None:
        _lib.check(_lib.load().dil_stream_synchronize(self.handle), "stream_sync")

    def __del__(self):
        try:
            if self.handle:
                _lib.load().dil_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _lib.check(_lib.load().dil_event_create(ctypes.byref(p)), "event_create")
        self.handle = p.value

    def record(self, stream=None) -> None:
        _lib.check(_lib.load().dil_event_record(self.handle, stream), "event_record")

    def synchronize(self) -> None:
        _lib.check(_lib.load().dil_event_synchronize(self.handle), "event_sync")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _lib.check(_lib.load().dil_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle), "elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().dil_event_destroy(self.handle)
        except Exception:
            pass


def synchronize() -> None:
    _lib.check(_lib.load().dil_device_synchronize(), "device_sync")


def stream_sync(stream) -> None:
    if stream is None:
        synchronize()
    else:
        _lib.check(_lib.load().dil_stream_synchronize(stream), "stream_sync")


def device_count() -> int:
    n = ctypes.c_int(0)
    r = _lib.load().dil_device_count(ctypes.byref(n))
    return n.value if r == 0 else 0


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    _lib.check(_lib.load().dil_device_name(buf, 256), "device_name")
    return buf.value.decode()
