"""Device buffers through the library's own allocator (``st_device_alloc`` and friends).

Used by the Python-level pipelines that keep data resident on the GPU (the multi-job batch,
the learned prefilter) so they need no other GPU framework in the process.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libseqtree_hip.so"


class Status(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("message", ctypes.c_char * 256)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        L = ctypes.CDLL(str(_LIB_PATH))
        vp = ctypes.c_void_p
        L.st_device_alloc.restype = ctypes.c_int32
        L.st_device_alloc.argtypes = [ctypes.c_uint64, ctypes.POINTER(vp), ctypes.POINTER(Status)]
        L.st_device_free.restype = ctypes.c_int32
        L.st_device_free.argtypes = [vp]
        L.st_memcpy.restype = ctypes.c_int32
        L.st_memcpy.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(Status)]
        L.st_synchronize.restype = ctypes.c_int32
        L.st_synchronize.argtypes = [ctypes.POINTER(Status)]
        _lib = L
    return _lib


def check(rc: int, st: Status, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {st.message.decode(errors='replace')}")


def synchronize() -> None:
    st = Status()
    check(lib().st_synchronize(ctypes.byref(st)), st, "st_synchronize")


class DeviceArray:
    """A contiguous device buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        st = Status()
        p = ctypes.c_void_p()
        check(lib().st_device_alloc(self.nbytes, ctypes.byref(p), ctypes.byref(st)), st,
              "st_device_alloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        st = Status()
        check(lib().st_memcpy(d.ptr, a.ctypes.data, a.nbytes, 1, ctypes.byref(st)), st,
              "st_memcpy")
        return d

    def numpy(self, count: int | None = None) -> np.ndarray:
        """The buffer on the host; with `count`, only its first `count` elements (flat)."""
        if count is None:
            out = np.empty(self.shape, self.dtype)
        else:
            out = np.empty(min(int(count), self.nbytes // self.dtype.itemsize), self.dtype)
        st = Status()
        check(lib().st_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2, ctypes.byref(st)), st,
              "st_memcpy")
        return out

    def fill_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        st = Status()
        check(lib().st_memcpy(self.ptr, a.ctypes.data, a.nbytes, 1, ctypes.byref(st)), st,
              "st_memcpy")

    def free(self) -> None:
        if self.ptr:
            lib().st_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass
