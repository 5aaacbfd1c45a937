"""Device arrays and streams on the engine's own HIP runtime.

A process that loads libseismic_xcorr.so must not also initialise a second
HIP runtime (e.g. the one bundled inside a PyTorch wheel): two HSA runtimes in
one process do not share the GPU.  These helpers give tests, bench and
callers device memory without any other GPU library."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import last_error, load_engine


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(load_engine().sxc_device_count(ctypes.byref(n)), "sxc_device_count")
    return n.value


def synchronize() -> None:
    _check(load_engine().sxc_device_synchronize(), "sxc_device_synchronize")


class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype=np.float32):
        self.lib = load_engine()
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(self.lib.sxc_device_alloc(ctypes.byref(p), self.nbytes), "sxc_device_alloc")
        self._ptr = p

    @property
    def ptr(self) -> int:
        return self._ptr.value or 0

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.copy_from(a)
        return d

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(self.lib.sxc_memcpy_h2d(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                       self.nbytes), "sxc_memcpy_h2d")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(self.lib.sxc_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                       self.nbytes), "sxc_memcpy_d2h")
        return out

    def zero_(self) -> "DeviceArray":
        _check(self.lib.sxc_memset_d8(ctypes.c_void_p(self.ptr), 0, self.nbytes), "sxc_memset_d8")
        return self

    def free(self) -> None:
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            self.lib.sxc_device_free(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        self.lib = load_engine()
        p = ctypes.c_void_p()
        _check(self.lib.sxc_stream_create(ctypes.byref(p)), "sxc_stream_create")
        self._h = p

    @property
    def handle(self) -> int:
        return self._h.value or 0

    def synchronize(self) -> None:
        _check(self.lib.sxc_stream_synchronize(self._h), "sxc_stream_synchronize")

    def destroy(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.sxc_stream_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
