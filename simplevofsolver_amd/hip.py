"""Device buffers for callers that hand raw device pointers to the C ABI (the
tri-diagonal solver, the linear system).  The memory comes from the solver
library's own HIP runtime (vof_device_alloc / vof_device_free /
vof_device_copy), so it is valid for the library's kernels whatever other HIP
runtime (torch's) the process has loaded.  Plumbing only: no compute happens
here."""
from __future__ import annotations

import ctypes as C

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3


def _lib(ndims: int = 2):
    from . import load
    lib = load(ndims)
    if not getattr(lib, "_vof_dev_typed", False):
        lib.vof_device_alloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.vof_device_free.argtypes = [C.c_void_p]
        lib.vof_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib._vof_dev_typed = True
    return lib


def _chk(rc, what, lib):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib.vof_last_error().decode()}")


def copy(dst, src, nbytes: int, kind: int, ndims: int = 2) -> None:
    """Synchronous copy between raw pointers (kind 1: to device, 2: to host,
    3: device to device)."""
    lib = _lib(ndims)
    _chk(lib.vof_device_copy(dst, src, int(nbytes), int(kind)), "vof_device_copy", lib)


class DeviceBuffer:
    """Device bytes with synchronous host copies in and out."""

    def __init__(self, nbytes: int, ndims: int = 2):
        self.nbytes = int(nbytes)
        self.ndims = ndims
        self.ptr = C.c_void_p()
        lib = _lib(ndims)
        _chk(lib.vof_device_alloc(C.byref(self.ptr), self.nbytes), "vof_device_alloc", lib)

    @classmethod
    def from_array(cls, a: np.ndarray, ndims: int = 2) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, ndims)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError(f"{a.nbytes} bytes given, {self.nbytes} expected")
        copy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, _H2D, self.ndims)

    def download(self, dtype, shape=None) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        copy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, _D2H, self.ndims)
        return out.reshape(shape) if shape is not None else out

    def free(self) -> None:
        if self.ptr:
            lib = _lib(self.ndims)
            lib.vof_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
