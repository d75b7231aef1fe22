"""Device buffers through the HIP runtime (libamdhip64.so) for callers that
hand raw device pointers to the C ABI (the tri-diagonal solver, the linear
system), without importing torch.  Plumbing only: no compute happens here."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_RT = None


def runtime():
    """The HIP runtime library; raises if it cannot be loaded."""
    global _RT
    if _RT is not None:
        return _RT
    names = ["libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6"]
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    names += [os.path.join(rocm, "lib", n) for n in names[:]]
    last = None
    for n in names:
        try:
            rt = C.CDLL(n)
            break
        except OSError as e:  # try the next name
            last = e
    else:
        raise RuntimeError(f"cannot load the HIP runtime: {last}")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    rt.hipGetErrorString.restype = C.c_char_p
    rt.hipGetErrorString.argtypes = [C.c_int]
    _RT = rt
    return rt


_H2D, _D2H = 1, 2


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """hipMalloc'd bytes with host copies in and out (synchronous)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _chk(runtime().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError(f"{a.nbytes} bytes given, {self.nbytes} expected")
        _chk(runtime().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, _H2D), "hipMemcpy H2D")

    def download(self, dtype, shape=None) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _chk(runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out.reshape(shape) if shape is not None else out

    def free(self) -> None:
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _chk(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
