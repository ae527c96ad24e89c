"""Minimal ctypes access to the HIP runtime: device buffers and copies.

Only plumbing for the Python layer (tests, bench): the product's compute runs in
``libdoon.so``. Importing this module does not touch the GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        path = os.environ.get("DOON_HIP_LIB", "libamdhip64.so")
        try:
            _hip = ctypes.CDLL(path)
        except OSError:
            _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
    return _hip


def check(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(err).decode()} ({err})")


def device_count() -> int:
    n = ctypes.c_int(0)
    try:
        err = hip().hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        return 0
    return n.value if err == 0 else 0


def synchronize() -> None:
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
