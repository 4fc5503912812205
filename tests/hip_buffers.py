"""Device buffers for tests that call the library's low-level entry points
with device pointers: the HIP runtime's own hipMalloc / hipMemcpy /
hipFree through ctypes (no other GPU framework involved)."""

from __future__ import annotations

import ctypes

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(_hip, f).restype = ctypes.c_int
    return _hip


class DeviceBuffer:
    """A device allocation of nbytes (at least 1) freed with the object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            rc = hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed ({rc})")
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=dtype)
        if out.nbytes:
            rc = hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed ({rc})")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None


def synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")
