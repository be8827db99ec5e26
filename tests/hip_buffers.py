"""Device buffers for the GPU tests of the C-ABI entry points that take
device pointers (vlp_field_solve and its no-allocation variant).

The buffers come from the HIP runtime that libvlp4d.so itself is linked
against (libamdhip64.so.7 of /opt/rocm), called through ctypes: torch ships
its own HIP runtime, and two runtimes in one process do not share devices.
"""
import ctypes
import os

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        path = "/opt/rocm/lib/libamdhip64.so.7"
        lib = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


class DeviceArray:
    """A device allocation of nbytes, filled from / read into numpy arrays."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc")
        _ok(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
