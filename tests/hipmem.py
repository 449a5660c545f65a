"""Device buffers for GPU tests through the HIP runtime the library itself
links (test plumbing; torch ships a runtime of its own)."""
import ctypes
import ctypes.util
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import accelerint_amd as acc
        acc.lib()  # the library's runtime is then already in the process
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """A device copy of a numpy array (C order); ``get()`` copies it back."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(self.a.nbytes, 1)) == 0
        assert hip().hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host -> device

    @property
    def addr(self):
        return self.ptr.value

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device -> host
        return out

    def __del__(self):
        if _hip is not None and self.ptr:
            _hip.hipFree(self.ptr)
