"""Device buffers for tests that call entry points taking device pointers.

The buffers come from the HIP runtime that libdnlgpu.so is already bound to
(opened with RTLD_NOLOAD, so no second runtime is loaded), which keeps one HIP
runtime per process.  torch ships its own copy of the runtime under the same
soname, and whichever of the two loads first serves the process.
"""
import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from redynela_amd import _lib
        _lib.lib()  # loads libdnlgpu.so and the runtime it needs
        _hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    """A device copy of a numpy array (H2D at creation, D2H with .get())."""

    def __init__(self, a):
        self.host = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        check(hip().hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 1)), "hipMalloc")
        check(hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1), "hipMemcpy H2D")

    def get(self):
        out = np.empty_like(self.host)
        check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()
