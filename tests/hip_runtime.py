"""The HIP runtime the library itself loaded, for tests that move their own
device buffers (a process may also hold the copy bundled with torch; device
pointers of one runtime mean nothing to the other)."""
import ctypes as C
import os

import phylocsf_amd as P  # noqa: F401  (loads the library, and with it its runtime)
from phylocsf_amd import _lib


def runtime():
    _lib.lib()
    paths = []
    for ln in open("/proc/self/maps"):
        f = ln.split()
        if len(f) >= 6 and "libamdhip64" in f[-1] and f[-1] not in paths:
            paths.append(f[-1])
    own = [p for p in paths if os.sep + "torch" + os.sep not in p] or paths
    hip = C.CDLL(own[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class DeviceBuffer:
    """hipMalloc'd bytes, freed on close."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(1, nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D) == 0

    def download(self, arr):
        assert self.hip.hipMemcpy(arr.ctypes.data, self.ptr, arr.nbytes, D2H) == 0
        return arr

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()
