"""Device buffers for callers without torch (bench.py): hipMalloc/hipMemcpy/hipFree from the HIP runtime that
libx2dev.so itself loaded, so pointers are valid in the library's calls."""
import ctypes
import os

import numpy as np

_hip = None


def _runtime():
    global _hip
    if _hip is not None:
        return _hip
    from . import abi
    abi.load()  # loads libamdhip64 as the library's dependency; reuse that very file
    path = None
    with open("/proc/self/maps") as maps:
        for line in maps:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    if path is None or not os.path.exists(path):
        raise RuntimeError("the HIP runtime is not loaded (libx2dev.so failed to load?)")
    lib = ctypes.CDLL(path)
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    for f in (lib.hipMalloc, lib.hipFree, lib.hipMemcpy, lib.hipMemset):
        f.restype = ctypes.c_int
    _hip = lib
    return lib


class DeviceBuffer:
    """nbytes of device memory on the current device."""

    def __init__(self, nbytes):
        self.lib = _runtime()
        self.nbytes = max(1, int(nbytes))
        p = ctypes.c_void_p()
        if self.lib.hipMalloc(ctypes.byref(p), self.nbytes) != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed")
        self.ptr = p.value

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        if self.lib.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to device failed")

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        if self.lib.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:  # device to host
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def zero(self):
        if self.lib.hipMemset(self.ptr, 0, self.nbytes) != 0:
            raise RuntimeError("hipMemset failed")

    def free(self):
        if self.ptr:
            self.lib.hipFree(self.ptr)
            self.ptr = None
