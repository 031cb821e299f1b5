"""Device memory for the GPU tests through the HIP runtime that
libalsvinn_hip.so itself links (/opt/rocm's libamdhip64): hipMalloc'd
buffers, synchronous copies and device synchronisation, by ctypes.

torch is not used for device memory: its wheel bundles a second HIP/HSA
runtime, and two runtimes in one process do not share the GPU.
"""
import ctypes

import numpy as np

_rt = None


def rt():
    global _rt
    if _rt is None:
        import alsvinn_amd
        alsvinn_amd._lib.load()  # loads libamdhip64.so.7 as its dependency
        L = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _rt = L
    return _rt


H2D, D2H = 1, 2


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    """n elements of dtype in one hipMalloc allocation."""

    def __init__(self, n, dtype=np.float64):
        self.n, self.dtype = int(n), np.dtype(dtype)
        self.ptr = ctypes.c_void_p()
        _ok(rt().hipMalloc(ctypes.byref(self.ptr), max(1, self.n * self.dtype.itemsize)), "hipMalloc")

    @property
    def address(self):
        return self.ptr.value

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype).ravel()
        assert host.size == self.n
        _ok(rt().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D), "hipMemcpy")

    def download(self):
        synchronize()
        out = np.empty(self.n, dtype=self.dtype)
        _ok(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
