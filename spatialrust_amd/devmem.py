"""Minimal device buffers through the HIP runtime (ctypes), for callers that want
inputs resident in HBM before calling the C ABI with SR_MEM_DEVICE views."""
import ctypes

import numpy as np

from .registration import device_view

_hip = None


def _rt():
    global _hip
    if _hip is None:
        # By soname: the HIP runtime already in the process (the library's, or
        # one a framework loaded first) is reused. A second runtime instance,
        # which loading another file would start, sees no device.
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(code, what):
    if code != 0:
        raise RuntimeError(f"{what} failed with HIP error {code}")


class DeviceArray:
    """A float32 array copied to device memory (freed on close / deletion)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float32)
        self.n = host.size
        self.ptr = ctypes.c_void_p()
        _ok(_rt().hipMalloc(ctypes.byref(self.ptr), max(host.nbytes, 4)), "hipMalloc")
        _ok(_rt().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1), "hipMemcpy H2D")

    def close(self):
        if self.ptr and self.ptr.value:
            _rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCloud:
    """xyz columns of an (N,3) cloud resident in device memory."""

    def __init__(self, points: np.ndarray):
        pts = np.asarray(points, dtype=np.float32)
        self.cols = [DeviceArray(pts[:, a]) for a in range(3)]
        self.n = len(pts)
        self.view = device_view(self.cols[0].ptr.value, self.cols[1].ptr.value,
                                self.cols[2].ptr.value, self.n)

    def close(self):
        for c in self.cols:
            c.close()


def copy_calibration(nbytes: int):
    """One device-to-device hipMemcpy of `nbytes` (the runtime's plain copy
    kernel): a known-size reference for hardware byte counters."""
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _ok(_rt().hipMalloc(ctypes.byref(a), nbytes), "hipMalloc")
    _ok(_rt().hipMalloc(ctypes.byref(b), nbytes), "hipMalloc")
    try:
        _ok(_rt().hipMemcpy(b, a, nbytes, 3), "hipMemcpy D2D")
        synchronize()
    finally:
        _rt().hipFree(a)
        _rt().hipFree(b)


def synchronize():
    _ok(_rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
