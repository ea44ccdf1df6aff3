"""Minimal HIP runtime access through ctypes: device buffers for inputs that
must already sit in HBM (benchmarks, device-pointer entry points)."""
import ctypes as C
import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        for name in ('libamdhip64.so', '/opt/rocm/lib/libamdhip64.so'):
            try:
                _HIP = C.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise OSError('libamdhip64.so not found')
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipFree.argtypes = [C.c_void_p]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipDeviceSynchronize.argtypes = []
        _HIP.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
    return _HIP


def device_count():
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


class DeviceBuffer:
    """A hipMalloc'ed buffer filled from a numpy array."""

    def __init__(self, arr):
        arr = np.ascontiguousarray(arr)
        self.nbytes = arr.nbytes
        self.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise MemoryError('hipMalloc failed')
        if hip().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), self.nbytes, 1) != 0:   # HostToDevice
            raise RuntimeError('hipMemcpy failed')

    def download(self, dtype, shape):
        """The buffer's first prod(shape) elements as a numpy array."""
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError('download larger than the buffer')
        if hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2) != 0:   # DeviceToHost
            raise RuntimeError('hipMemcpy failed')
        return out

    def __del__(self):
        if getattr(self, 'ptr', None) and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = None
