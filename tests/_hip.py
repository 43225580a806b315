"""Device buffers for tests of the _dev entry points, through the HIP runtime
the library itself links (libamdhip64.so.7), so no second runtime is loaded."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of a shape)."""

    def __init__(self, a=None, shape=None, dtype=np.float64):
        src = None if a is None else np.ascontiguousarray(a)
        self.dtype = np.dtype(dtype if src is None else src.dtype)
        self.shape = tuple(shape if src is None else src.shape)
        self.nbytes = max(int(np.prod(self.shape)) * self.dtype.itemsize, 1)
        p = C.c_void_p()
        if hip().hipMalloc(C.byref(p), self.nbytes) != 0:
            raise MemoryError("hipMalloc failed")
        self.ptr = p.value
        if src is not None and src.nbytes:
            if hip().hipMemcpy(self.ptr, src.ctypes.data, src.nbytes, 1) != 0:   # host to device
                raise RuntimeError("hipMemcpy failed")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        hip().hipDeviceSynchronize()
        if out.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = None
