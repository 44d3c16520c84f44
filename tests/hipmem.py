"""Device buffers for tests that call device-pointer entry points: hipMalloc,
hipMemcpy and hipFree from the HIP runtime instance the library is bound to
(opened by its soname after the library is loaded, so the dynamic linker
returns that same instance), through ctypes.  torch ships its own copy of
the runtime, which does not start in a process where another copy has
already started, so these tests do not use torch."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from dense_hodlr_amd import _lib
        _lib.lib()  # loads the runtime the library links; the soname below finds it
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A column-major n x ncols FP64 array in device memory."""

    def __init__(self, n, ncols, fill=None):
        self.n, self.ncols = n, ncols
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), n * ncols * 8)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        if fill is not None:
            self.upload(np.full((n, ncols), fill, order="F"))

    def upload(self, A):
        A = np.asfortranarray(A, dtype=np.float64)
        assert A.shape == (self.n, self.ncols)
        rc = hip().hipMemcpy(self.ptr, A.ctypes.data, A.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")

    def download(self):
        A = np.zeros((self.n, self.ncols), order="F")
        rc = hip().hipMemcpy(A.ctypes.data, self.ptr, A.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
        return A

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()
