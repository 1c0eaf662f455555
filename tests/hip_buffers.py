"""Device buffers for tests that call the C ABI with device pointers."""
import ctypes

import numpy as np

from lulesh_amd import native


class HipBuffers:
    """Device buffers of the test (the caller's inputs and outputs), from the
    HIP runtime the library itself is linked against.  Torch is not used:
    its wheel carries a second HIP runtime, which cannot open the GPU once
    the library's runtime holds it in the same process."""

    def __init__(self):
        native.gpu_lib()  # loads the library's runtime first; the soname below resolves to it
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def put(self, a: np.ndarray) -> ctypes.c_void_p:
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p: ctypes.c_void_p, n: int, dtype) -> np.ndarray:
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []
