"""Device memory through the ROCm HIP runtime that libragdoll_hip.so itself links.

One HIP runtime per process: a second copy (such as the one bundled with PyTorch) must not be
initialised beside it, so device buffers for the library's device-pointer entry points are
allocated here rather than through torch.
"""
import ctypes

import numpy as np


class Hip:
    def __init__(self):
        self.L = None
        for name in ("/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so.7"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.L is None:
            raise RuntimeError("libamdhip64.so.7 not found")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        self.L.hipHostFree.argtypes = [ctypes.c_void_p]

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed with HIP error %d" % (what, rc))

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)), "hipMalloc")
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy")  # host to device
        return p

    def upload_to(self, ptr, arr):
        """Copies the contiguous array arr into device memory at ptr."""
        arr = np.ascontiguousarray(arr)
        self.check(self.L.hipMemcpy(ptr, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy")  # host to device

    def download(self, ptr, out):
        """Copies out.nbytes from device address ptr into the contiguous array out."""
        self.check(self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2), "hipMemcpy")  # device to host
        return out

    def memset(self, ptr, value, nbytes):
        self.check(self.L.hipMemset(ptr, value, nbytes), "hipMemset")

    def copy(self, dst, src, nbytes):
        self.check(self.L.hipMemcpy(dst, src, nbytes, 3), "hipMemcpy")  # device to device

    def free(self, ptr):
        self.L.hipFree(ctypes.c_void_p(ptr))

    def host_alloc(self, nbytes):
        """Pinned host memory of nbytes, as a writable uint8 numpy array (release with host_free)."""
        p = ctypes.c_void_p()
        self.check(self.L.hipHostMalloc(ctypes.byref(p), max(int(nbytes), 1), 0), "hipHostMalloc")
        arr = np.ctypeslib.as_array((ctypes.c_uint8 * max(int(nbytes), 1)).from_address(p.value))
        return p.value, arr

    def host_free(self, ptr):
        self.L.hipHostFree(ctypes.c_void_p(ptr))

    def memcpy(self, dst, src, nbytes, kind):
        """Raw copy between addresses (kind 1 host to device, 2 device to host, 3 device to device)."""
        self.check(self.L.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(nbytes), kind), "hipMemcpy")

    def synchronize(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")
