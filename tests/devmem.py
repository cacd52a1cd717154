"""Device buffers for the tests of the device-resident entry points, through the
HIP runtime libsift_hip.so itself links (libamdhip64.so.7), so that the
pointers belong to the same runtime as the library's own allocations."""
import ctypes

import numpy as np

import sift_hip

_H2D, _D2H = 1, 2
_RT = None


def _rt():
    global _RT
    if _RT is None:
        sift_hip.lib()  # the library first: its runtime is the one to share
        rt = ctypes.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        rt.hipMalloc.restype = ctypes.c_int
        rt.hipFree.argtypes = [ctypes.c_void_p]
        rt.hipFree.restype = ctypes.c_int
        rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        rt.hipMemcpy.restype = ctypes.c_int
        rt.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        rt.hipMemGetInfo.restype = ctypes.c_int
        rt.hipDeviceSynchronize.restype = ctypes.c_int
        _RT = rt
    return _RT


class DeviceBuffer:
    """One hipMalloc allocation, freed by free() or at garbage collection."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        rc = _rt().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1))
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed ({rc})")

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        rc = _rt().hipMemcpy(self.addr + offset, a.ctypes.data, a.nbytes, _H2D)
        assert rc == 0, rc

    def free(self):
        if self.ptr.value:
            _rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download(addr, nbytes):
    """Bytes [addr, addr + nbytes) of device memory."""
    out = np.empty(int(nbytes), np.uint8)
    if nbytes:
        rc = _rt().hipMemcpy(out.ctypes.data, addr, int(nbytes), _D2H)
        assert rc == 0, rc
    return out


def upload_image(img, stride=None):
    """A device copy of a 2-D u8 image with `stride` bytes per row (default its
    width); returns (buffer, (ptr, w, h, stride)) for Sift.detect_batch_device."""
    h, w = img.shape
    stride = w if stride is None else stride
    assert stride >= w
    host = np.zeros((h, stride), np.uint8)
    host[:, :w] = img
    buf = DeviceBuffer(host.nbytes)
    buf.upload(host)
    return buf, (buf.addr, w, h, stride)


def mem_info():
    """(free, total) device bytes."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    rc = _rt().hipMemGetInfo(ctypes.byref(f), ctypes.byref(t))
    assert rc == 0, rc
    return f.value, t.value


def synchronize():
    assert _rt().hipDeviceSynchronize() == 0
