"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/sift_hip.h declares, and the host mirror of the reference API behaves
like the reference (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import sift_hip

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "sift_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(sift_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_expected_entry_points():
    assert declared_functions() == sorted(sift_hip.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = sift_hip.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_keypoint_layout_is_28_bytes():
    assert sift_hip.KEYPOINT_DTYPE.itemsize == 28
    assert ctypes.sizeof(sift_hip.SiftParams) == 28


def test_params_new_stores_contrast_over_intervals():
    # Sift::new stores contrast_threshold / num_intervals (src/sift.rs:76)
    p = sift_hip.params_struct(1.6, 5, 3, 0.5, 0.04, 10.0)
    assert p.contrast_threshold == np.float32(np.float32(0.04) / np.float32(3))
    assert p.image_border_width == 5 and p.num_octaves == 5
    d = sift_hip.SiftParams()
    sift_hip.lib().sift_params_default(ctypes.byref(d))
    assert (d.sigma, d.num_octaves, d.num_intervals) == (np.float32(1.6), 4, 3)
    assert d.assumed_blur == 0.5 and d.edge_threshold == 10.0
    assert sift_hip.Sift.default().contrast_threshold == p.contrast_threshold


def test_backend_aliases():
    # src/lib.rs:36-49 plus "hip"
    B = sift_hip.SiftBackend
    for s, b in [("cpu", B.Cpu), ("GPU", B.WebGpu), ("wgpu", B.WebGpu), (" auto ", B.WebGpuWithCpuFallback),
                 ("prefer-gpu", B.WebGpuWithCpuFallback), ("v2", B.WebGpuV2), ("hip", B.Hip)]:
        assert B.from_str(s) is b
    with pytest.raises(ValueError):
        B.from_str("cuda")


def test_other_backends_are_not_silently_served():
    s = sift_hip.Sift.default()
    with pytest.raises(sift_hip.SiftError):
        s.detect_and_compute_with_backend(np.zeros((32, 32), np.uint8), "cpu")


def test_invalid_arguments_are_rejected_without_a_device():
    L = sift_hip.lib()
    p = sift_hip.params_struct(1.6, 4, 3, 0.5, 0.04, 10.0)
    out = ctypes.c_void_p()
    assert L.sift_hip_create(None, 0, ctypes.byref(out)) == -1
    bad = sift_hip.params_struct(1.6, 0, 3, 0.5, 0.04, 10.0)
    assert L.sift_hip_create(ctypes.byref(bad), 0, ctypes.byref(out)) == -1
    assert L.sift_hip_detect(None, None, None) == -1
    assert "null" in sift_hip.last_error()
    r = sift_hip.SiftResult()
    L.sift_result_free(ctypes.byref(r))  # freeing an empty result is a no-op
    L.sift_result_free(None)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_create_without_device_reports_device_error():
    L = sift_hip.lib()
    p = sift_hip.params_struct(1.6, 4, 3, 0.5, 0.04, 10.0)
    out = ctypes.c_void_p()
    assert L.sift_hip_create(ctypes.byref(p), 0, ctypes.byref(out)) == -2
    assert "device" in sift_hip.last_error()


def test_result_arrays_view_the_library_buffers_and_free_them_once():
    """Results reach Python as arrays over the library's malloc'd buffers (no
    copy); sift_result_free runs once, after the last array or view is gone."""
    import gc
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    n = 3
    kps = np.zeros(n, sift_hip.KEYPOINT_DTYPE)
    kps["x"] = [1.5, 2.5, 3.5]
    kps["layer"] = [1, 2, 3]
    desc = np.arange(n * 128, dtype=np.float32).reshape(n, 128) / 7.0
    kp_mem, de_mem = libc.malloc(kps.nbytes), libc.malloc(desc.nbytes)
    ctypes.memmove(kp_mem, kps.ctypes.data, kps.nbytes)
    ctypes.memmove(de_mem, desc.ctypes.data, desc.nbytes)
    r = sift_hip.SiftResult(kp_mem, ctypes.cast(de_mem, ctypes.POINTER(ctypes.c_float)), n)
    freed = []

    def free(p):
        freed.append(1)
        sift_hip.lib().sift_result_free(p)

    k, d = sift_hip._result_arrays(r, free, np.float32)
    assert k.ctypes.data == kp_mem and d.ctypes.data == de_mem  # views, not copies
    assert np.array_equal(k, kps) and np.array_equal(d, desc)
    row = d[1]
    del k, d
    gc.collect()
    assert freed == []  # a view of the descriptors is still alive
    assert np.array_equal(row, desc[1])
    del row
    gc.collect()
    assert freed == [1]
    # an empty result is released at once and gives empty arrays
    e = sift_hip.SiftResult(None, None, 0)
    k, d = sift_hip._result_arrays(e, free, np.float32)
    assert len(k) == 0 and d.shape == (0, 128) and freed == [1, 1]
