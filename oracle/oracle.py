"""ctypes binding of the CPU oracle (oracle/libsift_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the parity reference or the timed CPU
baseline, never as the product path. The oracle restates src/sift.rs of the
reference (see sift_oracle.c for the per-function citations).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                           ("response", "<f4"), ("octave", "<i4"), ("layer", "<i4")])


class OracleParams(ctypes.Structure):
    _fields_ = [("sigma", ctypes.c_float), ("num_octaves", ctypes.c_uint32),
                ("num_intervals", ctypes.c_uint32), ("assumed_blur", ctypes.c_float),
                ("contrast_threshold", ctypes.c_float), ("edge_threshold", ctypes.c_float),
                ("image_border_width", ctypes.c_uint32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsift_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        u8p = P(ctypes.c_uint8)
        f32p = P(ctypes.c_float)
        L.oracle_params_new.argtypes = [P(OracleParams), ctypes.c_float, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float]
        L.oracle_gaussian_kernel.argtypes = [ctypes.c_float, f32p, ctypes.c_int]
        L.oracle_gaussian_kernel.restype = ctypes.c_int
        L.oracle_gaussian_blur.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float, u8p]
        L.oracle_resize_lanczos3.argtypes = [u8p, ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int,
                                             ctypes.c_int]
        L.oracle_lanczos3_table.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_int),
                                            P(ctypes.c_int), f32p, ctypes.c_int]
        L.oracle_lanczos3_table.restype = ctypes.c_int
        L.oracle_level_sigmas.argtypes = [P(OracleParams), f32p, f32p]
        L.oracle_pyramid_size.argtypes = [P(OracleParams), ctypes.c_int, ctypes.c_int]
        L.oracle_pyramid_size.restype = ctypes.c_size_t
        L.oracle_dog_size.argtypes = [P(OracleParams), ctypes.c_int, ctypes.c_int]
        L.oracle_dog_size.restype = ctypes.c_size_t
        L.oracle_build_pyramid.argtypes = [P(OracleParams), u8p, ctypes.c_int, ctypes.c_int,
                                           u8p, f32p]
        kpp = ctypes.c_void_p
        L.oracle_detect_and_compute.argtypes = [
            P(OracleParams), u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            P(kpp), P(ctypes.c_size_t), P(kpp), P(ctypes.c_size_t),
            P(kpp), P(ctypes.c_void_p), P(ctypes.c_size_t), P(ctypes.c_double)]
        L.oracle_detect_and_compute.restype = ctypes.c_int
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_solve_linear_system.argtypes = [f32p, f32p, f32p]
        L.oracle_solve_linear_system.restype = ctypes.c_int
        L.oracle_gauss_pixel_bilinear.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                                  ctypes.c_float]
        L.oracle_gauss_pixel_bilinear.restype = ctypes.c_float
        L.oracle_normalize_and_clip_descriptor.argtypes = [f32p, ctypes.c_int]
        L.oracle_orientation_single.argtypes = [P(OracleParams), u8p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_void_p, f32p]
        L.oracle_orientation_single.restype = ctypes.c_int
        L.oracle_descriptor_single.argtypes = [P(OracleParams), u8p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_void_p, f32p]
        L.oracle_orientation_histogram_single.argtypes = [P(OracleParams), u8p, ctypes.c_int,
                                                          ctypes.c_int, ctypes.c_void_p, f32p]
        L.oracle_orientation_histogram_single.restype = ctypes.c_int
        L.oracle_smooth_orientation_histogram.argtypes = [f32p]
        L.oracle_orientation_peaks.argtypes = [f32p, f32p]
        L.oracle_orientation_peaks.restype = ctypes.c_int
        L.oracle_descriptor_raw_single.argtypes = [P(OracleParams), u8p, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_void_p, f32p]
        L.oracle_refine_single.argtypes = [P(OracleParams), f32p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_refine_single.restype = ctypes.c_int
        L.oracle_to_luma8.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                      u8p]
        L.oracle_descriptor_u8.argtypes = [f32p, ctypes.c_size_t, u8p]
        L.oracle_draw_keypoints.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                            ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, u8p,
                                            u8p]
        _LIB = L
    return _LIB


def _u8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _f32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def params(sigma=1.6, num_octaves=4, num_intervals=3, assumed_blur=0.5,
           contrast_threshold=0.04, edge_threshold=10.0):
    """Sift::new(...) of the reference (contrast stored divided by num_intervals)."""
    p = OracleParams()
    lib().oracle_params_new(ctypes.byref(p), sigma, num_octaves, num_intervals, assumed_blur,
                            contrast_threshold, edge_threshold)
    return p


def gaussian_kernel(sigma):
    n = lib().oracle_gaussian_kernel(sigma, None, 0)
    k = np.zeros(n, np.float32)
    lib().oracle_gaussian_kernel(sigma, _f32(k), n)
    return k


def gaussian_blur(img, sigma):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.empty_like(img)
    lib().oracle_gaussian_blur(_u8(img), img.shape[1], img.shape[0], sigma, _u8(out))
    return out


def resize_lanczos3(img, nw, nh):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.zeros((nh, nw), np.uint8)
    lib().oracle_resize_lanczos3(_u8(img), img.shape[1], img.shape[0], _u8(out), nw, nh)
    return out


def lanczos3_table(n_in, n_out):
    L = lib()
    m = L.oracle_lanczos3_table(n_in, n_out, None, None, None, 0)
    left = np.zeros(n_out, np.int32)
    cnt = np.zeros(n_out, np.int32)
    w = np.zeros((n_out, m), np.float32)
    L.oracle_lanczos3_table(n_in, n_out, left.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _f32(w), m)
    return left, cnt, w


def level_sigmas(p):
    L = p.num_intervals + 3
    b = np.zeros(p.num_octaves * L, np.float32)
    s0 = ctypes.c_float()
    lib().oracle_level_sigmas(ctypes.byref(p), _f32(b), ctypes.byref(s0))
    return b.reshape(p.num_octaves, L), s0.value


def octave_shapes(p, w, h):
    out = []
    for _ in range(p.num_octaves):
        out.append((w, h))
        w //= 2
        h //= 2
    return out


def build_pyramid(p, img):
    """Returns (gauss, dog): lists over octaves of arrays (L,H,W) u8 / (L-1,H,W) f32."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    L = lib()
    g = np.zeros(L.oracle_pyramid_size(ctypes.byref(p), w, h), np.uint8)
    d = np.zeros(L.oracle_dog_size(ctypes.byref(p), w, h), np.float32)
    L.oracle_build_pyramid(ctypes.byref(p), _u8(img), w, h, _u8(g), _f32(d))
    gs, ds = [], []
    go = do = 0
    nl = p.num_intervals + 3
    for (ww, hh) in octave_shapes(p, w, h):
        a = ww * hh
        gs.append(g[go:go + nl * a].reshape(nl, hh, ww))
        ds.append(d[do:do + (nl - 1) * a].reshape(nl - 1, hh, ww))
        go += nl * a
        do += (nl - 1) * a
    return gs, ds


def _take_kps(ptr, n):
    if n == 0:
        if ptr:
            lib().oracle_free(ptr)
        return np.zeros(0, KEYPOINT_DTYPE)
    buf = (ctypes.c_char * (n * KEYPOINT_DTYPE.itemsize)).from_address(ptr)
    arr = np.frombuffer(bytes(buf), KEYPOINT_DTYPE).copy()
    lib().oracle_free(ptr)
    return arr


def detect_and_compute(p, img, nthreads=0, stages=False):
    """detect_and_compute_cpu of the reference. Returns (kps, desc[n,128]) and,
    with stages=True, also (candidates, refined, stage_ms[6])."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    cand = ctypes.c_void_p()
    nc = ctypes.c_size_t()
    ref = ctypes.c_void_p()
    nr = ctypes.c_size_t()
    kp = ctypes.c_void_p()
    de = ctypes.c_void_p()
    nk = ctypes.c_size_t()
    ms = (ctypes.c_double * 6)()
    lib().oracle_detect_and_compute(ctypes.byref(p), _u8(img), w, h, nthreads,
                                    ctypes.byref(cand), ctypes.byref(nc), ctypes.byref(ref),
                                    ctypes.byref(nr), ctypes.byref(kp), ctypes.byref(de),
                                    ctypes.byref(nk), ms)
    n = nk.value
    kps = _take_kps(kp.value, n)
    if n:
        buf = (ctypes.c_float * (n * 128)).from_address(de.value)
        desc = np.frombuffer(buf, np.float32).reshape(n, 128).copy()
    else:
        desc = np.zeros((0, 128), np.float32)
    if de.value:
        lib().oracle_free(de.value)
    c = _take_kps(cand.value, nc.value)
    r = _take_kps(ref.value, nr.value)
    if stages:
        return kps, desc, c, r, list(ms)
    return kps, desc


# pixel formats (the SIFT_PIXEL_* values of the C ABI)
GRAY8, RGB8, RGBA8, LUMAA8, RGBA8_CANVAS = 0, 1, 2, 3, 4
_BPP = {GRAY8: 1, RGB8: 3, RGBA8: 4, LUMAA8: 2, RGBA8_CANVAS: 4}


def to_luma8(img, fmt):
    """Grey image of an (H, W, C) u8 array in pixel format `fmt`: to_luma8 of
    image 0.25.6 (src/sift.rs:1202) or the wasm canvas rule (src/wasm.rs:79-94)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    if img.size != h * w * _BPP[fmt]:
        raise ValueError("array shape does not match the pixel format")
    out = np.empty((h, w), np.uint8)
    lib().oracle_to_luma8(_u8(img), w, h, w * _BPP[fmt], fmt, _u8(out))
    return out


def draw_keypoints(img, fmt, kps, color=(255, 0, 0)):
    """draw_keypoints_to_image (src/lib.rs:70-116) of an (H, W[, C]) u8 image."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    out = np.empty((h, w, 3), np.uint8)
    kps = np.ascontiguousarray(kps)
    col = np.array(color, np.uint8)
    lib().oracle_draw_keypoints(_u8(img), w, h, w * _BPP[fmt], fmt,
                                kps.ctypes.data if len(kps) else None, len(kps), _u8(col), _u8(out))
    return out


def descriptor_u8(desc):
    """u8 descriptors of the WebGPU backends (src/shaders/descriptor.wgsl:256-262)."""
    d = np.ascontiguousarray(desc, np.float32)
    out = np.empty(d.shape, np.uint8)
    lib().oracle_descriptor_u8(_f32(d), d.size, _u8(out))
    return out


def solve_linear_system(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(9)
    b = np.ascontiguousarray(b, np.float32).reshape(3)
    x = np.zeros(3, np.float32)
    ok = lib().oracle_solve_linear_system(_f32(a), _f32(b), _f32(x))
    return x if ok else None


def gauss_pixel_bilinear(img, x, y):
    img = np.ascontiguousarray(img, np.uint8)
    return lib().oracle_gauss_pixel_bilinear(_u8(img), img.shape[1], img.shape[0], x, y)


def normalize_and_clip_descriptor(d):
    d = np.ascontiguousarray(d, np.float32).copy()
    lib().oracle_normalize_and_clip_descriptor(_f32(d), d.size)
    return d


def _kp_struct(x, y, size, angle=0.0, layer=1):
    k = np.zeros(1, KEYPOINT_DTYPE)
    k["x"], k["y"], k["size"], k["angle"], k["layer"] = x, y, size, angle, layer
    return k


def orientation_single(p, g, x, y, size, layer=1):
    """assign_orientations for one keypoint on level image g (octave 0)."""
    g = np.ascontiguousarray(g, np.uint8)
    k = _kp_struct(x, y, size, 0.0, layer)
    out = np.zeros(36, np.float32)
    n = lib().oracle_orientation_single(ctypes.byref(p), _u8(g), g.shape[1], g.shape[0],
                                        k.ctypes.data, _f32(out))
    return out[:n].copy()


def descriptor_single(p, g, x, y, size, angle, layer=1):
    """compute (descriptor) for one keypoint on level image g (octave 0)."""
    g = np.ascontiguousarray(g, np.uint8)
    k = _kp_struct(x, y, size, angle, layer)
    out = np.zeros(128, np.float32)
    lib().oracle_descriptor_single(ctypes.byref(p), _u8(g), g.shape[1], g.shape[0],
                                   k.ctypes.data, _f32(out))
    return out


def orientation_histogram_single(p, g, x, y, size, layer=1):
    """The raw 36-bin orientation histogram of one keypoint (before smoothing)."""
    g = np.ascontiguousarray(g, np.uint8)
    k = _kp_struct(x, y, size, 0.0, layer)
    out = np.zeros(36, np.float32)
    lib().oracle_orientation_histogram_single(ctypes.byref(p), _u8(g), g.shape[1], g.shape[0],
                                              k.ctypes.data, _f32(out))
    return out


def smooth_orientation_histogram(hist):
    h = np.ascontiguousarray(hist, np.float32).copy()
    assert h.size == 36
    lib().oracle_smooth_orientation_histogram(_f32(h))
    return h


def orientation_peaks(hist):
    h = np.ascontiguousarray(hist, np.float32)
    out = np.zeros(36, np.float32)
    n = lib().oracle_orientation_peaks(_f32(h), _f32(out))
    return out[:n].copy()


def descriptor_raw_single(p, g, x, y, size, angle, layer=1):
    """The 4 x 4 x 8 descriptor histogram of one keypoint before normalisation,
    indexed [(iy * 4 + ix) * 8 + ia]."""
    g = np.ascontiguousarray(g, np.uint8)
    k = _kp_struct(x, y, size, angle, layer)
    out = np.zeros(128, np.float32)
    lib().oracle_descriptor_raw_single(ctypes.byref(p), _u8(g), g.shape[1], g.shape[0],
                                       k.ctypes.data, _f32(out))
    return out


def refine_single(p, dog, x, y, layer, octave=0):
    """Refinement of one candidate at octave-pixel (x, y) -- placed as the
    extremum search places it, ((x + 0.5) * 2^octave) -- on DoG layers
    dog[nd, h, w]. Returns the refined keypoint or None."""
    dog = np.ascontiguousarray(dog, np.float32)
    nd, h, w = dog.shape
    c = np.zeros(1, KEYPOINT_DTYPE)
    sc = np.float32(2.0) ** octave
    c["x"] = (np.float32(x) + np.float32(0.5)) * sc
    c["y"] = (np.float32(y) + np.float32(0.5)) * sc
    c["layer"], c["octave"] = layer, octave
    out = np.zeros(1, KEYPOINT_DTYPE)
    ok = lib().oracle_refine_single(ctypes.byref(p), _f32(dog), w, h, nd, c.ctypes.data,
                                    out.ctypes.data)
    return out[0] if ok else None
