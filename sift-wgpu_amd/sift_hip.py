"""Python mirror of the reference's SIFT API over the C ABI of libsift_hip.so.

Mirrors (stevel705/sift-wgpu):
  Sift, Sift.new, Sift.default        src/sift.rs:36-80
  Sift.detect_and_compute_with_backend src/sift.rs:1244-1355
  SiftBackend + from_str aliases       src/lib.rs:18-49 (plus "hip")
  KeyPoint                             src/keypoints.rs:1-10
  colour input (to_luma8 on the device) src/sift.rs:1202, src/wasm.rs:79-103
  u8 descriptors / V2 keypoint view    src/shaders/descriptor.wgsl:256-262,
                                       src/sift.rs:1288-1291, src/gpu_sift_v2.rs:1699-1707

Only the Hip backend is implemented here; the reference's Cpu / WebGpu arms
live in the Rust crate and are not reachable from this library, so asking for
them raises SiftError (no silent CPU fallback). The library must be built
(__graft_entry__.build()); importing works without a GPU, but creating a
context on a machine without a HIP device raises SiftError.
"""
import ctypes
import enum
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsift_hip.so")

KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                           ("response", "<f4"), ("octave", "<i4"), ("layer", "<i4")])

SIFT_OK = 0


class SiftError(RuntimeError):
    pass


class SiftParams(ctypes.Structure):
    _fields_ = [("sigma", ctypes.c_float), ("num_octaves", ctypes.c_uint32),
                ("num_intervals", ctypes.c_uint32), ("assumed_blur", ctypes.c_float),
                ("contrast_threshold", ctypes.c_float), ("edge_threshold", ctypes.c_float),
                ("image_border_width", ctypes.c_uint32)]


class SiftImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("stride", ctypes.c_uint32)]


class SiftResult(ctypes.Structure):
    _fields_ = [("keypoints", ctypes.c_void_p), ("descriptors", ctypes.POINTER(ctypes.c_float)),
                ("count", ctypes.c_size_t)]


class SiftColorImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("stride", ctypes.c_uint32),
                ("format", ctypes.c_uint32)]


class SiftResultU8(ctypes.Structure):
    _fields_ = [("keypoints", ctypes.c_void_p), ("descriptors", ctypes.POINTER(ctypes.c_uint8)),
                ("count", ctypes.c_size_t)]


# pixel formats of SiftColorImage (SIFT_PIXEL_* in the C header)
PIXEL_GRAY8, PIXEL_RGB8, PIXEL_RGBA8, PIXEL_LUMAA8, PIXEL_RGBA8_CANVAS = 0, 1, 2, 3, 4
PIXEL_BYTES = {PIXEL_GRAY8: 1, PIXEL_RGB8: 3, PIXEL_RGBA8: 4, PIXEL_LUMAA8: 2,
               PIXEL_RGBA8_CANVAS: 4}
# keypoint conventions of detect_batch_u8 (SIFT_KEYPOINTS_* in the C header)
KEYPOINTS_CPU, KEYPOINTS_WGPU_V2 = 0, 1


class SiftStages(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_void_p), ("n_candidates", ctypes.c_size_t),
                ("refined", ctypes.c_void_p), ("n_refined", ctypes.c_size_t),
                ("result", SiftResult)]


class SiftKernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64),
                ("total_ms", ctypes.c_double), ("bytes", ctypes.c_double),
                ("ops", ctypes.c_double)]


class SiftServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch", ctypes.c_uint32),
                ("max_queue", ctypes.c_uint32)]


class SiftCompletion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("status", ctypes.c_int), ("result", SiftResult),
                ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]


# Every symbol include/sift_hip.h declares (checked by the CPU test suite).
EXPORTED_SYMBOLS = [
    "sift_params_default", "sift_params_new", "sift_hip_create", "sift_hip_destroy",
    "sift_hip_last_error", "sift_hip_detect", "sift_hip_detect_batch",
    "sift_hip_detect_batch_device", "sift_result_free", "sift_hip_pyramid",
    "sift_hip_detect_stages", "sift_stages_free", "sift_hip_profile", "sift_hip_profile_filter",
    "sift_hip_profile_read",
    "sift_hip_server_start", "sift_hip_submit", "sift_hip_poll", "sift_hip_server_stop",
    "sift_hip_to_luma8", "sift_hip_detect_color", "sift_hip_detect_batch_color",
    "sift_hip_detect_batch_u8", "sift_result_u8_free", "sift_hip_submit_color",
    "sift_draw_keypoints", "sift_hip_detect_batch_device_noalloc", "sift_hip_tune_blur",
]

# Bytes of the per-GPU scratch of sift_hip_detect_batch_device_noalloc.
SCRATCH_BYTES = 256 << 20

_LIB = None
_LIB_LOCK = threading.Lock()


def lib():
    """Loads libsift_hip.so; raises SiftError if it is not built."""
    global _LIB
    with _LIB_LOCK:
        if _LIB is not None:
            return _LIB
        if not os.path.exists(LIB_PATH):
            raise SiftError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.sift_params_default.argtypes = [P(SiftParams)]
        L.sift_params_new.argtypes = [P(SiftParams), ctypes.c_float, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float]
        L.sift_hip_create.argtypes = [P(SiftParams), ctypes.c_int, P(ctypes.c_void_p)]
        L.sift_hip_create.restype = ctypes.c_int
        L.sift_hip_destroy.argtypes = [ctypes.c_void_p]
        L.sift_hip_last_error.restype = ctypes.c_char_p
        L.sift_hip_detect.argtypes = [ctypes.c_void_p, P(SiftImage), P(SiftResult)]
        L.sift_hip_detect.restype = ctypes.c_int
        L.sift_hip_detect_batch.argtypes = [ctypes.c_void_p, P(SiftImage), P(SiftParams),
                                            ctypes.c_size_t, P(SiftResult)]
        L.sift_hip_detect_batch.restype = ctypes.c_int
        L.sift_hip_detect_batch_device.argtypes = [
            ctypes.c_void_p, P(SiftImage), P(SiftParams), ctypes.c_size_t, P(ctypes.c_size_t),
            P(ctypes.c_void_p), P(ctypes.c_void_p), P(ctypes.c_void_p)]
        L.sift_hip_detect_batch_device.restype = ctypes.c_int
        L.sift_hip_detect_batch_device_noalloc.argtypes = [
            ctypes.c_void_p, P(SiftImage), P(SiftParams), ctypes.c_size_t, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, P(ctypes.c_size_t)]
        L.sift_hip_detect_batch_device_noalloc.restype = ctypes.c_int
        L.sift_hip_tune_blur.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_float, ctypes.c_uint32, ctypes.c_char_p,
                                         ctypes.c_char_p, ctypes.c_uint32, P(ctypes.c_double),
                                         P(ctypes.c_int)]
        L.sift_hip_tune_blur.restype = ctypes.c_int
        L.sift_result_free.argtypes = [P(SiftResult)]
        L.sift_hip_pyramid.argtypes = [ctypes.c_void_p, P(SiftImage), P(ctypes.c_uint8),
                                       P(ctypes.c_float)]
        L.sift_hip_pyramid.restype = ctypes.c_int
        L.sift_hip_detect_stages.argtypes = [ctypes.c_void_p, P(SiftImage), P(SiftStages)]
        L.sift_hip_detect_stages.restype = ctypes.c_int
        L.sift_stages_free.argtypes = [P(SiftStages)]
        L.sift_hip_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.sift_hip_profile.restype = ctypes.c_int
        L.sift_hip_profile_filter.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.sift_hip_profile_filter.restype = ctypes.c_int
        L.sift_hip_profile_read.argtypes = [ctypes.c_void_p, P(SiftKernelStat), ctypes.c_size_t,
                                            P(ctypes.c_size_t)]
        L.sift_hip_profile_read.restype = ctypes.c_int
        L.sift_hip_server_start.argtypes = [ctypes.c_void_p, P(SiftServerConfig)]
        L.sift_hip_server_start.restype = ctypes.c_int
        L.sift_hip_submit.argtypes = [ctypes.c_void_p, P(SiftImage), P(SiftParams), ctypes.c_uint64]
        L.sift_hip_submit.restype = ctypes.c_int
        L.sift_hip_poll.argtypes = [ctypes.c_void_p, P(SiftCompletion), ctypes.c_size_t,
                                    P(ctypes.c_size_t), ctypes.c_int]
        L.sift_hip_poll.restype = ctypes.c_int
        L.sift_hip_server_stop.argtypes = [ctypes.c_void_p]
        L.sift_hip_server_stop.restype = ctypes.c_int
        L.sift_hip_to_luma8.argtypes = [ctypes.c_void_p, P(SiftColorImage), P(ctypes.c_uint8)]
        L.sift_hip_to_luma8.restype = ctypes.c_int
        L.sift_hip_detect_color.argtypes = [ctypes.c_void_p, P(SiftColorImage), P(SiftResult)]
        L.sift_hip_detect_color.restype = ctypes.c_int
        L.sift_hip_detect_batch_color.argtypes = [ctypes.c_void_p, P(SiftColorImage), P(SiftParams),
                                                  ctypes.c_size_t, P(SiftResult)]
        L.sift_hip_detect_batch_color.restype = ctypes.c_int
        L.sift_hip_detect_batch_u8.argtypes = [ctypes.c_void_p, P(SiftColorImage), P(SiftParams),
                                               ctypes.c_size_t, ctypes.c_uint32, P(SiftResultU8)]
        L.sift_hip_detect_batch_u8.restype = ctypes.c_int
        L.sift_result_u8_free.argtypes = [P(SiftResultU8)]
        L.sift_hip_submit_color.argtypes = [ctypes.c_void_p, P(SiftColorImage), P(SiftParams),
                                            ctypes.c_uint64]
        L.sift_hip_submit_color.restype = ctypes.c_int
        L.sift_draw_keypoints.argtypes = [P(SiftColorImage), ctypes.c_void_p, ctypes.c_size_t,
                                          P(ctypes.c_uint8), P(ctypes.c_uint8)]
        L.sift_draw_keypoints.restype = ctypes.c_int
        _LIB = L
        return L


def last_error():
    return lib().sift_hip_last_error().decode(errors="replace")


def _check(rc, what):
    if rc != SIFT_OK:
        raise SiftError(f"{what} failed ({rc}): {last_error()}")


class SiftBackend(enum.Enum):
    """src/lib.rs:18-28 plus the Hip arm this library provides."""
    Cpu = "cpu"
    WebGpu = "webgpu"
    WebGpuWithCpuFallback = "auto"
    WebGpuV2 = "gpuv2"
    Hip = "hip"

    @staticmethod
    def from_str(value):
        """FromStr aliases of src/lib.rs:36-49, plus "hip" / "rocm"."""
        v = value.strip().lower()
        table = {"cpu": SiftBackend.Cpu,
                 "webgpu": SiftBackend.WebGpu, "wgpu": SiftBackend.WebGpu, "gpu": SiftBackend.WebGpu,
                 "auto": SiftBackend.WebGpuWithCpuFallback,
                 "fallback": SiftBackend.WebGpuWithCpuFallback,
                 "prefer-gpu": SiftBackend.WebGpuWithCpuFallback,
                 "gpuv2": SiftBackend.WebGpuV2, "gpu-v2": SiftBackend.WebGpuV2,
                 "webgpu-v2": SiftBackend.WebGpuV2, "v2": SiftBackend.WebGpuV2,
                 "hip": SiftBackend.Hip, "rocm": SiftBackend.Hip}
        if v not in table:
            raise ValueError(f"Unknown SIFT backend: {value}")
        return table[v]


def params_struct(sigma, num_octaves, num_intervals, assumed_blur, contrast_threshold,
                  edge_threshold):
    p = SiftParams()
    lib().sift_params_new(ctypes.byref(p), sigma, num_octaves, num_intervals, assumed_blur,
                          contrast_threshold, edge_threshold)
    return p


class _Ctx:
    """Process-global context per device (the Rust arm keeps one in a OnceLock)."""
    _lock = threading.Lock()
    _ctx = {}

    @classmethod
    def get(cls, device=0):
        with cls._lock:
            if device not in cls._ctx:
                p = SiftParams()
                lib().sift_params_default(ctypes.byref(p))
                h = ctypes.c_void_p()
                _check(lib().sift_hip_create(ctypes.byref(p), device, ctypes.byref(h)),
                       "sift_hip_create")
                cls._ctx[device] = h
            return cls._ctx[device]


def _image_struct(img):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    if img.ndim != 2:
        raise ValueError("expected a 2-D u8 grey image")
    h, w = img.shape
    return img, SiftImage(img.ctypes.data, w, h, w)


def _color_struct(img, fmt=None):
    """(array kept alive, SiftColorImage). 2-D arrays are grey; (H, W, 3) RGB8,
    (H, W, 4) RGBA8 (or fmt=PIXEL_RGBA8_CANVAS), (H, W, 2) LumaA8."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    if fmt is None:
        if img.ndim == 2:
            fmt = PIXEL_GRAY8
        elif img.ndim == 3 and img.shape[2] in (2, 3, 4):
            fmt = {2: PIXEL_LUMAA8, 3: PIXEL_RGB8, 4: PIXEL_RGBA8}[img.shape[2]]
        else:
            raise ValueError("expected (H, W) grey or (H, W, 2|3|4) u8 pixels")
    if fmt not in PIXEL_BYTES:
        raise ValueError(f"unknown pixel format {fmt}")
    h, w = img.shape[:2]
    if img.size != h * w * PIXEL_BYTES[fmt]:
        raise ValueError("array shape does not match the pixel format")
    return img, SiftColorImage(img.ctypes.data, w, h, w * PIXEL_BYTES[fmt], fmt)


def to_luma8(img, fmt=None, device=0):
    """Grey u8 image of a colour image, converted on the device (to_luma8,
    src/sift.rs:1202; PIXEL_RGBA8_CANVAS: src/wasm.rs:79-94)."""
    ctx = _Ctx.get(device)
    a, s = _color_struct(img, fmt)
    out = np.empty((s.height, s.width), np.uint8)
    _check(lib().sift_hip_to_luma8(ctx, ctypes.byref(s),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
           "sift_hip_to_luma8")
    return out


def draw_keypoints_to_image(img, keypoints, color=(255, 0, 0), fmt=None):
    """draw_keypoints_to_image (src/lib.rs:70-116): an (H, W, 3) u8 RGB image with
    every keypoint drawn in `color` (host code of the library, no device work)."""
    a, s = _color_struct(img, fmt)
    kps = np.ascontiguousarray(keypoints, KEYPOINT_DTYPE)
    out = np.empty((s.height, s.width, 3), np.uint8)
    col = (ctypes.c_uint8 * 3)(*color)
    _check(lib().sift_draw_keypoints(ctypes.byref(s), kps.ctypes.data if len(kps) else None,
                                     len(kps), col,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
           "sift_draw_keypoints")
    return out


def wgpu_descriptor_view(desc_u8):
    """The f32 form the reference's WebGPU arms hand to callers: u8 / 255
    (src/sift.rs:1288-1291)."""
    return np.asarray(desc_u8, np.uint8).astype(np.float32) / np.float32(255.0)


class _ResultOwner:
    """Owns a result's C buffers (malloc'd by the library) and frees them once
    the last array over them is gone."""

    def __init__(self, r, free):
        self._r, self._free = r, free

    def __del__(self):
        self._free(ctypes.byref(self._r))


def _result_arrays(r, free, dtype):
    """(keypoints, descriptors) of a sift_result_t / sift_result_u8_t as numpy
    arrays over the library's own buffers (no copy); the buffers are released by
    `free` when both arrays and every view of them are gone."""
    n = int(r.count)
    if n == 0:
        free(ctypes.byref(r))
        return np.zeros(0, KEYPOINT_DTYPE), np.zeros((0, 128), dtype)
    owner = _ResultOwner(type(r).from_buffer_copy(r), free)
    kb = (ctypes.c_char * (n * KEYPOINT_DTYPE.itemsize)).from_address(r.keypoints)
    db = (ctypes.c_char * (n * 128 * np.dtype(dtype).itemsize)).from_address(
        ctypes.cast(r.descriptors, ctypes.c_void_p).value)
    kb._owner = owner
    db._owner = owner
    return np.frombuffer(kb, KEYPOINT_DTYPE), np.frombuffer(db, dtype).reshape(n, 128)


def _take_result_u8(r):
    return _result_arrays(r, lib().sift_result_u8_free, np.uint8)


def _take_result(r):
    return _result_arrays(r, lib().sift_result_free, np.float32)


def _take_kps(ptr, n):
    if n == 0 or not ptr:
        return np.zeros(0, KEYPOINT_DTYPE)
    kb = (ctypes.c_char * (n * KEYPOINT_DTYPE.itemsize)).from_address(ptr)
    return np.frombuffer(bytes(kb), KEYPOINT_DTYPE).copy()


class Sift:
    """`Sift` of src/sift.rs:36-80. contrast_threshold is the stored value
    (divided by num_intervals in new(), as the reference does)."""

    def __init__(self, sigma=1.6, num_octaves=4, num_intervals=3, assumed_blur=0.5,
                 contrast_threshold=0.04, edge_threshold=10.0, device=0):
        self.sigma = float(sigma)
        self.num_octaves = int(num_octaves)
        self.num_intervals = int(num_intervals)
        self.assumed_blur = float(assumed_blur)
        self.contrast_threshold = float(np.float32(contrast_threshold) / np.float32(num_intervals))
        self.edge_threshold = float(edge_threshold)
        self.image_border_width = 5
        self.device = device

    @classmethod
    def new(cls, sigma, num_octaves, num_intervals, assumed_blur, contrast_threshold,
            edge_threshold):
        return cls(sigma, num_octaves, num_intervals, assumed_blur, contrast_threshold,
                   edge_threshold)

    @classmethod
    def default(cls):
        return cls()

    def params(self):
        p = SiftParams()
        p.sigma = self.sigma
        p.num_octaves = self.num_octaves
        p.num_intervals = self.num_intervals
        p.assumed_blur = self.assumed_blur
        p.contrast_threshold = self.contrast_threshold
        p.edge_threshold = self.edge_threshold
        p.image_border_width = self.image_border_width
        return p

    # -- the reference's entry points ---------------------------------------
    def detect_and_compute_with_backend(self, img, backend):
        if isinstance(backend, str):
            backend = SiftBackend.from_str(backend)
        if backend is not SiftBackend.Hip:
            raise SiftError(f"backend {backend.name} is not provided by libsift_hip "
                            "(only SiftBackend.Hip is)")
        return self.detect_batch([img])[0]

    def detect_and_compute(self, img):
        return self.detect_and_compute_with_backend(img, SiftBackend.Hip)

    # -- batched / device / introspection ------------------------------------
    def _params_array(self, n, params):
        pa = (SiftParams * n)()
        for i in range(n):
            p = params[i] if params is not None else self
            pa[i] = p.params() if isinstance(p, Sift) else p
        return pa

    def detect_batch(self, images, params=None, formats=None):
        """images: grey (H, W) arrays, or colour arrays (see _color_struct; formats
        may name PIXEL_RGBA8_CANVAS per image), converted to grey on the device."""
        ctx = _Ctx.get(self.device)
        n = len(images)
        pa = self._params_array(n, params)
        outs = (SiftResult * n)()
        keep = []
        if formats is None and all(np.ndim(im) == 2 for im in images):
            ims = (SiftImage * n)()
            for i, im in enumerate(images):
                a, s = _image_struct(im)
                keep.append(a)
                ims[i] = s
            _check(lib().sift_hip_detect_batch(ctx, ims, pa, n, outs), "sift_hip_detect_batch")
        else:
            ims = (SiftColorImage * n)()
            for i, im in enumerate(images):
                a, s = _color_struct(im, None if formats is None else formats[i])
                keep.append(a)
                ims[i] = s
            _check(lib().sift_hip_detect_batch_color(ctx, ims, pa, n, outs),
                   "sift_hip_detect_batch_color")
        return [_take_result(outs[i]) for i in range(n)]

    def detect_batch_u8(self, images, params=None, keypoint_convention=KEYPOINTS_CPU,
                        formats=None):
        """Keypoints and u8 descriptors as the reference's WebGPU backends produce
        them (src/shaders/descriptor.wgsl:256-262); KEYPOINTS_WGPU_V2 reports the
        keypoints as the V2 backend does (src/gpu_sift_v2.rs:1699-1707)."""
        ctx = _Ctx.get(self.device)
        n = len(images)
        pa = self._params_array(n, params)
        ims = (SiftColorImage * n)()
        keep = []
        for i, im in enumerate(images):
            a, s = _color_struct(im, None if formats is None else formats[i])
            keep.append(a)
            ims[i] = s
        outs = (SiftResultU8 * n)()
        _check(lib().sift_hip_detect_batch_u8(ctx, ims, pa, n, keypoint_convention, outs),
               "sift_hip_detect_batch_u8")
        return [_take_result_u8(outs[i]) for i in range(n)]

    def detect_batch_device(self, images, params=None, ctx=None):
        """images: list of (device_ptr, w, h, stride). Returns (counts, kp_ptr,
        desc_ptr, offsets_ptr); results stay on the device. ctx: a context from
        new_context() (default: the process-global one)."""
        ctx = ctx if ctx is not None else _Ctx.get(self.device)
        n = len(images)
        ims = (SiftImage * n)()
        for i, (ptr, w, h, stride) in enumerate(images):
            ims[i] = SiftImage(ptr, w, h, stride)
        pa = (SiftParams * n)()
        for i in range(n):
            p = params[i] if params is not None else self
            pa[i] = p.params() if isinstance(p, Sift) else p
        counts = (ctypes.c_size_t * n)()
        kp = ctypes.c_void_p()
        de = ctypes.c_void_p()
        of = ctypes.c_void_p()
        _check(lib().sift_hip_detect_batch_device(ctx, ims, pa, n, counts, ctypes.byref(kp),
                                                  ctypes.byref(de), ctypes.byref(of)),
               "sift_hip_detect_batch_device")
        return list(counts), kp.value, de.value, of.value

    def detect_batch_device_noalloc(self, images, d_keypoints, d_descriptors, capacity,
                                    d_offsets, params=None):
        """sift_hip_detect_batch_device_noalloc: images as for detect_batch_device;
        results into the caller's device buffers (capacity keypoints of 28 B and
        descriptors of 512 B, n + 1 uint32 offsets). Returns (rc, counts, message):
        rc SIFT_OK, or an error code with counts the keypoints found (when the
        capacity was too small) and the library's message."""
        ctx = _Ctx.get(self.device)
        n = len(images)
        ims = (SiftImage * n)()
        for i, (ptr, w, h, stride) in enumerate(images):
            ims[i] = SiftImage(ptr, w, h, stride)
        pa = self._params_array(n, params)
        counts = (ctypes.c_size_t * n)()
        rc = lib().sift_hip_detect_batch_device_noalloc(ctx, ims, pa, n, d_keypoints, d_descriptors,
                                                         capacity, d_offsets, counts)
        return rc, list(counts), (last_error() if rc != SIFT_OK else "")

    def pyramid(self, img):
        """Gaussian (u8) and DoG (f32) pyramids: lists over octaves of arrays
        (S+3, H, W) and (S+2, H, W)."""
        ctx = _Ctx.get(self.device)
        a, s = _image_struct(img)
        h, w = a.shape
        L = self.num_intervals + 3
        shapes = []
        ww, hh = w, h
        for _ in range(self.num_octaves):
            shapes.append((ww, hh))
            ww //= 2
            hh //= 2
        gsz = sum(L * x * y for x, y in shapes)
        dsz = sum((L - 1) * x * y for x, y in shapes)
        g = np.zeros(max(gsz, 1), np.uint8)
        d = np.zeros(max(dsz, 1), np.float32)
        # the context's params are used for the pyramid: set them per call
        hh_ = ctypes.c_void_p()
        p = self.params()
        _check(lib().sift_hip_create(ctypes.byref(p), self.device, ctypes.byref(hh_)),
               "sift_hip_create")
        try:
            _check(lib().sift_hip_pyramid(hh_, ctypes.byref(s),
                                          g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                          d.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "sift_hip_pyramid")
        finally:
            lib().sift_hip_destroy(hh_)
        gs, ds = [], []
        go = do = 0
        for (x, y) in shapes:
            gs.append(g[go:go + L * x * y].reshape(L, y, x))
            ds.append(d[do:do + (L - 1) * x * y].reshape(L - 1, y, x))
            go += L * x * y
            do += (L - 1) * x * y
        return gs, ds

    def detect_stages(self, img):
        """(candidates, refined, keypoints, descriptors) of one image."""
        a, s = _image_struct(img)
        h_ = ctypes.c_void_p()
        p = self.params()
        _check(lib().sift_hip_create(ctypes.byref(p), self.device, ctypes.byref(h_)),
               "sift_hip_create")
        st = SiftStages()
        try:
            _check(lib().sift_hip_detect_stages(h_, ctypes.byref(s), ctypes.byref(st)),
                   "sift_hip_detect_stages")
            cand = _take_kps(st.candidates, int(st.n_candidates))
            ref = _take_kps(st.refined, int(st.n_refined))
            n = int(st.result.count)
            if n:
                kps = _take_kps(st.result.keypoints, n)
                desc = np.ctypeslib.as_array(st.result.descriptors, shape=(n * 128,)).copy()
                desc = desc.reshape(n, 128)
            else:
                kps = np.zeros(0, KEYPOINT_DTYPE)
                desc = np.zeros((0, 128), np.float32)
            lib().sift_stages_free(ctypes.byref(st))
        finally:
            lib().sift_hip_destroy(h_)
        return cand, ref, kps, desc


# blur forms of the tuning table (sift_hip_tune_blur)
BLUR_H_FORMS = ("h8", "h2", "hsmall", "hgeneric", "h16")
BLUR_V_FORMS = ("v", "vnarrow", "vsmall", "vgeneric", "vsmall8", "v8")


def tune_blur(width, height, sigma, jobs, h_form=None, v_form=None, reps=10, device=0):
    """(horizontal ms, vertical ms, horizontal form, vertical form) of one blur of
    sigma on `jobs` levels of width x height with the given forms (None = the
    built-in choice), or None when a form is not valid for the level."""
    ctx = _Ctx.get(device)
    out = (ctypes.c_double * 2)()
    forms = (ctypes.c_int * 2)()
    rc = lib().sift_hip_tune_blur(ctx, width, height, sigma, jobs,
                                  h_form.encode() if h_form else None,
                                  v_form.encode() if v_form else None, reps, out, forms)
    if rc == -1 and "not valid" in last_error():
        return None
    _check(rc, "sift_hip_tune_blur")
    return out[0], out[1], BLUR_H_FORMS[forms[0]], BLUR_V_FORMS[forms[1]]


def new_context(device=0):
    """A context of its own (sift_hip_create with the default params; it reads
    the tuning table as it is at this call); free it with free_context."""
    p = SiftParams()
    lib().sift_params_default(ctypes.byref(p))
    h = ctypes.c_void_p()
    _check(lib().sift_hip_create(ctypes.byref(p), device, ctypes.byref(h)), "sift_hip_create")
    return h


def free_context(h):
    lib().sift_hip_destroy(h)


def profile(enable=True, device=0, only=None):
    """sift_hip_profile: per-kernel HIP-event timing of the context's launches
    (resets the counters). only: time just that kernel's launches (e.g.
    "k_blur_h<8>"), so that the events add little to the timed work."""
    ctx = _Ctx.get(device)
    _check(lib().sift_hip_profile(ctx, 1 if enable else 0), "sift_hip_profile")
    _check(lib().sift_hip_profile_filter(ctx, only.encode() if only else None), "sift_hip_profile_filter")


def profile_read(device=0):
    """{kernel: (launches, total_ms, algorithmic_bytes, prescribed_f32_ops)} since
    profile(True)."""
    ctx = _Ctx.get(device)
    n = ctypes.c_size_t()
    _check(lib().sift_hip_profile_read(ctx, None, 0, ctypes.byref(n)), "sift_hip_profile_read")
    arr = (SiftKernelStat * max(n.value, 1))()
    _check(lib().sift_hip_profile_read(ctx, arr, n.value, ctypes.byref(n)), "sift_hip_profile_read")
    return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms), float(arr[i].bytes),
                                   float(arr[i].ops))
            for i in range(n.value)}


class SiftServer:
    """Continuous-batching request loop (sift_hip_server_start / submit / poll).

    submit() copies the frame and returns at once; poll() returns completions
    as (tag, status, keypoints, descriptors, submit_ms, start_ms, done_ms)."""

    def __init__(self, sift=None, hbm_budget_bytes=32 << 30, max_batch=64, max_queue=4096,
                 device=0):
        sift = sift or Sift.default()
        p = sift.params()
        self._ctx = ctypes.c_void_p()
        _check(lib().sift_hip_create(ctypes.byref(p), device, ctypes.byref(self._ctx)),
               "sift_hip_create")
        cfg = SiftServerConfig(hbm_budget_bytes, max_batch, max_queue)
        rc = lib().sift_hip_server_start(self._ctx, ctypes.byref(cfg))
        if rc != SIFT_OK:
            msg = last_error()
            lib().sift_hip_destroy(self._ctx)
            raise SiftError(f"sift_hip_server_start failed ({rc}): {msg}")
        self._keep = {}

    def submit(self, img, tag, sift=None, fmt=None):
        """Queues a frame (copied before returning); `sift` optionally gives this
        frame's parameters (a Sift), else the server's. Colour frames ((H, W, C)
        arrays, or fmt) are converted to grey on the device."""
        pp = ctypes.byref(sift.params()) if sift is not None else None
        if fmt is None and np.ndim(img) == 2:
            img = np.ascontiguousarray(img, dtype=np.uint8)
            h, w = img.shape
            s = SiftImage(img.ctypes.data, w, h, w)
            rc = lib().sift_hip_submit(self._ctx, ctypes.byref(s), pp, tag)
        else:
            img, s = _color_struct(img, fmt)
            rc = lib().sift_hip_submit_color(self._ctx, ctypes.byref(s), pp, tag)
        if rc != SIFT_OK:
            raise SiftError(f"sift_hip_submit failed ({rc}): {last_error()}")

    def poll(self, cap=256, timeout_ms=100):
        arr = (SiftCompletion * cap)()
        n = ctypes.c_size_t()
        lib().sift_hip_poll(self._ctx, arr, cap, ctypes.byref(n), timeout_ms)
        out = []
        for i in range(n.value):
            c = arr[i]
            kps, desc = _take_result(c.result) if c.status == SIFT_OK else (None, None)
            out.append((int(c.tag), int(c.status), kps, desc, c.submit_ms, c.start_ms, c.done_ms))
        return out

    def close(self):
        if self._ctx:
            lib().sift_hip_server_stop(self._ctx)
            lib().sift_hip_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
