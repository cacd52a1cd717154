"""Seeded synthetic grey-scale images (ctypes binding of csrc/imgen.c).

Workloads (the shapes, counts and seeds the tests and the benchmark use):
  512x512     200 blobs,  50 rects, noise 0.01, seed 1, 4 octaves
  1920x1080   800 blobs, 200 rects, noise 0.01, seed 2, 5 octaves
  3840x2160   3000 blobs, 800 rects, noise 0.02, seed 3, 6 octaves, contrast 0.03
  batch 64 images, sizes uniform over 640x480/1280x720/1920x1080/2048x1536,
        image i seed 1000+i, octaves floor(log2(min(w,h)))-3
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsift_imgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.sift_synth_image.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.c_uint8)]
        L.sift_synth_image.restype = ctypes.c_int
        _LIB = L
    return _LIB


def synth_image(w, h, n_blobs, n_rects, noise_sigma, seed):
    out = np.empty((h, w), np.uint8)
    rc = _lib().sift_synth_image(w, h, n_blobs, n_rects, noise_sigma, seed,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if rc != 0:
        raise ValueError(f"sift_synth_image failed ({rc})")
    return out


# name -> (w, h, blobs, rects, noise, seed, num_octaves, contrast_threshold)
CONFIGS = {
    "512x512": (512, 512, 200, 50, 0.01, 1, 4, 0.04),
    "1920x1080": (1920, 1080, 800, 200, 0.01, 2, 5, 0.04),
    "3840x2160": (3840, 2160, 3000, 800, 0.02, 3, 6, 0.03),
}

BATCH_SHAPES = [(640, 480), (1280, 720), (1920, 1080), (2048, 1536)]


def config_image(name):
    w, h, b, r, n, seed, octv, ct = CONFIGS[name]
    return synth_image(w, h, b, r, n, seed), octv, ct


def octaves_for(w, h):
    """floor(log2(min(w, h))) - 3, the per-size octave rule of the batched workload."""
    return int(math.floor(math.log2(min(w, h)))) - 3


def generic_image(w, h, seed, n_blobs=None, n_rects=None, noise=0.01):
    """The same generator scaled by area (200 blobs / 50 rects per 512x512)."""
    a = w * h / (512.0 * 512.0)
    nb = int(round(200 * a)) if n_blobs is None else n_blobs
    nr = int(round(50 * a)) if n_rects is None else n_rects
    return synth_image(w, h, nb, nr, noise, seed)


def batch_shapes(n=64, seed=77):
    """Seeded uniform draw over BATCH_SHAPES for the batched workload."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(BATCH_SHAPES), size=n)
    return [BATCH_SHAPES[i] for i in idx]


# The served workload: an open-loop Poisson stream with sizes in a 40/30/20/10% mix;
# request j's frame is generic_image(w, h, seed=100000 + j), octaves floor(log2 min) - 3
# (8 for 3840x2160).
SERVED_SIZES = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
SERVED_MIX = [0.4, 0.3, 0.2, 0.1]


def served_stream(n, rate=1000.0, seed=7):
    """(arrival times in s, size index per request) of requests 0..n-1."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    which = rng.choice(len(SERVED_SIZES), size=n, p=SERVED_MIX)
    return arrivals, which


def served_frames(which, per_size):
    """A rotating pool of distinct frames for a served stream: the first
    per_size[k] requests of size k get their own frame (seed 100000 + j); a later
    request of that size reuses the pooled frame of rank (its rank % per_size[k]).
    Returns (source request index per request, {source index: image})."""
    seen = [[] for _ in SERVED_SIZES]
    rank = [0] * len(SERVED_SIZES)
    src = np.empty(len(which), np.int64)
    for j, k in enumerate(which):
        pool = seen[k]
        if len(pool) < per_size[k]:
            pool.append(j)
            src[j] = j
        else:
            src[j] = pool[rank[k] % per_size[k]]
        rank[k] += 1
    frames = {}
    for k, pool in enumerate(seen):
        w, h = SERVED_SIZES[k]
        for j in pool:
            frames[j] = generic_image(w, h, seed=100000 + j)
    return src, frames
