"""sift_hip_detect_batch_device: frames already in device memory, results left
in device memory (the entry point the benchmark times), against the host-buffer
path (sift_hip_detect_batch) on the same frames and against the oracle.

Its device-input form differs from the host path: the initial blur reads the
caller's pointer with the caller's stride, no input copy is carved from the
arena, and the results come back through the returned device pointers and the
per-image keypoint offsets. Everything read back here goes through those
pointers and offsets.
"""
import numpy as np
import pytest

import devmem
import imgen
import oracle
import sift_hip
from parity import DESC_L2_MARGIN, DESC_L2_TOL, MIN_PR, match_keypoints

pytestmark = pytest.mark.gpu


def _read_back(n, counts, kp_ptr, desc_ptr, off_ptr):
    """Per-image (keypoints, descriptors) through the device pointers and the
    n + 1 keypoint offsets."""
    offs = devmem.download(off_ptr, 4 * (n + 1)).view(np.uint32).astype(np.int64)
    assert offs[0] == 0
    assert np.all(np.diff(offs) >= 0)
    assert list(np.diff(offs)) == list(counts)
    total = int(offs[-1])
    kps = devmem.download(kp_ptr, total * sift_hip.KEYPOINT_DTYPE.itemsize).view(sift_hip.KEYPOINT_DTYPE)
    desc = devmem.download(desc_ptr, total * 128 * 4).view(np.float32).reshape(total, 128)
    return [(kps[offs[i]:offs[i + 1]], desc[offs[i]:offs[i + 1]]) for i in range(n)]


def _device_detect(sift, imgs, params=None, strides=None):
    bufs, items = [], []
    for i, img in enumerate(imgs):
        b, it = devmem.upload_image(img, None if strides is None else strides[i])
        bufs.append(b)
        items.append(it)
    counts, kp, de, of = sift.detect_batch_device(items, params)
    got = _read_back(len(imgs), counts, kp, de, of)
    for b in bufs:
        b.free()
    return got


def _same(a, b):
    return a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()


def test_bench_step_64_frames_1080p():
    """The benchmark's step: 64 distinct 1920x1080 frames (seeds 2 + 1000 i), 5
    octaves, default thresholds, in one device-resident call; every frame equals
    the host path bit for bit, and the first frame matches the oracle."""
    w, h, b, r, noise, seed, octv, ct = imgen.CONFIGS["1920x1080"]
    imgs = [imgen.synth_image(w, h, b, r, noise, seed + 1000 * i) for i in range(64)]
    s = sift_hip.Sift(1.6, octv, 3, 0.5, ct, 10.0)
    got = _device_detect(s, imgs)
    host = s.detect_batch(imgs)
    for i in range(len(imgs)):
        assert len(got[i][0]) > 0, i
        assert _same(got[i], host[i]), i
    p = oracle.params(1.6, octv, 3, 0.5, ct, 10.0)
    okps, odesc = oracle.detect_and_compute(p, imgs[0])
    m = match_keypoints(okps, got[0][0], odesc, got[0][1])
    print(f"1080p frame 0: oracle {m['n_ref']} hip {m['n_got']} P {m['precision']:.5f} "
          f"R {m['recall']:.5f} desc_l2_max {m['desc_l2_max']:.2e}")
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
    assert m["desc_l2_max"] <= DESC_L2_TOL
    assert m["desc_l2_max"] <= DESC_L2_MARGIN


def test_uneven_batch_config():
    """The batched workload: 64 frames with sizes drawn from 640x480 / 1280x720 /
    1920x1080 / 2048x1536, frame i seed 1000 + i, octaves floor(log2(min)) - 3
    (5 / 6 / 7 / 7), in one device-resident call. Every frame equals the host
    path bit for bit; one frame of each size matches the oracle."""
    shapes = imgen.batch_shapes(64)
    assert set(shapes) == set(imgen.BATCH_SHAPES)
    imgs = [imgen.generic_image(w, h, seed=1000 + i) for i, (w, h) in enumerate(shapes)]
    params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h), 3, 0.5, 0.04, 10.0) for (w, h) in shapes]
    s = sift_hip.Sift.default()
    got = _device_detect(s, imgs, params)
    host = s.detect_batch(imgs, params)
    for i in range(len(imgs)):
        assert len(got[i][0]) > 0, i
        assert _same(got[i], host[i]), i
    for sz in imgen.BATCH_SHAPES:
        i = shapes.index(sz)
        p = oracle.params(1.6, imgen.octaves_for(*sz), 3, 0.5, 0.04, 10.0)
        okps, odesc = oracle.detect_and_compute(p, imgs[i])
        m = match_keypoints(okps, got[i][0], odesc, got[i][1])
        print(f"{sz[0]}x{sz[1]} frame {i}: oracle {m['n_ref']} hip {m['n_got']} "
              f"P {m['precision']:.5f} R {m['recall']:.5f} desc_l2_max {m['desc_l2_max']:.2e}")
        assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
        assert m["desc_l2_max"] <= DESC_L2_TOL
        assert m["desc_l2_max"] <= DESC_L2_MARGIN


def test_odd_strides_and_mixed_sizes():
    """Device frames whose rows are an odd number of bytes apart (unaligned rows
    for the initial blur's loads), next to frames with aligned strides and sizes
    that take every horizontal-pass form at octave 0."""
    shapes = [(640, 480), (97, 61), (1280, 720), (333, 45), (1920, 1080)]
    strides = [640 + 13, 97, 1280 + 64, 333 + 1, 1920 + 7]
    imgs = [imgen.generic_image(w, h, seed=300 + i) for i, (w, h) in enumerate(shapes)]
    params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h) if min(w, h) >= 64 else 3, 3, 0.5, 0.04, 10.0)
              for (w, h) in shapes]
    s = sift_hip.Sift.default()
    got = _device_detect(s, imgs, params, strides)
    host = s.detect_batch(imgs, params)
    for i in range(len(imgs)):
        assert _same(got[i], host[i]), i


def test_results_valid_until_next_call():
    """Two calls in a row on one context: each call's read-back equals the host
    path for its own frames (the second call reuses the arena)."""
    s = sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)
    a = [imgen.generic_image(800, 600, seed=400 + i) for i in range(3)]
    b = [imgen.generic_image(640, 480, seed=410 + i) for i in range(2)]
    ga = _device_detect(s, a)
    gb = _device_detect(s, b)
    ha, hb = s.detect_batch(a), s.detect_batch(b)
    for i in range(len(a)):
        assert _same(ga[i], ha[i]), i
    for i in range(len(b)):
        assert _same(gb[i], hb[i]), i
