"""Host results and served frames in the pinned host pool, in plain host memory
(the pool turned off or full) and in a mix of both within one batch: the same
keypoints and descriptors, bit for bit."""
import gc
import os
import time

import pytest

import imgen
import sift_hip

pytestmark = pytest.mark.gpu

ENV = "SIFT_HIP_HOST_POOL_BYTES"


def _with_pool_limit(value, fn):
    old = os.environ.get(ENV)
    if value is None:
        os.environ.pop(ENV, None)
    else:
        os.environ[ENV] = str(value)
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop(ENV, None)
        else:
            os.environ[ENV] = old


def _same(a, b):
    assert len(a) == len(b)
    for (ka, da), (kb, db) in zip(a, b):
        assert ka.tobytes() == kb.tobytes()
        assert da.tobytes() == db.tobytes()


def _copies(results):
    """Owned copies, so that no pool block stays referenced."""
    return [(k.copy(), d.copy()) for k, d in results]


def test_batch_results_pooled_plain_and_mixed():
    # result blocks: about 1 MiB -> 2 MiB class at 640x480, 4 MiB at 1920x1080
    shapes = [(640, 480), (1920, 1080), (640, 480)]
    imgs = [imgen.generic_image(w, h, seed=900 + i) for i, (w, h) in enumerate(shapes)]
    sift = sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)
    pooled = _copies(_with_pool_limit(None, lambda: sift.detect_batch(imgs)))
    assert all(len(k) > 1000 for k, _ in pooled)
    gc.collect()
    plain = _copies(_with_pool_limit(0, lambda: sift.detect_batch(imgs)))
    _same(pooled, plain)
    gc.collect()
    # 3 MiB: the first image's block fits, the other two go through the staging
    mixed = _copies(_with_pool_limit(3 << 20, lambda: sift.detect_batch(imgs)))
    _same(pooled, mixed)
    # single-image entry point, pool on and off
    one = _with_pool_limit(None, lambda: sift.detect_and_compute(imgs[1]))
    _same([pooled[1]], [one])
    one = _with_pool_limit(0, lambda: sift.detect_and_compute(imgs[1]))
    _same([pooled[1]], [one])


def test_served_frames_without_the_pool():
    shapes = [(640, 480), (1280, 720), (320, 240), (1920, 1080)]
    imgs = [imgen.generic_image(w, h, seed=950 + i) for i, (w, h) in enumerate(shapes)]
    sift = sift_hip.Sift.default()
    alone = _copies([sift.detect_and_compute(im) for im in imgs])

    def serve():
        got = {}
        with sift_hip.SiftServer(sift, max_batch=4) as srv:
            for i, im in enumerate(imgs):
                srv.submit(im, i)
            deadline = time.time() + 60
            while len(got) < len(imgs) and time.time() < deadline:
                for tag, st, kps, desc, *_ in srv.poll(timeout_ms=200):
                    assert st == 0
                    got[tag] = (kps.copy(), desc.copy())
        return [got[i] for i in range(len(imgs))]

    _same(alone, _with_pool_limit(0, serve))
    _same(alone, _with_pool_limit(None, serve))
