"""Serving path: frames submitted concurrently (mixed sizes, continuous
batching over two device lanes) give the same result as each frame alone."""
import threading
import time

import numpy as np
import pytest

import imgen
import sift_hip

pytestmark = pytest.mark.gpu


def test_result_same_alone_and_in_the_mix():
    shapes = [(640, 480), (1280, 720), (320, 240), (1920, 1080), (97, 61), (640, 480)] * 3
    imgs = [imgen.generic_image(w, h, seed=500 + i) for i, (w, h) in enumerate(shapes)]
    sift = sift_hip.Sift.default()
    alone = [sift.detect_and_compute(im) for im in imgs]
    got = {}
    with sift_hip.SiftServer(sift, max_batch=8) as srv:
        def producer():
            for i, im in enumerate(imgs):
                srv.submit(im, i)
                time.sleep(0.0005 * (i % 3))
        t = threading.Thread(target=producer)
        t.start()
        deadline = time.time() + 60
        while len(got) < len(imgs) and time.time() < deadline:
            for tag, st, kps, desc, ts, t0, t1 in srv.poll(timeout_ms=200):
                assert st == 0
                assert t1 >= t0 >= ts
                got[tag] = (kps, desc)
        t.join()
    assert len(got) == len(imgs)
    for i in range(len(imgs)):
        assert got[i][0].tobytes() == alone[i][0].tobytes(), i
        assert got[i][1].tobytes() == alone[i][1].tobytes(), i


def test_small_budget_still_completes():
    """An HBM budget that admits only one frame at a time serialises the work
    without failing or mixing results."""
    sift = sift_hip.Sift.default()
    imgs = [imgen.generic_image(640, 480, seed=600 + i) for i in range(4)]
    with sift_hip.SiftServer(sift, hbm_budget_bytes=20 << 20, max_batch=4) as srv:
        for i, im in enumerate(imgs):
            srv.submit(im, i)
        got = {}
        deadline = time.time() + 60
        while len(got) < len(imgs) and time.time() < deadline:
            for tag, st, kps, desc, *_ in srv.poll(timeout_ms=200):
                assert st == 0
                got[tag] = kps
    assert sorted(got) == list(range(4))
    for i, im in enumerate(imgs):
        assert got[i].tobytes() == sift.detect_and_compute(im)[0].tobytes()


def test_oversized_frame_rejected():
    sift = sift_hip.Sift.default()
    with sift_hip.SiftServer(sift, hbm_budget_bytes=1 << 20) as srv:
        with pytest.raises(sift_hip.SiftError):
            srv.submit(np.zeros((1080, 1920), np.uint8), 1)
