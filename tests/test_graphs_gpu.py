"""Replayed launch sequences (HIP graphs) give the same results as eager launches.

The library captures the launch sequence of a batch layout once and replays it
for later calls with the same layout; with kernel timing enabled it launches
eagerly. Each replay here is checked bit for bit against the eager result, with
other device work and synchronisations between the replays.
"""
import numpy as np
import pytest

import imgen
import sift_hip

pytestmark = pytest.mark.gpu


def _eager(fn):
    sift_hip.profile(True)
    try:
        return fn()
    finally:
        sift_hip.profile(False)


def _same(a, b):
    return a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()


@pytest.mark.parametrize("queue", ["0", "1"])
def test_replays_match_eager(queue, monkeypatch):
    monkeypatch.setenv("SIFT_HIP_QUEUE", queue)
    s = sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)
    a = imgen.generic_image(640, 480, seed=41)
    b = imgen.generic_image(640, 480, seed=42)
    c = imgen.generic_image(333, 221, seed=43)
    ref = {k: _eager(lambda im=im: s.detect_and_compute(im)) for k, im in (("a", a), ("b", b), ("c", c))}
    assert len(ref["a"][0]) > 0 and len(ref["b"][0]) > 0
    rgb = np.random.default_rng(44).integers(0, 256, (700, 900, 3), dtype=np.uint8)
    luma = sift_hip.to_luma8(rgb)
    for rep in range(3):
        for k, im in (("a", a), ("c", c), ("b", b)):
            got = s.detect_and_compute(im)
            assert _same(got, ref[k]), (rep, k)
            # unrelated device work (an eager launch on the same stream, with a
            # synchronisation) between replays
            assert np.array_equal(sift_hip.to_luma8(rgb), luma)


@pytest.mark.parametrize("queue", ["0", "1"])
def test_batch_replays_match_eager(queue, monkeypatch):
    monkeypatch.setenv("SIFT_HIP_QUEUE", queue)
    shapes = [(640, 480), (97, 61), (1280, 720)]
    imgs = [imgen.generic_image(w, h, seed=200 + i) for i, (w, h) in enumerate(shapes)]
    s = sift_hip.Sift.default()
    ref = _eager(lambda: s.detect_batch(imgs))
    for rep in range(3):
        got = s.detect_batch(imgs)
        for i in range(len(imgs)):
            assert _same(got[i], ref[i]), (rep, i)
        # a different layout in between evicts nothing and must not disturb the replay
        s.detect_and_compute(imgs[1])
