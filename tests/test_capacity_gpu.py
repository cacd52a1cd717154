"""Candidate and keypoint capacities start small here, so the first pass of a
fresh context overflows them and the library grows them and runs again; the
result must equal the one of a context that never overflowed.
"""
import ctypes
import os

import numpy as np
import pytest

import imgen
import sift_hip

pytestmark = pytest.mark.gpu


def _fresh_detect(img, octaves):
    L = sift_hip.lib()
    p = sift_hip.SiftParams()
    L.sift_params_new(ctypes.byref(p), 1.6, octaves, 3, 0.5, 0.04, 10.0)
    h = ctypes.c_void_p()
    assert L.sift_hip_create(ctypes.byref(p), 0, ctypes.byref(h)) == 0
    try:
        keep, im = sift_hip._image_struct(img)
        r = sift_hip.SiftResult()
        rc = L.sift_hip_detect(h, ctypes.byref(im), ctypes.byref(r))
        assert rc == 0, sift_hip.last_error()
        return sift_hip._take_result(r)
    finally:
        L.sift_hip_destroy(h)


@pytest.mark.parametrize("queue", ["0", "1"])
@pytest.mark.parametrize("case", ["512x512", "1920x1080"])
def test_overflow_and_retry_gives_the_same_result(case, queue, monkeypatch):
    """Run with the static keypoint assignment and with the device work queue
    (SIFT_HIP_QUEUE=1), whose counters must be cleared again for the pass that
    follows an overflow."""
    monkeypatch.setenv("SIFT_HIP_QUEUE", queue)
    img, octv, ct = imgen.config_image(case)
    ref = sift_hip.Sift(1.6, octv, 3, 0.5, ct, 10.0).detect_and_compute(img)
    assert len(ref[0]) > 512
    monkeypatch.setenv("SIFT_HIP_CAP_MIN", "256")
    monkeypatch.setenv("SIFT_HIP_CAP_DIVISOR", "100000000")
    got = _fresh_detect(img, octv)
    assert got[0].tobytes() == ref[0].tobytes()
    assert got[1].tobytes() == ref[1].tobytes()
