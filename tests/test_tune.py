"""The blur tuning table and script, on the CPU: the script's level schedule is
the oracle's, and the committed table lists only known forms in the layout the
library reads."""
import json
import os

import numpy as np

import oracle
import sift_hip
import tune_blur

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(REPO, "sift-wgpu_amd", "blur_tune.json")


def test_level_sigmas_match_the_oracle():
    for (S, O, sigma, blur) in [(3, 5, 1.6, 0.5), (2, 4, 1.6, 0.5), (4, 3, 1.4, 0.6), (3, 8, 1.6, 0.5)]:
        p = oracle.params(sigma, O, S, blur, 0.04, 10.0)
        ob, os0 = oracle.level_sigmas(p)
        s0, lev = tune_blur.level_sigmas(sigma, S, O, blur)
        assert np.float32(s0) == np.float32(os0)
        for (o, s, b) in lev:
            assert np.float32(b) == ob[o, s], (S, O, o, s)


def test_tap_counts():
    # r = ceil(2 sigma), 2r + 1 taps (imageproc gaussian_kernel_f32)
    for sigma in (1.2263, 1.5199, 2.0, 12.7, 198.0):
        assert tune_blur.taps_of(sigma) == len(oracle.gaussian_kernel(sigma))


def test_committed_table_layout():
    if not os.path.exists(TABLE):
        return
    t = json.load(open(TABLE))
    assert list(t)[-1] == "entries"
    keys = set()
    for e in t["entries"]:
        assert len(e) == 6
        w, h, nt, many, hf, vf = e
        assert w > 0 and h > 0 and nt % 2 == 1 and many in (0, 1)
        assert hf in sift_hip.BLUR_H_FORMS and vf in sift_hip.BLUR_V_FORMS
        keys.add((w, h, nt, many))
    assert len(keys) == len(t["entries"])
    # the 1080p batched workload's levels are all listed
    for (w, h, s) in tune_blur.level_shapes(1920, 1080, 5):
        assert (w, h, tune_blur.taps_of(s), 1) in keys
