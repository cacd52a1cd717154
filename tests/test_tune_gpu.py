"""Every blur form gives the reference's levels: pyramids built with one
horizontal and one vertical form forced on every level they are valid for
(SIFT_HIP_BLUR_FORCE) equal the oracle's bit for bit, so any choice the tuning
table makes leaves the pyramid unchanged; the tuning measurement runs for every
valid form."""
import numpy as np
import pytest

import imgen
import oracle
import sift_hip

pytestmark = pytest.mark.gpu

SHAPES = [(640, 480, 4), (333, 221, 4), (1920, 1080, 3), (97, 61, 3)]
_REF = {}


def _ref(w, h, octv):
    if (w, h) not in _REF:
        img = imgen.generic_image(w, h, seed=w + h)
        _REF[(w, h)] = (img, oracle.build_pyramid(oracle.params(1.6, octv, 3, 0.5, 0.04, 10.0), img)[0])
    return _REF[(w, h)]


@pytest.mark.parametrize("hform", sift_hip.BLUR_H_FORMS)
@pytest.mark.parametrize("vform", sift_hip.BLUR_V_FORMS)
def test_forced_forms_bit_exact(hform, vform, monkeypatch):
    monkeypatch.setenv("SIFT_HIP_BLUR_FORCE", f"{hform},{vform}")
    for (w, h, octv) in SHAPES:
        img, ref = _ref(w, h, octv)
        g, _ = sift_hip.Sift(1.6, octv, 3, 0.5, 0.04, 10.0).pyramid(img)
        for o in range(octv):
            assert np.array_equal(g[o], ref[o]), (w, h, o)


def test_tune_measurement_every_form():
    """sift_hip_tune_blur times every valid form, reports the forms it used and
    refuses invalid ones."""
    for (w, h, sigma, jobs) in [(1920, 1080, 1.2263, 4), (120, 67, 12.7, 2), (30, 16, 198.0, 1)]:
        base = sift_hip.tune_blur(w, h, sigma, jobs, reps=3)
        assert base is not None and base[0] > 0 and base[1] > 0
        assert base[2] in sift_hip.BLUR_H_FORMS and base[3] in sift_hip.BLUR_V_FORMS
        for f in sift_hip.BLUR_H_FORMS:
            r = sift_hip.tune_blur(w, h, sigma, jobs, h_form=f, reps=3)
            assert r is None or (r[0] > 0 and r[2] == f)
        for f in sift_hip.BLUR_V_FORMS:
            r = sift_hip.tune_blur(w, h, sigma, jobs, v_form=f, reps=3)
            assert r is None or (r[1] > 0 and r[3] == f)
    # the small vertical form stages at most 128 rows
    assert sift_hip.tune_blur(64, 2000, 250.0, 1, v_form="vsmall", reps=1) is None
