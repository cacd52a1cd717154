"""HIP path vs the CPU oracle, through the C ABI (libsift_hip.so).

Bars: u8 pyramid levels and f32 DoG levels bit-exact (<= 1e-5 max-abs on the
[0,1] scale is the stated tolerance); extremum candidates and refined keypoints
bit-exact (pure + - * / arithmetic in the reference order); final keypoints
matched within 0.05 px / 0.05 size / 1 degree with precision and recall
>= 99.5%; matched descriptors within 1e-3 L2 (orientation and descriptor use
the device's atan2f/expf/sinf/cosf, which may differ from glibc's by an ulp).
"""
import numpy as np
import pytest

import imgen
import oracle
import sift_hip
from parity import DESC_L2_MARGIN, DESC_L2_TOL, MIN_PR, check_keypoint_properties, match_keypoints

pytestmark = pytest.mark.gpu


def _sift(octaves=4, contrast=0.04, S=3, edge=10.0, sigma=1.6, blur=0.5):
    return (sift_hip.Sift(sigma, octaves, S, blur, contrast, edge),
            oracle.params(sigma, octaves, S, blur, contrast, edge))


CASES = {
    "512x512": lambda: imgen.config_image("512x512"),
    "1920x1080": lambda: imgen.config_image("1920x1080"),
    "3840x2160": lambda: imgen.config_image("3840x2160"),  # 3840x2160, 6 octaves, contrast 0.03
    "odd_97x61": lambda: (imgen.generic_image(97, 61, seed=9), 3, 0.04),
    "wide_333x45": lambda: (imgen.generic_image(333, 45, seed=10), 4, 0.04),
    # rows of 65 bitmask words (wider than 4096 px): the compaction's chunked path
    "wide_4160x200": lambda: (imgen.generic_image(4160, 200, seed=11), 4, 0.04),
}


@pytest.mark.parametrize("case", list(CASES))
def test_pyramid_and_dog_bit_exact(case):
    img, octv, ct = CASES[case]()
    s, p = _sift(octv, ct)
    g_hip, d_hip = s.pyramid(img)
    g_ref, d_ref = oracle.build_pyramid(p, img)
    for o in range(octv):
        assert np.array_equal(g_hip[o], g_ref[o]), f"gauss octave {o}"
        assert np.array_equal(d_hip[o], d_ref[o]), f"dog octave {o}"


@pytest.mark.parametrize("case", ["1920x1080", "odd_97x61"])
def test_two_pass_resize_equals_fused(case, monkeypatch):
    """The two-pass Lanczos3 octave step (f32 rows through HBM) and the fused one
    (rows in LDS) build the same pyramid."""
    img, octv, ct = CASES[case]()
    s, _ = _sift(octv, ct)
    g_fused, _ = s.pyramid(img)
    monkeypatch.setenv("SIFT_HIP_RESIZE_FUSED", "0")
    g_two, _ = s.pyramid(img)
    for o in range(octv):
        assert np.array_equal(g_fused[o], g_two[o]), f"gauss octave {o}"


@pytest.mark.parametrize("case", ["512x512", "1920x1080", "3840x2160", "odd_97x61", "wide_4160x200"])
def test_candidates_and_refined_bit_exact(case):
    img, octv, ct = CASES[case]()
    s, p = _sift(octv, ct)
    cand, ref, kps, desc = s.detect_stages(img)
    _, _, ocand, oref, _ = oracle.detect_and_compute(p, img, stages=True)
    assert len(cand) == len(ocand)
    assert cand.tobytes() == ocand.tobytes()
    assert len(ref) == len(oref)
    for f in ("x", "y", "response", "octave", "layer"):
        assert np.array_equal(ref[f], oref[f]), f
    # size = sigma * 2^(o + l/S): exp2 via a correctly rounded double path vs glibc
    assert np.allclose(ref["size"], oref["size"], rtol=2e-7, atol=0)


@pytest.mark.parametrize("S", [1, 5, 6])
def test_candidates_bit_exact_interval_counts(S):
    """The extremum kernel is instantiated per interval count (S + 2 DoG layers in
    LDS): candidates and refined keypoints bit-exact for the counts the other
    cases do not use, on single frames and, through a batch large enough for the
    multi-band strips, on 8 frames at once."""
    img = imgen.generic_image(640, 480, seed=50 + S)
    s, p = _sift(octaves=4, S=S)
    cand, ref, _, _ = s.detect_stages(img)
    _, _, ocand, oref, _ = oracle.detect_and_compute(p, img, stages=True)
    assert len(cand) > 0
    assert cand.tobytes() == ocand.tobytes()
    for f in ("x", "y", "response", "octave", "layer"):
        assert np.array_equal(ref[f], oref[f]), f
    frames = [imgen.generic_image(1920, 1080, seed=60 + 10 * S + i) for i in range(8)]
    together = s.detect_batch(frames, [s] * len(frames))
    for i, f in enumerate(frames):
        alone = s.detect_and_compute(f)
        assert alone[0].tobytes() == together[i][0].tobytes(), i
        assert alone[1].tobytes() == together[i][1].tobytes(), i


def test_row_wise_extrema_seven_intervals():
    """Seven intervals (nine DoG layers) are beyond the tiled extremum kernel's
    instantiations (1..6) and take the row-wise k_extrema: candidates and refined
    keypoints bit-exact against the oracle, final keypoints and descriptors
    within tolerance."""
    img = imgen.generic_image(640, 480, seed=57)
    s, p = _sift(octaves=4, S=7)
    cand, ref, kps, desc = s.detect_stages(img)
    _, _, ocand, oref, _ = oracle.detect_and_compute(p, img, stages=True)
    assert len(cand) > 0
    assert cand.tobytes() == ocand.tobytes()
    for f in ("x", "y", "response", "octave", "layer"):
        assert np.array_equal(ref[f], oref[f]), f
    _, _, _, m = _parity(img, s, p)
    assert m["n_ref"] > 0
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
    assert m["desc_l2_max"] <= DESC_L2_TOL
    assert m["desc_l2_max"] <= DESC_L2_MARGIN


def test_batch_mixing_interval_counts():
    """A batch whose frames use different interval counts (3 and 4) runs the
    row-wise extremum kernel for all of them; each frame equals its result alone
    (where the tiled kernel runs), bit for bit."""
    shapes = [(640, 480), (800, 600), (333, 221), (1280, 720)]
    imgs = [imgen.generic_image(w, h, seed=80 + i) for i, (w, h) in enumerate(shapes)]
    params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h), 3 + (i % 2), 0.5, 0.04, 10.0)
              for i, (w, h) in enumerate(shapes)]
    together = sift_hip.Sift.default().detect_batch(imgs, params)
    for i, img in enumerate(imgs):
        alone = params[i].detect_and_compute(img)
        assert len(alone[0]) > 0, i
        assert alone[0].tobytes() == together[i][0].tobytes(), i
        assert alone[1].tobytes() == together[i][1].tobytes(), i


def _parity(img, s, p):
    kps, desc = s.detect_and_compute(img)
    okps, odesc = oracle.detect_and_compute(p, img)
    m = match_keypoints(okps, kps, odesc, desc)
    return kps, desc, okps, m


@pytest.mark.parametrize("case", ["512x512", "1920x1080", "3840x2160", "odd_97x61", "wide_333x45"])
def test_keypoints_and_descriptors_match_oracle(case):
    img, octv, ct = CASES[case]()
    s, p = _sift(octv, ct)
    kps, desc, okps, m = _parity(img, s, p)
    print(f"{case}: oracle {m['n_ref']} hip {m['n_got']} matched {m['matched']} "
          f"P {m['precision']:.5f} R {m['recall']:.5f} desc_l2_max {m['desc_l2_max']:.2e}")
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
    assert m["desc_l2_max"] <= DESC_L2_TOL
    assert m["desc_l2_max"] <= DESC_L2_MARGIN
    check_keypoint_properties(kps, img.shape[1], img.shape[0])


@pytest.mark.parametrize("variant", [
    dict(octaves=5, contrast=0.03, S=3, edge=10.0),
    dict(octaves=4, contrast=0.04, S=2, edge=8.0),
    dict(octaves=3, contrast=0.02, S=4, edge=12.0, sigma=1.4, blur=0.6),
])
def test_parameter_variants(variant):
    img = imgen.generic_image(300, 220, seed=31)
    s, p = _sift(**variant)
    g_hip, _ = s.pyramid(img)
    g_ref, _ = oracle.build_pyramid(p, img)
    for a, b in zip(g_hip, g_ref):
        assert np.array_equal(a, b)
    _, _, _, m = _parity(img, s, p)
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
    assert m["desc_l2_max"] <= DESC_L2_TOL
    assert m["desc_l2_max"] <= DESC_L2_MARGIN


@pytest.mark.parametrize("sigma,S", [(1.9, 3), (1.6, 2)])
def test_pyramid_all_tap_paths(sigma, S):
    """Levels bit-exact on a frame wide enough for the 8-pixel horizontal pass at
    octave 0, with scale schedules whose kernels have 13 taps (sigma 1.9: 9, 7,
    9, 11, 13, 17 taps) and 17 or 21 taps (the block form; S = 2: 7, 9, 11, 15, 21
    taps), so every unrolled tap count, every word phase of the staged rows and
    the block form of both passes run."""
    img = imgen.generic_image(640, 480, seed=41)
    s, p = _sift(octaves=4, sigma=sigma, S=S)
    g_hip, _ = s.pyramid(img)
    g_ref, _ = oracle.build_pyramid(p, img)
    for o, (a, b) in enumerate(zip(g_hip, g_ref)):
        assert np.array_equal(a, b), f"gauss octave {o}"


def test_wide_windows_match_oracle():
    """sigma 2.0: 20% of the keypoints have descriptor windows wider than the LDS
    forms take (R > 31) and go through k_descriptor_wide; others pass through the
    R <= 28 and R <= 31 forms. All must match the oracle."""
    img = imgen.generic_image(640, 480, seed=70)
    s, p = _sift(octaves=4, sigma=2.0)
    kps, desc, okps, m = _parity(img, s, p)
    assert len(okps) > 0
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
    assert m["desc_l2_max"] <= DESC_L2_TOL
    assert m["desc_l2_max"] <= DESC_L2_MARGIN


def test_uniform_image_has_no_keypoints():
    # tests/integration_test.rs:205-226
    s, _ = _sift()
    kps, desc = s.detect_and_compute(np.full((100, 100), 128, np.uint8))
    assert len(kps) == 0 and len(desc) == 0


@pytest.mark.parametrize("shape", [(20, 15), (1, 1), (3, 2), (17, 9), (64, 1), (1, 64), (11, 11)])
def test_tiny_and_degenerate_images(shape):
    # tests/integration_test.rs:166-203 (20x15 gradient) and smaller: octaves
    # down to 0 pixels must not fault and must match the oracle exactly.
    w, h = shape
    img = np.tile((np.arange(w) * 10 % 256).astype(np.uint8), (h, 1))
    s, p = _sift(4)
    kps, desc = s.detect_and_compute(img)
    okps, odesc = oracle.detect_and_compute(p, img)
    assert len(kps) == len(desc) == len(okps)
    check_keypoint_properties(kps, w, h)
    g_hip, _ = s.pyramid(img)
    g_ref, _ = oracle.build_pyramid(p, img)
    for a, b in zip(g_hip, g_ref):
        assert np.array_equal(a, b)


def test_determinism():
    # tests/integration_test.rs:106-164: bitwise-equal repeated runs
    img, octv, ct = imgen.config_image("512x512")
    s, _ = _sift(octv, ct)
    k1, d1 = s.detect_and_compute(img)
    k2, d2 = s.detect_and_compute(img)
    assert k1.tobytes() == k2.tobytes() and d1.tobytes() == d2.tobytes()


def test_batch_equals_alone():
    """Each image's result is the same alone and inside a mixed-size batch."""
    shapes = [(640, 480), (97, 61), (1280, 720), (333, 45), (640, 480)]
    imgs = [imgen.generic_image(w, h, seed=100 + i) for i, (w, h) in enumerate(shapes)]
    params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h) if min(w, h) >= 64 else 3, 3, 0.5, 0.04, 10.0)
              for (w, h) in shapes]
    s = sift_hip.Sift.default()
    together = s.detect_batch(imgs, params)
    for i, img in enumerate(imgs):
        alone = params[i].detect_and_compute(img)
        assert alone[0].tobytes() == together[i][0].tobytes(), i
        assert alone[1].tobytes() == together[i][1].tobytes(), i


def test_large_batch_equals_alone():
    """A batch large enough for the extremum test's multi-band strips (heights that
    are not multiples of the strip or the band included) gives each image the
    result it gets alone, where the strips are single bands. Alone, each frame's
    candidate capacity takes the one-launch descriptor path; in the batch, the
    R <= 28 and R <= 31 launches: both give the same descriptors."""
    shapes = [(1920, 1080)] * 6 + [(1000, 1003), (1366, 777)]
    imgs = [imgen.generic_image(w, h, seed=200 + i) for i, (w, h) in enumerate(shapes)]
    params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h), 3, 0.5, 0.04, 10.0) for (w, h) in shapes]
    together = sift_hip.Sift.default().detect_batch(imgs, params)
    for i, img in enumerate(imgs):
        alone = params[i].detect_and_compute(img)
        assert len(alone[0]) > 0, i
        assert alone[0].tobytes() == together[i][0].tobytes(), i
        assert alone[1].tobytes() == together[i][1].tobytes(), i


def test_high_contrast_threshold():
    # tests/integration_test.rs:228-281
    img, octv, _ = imgen.config_image("512x512")
    kd, _ = sift_hip.Sift(1.6, octv, 3, 0.5, 0.04, 10.0).detect_and_compute(img)
    kh, dh = sift_hip.Sift(1.6, octv, 3, 0.5, 1.0, 10.0).detect_and_compute(img)
    assert len(kd) > 0
    assert len(kh) < len(kd) // 2 or len(kh) == 0
    assert len(kh) == len(dh)


def test_work_queue_equals_static_assignment(monkeypatch):
    """The orientation and descriptor kernels give the same results whether their
    work-groups take keypoints from the device work queue or by the static
    grid-stride assignment (the default), on a single frame and a batch."""
    imgs = [imgen.generic_image(w, h, seed=90 + i) for i, (w, h) in
            enumerate([(1280, 720), (640, 480), (1920, 1080), (333, 221)])]
    params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h), 3, 0.5, 0.04, 10.0) for (h, w) in
              (im.shape for im in imgs)]
    s = sift_hip.Sift.default()
    monkeypatch.setenv("SIFT_HIP_QUEUE", "1")
    queued = s.detect_batch(imgs, params)
    single = params[2].detect_and_compute(imgs[2])
    monkeypatch.setenv("SIFT_HIP_QUEUE", "0")
    static = s.detect_batch(imgs, params)
    single_static = params[2].detect_and_compute(imgs[2])
    for i in range(len(imgs)):
        assert len(queued[i][0]) > 0, i
        assert queued[i][0].tobytes() == static[i][0].tobytes(), i
        assert queued[i][1].tobytes() == static[i][1].tobytes(), i
    assert single[0].tobytes() == single_static[0].tobytes()
    assert single[1].tobytes() == single_static[1].tobytes()
