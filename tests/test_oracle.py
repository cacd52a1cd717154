"""Pins the CPU oracle (oracle/sift_oracle.c) before it is trusted as the parity
reference: the reference's own known answers, the reference's integration-test
invariants, closed forms, an independent library (PIL) and brute force."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import imgen
import oracle
from parity import check_keypoint_properties

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --- the reference's own unit tests (src/sift.rs:1424-1525) -----------------

def test_solve_linear_system_solvable():
    # src/sift.rs:1424-1438
    x = oracle.solve_linear_system([[3.0, 2.0, -1.0], [2.0, -2.0, 4.0], [-1.0, 0.5, -1.0]],
                                   [1.0, -2.0, 0.0])
    assert x is not None
    assert np.all(np.abs(x - np.array([1.0, -2.0, -2.0])) < 1e-5)


def test_solve_linear_system_singular():
    # src/sift.rs:1440-1454
    assert oracle.solve_linear_system([[1, 2, 3], [2, 4, 6], [0, 1, 1]], [1, 2, 3]) is None


def test_get_gauss_pixel_bilinear():
    # src/sift.rs:1456-1525 (all nine asserted values)
    img = np.array([[10, 20, 30], [40, 50, 60], [70, 80, 90]], np.uint8)
    cases = [((1.0, 1.0), 50), ((0.0, 0.0), 10), ((2.0, 2.0), 90), ((0.5, 0.0), 15),
             ((0.0, 0.5), 25), ((0.5, 0.5), 30), ((-0.5, 0.5), 25), ((-10.0, -10.0), 10),
             ((10.0, 10.0), 90)]
    for (x, y), v in cases:
        assert abs(oracle.gauss_pixel_bilinear(img, x, y) - v / 255.0) < 1e-6, (x, y)


# --- closed forms -------------------------------------------------------------

@pytest.mark.parametrize("sigma", [0.7, 1.2263, 1.5199, 3.0912, 24.73, 98.9])
def test_gaussian_taps_closed_form(sigma):
    k = oracle.gaussian_kernel(sigma)
    r = math.ceil(2 * np.float32(sigma))
    assert len(k) == 2 * r + 1
    i = np.arange(-r, r + 1, dtype=np.float64)
    ref = np.exp(-i * i / (2.0 * sigma * sigma)) / (math.sqrt(2 * math.pi) * sigma)
    assert np.allclose(k, ref, rtol=3e-6, atol=0)
    assert np.array_equal(k, k[::-1])
    # imageproc's taps are not renormalised: for sigma >= 1 the sum of the
    # truncated (+-2 sigma) samples stays below 1
    if sigma >= 1.0:
        assert k.sum() < 1.0


def _trunc_u8(v):
    return np.clip(np.floor(v), 0, 255)


@pytest.mark.parametrize("sigma", [1.2263, 1.5199, 2.0, 3.7])
def test_blur_impulse_response(sigma):
    """Impulse 255 at the centre: the H pass leaves trunc(255*w_dx) on the centre
    row, the V pass trunc(h(x) * w_dy): checks radius, tap placement, the H-then-V
    order and the intermediate u8 truncation."""
    k = oracle.gaussian_kernel(sigma).astype(np.float32)
    r = len(k) // 2
    n = 4 * r + 9
    c = n // 2
    img = np.zeros((n, n), np.uint8)
    img[c, c] = 255
    out = oracle.gaussian_blur(img, sigma)
    hrow = np.zeros(n, np.float32)
    hrow[c - r:c + r + 1] = _trunc_u8(np.float32(255.0) * k)
    expect = np.zeros((n, n), np.float32)
    expect[c - r:c + r + 1, :] = _trunc_u8(k[:, None] * hrow[None, :])
    assert np.array_equal(out, expect.astype(np.uint8))


@pytest.mark.parametrize("sigma", [1.2263, 1.5199, 3.0912, 12.36])
@pytest.mark.parametrize("value", [37, 128, 200, 255])
def test_blur_flat_field(sigma, value):
    k = oracle.gaussian_kernel(sigma).astype(np.float64)
    img = np.full((40, 50), value, np.uint8)
    out = oracle.gaussian_blur(img, sigma)
    h1 = value * k.sum()
    if abs(h1 - round(h1)) < 1e-3:
        pytest.skip("flat-field sum too close to an integer for a float64 closed form")
    h1 = math.floor(h1)
    v1 = h1 * k.sum()
    if abs(v1 - round(v1)) < 1e-3:
        pytest.skip("flat-field sum too close to an integer for a float64 closed form")
    assert np.all(out == math.floor(v1))


def test_level_sigmas_closed_form():
    p = oracle.params(1.6, 6, 3, 0.5, 0.04, 10.0)
    blur, s0 = oracle.level_sigmas(p)
    assert abs(s0 - math.sqrt(1.6 ** 2 - 0.25)) < 1e-6
    k = 2 ** (1 / 3)
    for o in range(6):
        for s in range(1, 6):
            t = 1.6 * k ** s * 2 ** o
            pv = 1.6 * k ** (s - 1) * 2 ** o
            assert abs(blur[o, s] - math.sqrt(t * t - pv * pv)) < 2e-6 * t


def test_lanczos_weights_normalised_and_centred():
    for n_in, n_out in [(512, 256), (135, 67), (1080, 540), (7, 3)]:
        left, cnt, w = oracle.lanczos3_table(n_in, n_out)
        for j in range(n_out):
            ws = w[j, :cnt[j]].astype(np.float64)
            assert abs(ws.sum() - 1.0) < 1e-5
            centre = (j + 0.5) * n_in / n_out - 0.5
            mean = float(np.sum(ws * (left[j] + np.arange(cnt[j]))))
            if 3 <= left[j] and left[j] + cnt[j] <= n_in - 3:
                assert abs(mean - centre) < 1e-2  # a centre off by half a pixel fails


@pytest.mark.parametrize("shape", [(512, 512), (135, 67), (1920, 1080), (31, 17), (7, 5), (241, 135)])
def test_lanczos_matches_pil(shape):
    """PIL's Lanczos resample uses the same window (support 3*scale, centre
    (j+0.5)*scale) and renormalised weights, but fixed-point arithmetic and a u8
    intermediate; on mid-range inputs (no overshoot clipping) it agrees within 1
    grey level. An off-by-one window or a transposed pass fails by tens."""
    from PIL import Image
    w, h = shape
    rng = np.random.default_rng(w * 7 + h)
    img = rng.integers(72, 184, (h, w)).astype(np.uint8)
    ours = oracle.resize_lanczos3(img, w // 2, h // 2).astype(int)
    pil = np.array(Image.fromarray(img).resize((w // 2, h // 2), Image.LANCZOS,
                                               reducing_gap=None)).astype(int)
    assert np.abs(ours - pil).max() <= 1
    assert np.abs(np.roll(ours, 1, axis=1) - pil).max() > 5


def test_lanczos_flat():
    img = np.full((61, 77), 143, np.uint8)
    assert np.all(oracle.resize_lanczos3(img, 38, 30) == 143)


def test_dog_identity():
    img = imgen.generic_image(96, 80, seed=11)
    p = oracle.params(1.6, 3, 3, 0.5, 0.04, 10.0)
    gs, ds = oracle.build_pyramid(p, img)
    for g, d in zip(gs, ds):
        q = g.astype(np.float32) / np.float32(255.0)
        assert np.array_equal(d, q[1:] - q[:-1])


def test_extrema_brute_force():
    """Strict 26-neighbour extrema over layers 1..S inside the 5-px border,
    recomputed with numpy from the oracle's DoG pyramid, in the reference order."""
    img = imgen.generic_image(120, 90, seed=5)
    p = oracle.params(1.6, 3, 3, 0.5, 0.04, 10.0)
    gs, ds = oracle.build_pyramid(p, img)
    _, _, cand, _, _ = oracle.detect_and_compute(p, img, stages=True)
    expect = []
    for o, d in enumerate(ds):
        L, H, W = d.shape
        for s in range(1, 4):
            for y in range(5, H - 5):
                for x in range(5, W - 5):
                    blk = d[s - 1:s + 2, y - 1:y + 2, x - 1:x + 2].ravel()
                    v = blk[13]
                    nb = np.delete(blk, 13)
                    if np.all(v > nb) or np.all(v < nb):
                        expect.append((o, s, y, x, v))
    got = [(int(k["octave"]), int(k["layer"]), int(k["y"] / 2 ** k["octave"]),
            int(k["x"] / 2 ** k["octave"]), k["response"]) for k in cand]
    assert len(got) == len(expect)
    for a, b in zip(got, expect):
        assert a[:4] == b[:4] and a[4] == b[4]


# --- single-keypoint closed forms for orientation and descriptor -------------

def _ramp(theta, n=81, slope=1.2):
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    v = 128 + slope * ((xx - n / 2) * math.cos(theta) + (yy - n / 2) * math.sin(theta))
    return np.clip(np.floor(v + 0.5), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("deg", [0, 30, 100, 200, 290])
def test_orientation_of_ramp(deg):
    """A linear ramp has one gradient direction: the dominant orientation is the
    ramp angle (bin interpolation error well under a bin)."""
    theta = math.radians(deg)
    g = _ramp(theta)
    p = oracle.params()
    ang = oracle.orientation_single(p, g, 40.0, 40.0, 2.0)
    assert len(ang) >= 1
    d = [abs((a - theta + math.pi) % (2 * math.pi) - math.pi) for a in ang]
    assert min(d) < math.radians(6.0)


@pytest.mark.parametrize("deg", [0, 45, 135])
def test_descriptor_of_ramp(deg):
    """Sampled relative to the ramp's own angle, all gradient mass falls in
    relative orientation bin 0 (and its trilinear neighbour bin 1 / 7)."""
    theta = math.radians(deg)
    g = _ramp(theta, n=121)
    p = oracle.params()
    d = oracle.descriptor_single(p, g, 60.0, 60.0, 2.0, theta).reshape(16, 8)
    mass = d.sum(axis=0)
    assert np.argmax(mass) in (0, 7)
    assert (mass[0] + mass[7] + mass[1]) / mass.sum() > 0.95
    assert abs(np.linalg.norm(d) - 1.0) < 1e-5


def test_normalize_and_clip():
    v = np.zeros(128, np.float32)
    v[0], v[1] = 3.0, 4.0
    out = oracle.normalize_and_clip_descriptor(v)
    # 0.6 and 0.8 clip to 0.2, renormalised: both become 1/sqrt(2)
    assert np.allclose(out[:2], [1 / math.sqrt(2)] * 2, atol=1e-6)
    assert np.all(out[2:] == 0)
    assert np.all(oracle.normalize_and_clip_descriptor(np.zeros(128, np.float32)) == 0)


# --- the reference's integration-test invariants (tests/integration_test.rs) --

def test_uniform_image_has_no_keypoints():
    # tests/integration_test.rs:205-226
    img = np.full((100, 100), 128, np.uint8)
    kps, desc = oracle.detect_and_compute(oracle.params(), img)
    assert len(kps) == 0 and len(desc) == 0


def test_small_gradient_image():
    # tests/integration_test.rs:166-203
    img = np.tile((np.arange(20) * 10).astype(np.uint8), (15, 1))
    kps, desc = oracle.detect_and_compute(oracle.params(), img)
    assert len(kps) == len(desc)
    check_keypoint_properties(kps, 20, 15)


def test_determinism_and_properties():
    # tests/integration_test.rs:64-164
    img, octv, ct = imgen.config_image("512x512")
    p = oracle.params(1.6, octv, 3, 0.5, ct, 10.0)
    k1, d1 = oracle.detect_and_compute(p, img, nthreads=4)
    k2, d2 = oracle.detect_and_compute(p, img, nthreads=1)
    assert len(k1) > 0 and len(k1) == len(d1)
    assert k1.tobytes() == k2.tobytes()
    assert np.array_equal(d1, d2)
    check_keypoint_properties(k1, 512, 512)
    assert d1.shape[1] == 128


def test_high_contrast_threshold():
    # tests/integration_test.rs:228-281
    img, octv, _ = imgen.config_image("512x512")
    kd, _ = oracle.detect_and_compute(oracle.params(1.6, octv, 3, 0.5, 0.04, 10.0), img)
    kh, dh = oracle.detect_and_compute(oracle.params(1.6, octv, 3, 0.5, 1.0, 10.0), img)
    assert len(kh) < len(kd) // 2 or len(kh) == 0
    assert len(kh) == len(dh)


# --- committed golden fixtures -----------------------------------------------

def test_generator_checksums():
    with open(os.path.join(GOLDEN, "images.json")) as f:
        ref = json.load(f)
    for name, sha in ref["workloads"].items():
        img, _, _ = imgen.config_image(name)
        assert hashlib.sha256(img.tobytes()).hexdigest() == sha, name


def test_oracle_golden_outputs():
    """Regression pin of the oracle's outputs on small seeded images
    (tests/golden/make_golden.py)."""
    with open(os.path.join(GOLDEN, "oracle_small.json")) as f:
        cases = json.load(f)
    for c in cases:
        img = imgen.generic_image(c["w"], c["h"], seed=c["seed"])
        p = oracle.params(*c["params"])
        kps, desc, cand, ref, _ = oracle.detect_and_compute(p, img, stages=True)
        gs, ds = oracle.build_pyramid(p, img)
        assert hashlib.sha256(b"".join(g.tobytes() for g in gs)).hexdigest() == c["pyramid_sha256"]
        assert len(cand) == c["n_candidates"] and len(ref) == c["n_refined"]
        assert len(kps) == c["n_keypoints"]
        assert hashlib.sha256(cand.tobytes()).hexdigest() == c["candidates_sha256"]
        assert hashlib.sha256(ref.tobytes()).hexdigest() == c["refined_sha256"]
        npz = np.load(os.path.join(GOLDEN, c["npz"]))
        assert np.array_equal(npz["keypoints"].view(oracle.KEYPOINT_DTYPE).ravel(), kps)
        assert np.array_equal(npz["descriptors"], desc)
