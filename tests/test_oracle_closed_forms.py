"""Closed-form and invariance checks of the oracle's refinement, orientation and
descriptor (src/sift.rs:299-474, 605-726, 917-1060), each chosen so that a
plausible slip in the restatement (a dropped term, a wrong offset, sign or
index, a wrong weight width) fails it:

  refinement   DoG stacks that are exact quadratics with a known sub-pixel and
               sub-scale extremum, a saddle, edge ratios either side of
               (r+1)^2/r, a contrast test that only the interpolated value passes;
  orientation  one-hot smoothing ([1,2,3,2,1]/9), the parabolic peak fit, the
               0.8 threshold and strict-peak rule, the bin-centre offset of an
               axis-aligned ramp;
  descriptor   the separable closed form of a constant-gradient image, a
               gradient confined to a half-plane, the relative-angle sign and
               invariance under 90-degree rotations of the image.
"""
import math

import numpy as np
import pytest

import imgen
import oracle

DEG = math.pi / 180.0


# --- refinement ----------------------------------------------------------------

def _quad_dog(x0, y0, l0, ax, ay, al, c, w=40, h=40, nd=5):
    """D(x, y, l) = c + (ax (x-x0)^2 + ay (y-y0)^2 + al (l-l0)^2) / 2: its extremum
    is at (x0, y0, l0) with value c and Hessian diag(ax, ay, al)."""
    l, y, x = np.mgrid[0:nd, 0:h, 0:w].astype(np.float64)
    d = c + 0.5 * (ax * (x - x0) ** 2 + ay * (y - y0) ** 2 + al * (l - l0) ** 2)
    return d.astype(np.float32)


P3 = oracle.params(1.6, 4, 3, 0.5, 0.04, 10.0)  # contrast 0.04 / 3, edge ratio 10


@pytest.mark.parametrize("start", [(20, 16, 2), (18, 13, 2), (23, 18, 3)])
def test_refine_quadratic_extremum(start):
    """A maximum at (20.3, 15.7, layer 2.2): from the nearest sample in one Newton
    step, from farther starts after the integer moves of the reference; x, y in
    input pixels are (cx + offset) * 2^octave, size = sigma * 2^(o + l/S)."""
    d = _quad_dog(20.3, 15.7, 2.2, -0.004, -0.004, -0.004, 0.05)
    for octave in (0, 1):
        k = oracle.refine_single(P3, d, *start, octave=octave)
        assert k is not None
        sc = 2.0 ** octave
        assert abs(k["x"] - 20.3 * sc) < 1e-3 * sc
        assert abs(k["y"] - 15.7 * sc) < 1e-3 * sc
        assert k["layer"] == 2
        assert abs(k["size"] - 1.6 * 2.0 ** (octave + 2.2 / 3.0)) < 1e-4
        assert abs(k["response"] - 0.05) < 1e-6
        assert k["octave"] == octave


def test_refine_layer_rounding():
    """The reported layer is round(l + offset); size uses the unrounded scale."""
    d = _quad_dog(12.0, 12.0, 1.6, -0.01, -0.01, -0.01, 0.05)
    k = oracle.refine_single(P3, d, 12, 12, 2)
    assert k is not None and k["layer"] == 2
    assert abs(k["size"] - 1.6 * 2.0 ** (1.6 / 3.0)) < 1e-4
    d = _quad_dog(12.0, 12.0, 1.4, -0.01, -0.01, -0.01, 0.05)
    k = oracle.refine_single(P3, d, 12, 12, 1)
    assert k is not None and k["layer"] == 1


def test_refine_rejects_saddle():
    """det of the 2x2 spatial Hessian <= 0 (a saddle) is rejected even though the
    offset converges and the contrast passes."""
    d = _quad_dog(20.2, 15.8, 2.1, -0.004, 0.004, -0.004, 0.05)
    assert oracle.refine_single(P3, d, 20, 16, 2) is None
    d = _quad_dog(20.2, 15.8, 2.1, -0.004, -0.004, -0.004, 0.05)
    assert oracle.refine_single(P3, d, 20, 16, 2) is not None


@pytest.mark.parametrize("rho,kept", [(9.8, True), (10.2, False), (1.0, True), (30.0, False)])
def test_refine_edge_ratio(rho, kept):
    """tr^2 / det = (1 + rho)^2 / rho for principal curvatures in ratio rho; kept
    iff it is below (r + 1)^2 / r = 12.1, i.e. iff rho < 10."""
    d = _quad_dog(20.2, 15.8, 2.1, -0.004 * rho, -0.004, -0.004, 0.05)
    assert (oracle.refine_single(P3, d, 20, 16, 2) is not None) == kept


def test_refine_contrast_uses_interpolated_value():
    """|D(x_hat)| = |D + g.offset / 2| against the stored threshold c / S: a peak
    of 1.2 thr whose nearest sample is below thr is kept, a peak of 0.9 thr is
    dropped."""
    thr = 0.04 / 3.0
    d = _quad_dog(20.3, 15.7, 2.2, -0.03, -0.03, -0.03, 1.2 * thr)
    assert abs(d[2, 16, 20]) < thr  # the sample itself fails the threshold
    k = oracle.refine_single(P3, d, 20, 16, 2)
    assert k is not None and abs(k["response"] - 1.2 * thr) < 1e-6
    d = _quad_dog(20.3, 15.7, 2.2, -0.03, -0.03, -0.03, 0.9 * thr)
    assert oracle.refine_single(P3, d, 20, 16, 2) is None


def test_refine_minimum():
    """A minimum (negative DoG peak) is refined like a maximum."""
    d = _quad_dog(20.3, 15.7, 2.2, 0.004, 0.004, 0.004, -0.05)
    k = oracle.refine_single(P3, d, 20, 16, 2)
    assert k is not None
    assert abs(k["x"] - 20.3) < 1e-3 and abs(k["response"] + 0.05) < 1e-6


# --- orientation -----------------------------------------------------------------

@pytest.mark.parametrize("k", [0, 17, 35])
def test_smoothing_of_one_hot_bin(k):
    """Two circular [1,1,1]/3 passes turn a one-hot bin into [1,2,3,2,1]/9."""
    h = np.zeros(36, np.float32)
    h[k] = 9.0
    s = oracle.smooth_orientation_histogram(h)
    expect = np.zeros(36, np.float32)
    for off, v in zip(range(-2, 3), (1, 2, 3, 2, 1)):
        expect[(k + off) % 36] = v
    assert np.array_equal(s, expect)


@pytest.mark.parametrize("i0,delta", [(3, 0.0), (3, 0.25), (20, -0.3), (30, 0.4)])
def test_peak_parabola(i0, delta):
    """For histogram values on a parabola with its vertex at bin i0 + delta, the
    three-point fit recovers delta, and the angle is (i0 + 0.5 + delta) * 10
    degrees (bin centres are half a bin in), mapped to (-180, 180]."""
    h = np.full(36, 0.1, np.float32)
    for i in (i0 - 1, i0, i0 + 1):
        h[i % 36] = 10.0 - 2.0 * (i - i0 - delta) ** 2
    a = oracle.orientation_peaks(h)
    assert len(a) == 1
    expect = (i0 + 0.5 + delta) * 10.0 * DEG
    expect = expect - 2 * math.pi if expect > math.pi else expect
    assert abs(a[0] - expect) < 2e-6


def test_peak_threshold_and_strictness():
    """Peaks at >= 0.8 of the maximum count, in bin order; a plateau (equal
    neighbour) is not a peak."""
    h = np.full(36, 0.1, np.float32)
    h[5], h[12], h[20], h[28], h[29] = 1.0, 0.81, 0.79, 0.9, 0.9
    a = oracle.orientation_peaks(h)
    assert len(a) == 2
    assert abs(a[0] - 55.0 * DEG) < 1e-5 and abs(a[1] - 125.0 * DEG) < 1e-5


def _ramp_x(n=81, slope=2):
    """I(x, y) = 128 + slope * (x - n // 2): every gradient is (2 slope / 255, 0)."""
    x = np.arange(n) - n // 2
    return np.tile((128 + slope * x).astype(np.uint8), (n, 1))


def test_axis_ramp_reports_bin_centre():
    """Gradients exactly along +x fall in bin 0; after smoothing bin 0 is a
    symmetric peak, so the reported angle is the bin-0 centre, 5 degrees, not 0."""
    g = _ramp_x()
    raw = oracle.orientation_histogram_single(P3, g, 40.0, 40.0, 2.0)
    assert raw[0] > 0 and np.all(raw[1:] == 0)
    a = oracle.orientation_single(P3, g, 40.0, 40.0, 2.0)
    assert len(a) == 1
    assert abs(a[0] - 5.0 * DEG) < 0.5 * DEG


def test_orientation_window_weight():
    """The raw bin-0 mass of the constant-gradient ramp is
    (4/255) * sum exp(-(dx^2+dy^2) / (2 (1.5 sigma)^2)) over the square window
    of radius round(4.5 sigma) (every sample inside the image)."""
    g = _ramp_x()
    for sigma in (1.7, 2.0, 2.6):
        raw = oracle.orientation_histogram_single(P3, g, 40.0, 40.0, sigma)
        R = int(math.floor(4.5 * sigma + 0.5))
        d = np.arange(-R, R + 1, dtype=np.float64)
        e = np.exp(-d ** 2 / (2 * (1.5 * sigma) ** 2))
        expect = (4.0 / 255.0) * e.sum() ** 2
        assert abs(raw[0] - expect) < 2e-5 * expect, sigma


# --- descriptor ------------------------------------------------------------------

def _tent_sums(sigma):
    """X[i] = sum over integer offsets p of the window of
    exp(-p^2 / (2 (6 sigma)^2)) * max(0, 1 - |p / (3 sigma) + 1.5 - i|), the 1-D
    factor of the descriptor of a constant gradient along x at angle 0 (bin
    width 3 sigma, 4 cells, window |xb - 1.5| < 2.5, weight sigma half the
    4-cell width)."""
    bw = 3.0 * sigma
    R = int(math.ceil(4 * bw * math.sqrt(2.0) * 0.5))
    p = np.arange(-R, R + 1, dtype=np.float64)
    xb = p / bw + 1.5
    inside = (xb > -1.0) & (xb < 4.0)
    e = np.exp(-p ** 2 / (2 * (2 * bw) ** 2)) * inside
    return np.array([(e * np.maximum(0.0, 1.0 - np.abs(xb - i))).sum() for i in range(4)])


@pytest.mark.parametrize("sigma", [1.7, 2.0, 2.6])
def test_descriptor_constant_gradient_closed_form(sigma):
    """A constant gradient (4/255, 0) at keypoint angle 0: every sample's relative
    angle is 0, so all mass is in orientation bin 0 and the spatial histogram is
    the outer product (4/255) X[iy] X[ix] of the 1-D tent sums."""
    g = _ramp_x(121)
    raw = oracle.descriptor_raw_single(P3, g, 60.0, 60.0, sigma, 0.0).reshape(4, 4, 8)
    assert np.all(raw[:, :, 1:] == 0)
    X = _tent_sums(sigma)
    model = (4.0 / 255.0) * np.outer(X, X)
    assert np.max(np.abs(raw[:, :, 0] - model)) < 2e-5 * model.max()


@pytest.mark.parametrize("angle,bin_", [(90.0, 6), (-90.0, 2), (180.0, 4)])
def test_descriptor_relative_angle(angle, bin_):
    """Gradients at 0 degrees seen from a keypoint at `angle` have relative angle
    rem_euclid(0 - angle, 360): 270 degrees -> bin 6 for +90, 90 -> bin 2 for -90."""
    g = _ramp_x(121)
    raw = oracle.descriptor_raw_single(P3, g, 60.0, 60.0, 2.0, angle * DEG).reshape(16, 8)
    mass = raw.sum(axis=0)
    assert mass[bin_] > 0.999 * mass.sum()


def test_descriptor_half_plane():
    """Gradients only at samples with x offset >= 1 (the ramp starts 2 px right
    of the keypoint): at angle 0 their bin x-coordinates are >= 1.5 + 1/bw, so
    cells of column 0 stay empty while column 1 receives mass; rows are all
    reached. At angle 180 the image is mirrored in the cell grid: column 3 is
    empty, column 2 not."""
    n = 121
    x = np.arange(n) - 60
    row = (128 + 2 * np.maximum(x - 1, 0)).astype(np.uint8)
    g = np.tile(row, (n, 1))
    for angle, empty, full in ((0.0, 0, 1), (180.0, 3, 2)):
        raw = oracle.descriptor_raw_single(P3, g, 60.0, 60.0, 2.0, angle * DEG).reshape(4, 4, 8)
        cells = raw.sum(axis=2)
        assert np.all(cells[:, empty] == 0), angle
        assert np.all(cells[:, full] > 0), angle
        assert np.all(cells.sum(axis=1) > 0), angle


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("angle", [0.3, 1.0, 2.5])
def test_descriptor_rotation_invariance(k, angle):
    """Rotating the image by k * 90 degrees about the keypoint (x towards y) and
    the keypoint angle with it permutes the window's samples, gradients and
    relative angles exactly: the descriptor is the same (up to f32 rounding of
    the angles)."""
    g = imgen.generic_image(81, 81, seed=12)
    d0 = oracle.descriptor_single(P3, g, 40.0, 40.0, 2.0, angle)
    gr = np.ascontiguousarray(np.rot90(g, -k))
    a = angle + k * math.pi / 2
    a = a - 2 * math.pi if a > math.pi else a
    d1 = oracle.descriptor_single(P3, gr, 40.0, 40.0, 2.0, a)
    assert np.linalg.norm(d0) > 0.99
    assert np.max(np.abs(d0 - d1)) < 2e-5
    # the other direction of rotation is not a symmetry
    d2 = oracle.descriptor_single(P3, np.ascontiguousarray(np.rot90(g, k)), 40.0, 40.0, 2.0, a)
    if k != 2:
        assert np.max(np.abs(d0 - d2)) > 1e-2
