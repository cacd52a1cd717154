"""draw_keypoints_to_image (src/lib.rs:70-116): the oracle's restatement of
imageproc's circle and line rasterisation is pinned by geometry, then the
library's host implementation (sift_draw_keypoints) must equal it pixel for
pixel. No device work."""
import math

import numpy as np
import pytest

import oracle
import sift_hip

RED = (255, 0, 0)


def kp(x, y, angle=0.0):
    k = np.zeros(1, sift_hip.KEYPOINT_DTYPE)
    k["x"], k["y"], k["angle"], k["size"] = x, y, angle, 2.0
    return k


def mask(rgb, base):
    return np.any(rgb != base, axis=-1)


def test_disc_radius_6_and_cross_of_radius_1():
    base = np.zeros((40, 40), np.uint8)
    # a keypoint pointing right: the line lies inside the radius-6 disc
    m = mask(oracle.draw_keypoints(base, oracle.GRAY8, kp(20.3, 20.7, 0.0)), 0)
    ys, xs = np.nonzero(m)
    d2 = (xs - 20) ** 2 + (ys - 20) ** 2
    assert d2.max() <= 6.5 ** 2
    yy, xx = np.mgrid[0:40, 0:40]
    inner = (xx - 20) ** 2 + (yy - 20) ** 2 <= 5.5 ** 2
    assert np.all(m[inner])
    # symmetric under both reflections and the transposition about (20, 20)
    sub = m[14:27, 14:27]
    assert np.array_equal(sub, sub[::-1, :]) and np.array_equal(sub, sub[:, ::-1])
    assert np.array_equal(sub, sub.T)


def test_disc_rows_of_the_midpoint_recurrence():
    """The radius-6 disc row by row, worked by hand from the midpoint circle
    recurrence that draw_filled_circle_mut follows (x = 0, y = r, p = 1 - r;
    while x <= y: fill rows cy +- y over cx +- x and rows cy +- x over cx +- y;
    x += 1; p < 0 ? p += 2x + 1 : (y -= 1, p += 2(x - y) + 1)):
    (x, y, p) = (0, 6, -5) (1, 6, -2) (2, 6, 3) (3, 5, 0) (4, 4, 1), so rows
    0..2 span +-6, row 3 +-5, row 4 +-4, row 5 +-3 and row 6 +-2. A keypoint
    pointing straight up keeps its line inside the disc."""
    m = mask(oracle.draw_keypoints(np.zeros((41, 41), np.uint8), oracle.GRAY8,
                                   kp(20.0, 20.0, -math.pi / 2)), 0)
    half = {0: 6, 1: 6, 2: 6, 3: 5, 4: 4, 5: 3, 6: 2}
    for dy in range(-8, 9):
        xs = np.nonzero(m[20 + dy])[0] - 20
        if abs(dy) > 6:
            assert len(xs) == 0, dy
        else:
            hw = half[abs(dy)]
            assert list(xs) == list(range(-hw, hw + 1)), dy


def test_line_pixels_follow_the_segment():
    base = np.zeros((60, 60), np.uint8)
    for a in np.linspace(-math.pi, math.pi, 37):
        with_line = mask(oracle.draw_keypoints(base, oracle.GRAY8, kp(30.25, 29.6, a)), 0)
        disc_only = mask(oracle.draw_keypoints(base, oracle.GRAY8, kp(30.25, 29.6, 0.0)), 0)
        x0, y0 = 30.25, 29.6
        x1, y1 = x0 + 6 * math.cos(a), y0 + 6 * math.sin(a)
        # every pixel drawn beyond the disc lies within 1.5 px of the segment (the
        # walk starts at truncated coordinates of whichever end has the smaller x)
        extra = np.argwhere(with_line & ~disc_only)
        for y, x in extra:
            t = max(0.0, min(1.0, ((x - x0) * (x1 - x0) + (y - y0) * (y1 - y0)) / 36.0))
            assert math.hypot(x - (x0 + t * (x1 - x0)), y - (y0 + t * (y1 - y0))) <= 1.5, (a, x, y)
        # the far end pixel is drawn
        ex, ey = int(x1), int(y1)
        assert with_line[ey, ex] or with_line[int(round(y1)), int(round(x1))], a


def test_axis_and_diagonal_lines_from_integer_points():
    # Bresenham from an integer point: the 45-degree line is the diagonal (it
    # stays inside the disc, so the drawing equals the disc-only one), and the
    # horizontal line is the plain span of its row
    base = np.zeros((40, 40), np.uint8)
    disc = mask(oracle.draw_keypoints(base, oracle.GRAY8, kp(20.0, 20.0, 0.0)), 0)
    diag = mask(oracle.draw_keypoints(base, oracle.GRAY8, kp(20.0, 20.0, math.pi / 4)), 0)
    assert np.array_equal(disc, diag)  # the diagonal to (24.24, 24.24) stays inside the disc
    for i in range(5):
        assert diag[20 + i, 20 + i]
    # a keypoint near the right edge pointing left: the line runs along its row
    m = mask(oracle.draw_keypoints(np.zeros((9, 40), np.uint8), oracle.GRAY8,
                                   kp(39.0, 4.0, math.pi)), 0)
    assert np.all(m[4, 33:40])


def test_points_outside_are_skipped_and_colour_kept():
    base = np.full((10, 12, 3), 7, np.uint8)
    out = oracle.draw_keypoints(base, oracle.RGB8, kp(0.5, 0.5, math.pi), color=(1, 2, 3))
    m = mask(out, np.array([7, 7, 7], np.uint8))
    assert m[0, 0] and m[6, 0] and m[0, 6] and not m[7, 0]
    assert np.all(out[m] == np.array([1, 2, 3], np.uint8))


@pytest.mark.parametrize("fmt,channels", [(oracle.GRAY8, 0), (oracle.RGB8, 3), (oracle.RGBA8, 4),
                                          (oracle.LUMAA8, 2)])
def test_library_matches_the_oracle(fmt, channels):
    rng = np.random.default_rng(fmt)
    h, w = 90, 131
    shape = (h, w) if channels == 0 else (h, w, channels)
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    n = 400
    kps = np.zeros(n, sift_hip.KEYPOINT_DTYPE)
    kps["x"] = rng.uniform(-3, w + 3, n)
    kps["y"] = rng.uniform(-3, h + 3, n)
    kps["angle"] = rng.uniform(-math.pi, math.pi, n)
    kps["size"] = rng.uniform(1, 20, n)
    got = sift_hip.draw_keypoints_to_image(img, kps, (255, 0, 0), fmt=fmt)
    ref = oracle.draw_keypoints(img, fmt, kps, (255, 0, 0))
    assert np.array_equal(got, ref)


def test_to_rgb8_without_keypoints():
    g = np.arange(12, dtype=np.uint8).reshape(3, 4)
    out = sift_hip.draw_keypoints_to_image(g, np.zeros(0, sift_hip.KEYPOINT_DTYPE))
    assert np.array_equal(out, np.repeat(g[:, :, None], 3, axis=-1))
