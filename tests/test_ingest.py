"""Pins the oracle's colour ingest and u8 descriptor view before they check the
device: to_luma8 of image 0.25.6 (called at src/sift.rs:1202), the wasm canvas
conversion (src/wasm.rs:79-94) and the WebGPU backends' u8 quantisation
(src/shaders/descriptor.wgsl:256-262), over every RGB triple where that is cheap."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
import sift_hip


def all_triples(bpp=3, alpha_seed=None):
    """A 4096 x 4096 image holding each of the 2^24 (R, G, B) triples once."""
    idx = np.arange(1 << 24, dtype=np.uint32)
    px = np.empty((1 << 24, bpp), np.uint8)
    px[:, 0] = idx & 255
    px[:, 1] = (idx >> 8) & 255
    px[:, 2] = idx >> 16
    if bpp == 4:
        px[:, 3] = np.random.default_rng(alpha_seed).integers(0, 256, 1 << 24, dtype=np.uint8)
    return px.reshape(4096, 4096, bpp)


def test_srgb_luma_grey_inputs_map_to_themselves():
    v = np.arange(256, dtype=np.uint8)
    img = np.stack([v, v, v], axis=-1).reshape(1, 256, 3)
    assert np.array_equal(oracle.to_luma8(img, oracle.RGB8)[0], v)
    rgba = np.concatenate([img, np.full((1, 256, 1), 7, np.uint8)], axis=-1)
    assert np.array_equal(oracle.to_luma8(rgba, oracle.RGBA8)[0], v)


def test_srgb_luma_primaries():
    # Rec. 709 luma of a full-scale primary, truncated: 0.2126, 0.7152, 0.0722 x 255
    img = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]], np.uint8)
    assert oracle.to_luma8(img, oracle.RGB8)[0].tolist() == [54, 182, 18, 255, 0]


def test_srgb_luma_every_triple_is_the_truncated_rec709_value():
    img = all_triples()
    got = oracle.to_luma8(img, oracle.RGB8).reshape(-1).astype(np.int64)
    r, g, b = (img.reshape(-1, 3)[:, k].astype(np.float64) for k in range(3))
    y = 0.2126 * r + 0.7152 * g + 0.0722 * b
    # truncation, not rounding: got <= y < got + 1 (up to the float64 error of y)
    assert np.all(got <= y + 1e-9)
    assert np.all(y < got + 1 - 1e-9)


def test_lumaa_takes_the_luma_byte():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (17, 23, 2), dtype=np.uint8)
    assert np.array_equal(oracle.to_luma8(img, oracle.LUMAA8), img[:, :, 0])


def test_canvas_luma_every_triple_against_exact_arithmetic():
    # src/wasm.rs:88-94: luma = 0.299 * r + 0.587 * g + 0.114 * b in f32, then
    # `as u8`. With the f32 constants taken exactly, the real sum fixes the
    # result wherever it is not within the f32 rounding error of an integer.
    img = all_triples(4, alpha_seed=5)
    got = oracle.to_luma8(img, oracle.RGBA8_CANVAS).reshape(-1).astype(np.int64)
    c = [float(np.float32(v)) for v in (0.299, 0.587, 0.114)]
    r, g, b = (img.reshape(-1, 4)[:, k].astype(np.float64) for k in range(3))
    exact = c[0] * r + c[1] * g + c[2] * b  # products exact in f64; sum within 1e-13
    fl = np.floor(exact)
    near = np.abs(exact - np.round(exact)) < 1e-4
    assert np.array_equal(got[~near], np.minimum(fl[~near], 255).astype(np.int64))
    assert np.all(np.abs(got[near] - np.round(exact[near])) <= 1)
    # the near-integer cases follow the f32 operation order of the reference
    r32, g32, b32 = (x[near].astype(np.float32) for x in (r, g, b))
    c32 = [np.float32(v) for v in (0.299, 0.587, 0.114)]
    s32 = (c32[0] * r32 + c32[1] * g32) + c32[2] * b32
    assert np.array_equal(got[near], np.minimum(np.floor(s32), 255).astype(np.int64))


def test_canvas_luma_white_and_single_triples():
    # exact rational check of a few triples, independent of numpy's float paths
    for rgb in [(255, 255, 255), (1, 1, 1), (100, 50, 25), (0, 255, 0), (17, 0, 255)]:
        c = [Fraction(float(np.float32(v))) for v in (0.299, 0.587, 0.114)]
        exact = sum(ci * v for ci, v in zip(c, rgb))
        img = np.array([[list(rgb) + [255]]], np.uint8)
        got = int(oracle.to_luma8(img, oracle.RGBA8_CANVAS)[0, 0])
        assert abs(got - min(int(exact), 255)) <= 1, (rgb, got, float(exact))
        if abs(exact - round(exact)) > Fraction(1, 10000):
            assert got == min(int(exact), 255), (rgb, got, float(exact))


def test_descriptor_u8_closed_forms():
    d = np.array([0.0, 0.2, 255 / 512, 0.5, 1 / 512, np.nextafter(np.float32(1 / 512), 0),
                  0.1, 1.0, 0.0019], np.float32)
    # trunc(clamp(d * 512, 0, 255)): 0, 102.4, 255, 256 -> 255, 1, 0.99.., 51.2, 512, 0.97
    assert oracle.descriptor_u8(d).tolist() == [0, 102, 255, 255, 1, 0, 51, 255, 0]


def test_wgpu_descriptor_view_divides_by_255():
    u = np.array([0, 1, 51, 255], np.uint8)
    v = sift_hip.wgpu_descriptor_view(u)
    assert v.dtype == np.float32
    assert v.tolist() == [np.float32(x) / np.float32(255) for x in (0, 1, 51, 255)]


def test_color_struct_infers_formats():
    for shape, fmt in [((5, 7), sift_hip.PIXEL_GRAY8), ((5, 7, 3), sift_hip.PIXEL_RGB8),
                       ((5, 7, 4), sift_hip.PIXEL_RGBA8), ((5, 7, 2), sift_hip.PIXEL_LUMAA8)]:
        a, s = sift_hip._color_struct(np.zeros(shape, np.uint8))
        assert (s.width, s.height, s.format) == (7, 5, fmt)
        assert s.stride == 7 * sift_hip.PIXEL_BYTES[fmt]
    a, s = sift_hip._color_struct(np.zeros((5, 7, 4), np.uint8), sift_hip.PIXEL_RGBA8_CANVAS)
    assert s.format == sift_hip.PIXEL_RGBA8_CANVAS
    with pytest.raises(ValueError):
        sift_hip._color_struct(np.zeros((5, 7, 3), np.uint8), sift_hip.PIXEL_RGBA8)
    with pytest.raises(ValueError):
        sift_hip._color_struct(np.zeros((5, 7, 5), np.uint8))
