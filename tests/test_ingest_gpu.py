"""Colour ingest and the WebGPU output conventions on the device, against the
oracle: to_luma8 / the canvas rule bit for bit over every RGB triple, colour
detection equal to detection on the oracle's grey image, and u8 descriptors
equal to the quantisation of the f32 descriptors."""
import ctypes

import numpy as np
import pytest

import imgen
import oracle
import sift_hip
from parity import MIN_PR, match_keypoints
from test_ingest import all_triples

pytestmark = pytest.mark.gpu


def _rgb_of(gray, seed):
    """A colour image whose channels differ (so the weights matter)."""
    rng = np.random.default_rng(seed)
    g = gray.astype(np.int16)
    r = np.clip(g + rng.integers(-40, 41, g.shape), 0, 255)
    b = np.clip(g + rng.integers(-40, 41, g.shape), 0, 255)
    return np.stack([r, g, b], axis=-1).astype(np.uint8)


@pytest.mark.parametrize("fmt", [oracle.RGB8, oracle.RGBA8, oracle.RGBA8_CANVAS])
def test_to_luma8_every_triple(fmt):
    img = all_triples(4 if fmt != oracle.RGB8 else 3, alpha_seed=11)
    got = sift_hip.to_luma8(img, fmt)
    assert np.array_equal(got, oracle.to_luma8(img, fmt))


@pytest.mark.parametrize("w,h,pad", [(1, 1, 0), (3, 5, 1), (1001, 77, 5), (1024, 3, 0), (1029, 9, 3)])
@pytest.mark.parametrize("fmt", [oracle.GRAY8, oracle.RGB8, oracle.RGBA8, oracle.LUMAA8,
                                 oracle.RGBA8_CANVAS])
def test_to_luma8_ragged_strided(w, h, pad, fmt):
    bpp = oracle._BPP[fmt]
    rng = np.random.default_rng(w * 131 + h * 7 + fmt)
    buf = rng.integers(0, 256, (h, w * bpp + pad), dtype=np.uint8)
    s = sift_hip.SiftColorImage(buf.ctypes.data, w, h, w * bpp + pad, fmt)
    out = np.empty((h, w), np.uint8)
    rc = sift_hip.lib().sift_hip_to_luma8(sift_hip._Ctx.get(0), ctypes.byref(s),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert rc == 0, sift_hip.last_error()
    ref = oracle.to_luma8(np.ascontiguousarray(buf[:, :w * bpp]).reshape(h, w, bpp), fmt)
    assert np.array_equal(out, ref)


def test_to_luma8_rejects_bad_input():
    L = sift_hip.lib()
    ctx = sift_hip._Ctx.get(0)
    buf = np.zeros((4, 12), np.uint8)
    out = np.empty(64, np.uint8)
    op = out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    bad = [sift_hip.SiftColorImage(buf.ctypes.data, 4, 4, 12, 9),     # unknown format
           sift_hip.SiftColorImage(buf.ctypes.data, 4, 4, 11, 1),     # stride < 3 * width
           sift_hip.SiftColorImage(buf.ctypes.data, 0, 4, 12, 1),     # empty
           sift_hip.SiftColorImage(None, 4, 4, 12, 1)]                # null data
    for s in bad:
        assert L.sift_hip_to_luma8(ctx, ctypes.byref(s), op) == -1


@pytest.mark.parametrize("fmt", [oracle.RGB8, oracle.RGBA8_CANVAS])
def test_detect_color_equals_detect_on_the_grey_image(fmt):
    gray = imgen.generic_image(640, 480, seed=71)
    rgb = _rgb_of(gray, 72)
    img = rgb if fmt == oracle.RGB8 else np.concatenate(
        [rgb, np.full(rgb.shape[:2] + (1,), 200, np.uint8)], axis=-1)
    s = sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)
    got = s.detect_batch([img], formats=[fmt])[0]
    ref = s.detect_and_compute(oracle.to_luma8(img, fmt))
    assert len(ref[0]) > 50
    assert got[0].tobytes() == ref[0].tobytes()
    assert got[1].tobytes() == ref[1].tobytes()


def test_batch_of_mixed_formats():
    shapes = [(320, 240), (97, 61), (640, 360), (256, 256)]
    grays = [imgen.generic_image(w, h, seed=80 + i) for i, (w, h) in enumerate(shapes)]
    imgs = [grays[0], _rgb_of(grays[1], 1),
            np.concatenate([_rgb_of(grays[2], 2), np.zeros(grays[2].shape + (1,), np.uint8)], -1),
            np.stack([grays[3], np.full_like(grays[3], 9)], -1)]
    fmts = [oracle.GRAY8, oracle.RGB8, oracle.RGBA8, oracle.LUMAA8]
    s = sift_hip.Sift.default()
    got = s.detect_batch(imgs, formats=fmts)
    ref = s.detect_batch([oracle.to_luma8(im if im.ndim == 3 else im[:, :, None], f)
                          for im, f in zip(imgs, fmts)])
    for g, r in zip(got, ref):
        assert g[0].tobytes() == r[0].tobytes()
        assert g[1].tobytes() == r[1].tobytes()


@pytest.mark.parametrize("conv", [sift_hip.KEYPOINTS_CPU, sift_hip.KEYPOINTS_WGPU_V2])
def test_u8_descriptors_and_keypoint_views(conv):
    imgs = [imgen.generic_image(640, 480, seed=91), imgen.generic_image(333, 221, seed=92)]
    s = sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)
    ref = s.detect_batch(imgs)
    got = s.detect_batch_u8(imgs, keypoint_convention=conv)
    for (gk, gd), (rk, rd) in zip(got, ref):
        assert len(rk) > 0 and gd.dtype == np.uint8 and gd.shape == (len(rk), 128)
        assert np.array_equal(gd, oracle.descriptor_u8(rd))
        if conv == sift_hip.KEYPOINTS_CPU:
            assert gk.tobytes() == rk.tobytes()
        else:
            for f in ("x", "y", "angle"):
                assert np.array_equal(gk[f], rk[f])
            assert np.array_equal(gk["size"], rk["size"] * np.float32(2))
            assert not gk["response"].any() and not gk["octave"].any() and not gk["layer"].any()


def test_u8_descriptors_match_the_oracle():
    img = imgen.generic_image(256, 192, seed=4)
    p = oracle.params(1.6, 4, 3, 0.5, 0.04, 10.0)
    okps, odesc = oracle.detect_and_compute(p, img)
    kps, desc = sift_hip.Sift(1.6, 4, 3, 0.5, 0.04, 10.0).detect_batch_u8([img])[0]
    m = match_keypoints(okps, kps)
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR
    ii = np.array([a for a, _ in m["pairs"]])
    jj = np.array([b for _, b in m["pairs"]])
    # the f32 descriptors agree to 1e-3 in L2, so quantised they differ by at
    # most one step, and only in a few values
    diff = np.abs(desc[jj].astype(np.int16) - oracle.descriptor_u8(odesc[ii]).astype(np.int16))
    assert diff.max() <= 1
    assert (diff > 0).mean() < 0.01


def test_served_color_frames():
    frames = []
    for j in range(6):
        g = imgen.generic_image(320 + 16 * j, 240, seed=300 + j)
        frames.append(_rgb_of(g, 400 + j))
    s = sift_hip.Sift.default()
    ref = [s.detect_and_compute(oracle.to_luma8(f, oracle.RGB8)) for f in frames]
    with sift_hip.SiftServer(s, hbm_budget_bytes=4 << 30) as srv:
        for j, f in enumerate(frames):
            srv.submit(f, j)
        done = {}
        for _ in range(200):
            for tag, st, k, d, *_ in srv.poll(timeout_ms=50):
                assert st == 0
                done[tag] = (k, d)
            if len(done) == len(frames):
                break
    assert sorted(done) == list(range(len(frames)))
    for j in range(len(frames)):
        assert done[j][0].tobytes() == ref[j][0].tobytes()
        assert done[j][1].tobytes() == ref[j][1].tobytes()
