"""The command-line front end (sift-wgpu_amd/sift_cli.py) against the reference
binary's behaviour (src/main.rs:12-140): argument handling, the stdout lines
bench/benchmark.py:85-97 parses, and Rust's Debug formatting of the values it
prints. The end-to-end run needs the device and is marked gpu; it checks the
CLI's keypoint count and drawing against the oracle's."""
import os

import numpy as np
import pytest

import imgen
import oracle
import sift_cli
import sift_hip
from parity import DESC_L2_TOL, MIN_PR, match_keypoints
from sift_hip import SiftBackend


def test_parse_args(monkeypatch):
    monkeypatch.delenv("SIFT_BACKEND", raising=False)
    assert sift_cli.parse_args(["img.png"]) == (SiftBackend.Hip, "img.png")
    assert sift_cli.parse_args(["--backend", "HIP", "a.png"]) == (SiftBackend.Hip, "a.png")
    assert sift_cli.parse_args(["a.png", "--backend", "gpu"]) == (SiftBackend.WebGpu, "a.png")
    monkeypatch.setenv("SIFT_BACKEND", "v2")
    assert sift_cli.parse_args(["a.png"])[0] is SiftBackend.WebGpuV2
    monkeypatch.setenv("SIFT_BACKEND", "no-such")
    assert sift_cli.parse_args(["a.png"])[0] is SiftBackend.Hip
    for bad in (["--backend"], ["--backend", "x", "a.png"], ["-q", "a.png"], []):
        with pytest.raises(ValueError):
            sift_cli.parse_args(bad)


def test_rust_duration_debug():
    cases = {12_345_678: "12.345678ms", 1_500_000_000: "1.5s", 830: "830ns", 1_000: "1µs",
             1_000_000: "1ms", 2_000_000_000: "2s", 3_000_050: "3.00005ms", 999_999: "999.999µs"}
    for ns, s in cases.items():
        assert sift_cli.rust_duration_debug(ns) == s


def test_benchmark_script_reads_the_timing_line():
    # bench/benchmark.py:85-90 keeps the text after the last ':' and needs "ms"
    line = f"SIFT processing took: {sift_cli.rust_duration_debug(7_250_000)}"
    t = line.split(":")[-1].strip()
    assert t.endswith("ms") and float(t[:-2]) == 7.25


def test_rust_f32_debug():
    cases = [(0.0, "0.0"), (1.0, "1.0"), (0.1, "0.1"), (0.2, "0.2"), (1e-5, "1e-5"),
             (0.0001, "0.0001"), (0.123456789, "0.12345679"), (2.5e-6, "2.5e-6")]
    for v, s in cases:
        assert sift_cli.rust_f32_debug(v) == s, v


def test_usage_on_bad_arguments(capsys):
    assert sift_cli.main(["--frobnicate"]) == 1
    err = capsys.readouterr().err
    assert "Error: Unknown argument: --frobnicate" in err and "Usage: sift" in err


@pytest.mark.gpu
def test_end_to_end(tmp_path, monkeypatch, capsys):
    from PIL import Image
    gray = imgen.generic_image(512, 384, seed=12)
    rgb = np.stack([gray, gray // 2 + 60, 255 - gray], -1).astype(np.uint8)
    path = tmp_path / "in.png"
    Image.fromarray(rgb, "RGB").save(path)
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("SIFT_BACKEND", raising=False)
    assert sift_cli.main([str(path)]) == 0
    out = capsys.readouterr().out.splitlines()
    kps, desc = sift_hip.Sift.default().detect_batch([rgb])[0]
    # the oracle on the same picture (its own to_luma8, Sift::default()):
    # the same keypoints, and the same drawing of them
    p = oracle.params(1.6, 4, 3, 0.5, 0.04, 10.0)
    okps, odesc = oracle.detect_and_compute(p, oracle.to_luma8(rgb, oracle.RGB8))
    m = match_keypoints(okps, kps, odesc, desc)
    assert m["precision"] >= MIN_PR and m["recall"] >= MIN_PR and m["desc_l2_max"] <= DESC_L2_TOL
    assert "Image loaded successfully: 512x384" in out
    assert "Selected SIFT backend: Hip" in out
    assert f"Found {len(okps)} keypoints." in out
    assert f"Computed {len(okps)} descriptors." in out
    assert any(l.startswith("SIFT processing took: ") for l in out)
    drawn = np.asarray(Image.open(tmp_path / "data" / "output_Hip.png"))
    assert np.array_equal(drawn, oracle.draw_keypoints(rgb, oracle.RGB8, okps, (255, 0, 0)))
    # a backend this library does not provide is an error, not a fallback
    assert sift_cli.main(["--backend", "cpu", str(path)]) == 1
