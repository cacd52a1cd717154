"""Command-line front end of the Hip backend, mirroring the reference binary
(src/main.rs:12-140): same arguments, environment variable, stdout lines and
output file, so that scripts written for it (bench/benchmark.py:85-97 reads
"SIFT processing took: ...ms" and "Found N keypoints.") drive this one as well.

    python sift_cli.py [--backend hip] <image_path>

Image files are decoded on the host with Pillow (the reference uses the image
crate's decoders; decoding is outside the SIFT path and JPEG decoders may
differ in the last bit). Colour pixels go to the device as they are and are
converted there with to_luma8's rule (src/sift.rs:1202).
"""
import os
import sys
import time

import numpy as np

import sift_hip
from sift_hip import SiftBackend, SiftError


def print_usage():
    # src/main.rs:12-19, with the Hip backend this library provides
    e = sys.stderr
    print("Usage: sift [--backend cpu|gpu|gpu-fallback|hip] <image_path>", file=e)
    print("  --backend cpu          Use CPU backend (default)", file=e)
    print("  --backend gpu          Use GPU (WebGPU) backend", file=e)
    print("  --backend gpu-fallback Use GPU with CPU fallback", file=e)
    print("  --backend hip          Use the MI355X (HIP) backend", file=e)
    print(file=e)
    print("Environment variable SIFT_BACKEND can also be used.", file=e)


def parse_args(argv):
    """src/main.rs:21-57: --backend VALUE, -h/--help, one positional path;
    SIFT_BACKEND when no --backend is given. Without either, the Hip backend
    (the reference's default, WebGpuWithCpuFallback, is not in this library)."""
    backend = None
    path = None
    i = 0
    while i < len(argv):
        a = argv[i]
        if a == "--backend":
            if i + 1 >= len(argv):
                raise ValueError("--backend requires a value")
            backend = SiftBackend.from_str(argv[i + 1])
            i += 2
        elif a in ("--help", "-h"):
            print_usage()
            sys.exit(0)
        elif not a.startswith("-"):
            path = a
            i += 1
        else:
            raise ValueError(f"Unknown argument: {a}")
    if backend is None:
        env = os.environ.get("SIFT_BACKEND")
        try:
            backend = SiftBackend.from_str(env) if env else SiftBackend.Hip
        except ValueError:
            backend = SiftBackend.Hip
    if path is None:
        raise ValueError("No image path provided")
    return backend, path


def rust_duration_debug(ns):
    """`{:?}` of std::time::Duration: the largest unit among s / ms / us / ns
    with a non-zero integer part, then the remainder's digits without trailing
    zeros (e.g. 12.3456ms, 1.5s, 830ns)."""
    ns = int(ns)
    if ns >= 1_000_000_000:
        whole, frac, digits, unit = ns // 1_000_000_000, ns % 1_000_000_000, 9, "s"
    elif ns >= 1_000_000:
        whole, frac, digits, unit = ns // 1_000_000, ns % 1_000_000, 6, "ms"
    elif ns >= 1_000:
        whole, frac, digits, unit = ns // 1_000, ns % 1_000, 3, "µs"
    else:
        whole, frac, digits, unit = ns, 0, 0, "ns"
    s = str(whole)
    if frac:
        s += "." + str(frac).rjust(digits, "0").rstrip("0")
    return s + unit


def rust_f32_debug(v):
    """`{:?}` of an f32: the shortest digits that read back as the same f32,
    with ".0" on integers and Rust's exponent form below 1e-4 / from 1e16."""
    v = np.float32(v)
    a = abs(v)  # compared in f32, as Rust compares with 1e-4_f32 and 1e16_f32
    if a != 0 and (a < np.float32(1e-4) or a >= np.float32(1e16)):
        m, e = np.format_float_scientific(v, unique=True, trim="-").split("e")
        return f"{m}e{int(e)}"
    s = np.format_float_positional(v, unique=True, trim="-")
    return s if ("." in s or "inf" in s or "nan" in s) else s + ".0"


def load_image(path):
    """(pixels, SIFT_PIXEL format): grey, RGB, RGBA and grey+alpha files keep
    their channels; other modes are converted to RGB by Pillow."""
    from PIL import Image
    im = Image.open(path)
    im.load()
    if im.mode == "L":
        return np.asarray(im, np.uint8), sift_hip.PIXEL_GRAY8
    if im.mode == "LA":
        return np.asarray(im, np.uint8), sift_hip.PIXEL_LUMAA8
    if im.mode == "RGBA":
        return np.asarray(im, np.uint8), sift_hip.PIXEL_RGBA8
    return np.asarray(im.convert("RGB"), np.uint8), sift_hip.PIXEL_RGB8


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    try:
        backend, path = parse_args(argv)
    except ValueError as e:
        print(f"Error: {e}", file=sys.stderr)
        print_usage()
        return 1
    try:
        os.makedirs("data", exist_ok=True)
    except OSError:
        print("Warning: Could not create data directory. Ensure it exists or you have permissions.",
              file=sys.stderr)
    print(f"Loading image from: {path}")
    try:
        pixels, fmt = load_image(path)
    except Exception as e:  # decoder errors of any kind, as the reference reports them
        print(f"Error loading image '{path}': {e}", file=sys.stderr)
        return 1
    h, w = pixels.shape[:2]
    print(f"Image loaded successfully: {w}x{h}")
    print(f"Selected SIFT backend: {backend.name}")
    sift = sift_hip.Sift.default()
    print("Detecting SIFT keypoints and computing descriptors...")
    t0 = time.perf_counter_ns()
    try:
        if backend is not SiftBackend.Hip:
            raise SiftError(f"backend {backend.name} is not provided by libsift_hip "
                            "(only SiftBackend.Hip is)")
        keypoints, descriptors = sift.detect_batch([pixels], formats=[fmt])[0]
    except SiftError as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    dt = time.perf_counter_ns() - t0
    print(f"SIFT processing took: {rust_duration_debug(dt)}")
    print(f"Found {len(keypoints)} keypoints.")
    print(f"Computed {len(descriptors)} descriptors.")
    if len(keypoints):
        k = keypoints[0]
        deg = np.float32(k["angle"]) * np.float32(57.2957795130823208767981548141051703)
        print(f"Example keypoint [0]: x={float(k['x']):.2f}, y={float(k['y']):.2f}, "
              f"size={float(k['size']):.2f}, angle={float(deg):.2f}, "
              f"response={float(k['response']):.4f}, octave={int(k['octave'])}, "
              f"layer={int(k['layer'])}")
        print(f"Example descriptor [0] length: {descriptors.shape[1]}")
        vals = ", ".join(rust_f32_debug(v) for v in descriptors[0][:10])
        print(f"Example descriptor [0] values (first 10): [{vals}]")
        print("Drawing keypoints on image...")
        out = sift_hip.draw_keypoints_to_image(pixels, keypoints, (255, 0, 0), fmt=fmt)
        out_path = f"data/output_{backend.name}.png"
        print(f"Saving image with keypoints to: {out_path}")
        try:
            from PIL import Image
            Image.fromarray(out, "RGB").save(out_path)
            print("PNG output image saved successfully.")
        except Exception as e:
            print(f"Error saving output image: {e}", file=sys.stderr)
    else:
        print("No keypoints found.")
    return 0


if __name__ == "__main__":
    sys.exit(main())
