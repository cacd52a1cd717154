"""Blur tuning: measures every valid blur form for the pyramid levels of the
workloads below and writes the table libsift_hip.so loads at context creation
(blur_tune.json next to the library).

For each level shape (width, height, tap count) and launch size -- many = 1 for
launches carrying a batch of 8 or more frames (timed with `jobs` copies of the
level), 0 for single frames and the small batches of the served stream -- the
horizontal forms are timed with the built-in vertical form and the vertical
forms with the built-in horizontal form (the two passes are independent), by
sift_hip_tune_blur (median of repeated launches, HIP events). The fastest form
is kept only when it beats the built-in one by more than --margin; otherwise
the entry names the built-in form, so the table never trades a measured loss
for noise. Every form computes the same exact sums, so any choice gives the
same pyramid (tests/test_tune_gpu.py checks each form against the oracle).

    python tune_blur.py [--out blur_tune.json] [--reps 15] [--margin 0.03]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, _HERE)

import imgen  # noqa: E402
import sift_hip  # noqa: E402


def level_sigmas(sigma, S, O, assumed_blur):
    """The initial blur and the per-level blur sigmas (octave o, level s >= 1) in
    the f32 order of Sift's schedule (src/sift.rs:132-147, 1205-1209)."""
    f = np.float32
    sigma, assumed_blur = f(sigma), f(assumed_blur)
    s0 = f(np.sqrt(sigma * sigma - assumed_blur * assumed_blur)) if sigma > assumed_blur else f(0)
    k = f(np.power(f(2.0), f(1.0) / f(S)))

    def powi(a, b):
        r = f(1.0)
        while True:
            if b & 1:
                r = f(r * a)
            b //= 2
            if b == 0:
                break
            a = f(a * a)
        return r

    tgt = [f(sigma * powi(k, s)) for s in range(S + 3)]
    out = []
    for o in range(O):
        prev = f(sigma * powi(f(2.0), o))
        for s in range(1, S + 3):
            t = f(tgt[s] * powi(f(2.0), o))
            out.append((o, s, float(f(np.sqrt(f(t * t - prev * prev))))))
            prev = t
    return float(s0), out


def taps_of(sigma):
    return 2 * int(math.ceil(2.0 * np.float32(sigma))) + 1


def level_shapes(w, h, octaves, S=3, sigma=1.6, assumed_blur=0.5):
    """(width, height, sigma) of every blur of a frame's pyramid."""
    s0, sig = level_sigmas(sigma, S, octaves, assumed_blur)
    shapes = [(w, h, s0)] if s0 > 1e-4 else []
    ws = [(w >> o, h >> o) for o in range(octaves)]
    for o, s, b in sig:
        ww, hh = ws[o]
        if ww > 0 and hh > 0 and b >= 1e-4:
            shapes.append((ww, hh, b))
    return shapes


def workloads():
    """(shape list, jobs per launch, many) of the benchmark, batch and served
    workloads and the single-frame configurations."""
    out = []
    # batched 1920x1080, 5 octaves, 64 frames per launch
    out.append((level_shapes(1920, 1080, 5), 64, 1))
    # the uneven 64-frame batch: about 16 frames of each size per launch
    for (w, h) in imgen.BATCH_SHAPES:
        out.append((level_shapes(w, h, imgen.octaves_for(w, h)), 16, 1))
    # served frames (small batches) and single frames
    for (w, h) in imgen.SERVED_SIZES:
        out.append((level_shapes(w, h, imgen.octaves_for(w, h)), 1, 0))
    for name, (w, h, _, _, _, _, octv, _) in imgen.CONFIGS.items():
        out.append((level_shapes(w, h, octv), 1, 0))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(_HERE, "blur_tune.json"))
    ap.add_argument("--reps", type=int, default=15)
    ap.add_argument("--margin", type=float, default=0.03)
    args = ap.parse_args()
    t_start = time.time()
    seen = {}
    for shapes, jobs, many in workloads():
        for (w, h, sigma) in shapes:
            key = (w, h, taps_of(sigma), many)
            if key not in seen:
                seen[key] = (sigma, jobs)
    entries, measured = [], {}
    for (w, h, nt, many), (sigma, jobs) in sorted(seen.items()):
        base = sift_hip.tune_blur(w, h, sigma, jobs, reps=args.reps)
        hs, vs = {}, {}
        for f in sift_hip.BLUR_H_FORMS:
            r = sift_hip.tune_blur(w, h, sigma, jobs, h_form=f, reps=args.reps)
            if r is not None:
                hs[f] = r[0]
        for f in sift_hip.BLUR_V_FORMS:
            r = sift_hip.tune_blur(w, h, sigma, jobs, v_form=f, reps=args.reps)
            if r is not None:
                vs[f] = r[1]
        h_builtin, v_builtin = base[2], base[3]
        h_best = min(hs, key=hs.get)
        v_best = min(vs, key=vs.get)
        h_pick = h_best if hs[h_best] < hs[h_builtin] * (1.0 - args.margin) else h_builtin
        v_pick = v_best if vs[v_best] < vs[v_builtin] * (1.0 - args.margin) else v_builtin
        entries.append([w, h, nt, many, h_pick, v_pick])
        measured[f"{w}x{h} taps {nt} many {many} jobs {jobs}"] = {
            "h_ms": {k: round(v, 4) for k, v in hs.items()},
            "v_ms": {k: round(v, 4) for k, v in vs.items()},
            "builtin": [h_builtin, v_builtin], "chosen": [h_pick, v_pick]}
        print(f"{w}x{h} taps {nt} many {many}: H {h_pick} ({hs[h_pick]:.4f} ms, built-in "
              f"{h_builtin} {hs[h_builtin]:.4f}), V {v_pick} ({vs[v_pick]:.4f} ms, built-in "
              f"{v_builtin} {vs[v_builtin]:.4f})", flush=True)
    # "entries" last: the library reads the six-field arrays from there on
    table = {"device": "MI355X (gfx950)", "reps": args.reps, "margin": args.margin,
             "seconds": round(time.time() - t_start, 1),
             "measured": measured, "entries": entries}
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {len(entries)} entries to {args.out}")


if __name__ == "__main__":
    main()
