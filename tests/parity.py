"""Shared checks for HIP-vs-oracle parity.

Tolerances (from the hot path's contract):
  pyramid and DoG levels: max-abs <= 1e-5 on the [0,1] scale (asserted bit-exact
    here: u8 levels equal, DoG floats equal);
  keypoints, sorted by (octave, layer, y, x): matched within 0.05 px in x and y,
    0.05 in size and 1 degree in angle, precision and recall >= 99.5%;
  matched 128-D descriptors: L2 distance <= 1e-3.
The GPU tests also hold the descriptors to DESC_L2_MARGIN = 3e-4, a regression
bound well inside the tolerance: the descriptor's fixed-point sums are scaled
per keypoint by its patch's texel range, which keeps the rounding small next
to the window's norm even for low-contrast windows (measured max 1.4e-4).
"""
import math

import numpy as np

POS_TOL = 0.05
SIZE_TOL = 0.05
ANGLE_TOL = math.radians(1.0)
DESC_L2_TOL = 1e-3
DESC_L2_MARGIN = 3e-4
MIN_PR = 0.995


def sort_keys(kps):
    return np.lexsort((kps["x"], kps["y"], kps["layer"], kps["octave"]))


def match_keypoints(ref, got, ref_desc=None, got_desc=None):
    """Greedy one-to-one matching of `got` against `ref` within the tolerances.
    Returns dict(precision, recall, matched, desc_l2_max, pairs)."""
    buckets = {}
    for j in range(len(got)):
        k = got[j]
        key = (int(k["octave"]), int(k["layer"]), int(math.floor(k["x"])), int(math.floor(k["y"])))
        buckets.setdefault(key, []).append(j)
    used = np.zeros(len(got), bool)
    pairs = []
    for i in range(len(ref)):
        r = ref[i]
        o, l = int(r["octave"]), int(r["layer"])
        fx, fy = int(math.floor(r["x"])), int(math.floor(r["y"]))
        best, bestd = -1, None
        for ddx in (-1, 0, 1):
            for ddy in (-1, 0, 1):
                for j in buckets.get((o, l, fx + ddx, fy + ddy), ()):
                    if used[j]:
                        continue
                    g = got[j]
                    if abs(g["x"] - r["x"]) > POS_TOL or abs(g["y"] - r["y"]) > POS_TOL:
                        continue
                    if abs(g["size"] - r["size"]) > SIZE_TOL:
                        continue
                    da = abs(float(g["angle"]) - float(r["angle"])) % (2 * math.pi)
                    da = min(da, 2 * math.pi - da)
                    if da > ANGLE_TOL:
                        continue
                    d = abs(g["x"] - r["x"]) + abs(g["y"] - r["y"]) + da
                    if best < 0 or d < bestd:
                        best, bestd = j, d
        if best >= 0:
            used[best] = True
            pairs.append((i, best))
    m = len(pairs)
    prec = m / len(got) if len(got) else 1.0
    rec = m / len(ref) if len(ref) else 1.0
    l2max = 0.0
    if ref_desc is not None and got_desc is not None and pairs:
        ii = np.array([p[0] for p in pairs])
        jj = np.array([p[1] for p in pairs])
        l2 = np.linalg.norm(ref_desc[ii].astype(np.float64) - got_desc[jj].astype(np.float64), axis=1)
        l2max = float(l2.max())
    return dict(precision=prec, recall=rec, matched=m, desc_l2_max=l2max, pairs=pairs,
                n_ref=len(ref), n_got=len(got))


def check_keypoint_properties(kps, w, h, num_intervals=3):
    """tests/integration_test.rs:13-62 of the reference."""
    for k in kps:
        assert -1.0 <= k["x"] <= w + 1.0
        assert -1.0 <= k["y"] <= h + 1.0
        assert k["size"] > 0.0
        assert -math.pi <= k["angle"] <= math.pi
        assert k["octave"] >= 0
        assert -2 <= k["layer"] <= num_intervals + 4
