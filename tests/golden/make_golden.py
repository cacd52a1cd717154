"""Writes the committed golden fixtures for the CPU oracle:
  images.json        SHA-256 of the seeded workload images (512x512, 1920x1080, 3840x2160)
  oracle_small.json  pyramid / candidate / refined hashes and counts for small
                     seeded images, plus oracle_small_*.npz with the final
                     keypoints and descriptors.
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "sift-wgpu_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import imgen  # noqa: E402
import oracle  # noqa: E402

CASES = [
    dict(w=96, h=80, seed=21, params=[1.6, 3, 3, 0.5, 0.04, 10.0]),
    dict(w=160, h=121, seed=22, params=[1.6, 4, 3, 0.5, 0.04, 10.0]),
    dict(w=201, h=77, seed=23, params=[1.6, 4, 2, 0.5, 0.03, 8.0]),
]


def main():
    imgs = {}
    for name in ("512x512", "1920x1080", "3840x2160"):
        img, _, _ = imgen.config_image(name)
        imgs[name] = hashlib.sha256(img.tobytes()).hexdigest()
    with open(os.path.join(HERE, "images.json"), "w") as f:
        json.dump({"workloads": imgs}, f, indent=1)
    out = []
    for i, c in enumerate(CASES):
        img = imgen.generic_image(c["w"], c["h"], seed=c["seed"])
        p = oracle.params(*c["params"])
        kps, desc, cand, ref, _ = oracle.detect_and_compute(p, img, stages=True)
        gs, _ = oracle.build_pyramid(p, img)
        npz = f"oracle_small_{i}.npz"
        np.savez_compressed(os.path.join(HERE, npz),
                            keypoints=np.frombuffer(kps.tobytes(), np.uint8),
                            descriptors=desc)
        out.append(dict(c, pyramid_sha256=hashlib.sha256(b"".join(g.tobytes() for g in gs)).hexdigest(),
                        n_candidates=len(cand), n_refined=len(ref), n_keypoints=len(kps),
                        candidates_sha256=hashlib.sha256(cand.tobytes()).hexdigest(),
                        refined_sha256=hashlib.sha256(ref.tobytes()).hexdigest(), npz=npz))
    with open(os.path.join(HERE, "oracle_small.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps([(c["w"], c["h"], c["n_candidates"], c["n_refined"], c["n_keypoints"]) for c in out]))


if __name__ == "__main__":
    main()
