"""Exactness of the device kernels' f32 shortcuts, checked on the CPU with exact
rational arithmetic (each f32 operation rounded once, as the hardware does).

- q255f (kernels_detect.hip): v / 255 for a byte v as fma(v, c, v * cl) with
  c = RN(1/255) and cl = RN(1/255 - c). The extremum test, refinement inputs
  and orientation gradients use it in place of the IEEE division the reference
  performs (src/sift.rs:97,120: `as f32 / 255.0`), so it must equal that
  division for all 256 byte values.
- div_two_pi (kernels_detect.hip): x / 2pi as x * RN(1/2pi) plus one FMA
  correction, for the orientation bin floor(angle * 36 / 2pi) of
  src/sift.rs:665-668; its floor must equal that of the IEEE quotient for every
  f32 angle (tests/c/div_two_pi_check.c enumerates all of them).
"""
import os
import subprocess
from fractions import Fraction

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rn32(x: Fraction) -> float:
    """Rounds an exact rational to the nearest f32 (ties to even)."""
    c = np.float32(float(x))
    cands = {np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))}

    def key(t):
        bits = int(np.array(t, dtype=np.float32).view(np.uint32))
        return (abs(Fraction(float(t)) - x), bits & 1)

    return float(min(cands, key=key))


C_HI = float.fromhex("0x1.010102p-8")
C_LO = float.fromhex("-0x1.fdfdfep-33")


def _q255f(v: int) -> float:
    lo = Fraction(_rn32(v * Fraction(C_LO)))          # v * cl
    return _rn32(v * Fraction(C_HI) + lo)             # fma(v, c, v * cl)


def test_q255f_constants_split_the_reciprocal():
    assert _rn32(Fraction(1, 255)) == C_HI
    assert _rn32(Fraction(1, 255) - Fraction(C_HI)) == C_LO


def test_q255f_equals_ieee_division_for_every_byte():
    ieee = np.arange(256, dtype=np.float32) / np.float32(255.0)
    got = np.array([_q255f(v) for v in range(256)], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), ieee.view(np.uint32))


def test_plain_reciprocal_product_is_not_enough():
    # the low part matters: v * RN(1/255) alone differs on many bytes
    c = Fraction(C_HI)
    ieee = np.arange(256, dtype=np.float32) / np.float32(255.0)
    plain = np.array([_rn32(v * c) for v in range(256)], dtype=np.float32)
    assert (plain != ieee).sum() > 100


def test_orientation_bin_division_has_the_ieee_floor_for_every_f32(tmp_path):
    src = os.path.join(REPO, "tests", "c", "div_two_pi_check.c")
    exe = str(tmp_path / "div_two_pi_check")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("DIV_TWO_PI_OK")
