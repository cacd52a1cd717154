"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, and the C
replay of the reference's HIP backend arm.

- oracle/_san/san_check: the oracle (test infrastructure) built with both
  sanitizers, over 1x1 to 333x45 images, every public function (CPU).
- tests/c/arm_replay_asan: the library's host code and the replay program built
  with both sanitizers (device code as usual). Without a GPU it runs the checks
  that need none (parameter helpers, argument validation, drawing); with one it
  makes the Rust arm's exact calls (sift_hip_detect_batch_color with n = 1,
  stride 0, the params struct the arm builds, for Luma8 / Rgb8 / Rgba8 / LumaA8
  buffers) and compares each result with the grey path bit for bit.
- tests/c/arm_replay: the same program against the library as shipped (GPU).

Leak checking is on; tests/c/lsan.supp names the ROCm runtime libraries, whose
allocations live until the process ends.
"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_CHECK = os.path.join(REPO, "oracle", "_san", "san_check")
REPLAY = os.path.join(REPO, "tests", "c", "arm_replay")
REPLAY_ASAN = REPLAY + "_asan"
SUPP = os.path.join(REPO, "tests", "c", "lsan.supp")


def _need(path, make_dir, target):
    """Builds `target` on the CPU if it is missing (build() normally has)."""
    if not os.path.exists(path):
        subprocess.run(["make", "-j8", target], cwd=make_dir, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    assert os.path.exists(path), path


def _run(path, timeout):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1"
    env["LSAN_OPTIONS"] = f"suppressions={SUPP}:print_suppressions=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([path], capture_output=True, text=True, timeout=timeout, env=env)
    out = r.stdout + r.stderr
    print(out[-4000:])
    return r.returncode, out


def test_oracle_under_sanitizers():
    _need(SAN_CHECK, os.path.join(REPO, "oracle"), "san")
    rc, out = _run(SAN_CHECK, 300)
    assert rc == 0, out[-2000:]
    assert "SAN_CHECK_OK" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out


def test_abi_host_checks_under_sanitizers():
    """On a machine without a GPU: the checks that need none; with one, all."""
    _need(REPLAY_ASAN, os.path.join(REPO, "sift-wgpu_amd"), "arm_replay")
    rc, out = _run(REPLAY_ASAN, 300)
    assert rc == 0, out[-2000:]
    assert "ARM_REPLAY_NO_DEVICE" in out or "ARM_REPLAY_OK" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out


@pytest.mark.gpu
def test_arm_replay_on_gpu():
    assert os.path.exists(REPLAY), "build the replay first (make -C sift-wgpu_amd arm_replay)"
    rc, out = _run(REPLAY, 120)
    assert rc == 0, out[-2000:]
    assert out.strip().splitlines()[-1] == "ARM_REPLAY_OK"


@pytest.mark.gpu
def test_arm_replay_under_sanitizers_on_gpu():
    assert os.path.exists(REPLAY_ASAN), "build the replay first (make -C sift-wgpu_amd arm_replay)"
    rc, out = _run(REPLAY_ASAN, 300)
    assert rc == 0, out[-2000:]
    assert "ARM_REPLAY_OK" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
