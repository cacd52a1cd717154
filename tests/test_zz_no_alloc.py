"""sift_hip_detect_batch_device_noalloc against sift_hip_detect_batch_device.

The no-allocation entry point works in the library's per-GPU static scratch
array and writes to the caller's buffers. Its results must equal those of the
device entry point it varies (keypoints, descriptors, per-image offsets and
counts bit for bit) on two calls in a row, the second with a smaller input; it
must allocate no device memory (free device memory is unchanged across calls);
it must write nothing past the caller's capacity, and it must refuse a batch
whose working set exceeds the scratch, with a message.
"""
import numpy as np
import pytest

import devmem
import imgen
import sift_hip

pytestmark = pytest.mark.gpu

KP = sift_hip.KEYPOINT_DTYPE.itemsize
GUARD = 4096  # bytes after each output buffer that must stay untouched
FILL = 0xA5


class _Out:
    """Caller-owned device outputs for `capacity` keypoints, with guard bytes."""

    def __init__(self, n, capacity):
        self.n, self.capacity = n, capacity
        self.kps = devmem.DeviceBuffer(capacity * KP + GUARD)
        self.desc = devmem.DeviceBuffer(capacity * 512 + GUARD)
        self.offs = devmem.DeviceBuffer(4 * (n + 1) + GUARD)
        for b in (self.kps, self.desc, self.offs):
            b.upload(np.full(b.nbytes, FILL, np.uint8))

    def guards_intact(self, used_kps):
        for b, used in ((self.kps, used_kps * KP), (self.desc, used_kps * 512),
                        (self.offs, 4 * (self.n + 1))):
            tail = devmem.download(b.addr + used, b.nbytes - used)
            if not np.all(tail == FILL):
                return False
        return True

    def read(self, counts):
        offs = devmem.download(self.offs.addr, 4 * (self.n + 1)).view(np.uint32).astype(np.int64)
        total = int(offs[-1])
        kps = devmem.download(self.kps.addr, total * KP).view(sift_hip.KEYPOINT_DTYPE)
        desc = devmem.download(self.desc.addr, total * 512).view(np.float32).reshape(total, 128)
        assert list(np.diff(offs)) == list(counts)
        return offs, kps, desc

    def free(self):
        for b in (self.kps, self.desc, self.offs):
            b.free()


def _reference(sift, items, params):
    """Results of sift_hip_detect_batch_device (the context's own arena)."""
    counts, kp, de, of = sift.detect_batch_device(items, params)
    n = len(items)
    offs = devmem.download(of, 4 * (n + 1)).view(np.uint32).astype(np.int64)
    total = int(offs[-1])
    kps = devmem.download(kp, total * KP).view(sift_hip.KEYPOINT_DTYPE)
    desc = devmem.download(de, total * 512).view(np.float32).reshape(total, 128)
    return counts, offs, kps, desc


def _upload(imgs, strides=None):
    bufs, items = [], []
    for i, img in enumerate(imgs):
        b, it = devmem.upload_image(img, None if strides is None else strides[i])
        bufs.append(b)
        items.append(it)
    return bufs, items


def test_noalloc_equals_device_entry_point_twice():
    big_shapes = [(1920, 1080), (1280, 720), (640, 480)]
    big = [imgen.generic_image(w, h, seed=900 + i) for i, (w, h) in enumerate(big_shapes)]
    big_params = [sift_hip.Sift(1.6, imgen.octaves_for(w, h), 3, 0.5, 0.04, 10.0) for (w, h) in big_shapes]
    small = [imgen.generic_image(640, 480, seed=950)]
    small_params = [sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)]
    s = sift_hip.Sift.default()
    bb, big_items = _upload(big, [1920, 1280 + 3, 640])
    sb, small_items = _upload(small)
    ref_big = _reference(s, big_items, big_params)
    ref_small = _reference(s, small_items, small_params)
    assert sum(ref_big[0]) > 0 and sum(ref_small[0]) > 0

    out_big = _Out(len(big), sum(ref_big[0]) + 100)
    out_small = _Out(len(small), sum(ref_small[0]) + 10)
    # one host-in / host-out call first, so that the HIP runtime's own one-time
    # set-up of its pinned copy paths happens before free memory is read
    s.detect_and_compute(small[0])
    devmem.synchronize()
    free_before, _ = devmem.mem_info()
    for rep in range(2):
        for items, params, out, ref in ((big_items, big_params, out_big, ref_big),
                                        (small_items, small_params, out_small, ref_small)):
            rc, counts, msg = s.detect_batch_device_noalloc(
                items, out.kps.addr, out.desc.addr, out.capacity, out.offs.addr, params)
            assert rc == 0, msg
            assert counts == list(ref[0])
            offs, kps, desc = out.read(counts)
            assert np.array_equal(offs, ref[1])
            assert kps.tobytes() == ref[2].tobytes()
            assert desc.tobytes() == ref[3].tobytes()
            assert out.guards_intact(int(offs[-1]))
            devmem.synchronize()
            free_now, _ = devmem.mem_info()
            assert free_now == free_before, (rep, free_before - free_now)
    for b in bb + sb:
        b.free()
    out_big.free()
    out_small.free()


def test_noalloc_capacity_too_small():
    """A capacity below the keypoint count: SIFT_ERR_NOMEM, the counts needed,
    nothing written past the capacity."""
    img = [imgen.generic_image(800, 600, seed=960)]
    s = sift_hip.Sift(1.6, 5, 3, 0.5, 0.04, 10.0)
    b, items = _upload(img)
    ref = _reference(s, items, None)
    need = sum(ref[0])
    assert need > 20
    out = _Out(1, need // 2)
    rc, counts, msg = s.detect_batch_device_noalloc(items, out.kps.addr, out.desc.addr,
                                                    out.capacity, out.offs.addr)
    assert rc == -3, (rc, msg)  # SIFT_ERR_NOMEM
    assert "keypoint_capacity" in msg
    assert counts == list(ref[0])
    # the offsets are those found; keypoints and descriptors stop at the capacity
    offs = devmem.download(out.offs.addr, 8).view(np.uint32).astype(np.int64)
    assert np.array_equal(offs, ref[1])
    assert out.guards_intact(out.capacity)
    for x in b:
        x.free()
    out.free()


def test_noalloc_refuses_batch_beyond_scratch():
    """Four 3840x2160 frames at 8 octaves need more than the 256 MiB scratch: a
    clear SIFT_ERR_NOMEM before any launch; a smaller call afterwards works."""
    frames = [np.zeros((2160, 3840), np.uint8) for _ in range(4)]
    s = sift_hip.Sift(1.6, 8, 3, 0.5, 0.04, 10.0)
    b, items = _upload(frames)
    out = _Out(4, 1024)
    rc, counts, msg = s.detect_batch_device_noalloc(items, out.kps.addr, out.desc.addr,
                                                    out.capacity, out.offs.addr)
    assert rc == -3, (rc, msg)  # SIFT_ERR_NOMEM
    assert "scratch" in msg
    rc, counts, msg = s.detect_batch_device_noalloc(items[:1], out.kps.addr, out.desc.addr,
                                                    out.capacity, out.offs.addr)
    assert rc == 0, msg
    assert counts == [0]  # a flat frame has no keypoints
    for x in b:
        x.free()
    out.free()
