"""Serving path: frames submitted concurrently (mixed sizes, continuous batching
over the server's device lanes, four by default) give the same result as each
frame alone."""
import threading
import time

import numpy as np
import pytest

import devmem
import imgen
import sift_hip

pytestmark = pytest.mark.gpu


def test_result_same_alone_and_in_the_mix():
    shapes = [(640, 480), (1280, 720), (320, 240), (1920, 1080), (97, 61), (640, 480)] * 3
    imgs = [imgen.generic_image(w, h, seed=500 + i) for i, (w, h) in enumerate(shapes)]
    sift = sift_hip.Sift.default()
    alone = [sift.detect_and_compute(im) for im in imgs]
    got = {}
    with sift_hip.SiftServer(sift, max_batch=8) as srv:
        def producer():
            for i, im in enumerate(imgs):
                srv.submit(im, i)
                time.sleep(0.0005 * (i % 3))
        t = threading.Thread(target=producer)
        t.start()
        deadline = time.time() + 60
        while len(got) < len(imgs) and time.time() < deadline:
            for tag, st, kps, desc, ts, t0, t1 in srv.poll(timeout_ms=200):
                assert st == 0
                assert t1 >= t0 >= ts
                got[tag] = (kps, desc)
        t.join()
    assert len(got) == len(imgs)
    for i in range(len(imgs)):
        assert got[i][0].tobytes() == alone[i][0].tobytes(), i
        assert got[i][1].tobytes() == alone[i][1].tobytes(), i


def test_served_config_stream():
    """The served workload, shortened: an open-loop Poisson stream at 1000 req/s
    with the 40/30/20/10% mix of 640x480 / 1280x720 / 1920x1080 / 3840x2160,
    octaves floor(log2 min) - 3 (3840x2160: 8), request j's frame seeded
    100000 + j (a rotating pool of 24 distinct frames), 32 GB HBM budget. Every
    completion equals its frame's result alone, bit for bit."""
    n = 400
    arrivals, which = imgen.served_stream(n, rate=1000.0, seed=7)
    assert set(which) == {0, 1, 2, 3}
    src, frames = imgen.served_frames(which, per_size=[8, 6, 6, 4])
    per_size = [sift_hip.Sift(1.6, imgen.octaves_for(w, h), 3, 0.5, 0.04, 10.0)
                for (w, h) in imgen.SERVED_SIZES]
    assert per_size[3].num_octaves == 8
    size_of = {}
    for j in range(n):
        size_of[int(src[j])] = int(which[j])
    alone = {j: per_size[size_of[j]].detect_and_compute(img) for j, img in frames.items()}
    got = {}
    with sift_hip.SiftServer(sift_hip.Sift.default(), hbm_budget_bytes=32 << 30, max_batch=64) as srv:
        t_start = time.perf_counter()

        def producer():
            for j in range(n):
                delay = t_start + arrivals[j] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                k = int(which[j])
                srv.submit(frames[int(src[j])], j, per_size[k])
        th = threading.Thread(target=producer)
        th.start()
        deadline = time.time() + 120
        while len(got) < n and time.time() < deadline:
            for tag, st, kps, desc, ts, t0, t1 in srv.poll(timeout_ms=200):
                assert st == 0, tag
                got[tag] = (kps, desc)
        th.join()
    assert len(got) == n
    for j in range(n):
        a = alone[int(src[j])]
        assert got[j][0].tobytes() == a[0].tobytes(), j
        assert got[j][1].tobytes() == a[1].tobytes(), j


def test_small_budget_still_completes():
    """An HBM budget that admits only one frame at a time serialises the work
    without failing or mixing results."""
    sift = sift_hip.Sift.default()
    imgs = [imgen.generic_image(640, 480, seed=600 + i) for i in range(4)]
    with sift_hip.SiftServer(sift, hbm_budget_bytes=20 << 20, max_batch=4) as srv:
        for i, im in enumerate(imgs):
            srv.submit(im, i)
        got = {}
        deadline = time.time() + 60
        while len(got) < len(imgs) and time.time() < deadline:
            for tag, st, kps, desc, *_ in srv.poll(timeout_ms=200):
                assert st == 0
                got[tag] = kps
    assert sorted(got) == list(range(4))
    for i, im in enumerate(imgs):
        assert got[i].tobytes() == sift.detect_and_compute(im)[0].tobytes()


def test_budget_bounds_device_memory():
    """The lanes' arenas together take at most the HBM budget (device memory read
    with hipMemGetInfo before the server starts, after it started and after a
    run), and frames larger than a lane's share are refused at submit."""
    sift = sift_hip.Sift.default()
    imgs = [imgen.generic_image(640, 480, seed=700 + i) for i in range(6)]
    alone = [sift.detect_and_compute(im) for im in imgs]  # the global context first
    devmem.synchronize()
    slack = 24 << 20  # weight tables (2 MiB per lane) and the runtime's own allocations
    for budget in (48 << 20, 5 << 30):
        free0, _ = devmem.mem_info()
        with sift_hip.SiftServer(sift, hbm_budget_bytes=budget, max_batch=8) as srv:
            free1, _ = devmem.mem_info()
            assert free0 - free1 <= budget + slack, (budget, free0 - free1)
            for i, im in enumerate(imgs):
                srv.submit(im, i)
            got = {}
            deadline = time.time() + 60
            while len(got) < len(imgs) and time.time() < deadline:
                for tag, st, kps, desc, *_ in srv.poll(timeout_ms=200):
                    assert st == 0
                    got[tag] = (kps, desc)
            free2, _ = devmem.mem_info()
            assert free0 - free2 <= budget + slack, (budget, free0 - free2)
            if budget < (1 << 30):
                with pytest.raises(sift_hip.SiftError):
                    srv.submit(np.zeros((2160, 3840), np.uint8), 99)
        assert sorted(got) == list(range(len(imgs)))
        for i in range(len(imgs)):
            assert got[i][0].tobytes() == alone[i][0].tobytes(), i
            assert got[i][1].tobytes() == alone[i][1].tobytes(), i


def test_oversized_frame_rejected():
    sift = sift_hip.Sift.default()
    with sift_hip.SiftServer(sift, hbm_budget_bytes=1 << 20) as srv:
        with pytest.raises(sift_hip.SiftError):
            srv.submit(np.zeros((1080, 1920), np.uint8), 1)
