"""Frame plans replayed from captured hipGraphs must give exactly the frames
that direct launches give: every replay is read back (a synchronisation and
host work between replays) and compared bit for bit with the eager frame of
the same camera; the uploaded per-frame data (camera-dependent draws) must
reach the replayed graph."""
import ctypes as C

import numpy as np
import pytest

from mop_amd import scenes

pytestmark = pytest.mark.gpu


def _stats(dev):
    r, c = C.c_uint64(), C.c_uint64()
    dev.lib.mop_hip_graph_stats(dev.dev, C.byref(r), C.byref(c))
    return r.value, c.value


def _same(a, b):
    return all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))
               for k in ("color", "depth", "ids", "hdr"))


def test_graph_replays_match_direct_frames():
    from mop_amd import driver
    dev = driver.Device()
    try:
        sc = scenes.small_scene(seed=71, width=160, height=120, n=6, with_env=True)
        drv = driver.SceneDriver(dev, sc)
        fb = dev.framebuffer(sc.width, sc.height)
        cams = [driver.camera_of(sc, 0.0, 0.0), driver.camera_of(sc, 1.5, -1.0)]
        dev.lib.mop_hip_set_graphs(dev.dev, 0)
        eager = []
        for cam in cams:
            drv.render(fb, cam)
            eager.append(fb.read(hdr=True))
        assert not _same(eager[0], eager[1])  # the cameras differ
        dev.lib.mop_hip_set_graphs(dev.dev, 1)
        r0, c0 = _stats(dev)
        for i in range(8):
            k = i % 2
            drv.render(fb, cams[k])
            got = fb.read(hdr=True)
            print(f"[graph] frame {i} camera {k} stats {_stats(dev)}", flush=True)
            assert _same(got, eager[k]), f"frame {i}"
        r1, c1 = _stats(dev)
        assert r1 - r0 >= 4 and c1 - c0 >= 1
        fb.close()
        drv.close()
    finally:
        dev.close()


def test_graph_replays_of_a_batch():
    from mop_amd import driver
    dev = driver.Device()
    try:
        scs = [scenes.small_scene(seed=80 + k, width=96 + 16 * k, height=72, n=4 + k)
               for k in range(4)]
        drv = [driver.SceneDriver(dev, sc, shadow_cache=True) for sc in scs]
        fbs = [dev.framebuffer(sc.width, sc.height) for sc in scs]
        dev.lib.mop_hip_set_graphs(dev.dev, 0)
        for d, fb in zip(drv, fbs):
            d.record(fb)
        dev.execute(fbs)
        eager = [fb.read(hdr=True) for fb in fbs]
        dev.lib.mop_hip_set_graphs(dev.dev, 1)
        r0, _ = _stats(dev)
        for rnd in range(5):
            for d, fb in zip(drv, fbs):
                d.record(fb)
            dev.execute(fbs)
            for k, fb in enumerate(fbs):
                assert _same(fb.read(hdr=True), eager[k]), (rnd, k)
        assert _stats(dev)[0] - r0 >= 3
        for fb in fbs:
            fb.close()
        for d in drv:
            d.close()
    finally:
        dev.close()
