"""Chunk culling of large order-free draws (a 256-triangle chunk whose
object-space box fails one trivial-reject plane is skipped, and only the
vertex workgroups of kept chunks are transformed) must not change a frame:
frames of a device with culling (the default) equal, bit for bit, the frames
of a device created with MOP_HIP_CHUNK_CULL=0, for cameras that cull many,
few and no chunks, for direct launches and for a multi-frame execution
replayed from a captured graph (the first execution computes the chunk
boxes, the next ones reuse them), and after the vertex buffer is rewritten
(the boxes are recomputed).  The full-size terrain parity tests compare
culled frames with the reference CPU backend."""
import os

import numpy as np
import pytest

from mop_amd import scenes

pytestmark = pytest.mark.gpu

KEYS = ("color", "depth", "ids", "hdr")


def _same(a, b):
    return {k: bool(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8))) for k in KEYS}


def _device(cull):
    from mop_amd import driver
    old = os.environ.get("MOP_HIP_CHUNK_CULL")
    os.environ["MOP_HIP_CHUNK_CULL"] = "1" if cull else "0"
    try:
        return driver.Device()
    finally:
        if old is None:
            del os.environ["MOP_HIP_CHUNK_CULL"]
        else:
            os.environ["MOP_HIP_CHUNK_CULL"] = old


# (yaw, pitch) orbits of the camera about its target: the view from inside
# the terrain, turned, and raised to look down on it
CAMS = [(0.0, 0.0), (90.0, 0.0), (200.0, 10.0), (30.0, 60.0)]


def _scene():
    # 384 x 384 quads: a 294,912-triangle terrain draw (culled by chunks)
    # and eight 5120-triangle spheres (not)
    return scenes.config3(320, 180, n_quads=384, n_spheres=8, shadow_size=1024)


def test_culled_frames_equal_unculled_frames():
    from mop_amd import driver
    sc = _scene()
    out = {}
    for cull in (True, False):
        dev = _device(cull)
        try:
            drv = driver.SceneDriver(dev, sc, shadow_size=1024)
            fb = dev.framebuffer(sc.width, sc.height)
            frames, culled = [], []
            for yaw, pitch in CAMS:
                drv.render(fb, driver.camera_of(sc, yaw, pitch))
                frames.append(fb.read(hdr=True))
                culled.append(dev.frame_stats()["culled_chunks"])
            assert dev.overflow_flags(reset=True) == 0
            fb.close()
            drv.close()
        finally:
            dev.close()
        print(f"[cull] culling {cull}: culled chunks per camera {culled}", flush=True)
        out[cull] = (frames, culled)
    n_chunks = (2 * 384 * 384 + 255) // 256
    on, off = out[True], out[False]
    assert off[1] == [0] * len(CAMS)
    assert on[1][0] > n_chunks // 4  # the view from inside skips most of the terrain
    for k in range(len(CAMS)):
        same = _same(on[0][k], off[0][k])
        assert all(same.values()), (CAMS[k], same)
        assert (on[0][k]["ids"] != 0).any()


def test_culled_graph_replays_and_rewritten_vertices():
    from mop_amd import driver
    sc = _scene()
    results = {}
    for cull in (True, False):
        dev = _device(cull)
        try:
            drv = driver.SceneDriver(dev, sc, shadow_size=1024)
            fbs = [dev.framebuffer(sc.width, sc.height) for _ in CAMS]
            dev.lib.mop_hip_set_graphs(dev.dev, 1)
            got = []
            for rnd in range(3):
                for fb, (yaw, pitch) in zip(fbs, CAMS):
                    drv.record(fb, driver.camera_of(sc, yaw + rnd, pitch))
                dev.execute(fbs)
                got.append([fb.read(hdr=True) for fb in fbs])
            # rewrite the terrain's vertex buffer (lowered by 2 units): the
            # chunk boxes of the old contents must not be used again
            mesh = sc.meshes[0]
            v = mesh.vertices.copy()
            v[:, 1] -= 2.0
            drv.update_vertices(0, v)
            for fb, (yaw, pitch) in zip(fbs, CAMS):
                drv.record(fb, driver.camera_of(sc, yaw, pitch))
            dev.execute(fbs)
            got.append([fb.read(hdr=True) for fb in fbs])
            assert dev.overflow_flags(reset=True) == 0
            for fb in fbs:
                fb.close()
            drv.close()
        finally:
            dev.close()
        results[cull] = got
    for rnd in range(len(results[True])):
        for k in range(len(CAMS)):
            same = _same(results[True][rnd][k], results[False][rnd][k])
            assert all(same.values()), (rnd, CAMS[k], same)
    # the rewrite changed the frames
    assert not np.array_equal(results[True][0][0]["depth"], results[True][3][0]["depth"])
