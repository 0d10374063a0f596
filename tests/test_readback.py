"""Asynchronous readback (mop_hip_framebuffer_readback_async): frames
rendered into two framebuffers in turn, each copied to the host behind the
next frame's kernels, read back equal to the same frames rendered and read
synchronously -- including a framebuffer re-rendered while its previous
copy may still be in flight."""
import numpy as np
import pytest

from mop_amd import scenes

pytestmark = pytest.mark.gpu


def test_pipelined_readback_equals_synchronous():
    from mop_amd import driver
    sc = scenes.small_scene(seed=81, width=200, height=120, n=8, with_env=True)
    dev = driver.Device()
    try:
        drv = driver.SceneDriver(dev, sc)
        cams = [driver.camera_of(sc, yaw_delta=7.0 * k) for k in range(6)]
        fb = dev.framebuffer(sc.width, sc.height)
        alone = []
        for c in cams:
            drv.render(fb, c)
            alone.append(fb.read(hdr=True))
        fb.close()
        fbs = [dev.framebuffer(sc.width, sc.height) for _ in range(2)]
        got = [None] * len(cams)
        for k, c in enumerate(cams):
            f = fbs[k & 1]
            drv.render(f, c)
            f.readback_async(hdr=True)
            if k >= 1:
                got[k - 1] = fbs[(k - 1) & 1].read(hdr=True)
        got[-1] = fbs[(len(cams) - 1) & 1].read(hdr=True)
        # re-render into a framebuffer right after starting its copy: the new
        # frame waits for the copy, and reads return the new frame
        drv.render(fbs[0], cams[0])
        fbs[0].readback_async()
        drv.render(fbs[0], cams[3])
        again = fbs[0].read(hdr=True)
        for f in fbs:
            f.close()
        drv.close()
    finally:
        dev.close()
    for k in range(len(cams)):
        for key in ("color", "depth", "ids", "hdr"):
            assert np.array_equal(got[k][key].view(np.uint8), alone[k][key].view(np.uint8)), (k, key)
    for key in ("color", "depth", "ids", "hdr"):
        assert np.array_equal(again[key].view(np.uint8), alone[3][key].view(np.uint8)), key
