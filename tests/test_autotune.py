"""The launch shapes the autotuner picks must never change a result: every
candidate (and every waves-per-SIMD variant of the shading pass) renders a mixed batch (large ground-plane triangles, near-plane
clipping, shadows, IBL) bit-identical to the default shapes, and the tuner
leaves a valid candidate set."""
import numpy as np
import pytest

from mop_amd import autotune, scenes

pytestmark = pytest.mark.gpu


def test_every_tuning_renders_the_same_frames():
    from mop_amd import driver
    dev = driver.Device()
    try:
        scs = [scenes.config1(200, 150),
               scenes.small_scene(seed=91, width=160, height=120, n=6, with_env=True),
               scenes.config3(160, 90, n_quads=48, n_spheres=4, shadow_size=512)]
        drv = [driver.SceneDriver(dev, sc) for sc in scs]
        fbs = [dev.framebuffer(sc.width, sc.height) for sc in scs]

        def one_batch():
            for d, fb in zip(drv, fbs):
                d.record(fb)
            dev.execute(fbs)
            dev.sync()

        autotune.set_tuning(dev, 64, 16, 8192)
        one_batch()
        ref = [fb.read(hdr=True) for fb in fbs]
        for cand in autotune.candidates():
            autotune.set_tuning(dev, *cand)
            one_batch()
            for k, fb in enumerate(fbs):
                got = fb.read(hdr=True)
                for key in ("color", "depth", "ids", "hdr"):
                    assert np.array_equal(got[key].view(np.uint8), ref[k][key].view(np.uint8)), \
                        (cand, scs[k].name, key)
        for w in autotune.SHADE_WAVES:
            autotune.set_tuning(dev, 64, 16, 8192)
            autotune.set_shade_waves(dev, w)
            one_batch()
            for k, fb in enumerate(fbs):
                got = fb.read(hdr=True)
                for key in ("color", "depth", "ids", "hdr"):
                    assert np.array_equal(got[key].view(np.uint8), ref[k][key].view(np.uint8)), \
                        ("shade waves", w, scs[k].name, key)
        best, table, best_w, wtable = autotune.autotune(dev, one_batch, reps=1)
        print("autotune:", best, table, best_w, wtable, flush=True)
        assert best in table and len(table) == len(autotune.candidates())
        assert best_w in autotune.SHADE_WAVES and len(wtable) == len(autotune.SHADE_WAVES)
        for fb in fbs:
            fb.close()
        for d in drv:
            d.close()
    finally:
        dev.close()
