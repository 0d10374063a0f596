"""Batched form: frames rendered together in one set of launches must equal
the same frames rendered alone (bit for bit, every buffer), and the
serving loop must complete every admitted request."""
import numpy as np
import pytest

from mop_amd import scenes

pytestmark = pytest.mark.gpu


def _scenes():
    out = []
    for k, (w, h) in enumerate([(96, 72), (160, 90), (64, 64), (200, 120), (128, 96)]):
        sc = scenes.small_scene(seed=40 + k, width=w, height=h, n=6 + k, with_env=(k % 2 == 1))
        out.append(sc)
    out.append(scenes.config3(120, 68, n_quads=48, n_spheres=4, shadow_size=512))
    return out


def test_batch_equals_alone():
    from mop_amd import driver
    dev = driver.Device()
    try:
        scs = _scenes()
        drv = [driver.SceneDriver(dev, sc) for sc in scs]
        fbs = [dev.framebuffer(sc.width, sc.height) for sc in scs]
        alone = []
        for d, fb in zip(drv, fbs):
            d.render(fb)
            alone.append(fb.read(hdr=True))
        for d, fb in zip(drv, fbs):
            d.record(fb)
        dev.execute(fbs)
        for k, fb in enumerate(fbs):
            got = fb.read(hdr=True)
            for key in ("color", "depth", "ids", "hdr"):
                assert np.array_equal(got[key].view(np.uint8), alone[k][key].view(np.uint8)), \
                    (scs[k].name, key)
            assert (got["ids"] > 0).any()
        for fb in fbs:
            fb.close()
        for d in drv:
            d.close()
    finally:
        dev.close()


def test_serving_loop_completes_requests():
    from mop_amd import driver, serve
    dev = driver.Device()
    try:
        pool = []
        for k in range(3):
            sc = scenes.small_scene(seed=60 + k, width=64, height=48, n=5)
            pool.append((driver.SceneDriver(dev, sc, shadow_cache=True), sc))
        srv = serve.Server(dev, pool, budget_bytes=4 << 30, max_batch=8)
        reqs = serve.make_requests(rate=400.0, duration=0.5, n_scenes=len(pool))
        res = srv.run(reqs, time_limit=20.0)
        print("serving:", res)
        assert res["completed"] + res["rejected"] == len(reqs)
        assert res["completed"] > 0 and res["p99_ms"] is not None
        plain = srv.run(reqs, plain=True, time_limit=20.0)
        print("plain:", plain)
        assert plain["completed"] + plain["rejected"] == len(reqs)
        srv.close()
        for d, _ in pool:
            d.close()
    finally:
        dev.close()


def test_serving_results_equal_alone():
    # continuous batching: requests recorded into free slots while earlier
    # batches are still on the device; every kept result must equal the same
    # request rendered alone, bit for bit
    from mop_amd import driver, serve
    dev = driver.Device()
    try:
        pool = []
        for k in range(3):
            sc = scenes.small_scene(seed=70 + k, width=64, height=48, n=5, with_env=(k == 1))
            pool.append((driver.SceneDriver(dev, sc, shadow_cache=True), sc))
        srv = serve.Server(dev, pool, budget_bytes=4 << 30, max_batch=8, inflight=2)
        reqs = serve.make_requests(rate=600.0, duration=0.25, n_scenes=len(pool), seed=1234)
        keep = list(range(0, min(len(reqs), 24), 3))
        res = srv.run(reqs, time_limit=20.0, keep=keep)
        print("serving:", {k: v for k, v in res.items() if k != "outputs"}, flush=True)
        assert res["batches"] < res["completed"]  # requests really shared launches
        assert len(res["outputs"]) == len(keep)
        for k in keep:
            t, si, yaw, pitch, cls = reqs[k]
            (w, h), _ = serve.RES_CLASSES[cls]
            fb = dev.framebuffer(w, h)
            pool[si][0].render(fb, driver.camera_of(pool[si][1], yaw, pitch))
            alone = fb.read(hdr=True)
            fb.close()
            for key in ("color", "depth", "ids", "hdr"):
                assert np.array_equal(alone[key].view(np.uint8),
                                      res["outputs"][k][key].view(np.uint8)), (k, key)
        srv.close()
        for d, _ in pool:
            d.close()
    finally:
        dev.close()
