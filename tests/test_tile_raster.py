"""The tile raster path for opaque draws (mop_hip_set_opaque_raster): the
order-free draws are binned to 16x16 screen tiles by prefix scan in
submission order and each tile's list is walked in order with its colour,
depth and id held in LDS -- the path the blended draws take -- instead of
the per-pixel (depth, submission) minimum.  Its frames must equal the
reference CPU backend's (cpu_backend.c:588-620, rasterizer.c:1926-2108)
exactly as the default path's do, on the golden scenes and on the full-size
1080p PBR workload; forcing tiny tile bins shows the opaque draws really
went through the bins and that the re-run frame still matches."""
import pytest

import mop_amd
import refcore
from mop_amd import scenes
from test_golden import GOLD, NAMES, load
from golden.gen_golden import small_cases

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")]


def _render_tiled(sc, shadow_size, bins=0):
    from mop_amd import driver
    lib = mop_amd.lib()
    dev = driver.Device()
    try:
        dev.opaque_raster(True)
        before = lib.mop_hip_overflow_reruns()
        dev.overflow_flags(reset=True)
        if bins:
            dev.scratch_limits(bins=bins)
        drv = driver.SceneDriver(dev, sc, shadow_size=shadow_size)
        fb = dev.framebuffer(sc.width, sc.height)
        try:
            drv.render(fb)
            flags = dev.overflow_flags(reset=True)  # the first run, before readback
            out = fb.read(hdr=True)
            out["exposure"] = drv.exposure()
            out["flags"] = flags
            out["reruns"] = lib.mop_hip_overflow_reruns() - before
        finally:
            fb.close()
            drv.close()
        return out
    finally:
        dev.close()


@pytest.mark.parametrize("name", NAMES)
def test_tiled_opaque_frame_matches_golden(name):
    sc = small_cases()[name]
    print(f"[case] tiled opaque raster, driver {name}", flush=True)
    r = _render_tiled(sc, 1024)
    cmp = refcore.compare(r, load(name))
    print(f"[parity] tiled {name}: {cmp}", flush=True)
    assert abs(r["exposure"] - GOLD[name]["exposure"]) <= 1e-6 * GOLD[name]["exposure"]
    refcore.assert_parity(cmp)


def test_tiled_opaque_forced_bin_overflow_reruns_and_matches():
    sc = scenes.config2(320, 180, n_instances=24, subdiv=2, shadow_size=1024,
                        env_size=(256, 128))
    print(f"[case] tiled opaque raster, bins=256, {sc.name}", flush=True)
    ref = refcore.render(sc, 1)
    r = _render_tiled(sc, 1024, bins=256)
    print(f"[overflow] tiled first run flags {r['flags']:#x}, re-runs {r['reruns']}", flush=True)
    assert r["flags"] & 16 and r["reruns"] >= 1
    cmp = refcore.compare(r, ref)
    print(f"[parity] tiled forced bins {sc.name}: {cmp}", flush=True)
    refcore.assert_parity(cmp)


def test_tiled_opaque_config2_full_size():
    sc = scenes.config2()
    print(f"[case] tiled opaque raster, driver {sc.name} tris={sc.triangle_count} shadow 2048",
          flush=True)
    ref = refcore.render(sc, 1, shadow_size=2048)
    r = _render_tiled(sc, 2048)
    cmp = refcore.compare(r, ref)
    print(f"[parity] tiled {sc.name}: {cmp}", flush=True)
    assert r["flags"] == 0
    assert abs(r["exposure"] - ref["exposure"]) <= 1e-6 * ref["exposure"]
    refcore.assert_parity(cmp)
