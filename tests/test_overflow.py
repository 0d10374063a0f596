"""Scratch overflow: a frame whose clip queue, big-triangle checkpoint store
or tile bins overflowed is detected (the device's overflow word), grown and
run again at readback, and the re-run frame matches the reference.  Small
capacities are forced with mop_hip_set_scratch_limits (frame driver) or
mop_hip_set_default_scratch_limits (the drop-in path, whose device the
viewport core creates).  Also the ordered path's long tile lists: tiles
overlapped by more than 4096 screen triangles."""
import pytest

import mop_amd
import refcore
from mop_amd import scenes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")]


def test_forced_small_clip_queue_and_checkpoints_frame_driver():
    from mop_amd import driver
    lib = mop_amd.lib()
    sc = scenes.config3(320, 180, n_quads=96, n_spheres=8, shadow_size=1024)
    print(f"[case] forced limits clip=8 chk=64 on {sc.name}", flush=True)
    ref = refcore.render(sc, 1)
    dev = driver.Device()
    try:
        before = lib.mop_hip_overflow_reruns()
        dev.overflow_flags(reset=True)
        dev.scratch_limits(clip=8, chk=64)
        drv = driver.SceneDriver(dev, sc, shadow_size=1024)
        fb = dev.framebuffer(sc.width, sc.height)
        drv.render(fb)
        flags = dev.overflow_flags(reset=True)  # the first run, before readback
        got = fb.read(hdr=True)                 # grows the stores, runs again
        st = dev.frame_stats()
        after_flags = dev.overflow_flags(reset=True)
        reruns = lib.mop_hip_overflow_reruns() - before
        fb.close()
        drv.close()
    finally:
        dev.close()
    print(f"[overflow] first run flags {flags:#x}, re-run flags {after_flags:#x}, "
          f"re-runs {reruns}, stats {st}", flush=True)
    assert st["clipped"] > 8 and st["big_subtris"] > 0
    assert flags & 1 and flags & 32
    assert after_flags == 0 and reruns >= 1
    cmp = refcore.compare(got, ref)
    print(f"[parity] forced limits {sc.name}: {cmp}", flush=True)
    refcore.assert_parity(cmp)


def test_forced_small_tile_bins_dropin():
    lib = mop_amd.lib()
    sc = scenes.ordered_scene(12, 480, 320)
    print(f"[case] forced limits bins=256 on {sc.name}", flush=True)
    ref = refcore.render(sc, 1)
    before = lib.mop_hip_overflow_reruns()
    lib.mop_hip_set_default_scratch_limits(0, 0, 256)
    try:
        hip = refcore.render(sc, 2)
    finally:
        lib.mop_hip_set_default_scratch_limits(0, 0, 0)
    reruns = lib.mop_hip_overflow_reruns() - before
    cmp = refcore.compare(hip, ref)
    print(f"[parity] forced bins {sc.name}: {cmp}, re-runs {reruns}", flush=True)
    assert reruns >= 1
    refcore.assert_parity(cmp)


def test_dense_ordered_long_tile_lists():
    sc = scenes.dense_ordered_scene()
    print(f"[case] {sc.name} tris={sc.triangle_count}", flush=True)
    ref = refcore.render(sc, 1)
    hip = refcore.render(sc, 2)
    cmp = refcore.compare(hip, ref)
    print(f"[parity] {sc.name}: {cmp}", flush=True)
    refcore.assert_parity(cmp)
