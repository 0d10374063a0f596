"""Autotuner for the batched frame launches.

Picks, per batch shape, the launch shapes the frame kernels take at run time
(`mop_hip_set_tuning`): the bbox area up to which one lane walks a screen
triangle (larger ones get one wave with a lane per row), the shading tile
edge (8, 16 or 32 pixels, i.e. 64/256/1024-lane workgroups) and the number
of waves working through the queue of large triangles; then, for the best of
those, the waves per SIMD the shading pass runs at (`mop_hip_set_shade_waves`:
4, 5, 6 or 8; the kernel is compiled for each).  The results never depend on
them -- the tests check that bit for bit -- only the time does.  Each
candidate is timed on the device (the batch's stream span measured with HIP
events by the backend) after one untimed run.
"""
import ctypes as C
import itertools

SMALL_AREA = (64, 256, 1024)
SHADE_TILE = (8, 16, 32)
RASTER_GRID = (2048, 8192, 32768)
SHADE_WAVES = (4, 5, 6, 8)


def candidates():
    return list(itertools.product(SMALL_AREA, SHADE_TILE, RASTER_GRID))


def set_tuning(dev, small_area, shade_tile, raster_grid):
    rc = dev.lib.mop_hip_set_tuning(dev.dev, small_area, shade_tile, raster_grid)
    if rc != 0:
        raise ValueError(f"invalid tuning {(small_area, shade_tile, raster_grid)}")


def set_shade_waves(dev, waves):
    if dev.lib.mop_hip_set_shade_waves(dev.dev, waves) != 0:
        raise ValueError(f"invalid shading waves per SIMD {waves}")


def device_ms(dev):
    f = dev.lib.mop_hip_frame_gpu_time_ms
    f.restype = C.c_float
    return float(f(dev.dev))


def _time(dev, run_batch, reps):
    run_batch()
    ms = []
    for _ in range(reps):
        run_batch()
        ms.append(device_ms(dev))
    return min(ms)


def autotune(dev, run_batch, cands=None, reps=2, waves=SHADE_WAVES):
    """run_batch() records and executes one batch and waits for it.
    Returns (best (small_area, shade_tile, raster_grid), {candidate: ms},
    best shading waves per SIMD, {waves: ms}) and leaves the best settings
    on the device."""
    set_shade_waves(dev, 5)
    table = {}
    for cand in cands or candidates():
        set_tuning(dev, *cand)
        table[cand] = _time(dev, run_batch, reps)
    best = min(table, key=table.get)
    set_tuning(dev, *best)
    wtable = {}
    for w in waves:
        set_shade_waves(dev, w)
        wtable[w] = _time(dev, run_batch, reps)
    best_w = min(wtable, key=wtable.get)
    set_shade_waves(dev, best_w)
    return best, table, best_w, wtable
