"""Autotuner for the batched frame launches.

Picks, per batch shape, the launch shapes the frame kernels take at run time
(`mop_hip_set_tuning`): the bbox area up to which one lane walks a screen
triangle (larger ones get one wave with a lane per row), the shading tile
edge (8, 16 or 32 pixels, i.e. 64/256/1024-lane workgroups) and the number
of waves working through the queue of large triangles.  The
results never depend on them -- the tests check that bit for bit -- only the
time does.  Each candidate is timed on the device (the batch's stream span
measured with HIP events by the backend) after one untimed run.
"""
import ctypes as C
import itertools

SMALL_AREA = (64, 256, 1024)
SHADE_TILE = (8, 16, 32)
RASTER_GRID = (2048, 8192, 32768)


def candidates():
    return list(itertools.product(SMALL_AREA, SHADE_TILE, RASTER_GRID))


def set_tuning(dev, small_area, shade_tile, raster_grid):
    rc = dev.lib.mop_hip_set_tuning(dev.dev, small_area, shade_tile, raster_grid)
    if rc != 0:
        raise ValueError(f"invalid tuning {(small_area, shade_tile, raster_grid)}")


def device_ms(dev):
    f = dev.lib.mop_hip_frame_gpu_time_ms
    f.restype = C.c_float
    return float(f(dev.dev))


def autotune(dev, run_batch, cands=None, reps=2):
    """run_batch() records and executes one batch and waits for it.
    Returns (best (small_area, shade_tile, raster_grid), {candidate: ms})
    and leaves the best tuning set on the device."""
    table = {}
    for cand in cands or candidates():
        set_tuning(dev, *cand)
        run_batch()
        ms = []
        for _ in range(reps):
            run_batch()
            ms.append(device_ms(dev))
        table[cand] = min(ms)
    best = min(table, key=table.get)
    set_tuning(dev, *best)
    return best, table
