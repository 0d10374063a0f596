"""Device IBL precompute (ibl.hip) against the tables the reference viewport
core computes on the host (environment.c:115-484) for the same equirect
environment.  The tables must be bit-identical: the precompute's sinf/cosf
arguments are evaluated on the host with the same C library, and its
atan2f/asinf are that library's algorithms (libmf.h); everything else is
IEEE single precision in the reference's order."""
import ctypes as C

import numpy as np
import pytest

import refcore
from mop_amd import scenes

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")]


def test_ibl_tables_match_reference():
    import mop_amd
    sc = scenes.small_scene(seed=5, with_env=True)
    vp = refcore.RefViewport(sc, 1)
    ref = vp.ibl_tables()
    vp.close()
    lib = mop_amd.lib()
    dev = lib.mop_hip_device_create()
    assert dev
    try:
        env = np.ascontiguousarray(sc.env, np.float32)
        ibl = lib.mop_hip_ibl_create(dev, env.ctypes.data_as(C.POINTER(C.c_float)),
                                     env.shape[1], env.shape[0])
        assert ibl
        irr = np.zeros_like(ref["irr"])
        pf = np.zeros_like(ref["pf"])
        lut = np.zeros_like(ref["lut"])
        F = C.POINTER(C.c_float)
        assert lib.mop_hip_ibl_download(dev, ibl, irr.ctypes.data_as(F), pf.ctypes.data_as(F),
                                        lut.ctypes.data_as(F)) == 0
        for name, a, b in (("irradiance", irr, ref["irr"]), ("prefiltered", pf, ref["pf"]),
                           ("brdf_lut", lut, ref["lut"])):
            err = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
            print(f"{name}: max rel {err.max():.3e}, exact {np.mean(a == b):.4f}", flush=True)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
        lib.mop_hip_ibl_destroy(dev, ibl)
    finally:
        lib.mop_hip_device_destroy(dev)
