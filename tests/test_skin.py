"""Skinning and morph targets (viewport.c:1572-1680, 1763-1805).

The reference core skins on its side and uploads the deformed vertices with
buffer_update; after a render those bytes are what buffer_read returns.  The
oracle restatement (mop_oracle_skin_apply / _morph_apply) and the device
kernels (mop_hip_skin_apply / _morph_apply) are given the core's own bind-pose
snapshot and bone matrices / morph targets and must reproduce those bytes
exactly.  The drop-in test renders the skinned (flexible-format) draws through
the core with both backends."""
import ctypes as C

import numpy as np
import pytest

import refcore
from mop_amd import abi, scenes

pytestmark = pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")


def _reference(sc):
    """Reference core (single-threaded CPU backend): per skinned mesh, the
    deformed vertex bytes after one render and the bind-pose snapshot."""
    vp = refcore.RefViewport(sc, 1)
    vp.render()
    out = [vp.skinned_vertex_bytes(k) for k in range(len(sc.skinned))]
    vp.close()
    return out


def _apply(lib, prefix, dev, sm, bind):
    """Run <prefix>skin_apply or <prefix>morph_apply on a fresh buffer holding
    the bind pose; returns the buffer's bytes afterwards."""
    fmt = abi.vertex_format(sm.attrs, sm.stride)
    bind = np.ascontiguousarray(bind)
    desc = abi.BufferDesc(bind.ctypes.data, bind.nbytes)
    create = getattr(lib, prefix + "buffer_create")
    create.restype = C.c_void_p
    create.argtypes = [C.c_void_p, C.c_void_p]
    read = getattr(lib, prefix + "buffer_read")
    read.restype = C.c_void_p
    read.argtypes = [C.c_void_p]
    destroy = getattr(lib, prefix + "buffer_destroy")
    destroy.argtypes = [C.c_void_p, C.c_void_p]
    vb = create(dev, C.byref(desc))
    assert vb
    n = len(bind)
    if sm.bones is not None:
        fn = getattr(lib, prefix + "skin_apply")
        fn.argtypes = [C.c_void_p] * 3 + [C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]
        bn = np.ascontiguousarray(sm.bones, np.float32)
        rc = fn(dev, vb, bind.ctypes.data, n, C.byref(fmt), bn.ctypes.data, len(bn))
    else:
        fn = getattr(lib, prefix + "morph_apply")
        fn.argtypes = [C.c_void_p] * 3 + [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_uint32]
        tg = np.ascontiguousarray(sm.morph_targets, np.float32)
        wt = np.ascontiguousarray(sm.morph_weights, np.float32)
        rc = fn(dev, vb, bind.ctypes.data, n, C.byref(fmt), tg.ctypes.data, wt.ctypes.data,
                len(wt))
    assert rc == 0
    p = read(vb)
    got = np.ctypeslib.as_array((C.c_uint8 * bind.nbytes).from_address(p)).reshape(bind.shape).copy()
    destroy(dev, vb)
    return got


@pytest.mark.parametrize("morph", [False, True])
def test_oracle_skinning_matches_reference_core(morph):
    sc = scenes.skinned_scene(51, morph=morph)
    ref = _reference(sc)
    lib = C.CDLL(refcore.ORACLE_LIB)
    for (deformed, bind), sm in zip(ref, sc.skinned):
        assert bind is not None
        assert not np.array_equal(deformed, bind)  # the core did deform it
        got = _apply(lib, "mop_oracle_", None, sm, bind)
        assert np.array_equal(got, deformed), sc.name


@pytest.mark.gpu
@pytest.mark.parametrize("morph", [False, True])
def test_hip_skinning_kernels_match_reference_core(morph):
    import mop_amd
    sc = scenes.skinned_scene(52, morph=morph)
    print(f"[case] {sc.name}", flush=True)
    ref = _reference(sc)
    lib = mop_amd.lib()
    dev = lib.mop_hip_device_create()
    assert dev, lib.mop_hip_last_error()
    try:
        for (deformed, bind), sm in zip(ref, sc.skinned):
            got = _apply(lib, "mop_hip_", dev, sm, bind)
            diff = int((got != deformed).any(axis=1).sum())
            print(f"[skin] {sc.name} object {sm.object_id}: {diff} differing vertices", flush=True)
            assert diff == 0
    finally:
        lib.mop_hip_device_destroy(dev)


@pytest.mark.gpu
@pytest.mark.parametrize("morph", [False, True])
def test_skinned_meshes_drop_in_parity(morph):
    sc = scenes.skinned_scene(53, morph=morph)
    print(f"[case] {sc.name}", flush=True)
    ref = refcore.render(sc, 1)
    hip = refcore.render(sc, 2)
    cmp = refcore.compare(hip, ref)
    print(f"[parity] {sc.name}: {cmp}", flush=True)
    refcore.assert_parity(cmp)
    assert (ref["ids"] >= 400).any()
