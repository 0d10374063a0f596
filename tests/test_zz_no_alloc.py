"""The allocation-free shadow entry point (mop_hip_shadow_render_noalloc)
against the shadow pass it shares kernels with (mop_hip_shadow_begin / draw
/ end): the same caster meshes and light matrix give bit-identical maps, on
two calls in a row (the second with a smaller input and map, so stale
scratch contents of the first would show), and oversized input is refused
with an error."""
import ctypes as C

import numpy as np
import pytest

import mop_amd
from mop_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _light_vp():
    def norm(v):
        return v / np.linalg.norm(v)
    eye, tgt, up = np.array([6.0, 10.0, 4.0]), np.zeros(3), np.array([0.0, 1.0, 0.0])
    f = norm(tgt - eye)
    s = norm(np.cross(f, up))
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[0, 3], view[1, 3], view[2, 3] = -s @ eye, -u @ eye, f @ eye
    l, r, b, t, n, fa = -8.0, 8.0, -8.0, 8.0, 0.1, 30.0
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1], proj[2, 2] = 2 / (r - l), 2 / (t - b), -2 / (fa - n)
    proj[0, 3], proj[1, 3], proj[2, 3], proj[3, 3] = -(r + l) / (r - l), -(t + b) / (t - b), \
        -(fa + n) / (fa - n), 1.0
    m = (proj @ view).astype(np.float32)
    return abi.Mat4((C.c_float * 16)(*m.T.reshape(-1).tolist()))  # column-major


def _mat(m):
    if m is None:
        return abi.Mat4((C.c_float * 16)(*np.eye(4, dtype=np.float32).reshape(-1).tolist()))
    return abi.Mat4((C.c_float * 16)(*np.asarray(m, np.float32).reshape(-1).tolist()))


def _buffers(lib, dev, sc, count):
    out = []
    for m in sc.meshes[:count]:
        v = np.ascontiguousarray(m.vertices, np.float32)
        i = np.ascontiguousarray(m.indices, np.uint32)
        vb = lib.mop_hip_buffer_create(dev, C.byref(abi.BufferDesc(v.ctypes.data, v.nbytes)))
        ib = lib.mop_hip_buffer_create(dev, C.byref(abi.BufferDesc(i.ctypes.data, i.nbytes)))
        assert vb and ib
        out.append((vb, len(v), ib, len(i), _mat(m.model)))
    return out


def _declare(lib):
    f = lib.mop_hip_shadow_render_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                  C.POINTER(abi.Mat4), C.POINTER(abi.Mat4), C.c_int]
    g = lib.mop_hip_shadow_map_read
    g.restype = C.c_int
    g.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
    return f, g


def _read(lib, dev, smap, size):
    out = np.zeros(size * size, np.float32)
    assert lib.mop_hip_shadow_map_read(dev, smap, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return out


def test_noalloc_shadow_matches_the_shadow_pass():
    lib = mop_amd.lib()
    noalloc, _ = _declare(lib)
    dev = lib.mop_hip_device_create()
    assert dev, lib.mop_hip_last_error()
    try:
        sc = scenes.small_scene(seed=3, n=12)
        lvp = _light_vp()
        bufs = _buffers(lib, dev, sc, len(sc.meshes))
        for call, (count, size) in enumerate([(len(bufs), 512), (3, 256)]):
            print(f"[case] no-alloc shadow call {call}: {count} casters, {size}^2", flush=True)
            ref_map = lib.mop_hip_shadow_map_create(dev, size)
            got_map = lib.mop_hip_shadow_map_create(dev, size)
            assert ref_map and got_map
            lib.mop_hip_shadow_begin(dev, ref_map)
            for vb, nv, ib, ni, model in bufs[:count]:
                lib.mop_hip_shadow_draw(dev, vb, nv, ib, ni, C.byref(model), C.byref(lvp))
            lib.mop_hip_shadow_end(dev, C.byref(lvp))
            for k, (vb, nv, ib, ni, model) in enumerate(bufs[:count]):
                rc = noalloc(dev, got_map, vb, nv, ib, ni, C.byref(model), C.byref(lvp),
                             1 if k == 0 else 0)
                assert rc == 0, lib.mop_hip_last_error()
            ref = _read(lib, dev, ref_map, size)
            got = _read(lib, dev, got_map, size)
            print(f"[no-alloc] call {call}: {int((ref < 1.0).sum())} texels drawn, "
                  f"{int((ref.view(np.uint32) != got.view(np.uint32)).sum())} differ", flush=True)
            assert (ref < 1.0).sum() > 100
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
            lib.mop_hip_shadow_map_destroy(dev, ref_map)
            lib.mop_hip_shadow_map_destroy(dev, got_map)
        # beyond the scratch maximum: refused with an error, nothing launched
        big = np.zeros(((1 << 19) + 1, 12), np.float32)
        idx = np.zeros(3, np.uint32)
        vb = lib.mop_hip_buffer_create(dev, C.byref(abi.BufferDesc(big.ctypes.data, big.nbytes)))
        ib = lib.mop_hip_buffer_create(dev, C.byref(abi.BufferDesc(idx.ctypes.data, idx.nbytes)))
        smap = lib.mop_hip_shadow_map_create(dev, 64)
        ident = _mat(None)
        assert noalloc(dev, smap, vb, len(big), ib, 3, C.byref(ident), C.byref(lvp), 1) == -1
        assert b"maximum" in lib.mop_hip_last_error()
        lib.mop_hip_shadow_map_destroy(dev, smap)
        lib.mop_hip_buffer_destroy(dev, vb)
        lib.mop_hip_buffer_destroy(dev, ib)
        for vb, _, ib, _, _ in bufs:
            lib.mop_hip_buffer_destroy(dev, vb)
            lib.mop_hip_buffer_destroy(dev, ib)
    finally:
        lib.mop_hip_device_destroy(dev)
