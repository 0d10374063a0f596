"""The .mop v2 scene file read into HBM (mop_hip_scene_file_*), against the
reference's own saver and loader (mop_scene_save / mop_scene_load /
mop_scene_get_mesh, src/loader/mop_scene.c) built under oracle/_ref.

CPU: the host side -- header and table-of-contents checks reject exactly the
files the reference rejects; counts, camera and lights read back as saved;
the committed fixture decodes (reference loader) to its committed arrays.
GPU: the device decode of quantized and raw sections (16- and 32-bit
indices) equals the reference's host decode bit for bit, and a scene whose
geometry comes from the file renders like the reference core fed the
reference-decoded meshes."""
import ctypes as C
import dataclasses
import os
import struct

import numpy as np
import pytest

import refcore
import mop_amd
from mop_amd import abi, scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built")


def _lib():
    return mop_amd.lib()


def _saved(tmp_path, quantize=True, big=False, name="s.mop"):
    sc = scenes.file_scene(big=big)
    path = str(tmp_path / name)
    refcore.save_scene_file(sc, path, quantize)
    return sc, path


def _ours_info(path):
    lib = _lib()
    f = lib.mop_hip_scene_file_load(path.encode())
    if not f:
        return None
    try:
        eye, tgt, up = abi.Vec3(), abi.Vec3(), abi.Vec3()
        fov, n, fa = C.c_float(), C.c_float(), C.c_float()
        assert lib.mop_hip_scene_file_get_camera(f, C.byref(eye), C.byref(tgt), C.byref(up),
                                                 C.byref(fov), C.byref(n), C.byref(fa))
        cam = np.array([eye.x, eye.y, eye.z, tgt.x, tgt.y, tgt.z, up.x, up.y, up.z,
                        fov.value, n.value, fa.value], np.float32)
        lights = []
        for k in range(lib.mop_hip_scene_file_light_count(f)):
            L = abi.Light()
            assert lib.mop_hip_scene_file_get_light(f, k, C.byref(L))
            lights.append(L)
        return (lib.mop_hip_scene_file_mesh_count(f), lib.mop_hip_scene_file_light_count(f),
                cam, lights)
    finally:
        lib.mop_hip_scene_file_free(f)


def test_counts_camera_lights_match_reference(tmp_path):
    sc, path = _saved(tmp_path)
    ref = refcore.scene_file_info(path)
    ours = _ours_info(path)
    assert ref is not None and ours is not None
    assert ours[:2] == ref[:2] == (len(sc.meshes), len(sc.lights))
    assert np.array_equal(ours[2], ref[2])
    for L, S in zip(ours[3], sc.lights):
        assert L.type == S.type and L.active and not L.cast_shadows
        f32 = lambda t: np.asarray(t, np.float32)  # noqa: E731
        assert np.array_equal(f32([L.intensity, L.color.r, L.color.g, L.color.b]),
                              f32([S.intensity, *S.color[:3]]))
        if S.type != 0:
            assert np.array_equal(f32([L.position.x, L.position.y, L.position.z]),
                                  f32(S.position))


def _corrupt(path, out, offset, fmt, value):
    b = bytearray(open(path, "rb").read())
    struct.pack_into(fmt, b, offset, value)
    open(out, "wb").write(bytes(b))


def test_rejects_what_the_reference_rejects(tmp_path):
    sc, path = _saved(tmp_path)
    cases = {}
    cases["magic"] = str(tmp_path / "magic.mop")
    _corrupt(path, cases["magic"], 0, "<I", 0x4D4F5001)
    cases["version"] = str(tmp_path / "version.mop")
    _corrupt(path, cases["version"], 4, "<I", 3)
    cases["sections"] = str(tmp_path / "sections.mop")
    _corrupt(path, cases["sections"], 8, "<I", 257)
    cases["toc"] = str(tmp_path / "toc.mop")  # the table of contents cut short
    open(cases["toc"], "wb").write(open(path, "rb").read()[:64 + 24 * 2])
    short = str(tmp_path / "short.mop")
    open(short, "wb").write(open(path, "rb").read()[:40])
    cases["short"] = short
    cases["missing"] = str(tmp_path / "does_not_exist.mop")
    for name, p in cases.items():
        assert refcore.scene_file_info(p) is None, name
        assert _ours_info(p) is None, name
        assert b"scene_file_load" in _lib().mop_hip_last_error(), name


def test_truncated_mesh_section_is_refused(tmp_path):
    # a mesh section running past the end of the file (mop_scene.c:407-410)
    sc, path = _saved(tmp_path)
    data = open(path, "rb").read()
    cut = str(tmp_path / "cut.mop")
    open(cut, "wb").write(data[:-64])
    lib = _lib()
    f = lib.mop_hip_scene_file_load(cut.encode())
    assert f
    try:
        last = lib.mop_hip_scene_file_mesh_count(f) - 1
        lm = abi.LoadedMesh()
        assert not lib.mop_hip_scene_file_get_mesh(None, f, last, C.byref(lm))
        assert b"out of bounds" in lib.mop_hip_last_error()
        assert not lm.vertex_buffer and not lm.index_buffer
    finally:
        lib.mop_hip_scene_file_free(f)
    with pytest.raises(RuntimeError):
        refcore.scene_file_mesh(cut, last)


@pytest.mark.parametrize("quantize", [True, False])
def test_writer_bytes_equal_reference_saver(tmp_path, quantize):
    # mop_amd.mopfile.save against mop_scene_save (mop_scene.c:544-768) for the
    # same meshes, lights and camera; the big terrain takes 32-bit indices
    from mop_amd import mopfile
    sc, path = _saved(tmp_path, quantize=quantize, big=True)
    ours = str(tmp_path / "ours.mop")
    mopfile.save(sc, ours, quantized=quantize)
    a, b = open(ours, "rb").read(), open(path, "rb").read()
    assert len(a) == len(b)
    diff = np.flatnonzero(np.frombuffer(a, np.uint8) != np.frombuffer(b, np.uint8))
    assert diff.size == 0, f"{diff.size} bytes differ, first at {diff[:8]}"


def test_golden_fixture_decodes_to_committed_arrays():
    # tests/golden/scene_quantized.mop was written by the reference's saver;
    # its decode by the reference loader is committed beside it
    path = os.path.join(GOLDEN, "scene_quantized.mop")
    want = np.load(os.path.join(GOLDEN, "scene_quantized_decoded.npz"))
    nm = refcore.scene_file_info(path)[0]
    assert nm == int(want["n_meshes"])
    for k in range(nm):
        v, i, box = refcore.scene_file_mesh(path, k)
        assert np.array_equal(v.view(np.uint32), want[f"v{k}"].view(np.uint32))
        assert np.array_equal(i, want[f"i{k}"])
        assert np.array_equal(box, want[f"box{k}"])


# ---------------------------------------------------------------- GPU ----

def _device_meshes(dev, path):
    lib = _lib()
    f = lib.mop_hip_scene_file_load(path.encode())
    assert f, lib.mop_hip_last_error()
    out = []
    try:
        for k in range(lib.mop_hip_scene_file_mesh_count(f)):
            lm = abi.LoadedMesh()
            ok = lib.mop_hip_scene_file_get_mesh(dev.dev, f, k, C.byref(lm))
            assert ok, lib.mop_hip_last_error()
            vp = lib.mop_hip_buffer_read(lm.vertex_buffer)
            ip = lib.mop_hip_buffer_read(lm.index_buffer)
            v = np.ctypeslib.as_array((C.c_float * (lm.vertex_count * 12)).from_address(vp))
            i = np.ctypeslib.as_array((C.c_uint32 * lm.index_count).from_address(ip))
            out.append({"v": v.reshape(-1, 12).copy(), "i": i.copy(),
                        "box": np.array([lm.bbox_min.x, lm.bbox_min.y, lm.bbox_min.z,
                                         lm.bbox_max.x, lm.bbox_max.y, lm.bbox_max.z],
                                        np.float32),
                        "aabb": np.array([lm.aabb_min.x, lm.aabb_min.y, lm.aabb_min.z,
                                          lm.aabb_max.x, lm.aabb_max.y, lm.aabb_max.z],
                                         np.float32)})
            lib.mop_hip_loaded_mesh_free(dev.dev, C.byref(lm))
    finally:
        lib.mop_hip_scene_file_free(f)
    return out


def _check_decode(path, got):
    for k, g in enumerate(got):
        v, i, box = refcore.scene_file_mesh(path, k)
        nbad = int((g["v"].view(np.uint32) != v.view(np.uint32)).any(axis=1).sum())
        print(f"[decode] {os.path.basename(path)} mesh {k}: {len(v)} vertices, "
              f"{len(i)} indices, {nbad} vertices differ", flush=True)
        assert nbad == 0
        assert np.array_equal(g["i"], i)
        assert np.array_equal(g["box"], box)
        assert np.array_equal(g["aabb"][:3], v[:, :3].min(0))
        assert np.array_equal(g["aabb"][3:], v[:, :3].max(0))


@pytest.mark.gpu
@pytest.mark.parametrize("quantize", [True, False])
def test_device_decode_matches_reference_decode(tmp_path, quantize):
    from mop_amd import driver
    sc, path = _saved(tmp_path, quantize=quantize, big=True)
    dev = driver.Device()
    try:
        got = _device_meshes(dev, path)
    finally:
        dev.close()
    assert len(got) == len(sc.meshes)
    assert any(len(g["v"]) >= 65536 for g in got)  # a 32-bit index section
    _check_decode(path, got)


@pytest.mark.gpu
def test_device_decode_of_golden_fixture():
    from mop_amd import driver
    path = os.path.join(GOLDEN, "scene_quantized.mop")
    want = np.load(os.path.join(GOLDEN, "scene_quantized_decoded.npz"))
    dev = driver.Device()
    try:
        got = _device_meshes(dev, path)
    finally:
        dev.close()
    for k, g in enumerate(got):
        assert np.array_equal(g["v"].view(np.uint32), want[f"v{k}"].view(np.uint32))
        assert np.array_equal(g["i"], want[f"i{k}"])


@pytest.mark.gpu
def test_scene_from_file_renders_like_reference(tmp_path):
    # geometry decoded on the device feeding the resident frame driver, against
    # the reference core drawing the reference-decoded meshes
    from mop_amd import driver
    sc, path = _saved(tmp_path, quantize=True, big=True)
    meshes = []
    for k, m in enumerate(sc.meshes):
        v, i, _ = refcore.scene_file_mesh(path, k)
        meshes.append(dataclasses.replace(m, vertices=v, indices=i))
    dec = dataclasses.replace(sc, meshes=meshes)
    print(f"[case] file scene {dec.name} tris={dec.triangle_count}", flush=True)
    ref = refcore.render(dec, 1)
    dev = driver.Device()
    try:
        drv = driver.SceneDriver(dev, dataclasses.replace(dec, mop_file=path), shadow_size=1024)
        fb = dev.framebuffer(dec.width, dec.height)
        drv.render(fb)
        got = fb.read(hdr=True)
        fb.close()
        drv.close()
    finally:
        dev.close()
    cmp = refcore.compare(got, ref)
    print(f"[parity] file scene: {cmp}", flush=True)
    refcore.assert_parity(cmp)
