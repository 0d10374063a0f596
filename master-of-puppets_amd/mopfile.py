"""Writer of the .mop v2 scene file (src/loader/mop_scene.c), so resident
scenes can be stored and streamed back through the device decoder
(mop_hip_scene_file_* in libmop_hip.so) without the reference core.

The bytes equal what mop_scene_save (mop_scene.c:544-768) writes for a
viewport holding the same meshes, lights and camera: the same section order
(camera, lights, meshes), 8-byte alignment, 16-bit indices below 65,536
vertices, and -- with quantize -- quantize_vertex / oct_encode
(mop_scene.c:121-196) in single precision: positions as u16 fractions of the
mesh box, octahedral normals as i16, colour as u8, uv as u16.  Conversions
to integers truncate toward zero through a 32-bit integer, as the x86-64
build does.
"""
import struct

import numpy as np

from .scenes import Scene

MAGIC = 0x4D4F5002
VERSION = 2
SAVE_QUANTIZE = 0x01
SECTION_MESH, SECTION_CAMERA, SECTION_LIGHT = 1, 2, 3
F32 = np.float32


def _to_int(x, dtype):
    """(dtype)(float) on x86-64: truncation to int32, then the low bits."""
    return np.trunc(x).astype(np.int64).astype(np.int32).astype(dtype)


def quantize(vertices, bmin, bmax):
    """18-byte records of quantize_vertex (mop_scene.c:165-196)."""
    v = np.ascontiguousarray(vertices, F32)
    n = len(v)
    rec = np.zeros(n, dtype=[("pos", "<u2", 3), ("nrm", "<i2", 2), ("col", "u1", 4),
                             ("uv", "<u2", 2)])
    assert rec.itemsize == 18
    for i in range(3):
        rng = F32(bmax[i] - bmin[i])
        if rng > F32(1e-8):
            t = (v[:, i] - F32(bmin[i])) / rng
        else:
            t = np.zeros(n, F32)
        t = np.minimum(np.maximum(t, F32(0.0)), F32(1.0))
        rec["pos"][:, i] = _to_int(t * F32(65535.0) + F32(0.5), np.uint16)
    # oct_encode (mop_scene.c:121-140)
    nx, ny, nz = v[:, 3], v[:, 4], v[:, 5]
    s = (np.abs(nx) + np.abs(ny)) + np.abs(nz)
    s = np.where(s < F32(1e-8), F32(1e-8), s).astype(F32)
    ox, oy = nx / s, ny / s
    sx = np.where(ox >= F32(0.0), F32(1.0), F32(-1.0)).astype(F32)
    sy = np.where(oy >= F32(0.0), F32(1.0), F32(-1.0)).astype(F32)
    fx, fy = (F32(1.0) - np.abs(oy)) * sx, (F32(1.0) - np.abs(ox)) * sy
    neg = nz < F32(0.0)
    ox, oy = np.where(neg, fx, ox).astype(F32), np.where(neg, fy, oy).astype(F32)
    rec["nrm"][:, 0] = _to_int(ox * F32(32767.0), np.int16)
    rec["nrm"][:, 1] = _to_int(oy * F32(32767.0), np.int16)
    for c in range(4):
        rec["col"][:, c] = _to_int(v[:, 6 + c] * F32(255.0) + F32(0.5), np.uint8)
    for c in range(2):
        rec["uv"][:, c] = _to_int(v[:, 10 + c] * F32(65535.0) + F32(0.5), np.uint16)
    return rec.tobytes()


def _pad8(b):
    return b + bytes((-len(b)) % 8)


def _camera_section(sc):
    fov = F32(sc.fov) * (F32(3.14159265358979323846) / F32(180.0))  # viewport.c:1068
    aspect = F32(sc.width) / F32(sc.height) if sc.height > 0 else F32(1.0)
    vals = [*sc.eye, *sc.target, *sc.up, fov, sc.znear, sc.zfar, aspect]
    return struct.pack("<13f", *[float(F32(x)) for x in vals]) + bytes(4)


def _light_section(L):
    vals = [*L.position, *L.direction, *L.color, L.intensity, L.range, L.inner_cos,
            L.outer_cos]
    return struct.pack("<I14f", L.type, *[float(F32(x)) for x in vals]) + bytes(4)


def _mesh_section(m, quantized):
    v = np.ascontiguousarray(m.vertices, F32)
    idx = np.ascontiguousarray(m.indices, np.uint32)
    bmin, bmax = v[:, 0:3].min(0), v[:, 0:3].max(0)
    nv, ni = len(v), len(idx)
    index_size = 2 if nv < 65536 else 4
    hdr = struct.pack("<4I6f", nv, ni, 1 if quantized else 0, index_size,
                      *[float(x) for x in bmin], *[float(x) for x in bmax]) + bytes(8)
    vb = quantize(v, bmin, bmax) if quantized else v.tobytes()
    ib = idx.astype(np.uint16).tobytes() if index_size == 2 else idx.tobytes()
    return hdr + _pad8(vb) + ib


def save(scene: Scene, path, quantized=True):
    """scene's meshes (in order), lights and camera as a .mop v2 file."""
    sections = [(SECTION_CAMERA, 0, _camera_section(scene))]
    sections += [(SECTION_LIGHT, 0, _light_section(L)) for L in scene.lights]
    sections += [(SECTION_MESH, 1 if quantized else 0, _mesh_section(m, quantized))
                 for m in scene.meshes if len(m.vertices)]
    n = len(sections)
    off = 64 + 24 * n
    off += (-off) % 8
    toc, body, at = [], [], off
    for kind, flags, data in sections:
        toc.append(struct.pack("<IIQQ", kind, flags, at, len(data)))
        body.append(_pad8(data))
        at += len(body[-1])
    head = struct.pack("<4I", MAGIC, VERSION, n, SAVE_QUANTIZE if quantized else 0) + bytes(48)
    out = head + b"".join(toc)
    out += bytes((-len(out)) % 8)
    with open(path, "wb") as f:
        f.write(out)
        for b in body:
            f.write(b)
