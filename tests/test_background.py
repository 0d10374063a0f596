"""The environment background on the device: the RHI draw_skybox slot
(inverse view-projection and camera position, the arguments the core passes
to GPU backends, viewport.c:2209-2219) against the CPU skybox form
(mop_hip_draw_background_hdri, bit-identical with the CPU backend -- see the
drop-in parity tests).  The two derive each pixel's ray by different
arithmetic, so they agree to rounding, not bit for bit."""
import ctypes as C
import math

import numpy as np
import pytest

import mop_amd
from mop_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _camera(eye, target, up, fov_deg, w, h, near=0.1, far=100.0):
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    t = math.tan(math.radians(fov_deg) / 2)
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = 1 / ((w / h) * t), 1 / t
    proj[2, 2], proj[2, 3], proj[3, 2] = -(far + near) / (far - near), \
        -2 * far * near / (far - near), -1.0
    inv = np.linalg.inv(proj @ view).astype(np.float32)
    return (C.c_float * 16)(*inv.T.reshape(-1).tolist())


def test_draw_skybox_matches_cpu_skybox_form():
    lib = mop_amd.lib()
    lib.mop_hip_draw_skybox.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float,
                                        C.c_float]
    lib.mop_hip_draw_background_hdri.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                                 abi.Vec3, abi.Vec3, abi.Vec3, C.c_float,
                                                 C.c_float]
    dev = lib.mop_hip_device_create()
    assert dev, lib.mop_hip_last_error()
    try:
        env = np.ascontiguousarray(scenes.sky_env(512, 256), np.float32)
        tex = lib.mop_hip_texture_create_hdr(dev, env.shape[1], env.shape[0],
                                             env.ctypes.data_as(C.POINTER(C.c_float)))
        assert tex
        W, H = 320, 200
        eye, target, up, fov = (3.0, 1.5, 6.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), 55.0
        inv_vp = _camera(eye, target, up, fov, W, H)
        cam = (C.c_float * 3)(*eye)
        out = []
        for form in ("skybox", "cpu"):
            fb = lib.mop_hip_framebuffer_create(dev, C.byref(abi.FramebufferDesc(W, H)))
            lib.mop_hip_frame_begin(dev, fb, abi.Color(0.1, 0.1, 0.11, 1.0))
            if form == "skybox":
                lib.mop_hip_draw_skybox(dev, fb, tex, inv_vp, cam, 0.3, 1.5)
            else:
                lib.mop_hip_draw_background_hdri(dev, fb, tex, math.radians(fov),
                                                 abi.Vec3(*eye), abi.Vec3(*target),
                                                 abi.Vec3(*up), 0.3, 1.5)
            lib.mop_hip_frame_end(dev, fb)
            w, h = C.c_int(), C.c_int()
            p = lib.mop_hip_framebuffer_read_hdr(dev, fb, C.byref(w), C.byref(h))
            out.append(np.ctypeslib.as_array(p, (W * H * 4,)).reshape(H, W, 4).copy())
            lib.mop_hip_framebuffer_destroy(dev, fb)
        lib.mop_hip_texture_destroy(dev, tex)
    finally:
        lib.mop_hip_device_destroy(dev)
    a, b = out
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
    print(f"[skybox] max rel {rel.max():.3e}, 99.9th pct {np.percentile(rel, 99.9):.3e}, "
          f"bit-identical {np.mean(a == b):.4f}", flush=True)
    assert (b[..., 3] == 1.0).all() and (b[..., :3] > 0).all()
    assert np.percentile(rel, 99.9) <= 1e-4 and rel.max() <= 1e-2
