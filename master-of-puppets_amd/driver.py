"""Python helpers over the C frame driver (include/mop_hip_scene.h): resident
scenes, framebuffers, single and batched frames, readback to numpy.  The
rendering itself happens in libmop_hip.so."""
import ctypes as C
import math

import numpy as np

from . import abi, lib as _lib


class Device:
    def __init__(self):
        self.lib = _lib()
        self.dev = self.lib.mop_hip_device_create()
        if not self.dev:
            raise RuntimeError("HIP device unavailable: " + self.lib.mop_hip_last_error().decode())

    def framebuffer(self, w, h):
        fb = self.lib.mop_hip_framebuffer_create(self.dev, C.byref(abi.FramebufferDesc(w, h)))
        if not fb:
            raise RuntimeError("framebuffer_create failed: " + self.lib.mop_hip_last_error().decode())
        return Framebuffer(self, fb, w, h)

    def sync(self):
        self.lib.mop_hip_device_sync(self.dev)

    def opaque_raster(self, tiled):
        """Raster path of the opaque draws recorded from now on: False the
        visibility-buffer minimum (default), True the binned tile path."""
        if self.lib.mop_hip_set_opaque_raster(self.dev, 1 if tiled else 0) != 0:
            raise ValueError("mop_hip_set_opaque_raster failed")

    def execute(self, fbs):
        arr = (C.c_void_p * len(fbs))(*[f.fb for f in fbs])
        self.lib.mop_hip_frames_execute(self.dev, arr, len(fbs))

    def overflow_flags(self, reset=False):
        """OR of the scratch-overflow bits of every batch executed since the
        last reset (waits for the device)."""
        return int(self.lib.mop_hip_overflow_flags(self.dev, 1 if reset else 0))

    def scratch_limits(self, clip=0, chk=0, bins=0):
        """Force small scratch capacities for the next executions."""
        if self.lib.mop_hip_set_scratch_limits(self.dev, clip, chk, bins) != 0:
            raise RuntimeError("set_scratch_limits failed")

    def frame_stats(self):
        st = abi.FrameStats()
        self.lib.mop_hip_get_frame_stats(self.dev, C.byref(st))
        return {"triangles_in": st.triangles_in, "subtris": st.subtris, "clipped": st.clipped,
                "big_subtris": st.big_subtris, "shadow_tris": st.shadow_tris,
                "gpu_ms": st.gpu_ms, "culled_chunks": st.culled_chunks}

    def timers(self, on):
        self.lib.mop_hip_timers_enable(self.dev, 1 if on else 0)

    def timer(self, name):
        tot, cnt = C.c_double(), C.c_uint64()
        if self.lib.mop_hip_timers_read(self.dev, name.encode(), C.byref(tot), C.byref(cnt)) != 0:
            return None, 0
        return tot.value, cnt.value

    def close(self):
        if self.dev:
            self.lib.mop_hip_device_destroy(self.dev)
            self.dev = None


class Framebuffer:
    def __init__(self, device, fb, w, h):
        self.d, self.fb, self.w, self.h = device, fb, w, h

    def done(self):
        return self.d.lib.mop_hip_frame_done(self.d.dev, self.fb) == 1

    def readback_done(self):
        """True when the copy readback_async started has reached the host."""
        return self.d.lib.mop_hip_readback_done(self.d.dev, self.fb) == 1

    def read(self, color=True, depth=True, ids=True, hdr=False):
        lib, dev, fb = self.d.lib, self.d.dev, self.fb
        w, h, n = self.w, self.h, self.w * self.h
        iw, ih = C.c_int(), C.c_int()
        out = {}
        if color:
            p = lib.mop_hip_framebuffer_read_color(dev, fb, C.byref(iw), C.byref(ih))
            out["color"] = np.ctypeslib.as_array(p, (n * 4,)).reshape(h, w, 4).copy()
        if depth:
            p = lib.mop_hip_framebuffer_read_depth(dev, fb, C.byref(iw), C.byref(ih))
            out["depth"] = np.ctypeslib.as_array(p, (n,)).reshape(h, w).copy()
        if ids:
            p = lib.mop_hip_framebuffer_read_object_id(dev, fb, C.byref(iw), C.byref(ih))
            out["ids"] = np.ctypeslib.as_array(p, (n,)).reshape(h, w).copy()
        if hdr:
            p = lib.mop_hip_framebuffer_read_hdr(dev, fb, C.byref(iw), C.byref(ih))
            out["hdr"] = np.ctypeslib.as_array(p, (n * 4,)).reshape(h, w, 4).copy()
        return out

    def readback_async(self, color=True, depth=True, ids=True, hdr=False):
        """Start copying this frame's buffers to host memory behind the
        device's later work; read() then waits for the copy."""
        mask = (1 if color else 0) | (2 if depth else 0) | (4 if ids else 0) | (8 if hdr else 0)
        self.d.lib.mop_hip_framebuffer_readback_async(self.d.dev, self.fb, mask)

    def read_color_raw(self):
        """Bring RGBA8 to host (pinned buffer owned by the framebuffer)."""
        iw, ih = C.c_int(), C.c_int()
        return self.d.lib.mop_hip_framebuffer_read_color(self.d.dev, self.fb, C.byref(iw),
                                                          C.byref(ih))

    def close(self):
        if self.fb:
            self.d.lib.mop_hip_framebuffer_destroy(self.d.dev, self.fb)
            self.fb = None


def camera_of(sc, yaw_delta=0.0, pitch_delta=0.0):
    """Camera of a Scene, optionally orbited about its target (degrees)."""
    ex, ey, ez = (sc.eye[i] - sc.target[i] for i in range(3))
    if yaw_delta or pitch_delta:
        dist = math.sqrt(ex * ex + ey * ey + ez * ez)
        yaw = math.atan2(ex, ez) + math.radians(yaw_delta)
        pitch = max(-1.5, min(1.5, math.asin(ey / dist) + math.radians(pitch_delta)))
        ex = dist * math.cos(pitch) * math.sin(yaw)
        ey = dist * math.sin(pitch)
        ez = dist * math.cos(pitch) * math.cos(yaw)
    eye = (sc.target[0] + ex, sc.target[1] + ey, sc.target[2] + ez)
    return abi.Camera(abi.Vec3(*eye), abi.Vec3(*sc.target), abi.Vec3(*sc.up), sc.fov, sc.znear,
                      sc.zfar)


class SceneDriver:
    """A resident scene on the device."""

    def __init__(self, device, sc, shadow_cache=False, shadow_size=None):
        lib = device.lib
        self.d, self.sc = device, sc
        self.s = lib.mop_hip_scene_create(device.dev)
        if not self.s:
            raise RuntimeError("scene_create failed")
        mop_file = getattr(sc, "mop_file", None)
        if mop_file is not None:
            self._add_file_meshes(lib, sc, mop_file)
        for m in (sc.meshes if mop_file is None else ()):
            v = np.ascontiguousarray(m.vertices, np.float32)
            i = np.ascontiguousarray(m.indices, np.uint32)
            mm = None if m.model is None else np.ascontiguousarray(m.model, np.float32)
            rc = lib.mop_hip_scene_add_mesh(self.s, v.ctypes.data, len(v), i.ctypes.data, len(i),
                                            m.object_id, None if mm is None else mm.ctypes.data,
                                            int(m.has_material), m.metallic, m.roughness)
            if rc < 0:
                raise RuntimeError("add_mesh failed: " + lib.mop_hip_last_error().decode())
        # blend mode and opacity as the core's mesh setters leave them
        for k, m in enumerate(sc.meshes):
            blend, op = getattr(m, "blend", 0), getattr(m, "opacity", 1.0)
            if blend != 0 or op < 1.0:
                if lib.mop_hip_scene_set_mesh_blend(self.s, k, blend, min(op, 1.0)) != 0:
                    raise RuntimeError("set_mesh_blend failed")
        for L in sc.lights:
            l = abi.Light()
            l.type = L.type
            l.position = abi.Vec3(*L.position)
            l.direction = abi.Vec3(*L.direction)
            l.color = abi.Color(*L.color)
            l.intensity = L.intensity
            l.range = L.range
            l.spot_inner_cos = L.inner_cos
            l.spot_outer_cos = L.outer_cos
            l.active = True
            l.cast_shadows = L.cast_shadows
            lib.mop_hip_scene_add_light(self.s, C.byref(l))
        if sc.env is not None:
            env = np.ascontiguousarray(sc.env, np.float32)
            if lib.mop_hip_scene_set_environment(
                    self.s, env.ctypes.data_as(C.POINTER(C.c_float)), env.shape[1], env.shape[0],
                    sc.env_rotation, sc.env_intensity) != 0:
                raise RuntimeError("set_environment failed")
        lib.mop_hip_scene_set_shadow_size(self.s, shadow_size or sc.shadow_size)
        lib.mop_hip_scene_set_post_effects(self.s, sc.post_effects)
        lib.mop_hip_scene_set_shadow_cache(self.s, 1 if shadow_cache else 0)
        self.cam = camera_of(sc)

    def _add_file_meshes(self, lib, sc, path):
        """Mesh k of the .mop v2 file decoded on the device (quantized
        sections never leave their 18-byte form on the host) and added with
        the object id, transform and material of sc.meshes[k]."""
        f = lib.mop_hip_scene_file_load(str(path).encode())
        if not f:
            raise RuntimeError("scene_file_load failed: " + lib.mop_hip_last_error().decode())
        try:
            n = lib.mop_hip_scene_file_mesh_count(f)
            if n != len(sc.meshes):
                raise RuntimeError(f"{path} holds {n} meshes, the scene {len(sc.meshes)}")
            for k, m in enumerate(sc.meshes):
                lm = abi.LoadedMesh()
                if not lib.mop_hip_scene_file_get_mesh(self.d.dev, f, k, C.byref(lm)):
                    raise RuntimeError("scene_file_get_mesh failed: " +
                                       lib.mop_hip_last_error().decode())
                mm = None if m.model is None else np.ascontiguousarray(m.model, np.float32)
                rc = lib.mop_hip_scene_add_loaded_mesh(
                    self.s, C.byref(lm), m.object_id, None if mm is None else mm.ctypes.data,
                    int(m.has_material), m.metallic, m.roughness)
                if rc < 0:
                    lib.mop_hip_loaded_mesh_free(self.d.dev, C.byref(lm))
                    raise RuntimeError("add_loaded_mesh failed")
        finally:
            lib.mop_hip_scene_file_free(f)

    def update_vertices(self, k, vertices):
        """Rewrite mesh k's vertices in place (same count)."""
        v = np.ascontiguousarray(vertices, np.float32)
        if self.d.lib.mop_hip_scene_update_mesh_vertices(self.s, k, v.ctypes.data, len(v)) != 0:
            raise RuntimeError("update_mesh_vertices failed")

    def render(self, fb, cam=None):
        return self.d.lib.mop_hip_scene_render(self.s, fb.fb, C.byref(cam or self.cam), fb.w, fb.h)

    def record(self, fb, cam=None):
        return self.d.lib.mop_hip_scene_record(self.s, fb.fb, C.byref(cam or self.cam), fb.w, fb.h)

    def exposure(self):
        return self.d.lib.mop_hip_scene_exposure(self.s)

    def close(self):
        if self.s:
            self.d.lib.mop_hip_scene_destroy(self.s)
            self.s = None
