"""Test helper: render a Scene through the *reference* viewport core.

Loads oracle/_ref/ref_harness.so (the reference library compiled from its own
sources plus oracle/ref_harness.c).  Modes: 0 = reference CPU backend
multithreaded, 1 = reference CPU backend single-threaded (goldens), 2 = the
unmodified reference core driving this repository's HIP backend, 3 = the same
core driving the CPU restatement in oracle/mop_oracle.c.
"""
import ctypes as C
import os
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness.so")

_h = None
_hip_loaded = False


def available():
    return os.path.exists(HARNESS)


def harness():
    global _h
    if _h is None:
        if not available():
            raise RuntimeError(f"{HARNESS} missing: run `make -C oracle`")
        h = C.CDLL(HARNESS, mode=C.RTLD_GLOBAL)
        VP, F = C.c_void_p, C.POINTER(C.c_float)
        h.mh_set_mode.argtypes = [C.c_int]
        h.mh_load_hip.argtypes = [C.c_char_p]
        h.mh_load_backend.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]
        h.mh_viewport.restype = VP
        h.mh_viewport.argtypes = [C.c_int, C.c_int]
        h.mh_viewport_ssaa.restype = VP
        h.mh_viewport_ssaa.argtypes = [C.c_int, C.c_int, C.c_int]
        h.mh_destroy.argtypes = [VP]
        h.mh_add_mesh.restype = VP
        h.mh_add_mesh.argtypes = [VP, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                  C.c_uint32, C.c_void_p, C.c_int, C.c_float, C.c_float,
                                  C.c_int, C.c_float]
        h.mh_add_light.argtypes = [VP, C.c_int, F, F, F, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_int]
        h.mh_add_instanced.argtypes = [VP, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                       C.c_uint32, C.c_void_p, C.c_uint32]
        h.mh_set_camera.argtypes = [VP, F, F, F, C.c_float, C.c_float, C.c_float]
        h.mh_resize.argtypes = [VP, C.c_int, C.c_int]
        h.mh_set_env_background.argtypes = [VP, C.c_int]
        h.mh_set_outline.argtypes = [VP, C.c_int, F, C.c_float, C.c_float]
        h.mh_select.argtypes = [VP, C.c_uint32, C.c_int]
        h.mh_save_scene.argtypes = [VP, C.c_char_p, C.c_uint32]
        h.mh_scene_file_info.argtypes = [C.c_char_p, C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32), C.c_void_p]
        h.mh_scene_file_mesh.argtypes = [C.c_char_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                         C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32), C.c_void_p]
        h.mh_set_shadow_size.argtypes = [C.c_int]
        h.mh_set_clear_color.argtypes = [VP, F]
        h.mh_set_post_effects.argtypes = [VP, C.c_uint32]
        h.mh_set_fog.argtypes = [VP, F, C.c_float, C.c_float]
        h.mh_set_ambient.argtypes = [VP, C.c_float]
        h.mh_set_exposure.argtypes = [VP, C.c_float]
        h.mh_set_shading.argtypes = [VP, C.c_int]
        h.mh_mesh_set_shading.argtypes = [VP, C.c_int]
        h.mh_set_light_dir.argtypes = [VP, F]
        h.mh_set_render_mode.argtypes = [VP, C.c_int]
        h.mh_add_mesh_flex.restype = VP
        h.mh_add_mesh_flex.argtypes = [VP, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        h.mh_mesh_set_bones.argtypes = [VP, VP, C.c_void_p, C.c_uint32]
        h.mh_mesh_set_morph.argtypes = [VP, VP, C.c_void_p, C.c_void_p, C.c_uint32]
        h.mh_mesh_vertex_bytes.argtypes = [VP, VP, C.c_void_p, C.c_size_t]
        h.mh_mesh_bind_pose.restype = C.c_void_p
        h.mh_mesh_bind_pose.argtypes = [VP]
        h.mh_set_environment.argtypes = [VP, C.c_void_p, C.c_int, C.c_int, C.c_char_p,
                                         C.c_float, C.c_float]
        h.mh_get_exposure.restype = C.c_float
        h.mh_get_exposure.argtypes = [VP]
        h.mh_render.restype = C.c_double
        h.mh_render.argtypes = [VP]
        h.mh_triangle_count.restype = C.c_uint32
        h.mh_triangle_count.argtypes = [VP]
        h.mh_read.argtypes = [VP, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        h.mh_pick.argtypes = [VP, C.c_int, C.c_int, C.POINTER(C.c_uint32),
                              C.POINTER(C.c_float)]
        FP = C.POINTER(F)
        IP = C.POINTER(C.c_int)
        h.mh_get_ibl.argtypes = [VP, FP, IP, IP, FP, IP, IP, IP, FP]
        _h = h
    return _h


ORACLE_LIB = os.path.join(REPO, "oracle", "_build", "liboracle.so")
_oracle_loaded = False


def use_oracle():
    """Make mode 3 available (loads the CPU restatement oracle/mop_oracle.c)."""
    global _oracle_loaded
    h = harness()
    if not _oracle_loaded:
        rc = h.mh_load_backend(3, ORACLE_LIB.encode(), b"mop_oracle_", b"oracle")
        if rc != 0:
            raise RuntimeError(f"mh_load_backend(oracle) failed ({rc})")
        _oracle_loaded = True


def use_hip():
    """Make mode 2 available (loads libmop_hip.so into the harness)."""
    global _hip_loaded
    h = harness()
    if not _hip_loaded:
        lib = os.path.join(REPO, "master-of-puppets_amd", "libmop_hip.so")
        rc = h.mh_load_hip(lib.encode())
        if rc != 0:
            raise RuntimeError(f"mh_load_hip failed ({rc})")
        _hip_loaded = True


def _f3(t):
    return (C.c_float * 4)(*(list(t) + [0.0] * (4 - len(t))))


class RefViewport:
    """A reference-core viewport holding one Scene.

    shadow_size: render the core's shadow pass into a map of this size
    instead of the fixed 1024^2 (the harness override, applied to the
    reference rasterizer in modes 0/1 and passed to the routed backend in
    modes 2/3); None keeps the core's size."""

    def __init__(self, scene, mode, env_path=None, shadow_size=None):
        h = harness()
        if mode == 2:
            use_hip()
        elif mode == 3:
            use_oracle()
        if h.mh_set_mode(mode) != 0:
            raise RuntimeError(f"mode {mode} unavailable")
        self.h, self.mode, self.scene = h, mode, scene
        self.shadow_size = shadow_size
        self.size = (scene.width, scene.height)
        self.vp = h.mh_viewport_ssaa(scene.width, scene.height, getattr(scene, "ssaa", 1))
        if not self.vp:
            raise RuntimeError("mop_viewport_create failed")
        self._keep = []
        for m in scene.meshes:
            v = np.ascontiguousarray(m.vertices, np.float32)
            i = np.ascontiguousarray(m.indices, np.uint32)
            model = None
            if m.model is not None:
                mm = np.ascontiguousarray(m.model, np.float32)
                self._keep.append(mm)
                model = mm.ctypes.data
            ok = h.mh_add_mesh(self.vp, v.ctypes.data, len(v), i.ctypes.data, len(i),
                               m.object_id, model, int(m.has_material), m.metallic,
                               m.roughness, getattr(m, "blend", 0), getattr(m, "opacity", 1.0))
            if not ok:
                raise RuntimeError("mop_viewport_add_mesh failed")
            h.mh_mesh_set_shading(ok, getattr(m, "shading", -1))
        for im in getattr(scene, "instanced", []):
            v = np.ascontiguousarray(im.vertices, np.float32)
            i = np.ascontiguousarray(im.indices, np.uint32)
            xf = np.ascontiguousarray(im.transforms, np.float32)
            self._keep += [v, i, xf]
            if h.mh_add_instanced(self.vp, v.ctypes.data, len(v), i.ctypes.data, len(i),
                                  im.object_id, xf.ctypes.data, len(xf)) != 0:
                raise RuntimeError("mop_viewport_add_instanced_mesh failed")
        self.skinned = []
        for sm in getattr(scene, "skinned", []):
            data = np.ascontiguousarray(sm.data, np.uint8)
            idx = np.ascontiguousarray(sm.indices, np.uint32)
            sem = np.asarray([a[0] for a in sm.attrs], np.int32)
            fmt = np.asarray([a[1] for a in sm.attrs], np.int32)
            off = np.asarray([a[2] for a in sm.attrs], np.uint32)
            mm = None if sm.model is None else np.ascontiguousarray(sm.model, np.float32)
            self._keep += [data, idx, sem, fmt, off, mm]
            mesh = h.mh_add_mesh_flex(self.vp, data.ctypes.data, len(data), sm.stride,
                                      idx.ctypes.data, len(idx), sm.object_id, len(sm.attrs),
                                      sem.ctypes.data, fmt.ctypes.data, off.ctypes.data,
                                      None if mm is None else mm.ctypes.data)
            if not mesh:
                raise RuntimeError("mop_viewport_add_mesh_ex failed")
            if sm.morph_targets is not None:
                tg = np.ascontiguousarray(sm.morph_targets, np.float32)
                wt = np.ascontiguousarray(sm.morph_weights, np.float32)
                self._keep += [tg, wt]
                h.mh_mesh_set_morph(self.vp, mesh, tg.ctypes.data, wt.ctypes.data, len(wt))
            if sm.bones is not None:
                bn = np.ascontiguousarray(sm.bones, np.float32)
                self._keep.append(bn)
                h.mh_mesh_set_bones(self.vp, mesh, bn.ctypes.data, len(bn))
            self.skinned.append((mesh, sm))
        # set_light_dir / set_ambient also rewrite lights[0] (viewport.c:1028-1049),
        # so they go before the scene's lights are added
        h.mh_set_light_dir(self.vp, _f3(getattr(scene, "light_dir", (0.3, 1.0, 0.5))))
        if getattr(scene, "ambient", None) is not None:
            h.mh_set_ambient(self.vp, scene.ambient)
        for L in scene.lights:
            h.mh_add_light(self.vp, L.type, _f3(L.position), _f3(L.direction), _f3(L.color),
                           L.intensity, L.range, L.inner_cos, L.outer_cos, int(L.cast_shadows))
        h.mh_set_camera(self.vp, _f3(scene.eye), _f3(scene.target), _f3(scene.up), scene.fov,
                        scene.znear, scene.zfar)
        h.mh_set_post_effects(self.vp, scene.post_effects)
        fog = getattr(scene, "fog", None)
        if fog is not None:  # (rgba, near, far)
            h.mh_set_fog(self.vp, (C.c_float * 4)(*fog[0]), fog[1], fog[2])
        h.mh_set_shading(self.vp, getattr(scene, "shading", 1))
        h.mh_set_render_mode(self.vp, getattr(scene, "render_mode", 0))
        if scene.env is not None:
            env = np.ascontiguousarray(scene.env, np.float32)
            path = env_path or os.path.join(tempfile.gettempdir(),
                                            f"mop_env_{os.getpid()}_{id(self)}.exr")
            rc = h.mh_set_environment(self.vp, env.ctypes.data, env.shape[1], env.shape[0],
                                      path.encode(), scene.env_rotation, scene.env_intensity)
            try:
                os.remove(path)
            except OSError:
                pass
            if rc != 0:
                raise RuntimeError(f"set_environment failed ({rc})")
            if getattr(scene, "env_background", False):
                h.mh_set_env_background(self.vp, 1)
        ol = getattr(scene, "outline", None)
        if ol is not None:
            # {"accent": rgb, "alpha_selected": a, "alpha_unselected": b,
            #  "selected": [ids]} -- the outline overlay with that theme
            h.mh_set_outline(self.vp, 1, _f3(ol["accent"]), ol["alpha_selected"],
                             ol["alpha_unselected"])
            for oid in ol.get("selected", ()):
                h.mh_select(self.vp, int(oid), 1)

    def render(self):
        if self.h.mh_set_mode(self.mode) != 0:
            raise RuntimeError("mode switch failed")
        self.h.mh_set_shadow_size(int(self.shadow_size or 0))
        try:
            return self.h.mh_render(self.vp)
        finally:
            self.h.mh_set_shadow_size(0)

    def set_view(self, width, height, eye):
        """Resize the viewport and move the camera eye (target, up, fov and
        planes stay the scene's)."""
        if (width, height) != self.size:
            self.h.mh_resize(self.vp, width, height)
            self.size = (width, height)
        sc = self.scene
        self.h.mh_set_camera(self.vp, _f3(eye), _f3(sc.target), _f3(sc.up), sc.fov,
                             sc.znear, sc.zfar)

    def read(self, hdr=True):
        """RGBA8 at the viewport size (the core averages SSAA samples,
        viewport.c:3385-3411); depth, ID and HDR at the internal size."""
        w, h = self.size
        k = getattr(self.scene, "ssaa", 1)
        color = np.zeros((h, w, 4), np.uint8)
        depth = np.zeros((h * k, w * k), np.float32)
        ids = np.zeros((h * k, w * k), np.uint32)
        hdr_a = np.zeros((h * k, w * k, 4), np.float32) if hdr else None
        rc = self.h.mh_read(self.vp, color.ctypes.data, depth.ctypes.data, ids.ctypes.data,
                            hdr_a.ctypes.data if hdr else None)
        if rc != 0:
            raise RuntimeError(f"mh_read failed ({rc})")
        return {"color": color, "depth": depth, "ids": ids, "hdr": hdr_a}

    def skinned_vertex_bytes(self, k):
        """Current vertex bytes of skinned mesh k as the backend returns them,
        and its bind-pose snapshot (both (n, stride) uint8)."""
        mesh, sm = self.skinned[k]
        out = np.zeros_like(np.ascontiguousarray(sm.data, np.uint8))
        if self.h.mh_mesh_vertex_bytes(self.vp, mesh, out.ctypes.data, out.nbytes) != 0:
            raise RuntimeError("buffer_read failed")
        bp = self.h.mh_mesh_bind_pose(mesh)
        bind = None
        if bp:
            bind = np.ctypeslib.as_array((C.c_uint8 * out.nbytes).from_address(bp)).reshape(
                out.shape).copy()
        return out, bind

    def exposure(self):
        return self.h.mh_get_exposure(self.vp)

    def triangle_count(self):
        return self.h.mh_triangle_count(self.vp)

    def pick(self, x, y):
        i, d = C.c_uint32(), C.c_float()
        hit = self.h.mh_pick(self.vp, x, y, C.byref(i), C.byref(d))
        return bool(hit), i.value, d.value

    def ibl_tables(self):
        F = C.POINTER(C.c_float)
        irr, pf, lut = F(), F(), F()
        iw, ih, pw, ph, pl = (C.c_int() for _ in range(5))
        if self.h.mh_get_ibl(self.vp, C.byref(irr), C.byref(iw), C.byref(ih), C.byref(pf),
                             C.byref(pw), C.byref(ph), C.byref(pl), C.byref(lut)) != 0:
            return None
        npf = sum(max(pw.value >> l, 1) * max(ph.value >> l, 1) for l in range(pl.value))
        return {
            "irr": np.ctypeslib.as_array(irr, (ih.value, iw.value, 4)).copy(),
            "pf": np.ctypeslib.as_array(pf, (npf * 4,)).copy(),
            "pf_w": pw.value, "pf_h": ph.value, "pf_levels": pl.value,
            "lut": np.ctypeslib.as_array(lut, (256, 256, 4)).copy(),
        }

    def close(self):
        if self.vp:
            self.h.mh_destroy(self.vp)
            self.vp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def save_scene_file(scene, path, quantize=True):
    """The scene's meshes, lights and camera written by the reference's
    mop_scene_save (.mop v2); quantize selects 18-byte quantized vertices."""
    v = RefViewport(scene, 1)
    try:
        rc = v.h.mh_save_scene(v.vp, str(path).encode(), 1 if quantize else 0)
    finally:
        v.close()
    if rc != 0:
        raise RuntimeError(f"mop_scene_save failed ({rc})")


def scene_file_info(path):
    """(mesh count, light count, camera floats) by the reference's loader, or
    None when mop_scene_load rejects the file."""
    h = harness()
    nm, nl = C.c_uint32(), C.c_uint32()
    cam = (C.c_float * 12)()
    if h.mh_scene_file_info(str(path).encode(), C.byref(nm), C.byref(nl), cam) != 0:
        return None
    return nm.value, nl.value, np.array(cam[:], np.float32)


def scene_file_mesh(path, idx):
    """The reference's host decode of mesh idx (mop_scene_get_mesh):
    vertices (n, 12) float32, indices uint32, header box (6,) float32."""
    h = harness()
    vc, ic = C.c_uint32(), C.c_uint32()
    if h.mh_scene_file_mesh(str(path).encode(), idx, None, 0, None, 0, C.byref(vc),
                            C.byref(ic), None) != 0:
        raise RuntimeError(f"mop_scene_get_mesh({idx}) failed")
    v = np.zeros((vc.value, 12), np.float32)
    i = np.zeros(ic.value, np.uint32)
    box = np.zeros(6, np.float32)
    if h.mh_scene_file_mesh(str(path).encode(), idx, v.ctypes.data, vc.value, i.ctypes.data,
                            ic.value, C.byref(vc), C.byref(ic), box.ctypes.data) != 0:
        raise RuntimeError(f"mop_scene_get_mesh({idx}) failed")
    return v, i, box


def render(scene, mode, frames=1, shadow_size=None):
    vp = RefViewport(scene, mode, shadow_size=shadow_size)
    try:
        for _ in range(frames):
            vp.render()
        out = vp.read()
        out["exposure"] = vp.exposure()
        out["triangles"] = vp.triangle_count()
        return out
    finally:
        vp.close()


def compare(a, b):
    """Parity summary between two renders (dicts from read())."""
    res = {}
    res["depth_diff_px"] = int((a["depth"].view(np.uint32) != b["depth"].view(np.uint32)).sum())
    res["id_diff_px"] = int((a["ids"] != b["ids"]).sum())
    dc = np.abs(a["color"].astype(np.int16) - b["color"].astype(np.int16))
    res["color_max_diff"] = int(dc.max())
    res["color_diff_px"] = int((dc.max(-1) > 0).sum())
    res["color_over_1_px"] = int((dc.max(-1) > 1).sum())
    if a.get("hdr") is not None and b.get("hdr") is not None:
        ha, hb = a["hdr"], b["hdr"]
        # relative error; |ref| below 1e-6 is compared against 1e-6 (an
        # absolute floor of 1e-10 at the 1e-4 bar)
        rel = np.abs(ha - hb) / np.maximum(np.abs(hb), HDR_FLOOR)
        res["hdr_max_rel"] = float(rel.max())
        res["hdr_over_tol_px"] = int((rel > HDR_REL_TOL).any(-1).sum())
        res["hdr_bitdiff_px"] = int((ha.view(np.uint32) != hb.view(np.uint32)).any(-1).sum())
    return res


# parity bar for the linear HDR buffer: |a - b| <= 1e-4 * max(|ref|, 1e-6)
HDR_REL_TOL = 1e-4
HDR_FLOOR = 1e-6


def assert_parity(cmp, exact=True):
    """The parity bar.  The required one: depth and object ID bit-exact,
    RGBA8 within 1/255, HDR within 1e-4 relative (when compared).  With
    exact=True (the default) RGBA8 and HDR must also be bit-identical: the
    backend evaluates every shading expression in the reference's order with
    IEEE single precision and the host C library's transcendental results,
    so any difference is a defect, not rounding."""
    assert cmp["id_diff_px"] == 0, cmp
    assert cmp["depth_diff_px"] == 0, cmp
    assert cmp["color_max_diff"] <= 1, cmp
    if "hdr_max_rel" in cmp:
        assert cmp["hdr_max_rel"] <= HDR_REL_TOL, cmp
    if exact:
        assert cmp["color_diff_px"] == 0, cmp
        if "hdr_bitdiff_px" in cmp:
            assert cmp["hdr_bitdiff_px"] == 0, cmp
