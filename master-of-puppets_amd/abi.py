"""ctypes declarations of include/mop_hip.h (mirror structs and entry points)."""
import ctypes as C

c_float_p = C.POINTER(C.c_float)
c_u32_p = C.POINTER(C.c_uint32)
c_u8_p = C.POINTER(C.c_uint8)
c_int_p = C.POINTER(C.c_int)


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Color(C.Structure):
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float), ("a", C.c_float)]


class Mat4(C.Structure):
    _fields_ = [("d", C.c_float * 16)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", Vec3), ("direction", Vec3),
                ("color", Color), ("intensity", C.c_float), ("range", C.c_float),
                ("spot_inner_cos", C.c_float), ("spot_outer_cos", C.c_float),
                ("active", C.c_bool), ("cast_shadows", C.c_bool), ("viewport", C.c_void_p)]


class FogParams(C.Structure):
    _fields_ = [("color", Color), ("near_dist", C.c_float), ("far_dist", C.c_float)]


class BufferDesc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t)]


class VertexAttrib(C.Structure):
    _fields_ = [("semantic", C.c_int32), ("format", C.c_int32), ("offset", C.c_uint32)]


class VertexFormat(C.Structure):
    _fields_ = [("attribs", VertexAttrib * 12), ("attrib_count", C.c_uint32),
                ("stride", C.c_uint32)]


def vertex_format(attrs, stride):
    """attrs: [(semantic, format, offset)] -> VertexFormat"""
    f = VertexFormat()
    for i, (sem, fmt, off) in enumerate(attrs):
        f.attribs[i] = VertexAttrib(sem, fmt, off)
    f.attrib_count = len(attrs)
    f.stride = stride
    return f


class FramebufferDesc(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int)]


class DrawCall(C.Structure):
    _fields_ = [
        ("vertex_buffer", C.c_void_p), ("index_buffer", C.c_void_p),
        ("vertex_count", C.c_uint32), ("index_count", C.c_uint32), ("object_id", C.c_uint32),
        ("model", Mat4), ("view", Mat4), ("projection", Mat4), ("mvp", Mat4),
        ("base_color", Color), ("opacity", C.c_float), ("light_dir", Vec3),
        ("ambient", C.c_float), ("shading_mode", C.c_int32), ("wireframe", C.c_bool),
        ("depth_test", C.c_bool), ("depth_write", C.c_bool), ("backface_cull", C.c_bool),
        ("texture", C.c_void_p), ("normal_map", C.c_void_p),
        ("metallic_roughness_map", C.c_void_p), ("ao_map", C.c_void_p),
        ("blend_mode", C.c_int32), ("metallic", C.c_float), ("roughness", C.c_float),
        ("emissive", Vec3), ("lights", C.POINTER(Light)), ("light_count", C.c_uint32),
        ("cam_eye", Vec3), ("vertex_format", C.c_void_p), ("line_width", C.c_float),
        ("depth_bias", C.c_float), ("aabb_min", Vec3), ("aabb_max", Vec3),
        ("cast_shadows", C.c_bool)]


class Camera(C.Structure):
    _fields_ = [("eye", Vec3), ("target", Vec3), ("up", Vec3), ("fov_deg", C.c_float),
                ("znear", C.c_float), ("zfar", C.c_float)]


class LoadedMesh(C.Structure):
    _fields_ = [("vertex_buffer", C.c_void_p), ("index_buffer", C.c_void_p),
                ("vertex_count", C.c_uint32), ("index_count", C.c_uint32),
                ("bbox_min", Vec3), ("bbox_max", Vec3), ("aabb_min", Vec3), ("aabb_max", Vec3)]


class FrameStats(C.Structure):
    _fields_ = [("triangles_in", C.c_uint32), ("subtris", C.c_uint32), ("clipped", C.c_uint32),
                ("big_subtris", C.c_uint32), ("shadow_tris", C.c_uint32), ("gpu_ms", C.c_float),
                ("culled_chunks", C.c_uint32)]


# entry points of include/mop_hip.h: name -> (restype, argtypes)
VP = C.c_void_p
SIGS = {
    "mop_hip_device_create": (VP, []),
    "mop_hip_device_destroy": (None, [VP]),
    "mop_hip_buffer_create": (VP, [VP, C.POINTER(BufferDesc)]),
    "mop_hip_buffer_destroy": (None, [VP, VP]),
    "mop_hip_buffer_update": (None, [VP, VP, VP, C.c_size_t, C.c_size_t]),
    "mop_hip_buffer_read": (VP, [VP]),
    "mop_hip_framebuffer_create": (VP, [VP, C.POINTER(FramebufferDesc)]),
    "mop_hip_framebuffer_destroy": (None, [VP, VP]),
    "mop_hip_framebuffer_resize": (None, [VP, VP, C.c_int, C.c_int]),
    "mop_hip_frame_begin": (None, [VP, VP, Color]),
    "mop_hip_frame_end": (None, [VP, VP]),
    "mop_hip_frame_submit": (None, [VP, VP]),
    "mop_hip_frame_wait_readback": (None, [VP, VP]),
    "mop_hip_draw": (None, [VP, VP, C.POINTER(DrawCall)]),
    "mop_hip_draw_instanced": (None, [VP, VP, C.POINTER(DrawCall), C.POINTER(Mat4), C.c_uint32]),
    "mop_hip_pick_read_id": (C.c_uint32, [VP, VP, C.c_int, C.c_int]),
    "mop_hip_pick_read_depth": (C.c_float, [VP, VP, C.c_int, C.c_int]),
    "mop_hip_framebuffer_read_color": (c_u8_p, [VP, VP, c_int_p, c_int_p]),
    "mop_hip_framebuffer_read_object_id": (c_u32_p, [VP, VP, c_int_p, c_int_p]),
    "mop_hip_framebuffer_read_depth": (c_float_p, [VP, VP, c_int_p, c_int_p]),
    "mop_hip_framebuffer_read_hdr": (c_float_p, [VP, VP, c_int_p, c_int_p]),
    "mop_hip_texture_create": (VP, [VP, C.c_int, C.c_int, c_u8_p]),
    "mop_hip_texture_create_hdr": (VP, [VP, C.c_int, C.c_int, c_float_p]),
    "mop_hip_texture_destroy": (None, [VP, VP]),
    "mop_hip_texture_read_rgba8": (C.c_bool, [VP, VP, c_u8_p, C.c_size_t]),
    "mop_hip_frame_gpu_time_ms": (C.c_float, [VP]),
    "mop_hip_set_exposure": (None, [VP, C.c_float]),
    "mop_rhi_backend_hip": (VP, []),
    "mop_hip_shadow_render_mesh": (None, [VP, VP, C.c_uint32, c_u32_p, C.c_uint32,
                                          C.POINTER(Mat4), C.POINTER(Mat4), C.c_int, C.c_int]),
    "mop_hip_shadow_set": (None, [VP, C.c_int, C.c_int, C.POINTER(Mat4)]),
    "mop_hip_shadow_clear": (None, [VP]),
    "mop_hip_ibl_set": (None, [VP, c_float_p, C.c_int, C.c_int, c_float_p, C.c_int, C.c_int,
                               C.c_int, c_float_p, C.c_int, C.c_float, C.c_float]),
    "mop_hip_ibl_clear": (None, [VP]),
    "mop_hip_hdr_resolve": (None, [VP, VP, C.c_float]),
    "mop_hip_postprocess_apply": (None, [VP, VP, C.c_uint32, C.POINTER(FogParams)]),
    "mop_hip_last_error": (C.c_char_p, []),
    "mop_hip_get_frame_stats": (None, [VP, C.POINTER(FrameStats)]),
    "mop_hip_timers_enable": (None, [VP, C.c_int]),
    "mop_hip_timers_read": (C.c_int, [VP, C.c_char_p, C.POINTER(C.c_double),
                                      C.POINTER(C.c_uint64)]),
    "mop_hip_probe_atomic_rate": (C.c_double, [C.c_int]),
    "mop_hip_shadow_map_create": (VP, [VP, C.c_int]),
    "mop_hip_shadow_map_destroy": (None, [VP, VP]),
    "mop_hip_shadow_begin": (None, [VP, VP]),
    "mop_hip_shadow_draw": (None, [VP, VP, C.c_uint32, VP, C.c_uint32, C.POINTER(Mat4),
                                   C.POINTER(Mat4)]),
    "mop_hip_shadow_end": (None, [VP, C.POINTER(Mat4)]),
    "mop_hip_shadow_bind": (None, [VP, VP, C.POINTER(Mat4)]),
    "mop_hip_ibl_create": (VP, [VP, c_float_p, C.c_int, C.c_int]),
    "mop_hip_ibl_destroy": (None, [VP, VP]),
    "mop_hip_ibl_bind": (None, [VP, VP, C.c_float, C.c_float]),
    "mop_hip_ibl_download": (C.c_int, [VP, VP, c_float_p, c_float_p, c_float_p]),
    "mop_hip_scene_create": (VP, [VP]),
    "mop_hip_scene_destroy": (None, [VP]),
    "mop_hip_scene_add_mesh": (C.c_int, [VP, VP, C.c_uint32, VP, C.c_uint32, C.c_uint32, VP,
                                         C.c_int, C.c_float, C.c_float]),
    "mop_hip_scene_add_light": (C.c_int, [VP, C.POINTER(Light)]),
    "mop_hip_scene_set_environment": (C.c_int, [VP, c_float_p, C.c_int, C.c_int, C.c_float,
                                                C.c_float]),
    "mop_hip_scene_set_shadow_size": (None, [VP, C.c_int]),
    "mop_hip_scene_set_shadow_cache": (None, [VP, C.c_int]),
    "mop_hip_scene_set_ambient": (None, [VP, C.c_float]),
    "mop_hip_scene_set_exposure": (None, [VP, C.c_float]),
    "mop_hip_scene_exposure": (C.c_float, [VP]),
    "mop_hip_scene_set_clear_color": (None, [VP, Color]),
    "mop_hip_scene_set_post_effects": (None, [VP, C.c_uint32]),
    "mop_hip_scene_set_mesh_blend": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_float]),
    "mop_hip_scene_update_mesh_vertices": (C.c_int, [VP, C.c_uint32, VP, C.c_uint32]),
    "mop_hip_scene_render": (C.c_uint32, [VP, VP, C.POINTER(Camera), C.c_int, C.c_int]),
    "mop_hip_scene_triangles": (C.c_uint32, [VP]),
    "mop_hip_scene_record": (C.c_uint32, [VP, VP, C.POINTER(Camera), C.c_int, C.c_int]),
    "mop_hip_frame_request_post": (None, [VP, VP, C.c_float, C.c_uint32, C.POINTER(FogParams)]),
    "mop_hip_frames_execute": (None, [VP, C.POINTER(VP), C.c_size_t]),
    "mop_hip_frame_done": (C.c_int, [VP, VP]),
    "mop_hip_set_tuning": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32]),
    "mop_hip_set_shade_waves": (C.c_int, [VP, C.c_uint32]),
    "mop_hip_set_opaque_raster": (C.c_int, [VP, C.c_uint32]),
    "mop_hip_set_graphs": (None, [VP, C.c_int]),
    "mop_hip_graph_stats": (None, [VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mop_hip_skin_apply": (C.c_int, [VP, VP, VP, C.c_uint32, VP, VP, C.c_uint32]),
    "mop_hip_morph_apply": (C.c_int, [VP, VP, VP, C.c_uint32, VP, VP, VP, C.c_uint32]),
    "mop_hip_device_sync": (None, [VP]),
    "mop_hip_overflow_flags": (C.c_uint32, [VP, C.c_int]),
    "mop_hip_set_scratch_limits": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32]),
    "mop_hip_set_default_scratch_limits": (None, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "mop_hip_overflow_reruns": (C.c_uint64, []),
    "mop_hip_framebuffer_readback_async": (None, [VP, VP, C.c_uint32]),
    "mop_hip_readback_done": (C.c_int, [VP, VP]),
    "mop_hip_shadow_map_read": (C.c_int, [VP, VP, c_float_p]),
    "mop_hip_framebuffer_read_color_downsampled": (c_u8_p, [VP, VP, C.c_int, c_int_p, c_int_p]),
    "mop_hip_draw_skybox": (None, [VP, VP, VP, c_float_p, c_float_p, C.c_float, C.c_float]),
    "mop_hip_draw_background_hdri": (None, [VP, VP, VP, C.c_float, Vec3, Vec3, Vec3, C.c_float,
                                            C.c_float]),
    "mop_hip_draw_outline": (None, [VP, VP, Color, C.c_float, C.c_float, c_u32_p, C.c_uint32]),
    "mop_hip_scene_file_load": (VP, [C.c_char_p]),
    "mop_hip_scene_file_free": (None, [VP]),
    "mop_hip_scene_file_mesh_count": (C.c_uint32, [VP]),
    "mop_hip_scene_file_light_count": (C.c_uint32, [VP]),
    "mop_hip_scene_file_get_camera": (C.c_bool, [VP, C.POINTER(Vec3), C.POINTER(Vec3),
                                                 C.POINTER(Vec3), c_float_p, c_float_p,
                                                 c_float_p]),
    "mop_hip_scene_file_get_light": (C.c_bool, [VP, C.c_uint32, C.POINTER(Light)]),
    "mop_hip_scene_file_get_mesh": (C.c_bool, [VP, VP, C.c_uint32, C.POINTER(LoadedMesh)]),
    "mop_hip_loaded_mesh_free": (None, [VP, C.POINTER(LoadedMesh)]),
    "mop_hip_scene_add_loaded_mesh": (C.c_int, [VP, C.POINTER(LoadedMesh), C.c_uint32, VP,
                                                C.c_int, C.c_float, C.c_float]),
    "mop_hip_shadow_render_noalloc": (C.c_int, [VP, VP, VP, C.c_uint32, VP, C.c_uint32,
                                                C.POINTER(Mat4), C.POINTER(Mat4), C.c_int]),
}


def declare(lib):
    for name, (res, args) in SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
