/* Compile-time proof that the HIP backend's mirror types and RHI table have
 * the reference's layout (built by tests/test_abi.py against the reference
 * headers; nothing is executed). */
#include <stddef.h>

#include "rhi/rhi.h"
#include <mop/render/postprocess.h>
#include "mop_hip_rhi.h"

#define SAME_SIZE(a, b) _Static_assert(sizeof(a) == sizeof(b), "sizeof " #a)
#define SAME_OFF(a, b, f) _Static_assert(offsetof(a, f) == offsetof(b, f), #a "." #f)

SAME_SIZE(MopVec3, MopHipVec3);
SAME_SIZE(MopVec4, MopHipVec4);
SAME_SIZE(MopMat4, MopHipMat4);
SAME_SIZE(MopColor, MopHipColor);
SAME_SIZE(MopVertex, MopHipVertex);
SAME_OFF(MopVertex, MopHipVertex, position);
SAME_OFF(MopVertex, MopHipVertex, normal);
SAME_OFF(MopVertex, MopHipVertex, color);
SAME_OFF(MopVertex, MopHipVertex, u);
SAME_OFF(MopVertex, MopHipVertex, v);

SAME_SIZE(MopLight, MopHipLight);
SAME_OFF(MopLight, MopHipLight, type);
SAME_OFF(MopLight, MopHipLight, position);
SAME_OFF(MopLight, MopHipLight, direction);
SAME_OFF(MopLight, MopHipLight, color);
SAME_OFF(MopLight, MopHipLight, intensity);
SAME_OFF(MopLight, MopHipLight, range);
SAME_OFF(MopLight, MopHipLight, spot_inner_cos);
SAME_OFF(MopLight, MopHipLight, spot_outer_cos);
SAME_OFF(MopLight, MopHipLight, active);
SAME_OFF(MopLight, MopHipLight, cast_shadows);
SAME_OFF(MopLight, MopHipLight, viewport);

SAME_SIZE(MopVertexAttrib, MopHipVertexAttrib);
SAME_SIZE(MopVertexFormat, MopHipVertexFormat);
SAME_OFF(MopVertexFormat, MopHipVertexFormat, attrib_count);
SAME_OFF(MopVertexFormat, MopHipVertexFormat, stride);
SAME_SIZE(MopFogParams, MopHipFogParams);
SAME_SIZE(MopRhiBufferDesc, MopHipBufferDesc);
SAME_SIZE(MopRhiFramebufferDesc, MopHipFramebufferDesc);

SAME_SIZE(MopRhiDrawCall, MopHipDrawCall);
#define DC(f) SAME_OFF(MopRhiDrawCall, MopHipDrawCall, f)
DC(vertex_buffer); DC(index_buffer); DC(vertex_count); DC(index_count); DC(object_id);
DC(model); DC(view); DC(projection); DC(mvp); DC(base_color); DC(opacity); DC(light_dir);
DC(ambient); DC(shading_mode); DC(wireframe); DC(depth_test); DC(depth_write);
DC(backface_cull); DC(texture); DC(normal_map); DC(metallic_roughness_map); DC(ao_map);
DC(blend_mode); DC(metallic); DC(roughness); DC(emissive); DC(lights); DC(light_count);
DC(cam_eye); DC(vertex_format); DC(line_width); DC(depth_bias); DC(aabb_min); DC(aabb_max);
DC(cast_shadows);

SAME_SIZE(MopRhiBackend, MopHipRhiBackend);
#define BK(f) SAME_OFF(MopRhiBackend, MopHipRhiBackend, f)
BK(name); BK(device_create); BK(device_destroy); BK(buffer_create); BK(buffer_destroy);
BK(framebuffer_create); BK(framebuffer_destroy); BK(framebuffer_resize);
BK(framebuffer_create_from_texture); BK(frame_begin); BK(frame_end); BK(frame_submit);
BK(frame_wait_readback); BK(draw); BK(pick_read_id); BK(pick_read_depth);
BK(framebuffer_read_color); BK(framebuffer_read_object_id); BK(framebuffer_read_depth);
BK(texture_create); BK(texture_create_hdr); BK(texture_create_ex); BK(texture_destroy);
BK(draw_instanced); BK(buffer_update); BK(buffer_read); BK(frame_gpu_time_ms);
BK(set_exposure); BK(set_bloom); BK(set_ssao); BK(set_ssr); BK(set_oit); BK(add_decal);
BK(remove_decal); BK(clear_decals); BK(set_volumetric); BK(set_taa_params);
BK(set_ibl_textures); BK(draw_skybox); BK(draw_overlays); BK(shader_create);
BK(shader_destroy); BK(framebuffer_copy_to_texture); BK(texture_read_rgba8);
BK(set_gpu_driven_rendering);

/* enum values the mirrors hard-code */
_Static_assert(MOP_LIGHT_SPOT == MOP_HIP_LIGHT_SPOT, "light enum");
_Static_assert(MOP_BLEND_MULTIPLY == MOP_HIP_BLEND_MULTIPLY, "blend enum");
_Static_assert(MOP_SHADING_SMOOTH == MOP_HIP_SHADING_SMOOTH, "shading enum");
_Static_assert(MOP_POST_FXAA == MOP_HIP_POST_FXAA, "post enum");
_Static_assert(MOP_FORMAT_UBYTE4 == MOP_HIP_FORMAT_UBYTE4, "format enum");
_Static_assert(MOP_ATTRIB_TEXCOORD0 == MOP_HIP_ATTRIB_TEXCOORD0, "attrib enum");

int mop_hip_abi_check;
