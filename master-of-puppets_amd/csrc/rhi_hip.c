/*
 * rhi_hip.c -- C11 adapter: the RHI function table of the HIP backend.
 *
 * This is the file a maintainer drops next to src/backend/cpu/cpu_backend.c;
 * it is the counterpart of the CPU table (cpu_backend.c:1296-1343).  Slots
 * the CPU backend leaves NULL for GPU-only effects stay NULL here too (every
 * caller NULL-checks them, rhi.h:115-129).
 */
#include "mop_hip_rhi.h"

static const MopHipRhiBackend s_hip_backend = {
    .name = "hip",
    .device_create = mop_hip_device_create,
    .device_destroy = mop_hip_device_destroy,
    .buffer_create = mop_hip_buffer_create,
    .buffer_destroy = mop_hip_buffer_destroy,
    .framebuffer_create = mop_hip_framebuffer_create,
    .framebuffer_destroy = mop_hip_framebuffer_destroy,
    .framebuffer_resize = mop_hip_framebuffer_resize,
    .framebuffer_create_from_texture = 0,
    .frame_begin = mop_hip_frame_begin,
    .frame_end = mop_hip_frame_end,
    .frame_submit = mop_hip_frame_submit,
    .frame_wait_readback = mop_hip_frame_wait_readback,
    .draw = mop_hip_draw,
    .pick_read_id = mop_hip_pick_read_id,
    .pick_read_depth = mop_hip_pick_read_depth,
    .framebuffer_read_color = mop_hip_framebuffer_read_color,
    .framebuffer_read_object_id = mop_hip_framebuffer_read_object_id,
    .framebuffer_read_depth = mop_hip_framebuffer_read_depth,
    .texture_create = mop_hip_texture_create,
    .texture_create_hdr = mop_hip_texture_create_hdr,
    .texture_create_ex = 0,
    .texture_destroy = mop_hip_texture_destroy,
    .draw_instanced = mop_hip_draw_instanced,
    .buffer_update = mop_hip_buffer_update,
    .buffer_read = mop_hip_buffer_read,
    .frame_gpu_time_ms = mop_hip_frame_gpu_time_ms,
    .set_exposure = mop_hip_set_exposure,
    .texture_read_rgba8 = mop_hip_texture_read_rgba8,
    .draw_skybox = mop_hip_draw_skybox,
};

const void *mop_rhi_backend_hip(void) { return &s_hip_backend; }
