/*
 * mop_hip_rhi.h -- layout mirror of the reference RHI function table
 * (`MopRhiBackend`, src/rhi/rhi.h:132-339) filled by mop_rhi_backend_hip().
 *
 * Field order, types and therefore offsets equal the reference table, so a
 * maintainer returns `(const MopRhiBackend *)mop_rhi_backend_hip()` from
 * mop_rhi_get_backend (rhi.c:50-71).  The handle types are this library's
 * opaque structs; the reference treats its handles as opaque too.
 */
#ifndef MOP_HIP_RHI_H
#define MOP_HIP_RHI_H

#include "mop_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct MopHipRhiBackend {
  const char *name;
  MopHipDevice *(*device_create)(void);
  void (*device_destroy)(MopHipDevice *device);
  MopHipBuffer *(*buffer_create)(MopHipDevice *device, const MopHipBufferDesc *desc);
  void (*buffer_destroy)(MopHipDevice *device, MopHipBuffer *buffer);
  MopHipFramebuffer *(*framebuffer_create)(MopHipDevice *device,
                                           const MopHipFramebufferDesc *desc);
  void (*framebuffer_destroy)(MopHipDevice *device, MopHipFramebuffer *fb);
  void (*framebuffer_resize)(MopHipDevice *device, MopHipFramebuffer *fb, int width,
                             int height);
  MopHipFramebuffer *(*framebuffer_create_from_texture)(MopHipDevice *device,
                                                        MopHipTexture *color, int width,
                                                        int height);
  void (*frame_begin)(MopHipDevice *device, MopHipFramebuffer *fb, MopHipColor clear_color);
  void (*frame_end)(MopHipDevice *device, MopHipFramebuffer *fb);
  void (*frame_submit)(MopHipDevice *device, MopHipFramebuffer *fb);
  void (*frame_wait_readback)(MopHipDevice *device, MopHipFramebuffer *fb);
  void (*draw)(MopHipDevice *device, MopHipFramebuffer *fb, const MopHipDrawCall *call);
  uint32_t (*pick_read_id)(MopHipDevice *device, MopHipFramebuffer *fb, int x, int y);
  float (*pick_read_depth)(MopHipDevice *device, MopHipFramebuffer *fb, int x, int y);
  const uint8_t *(*framebuffer_read_color)(MopHipDevice *device, MopHipFramebuffer *fb,
                                           int *out_width, int *out_height);
  const uint32_t *(*framebuffer_read_object_id)(MopHipDevice *device, MopHipFramebuffer *fb,
                                                int *out_width, int *out_height);
  const float *(*framebuffer_read_depth)(MopHipDevice *device, MopHipFramebuffer *fb,
                                         int *out_width, int *out_height);
  MopHipTexture *(*texture_create)(MopHipDevice *device, int width, int height,
                                   const uint8_t *rgba_data);
  MopHipTexture *(*texture_create_hdr)(MopHipDevice *device, int width, int height,
                                       const float *rgba_float_data);
  MopHipTexture *(*texture_create_ex)(MopHipDevice *device, int width, int height, int format,
                                      int mip_levels, const uint8_t *data, size_t data_size);
  void (*texture_destroy)(MopHipDevice *device, MopHipTexture *texture);
  void (*draw_instanced)(MopHipDevice *device, MopHipFramebuffer *fb, const MopHipDrawCall *call,
                         const MopHipMat4 *instance_transforms, uint32_t instance_count);
  void (*buffer_update)(MopHipDevice *device, MopHipBuffer *buffer, const void *data,
                        size_t offset, size_t size);
  const void *(*buffer_read)(MopHipBuffer *buffer);
  float (*frame_gpu_time_ms)(MopHipDevice *dev);
  void (*set_exposure)(MopHipDevice *dev, float exposure);
  void (*set_bloom)(MopHipDevice *dev, bool enabled, float threshold, float intensity);
  void (*set_ssao)(MopHipDevice *dev, bool enabled);
  void (*set_ssr)(MopHipDevice *dev, bool enabled, float intensity);
  void (*set_oit)(MopHipDevice *dev, bool enabled);
  int32_t (*add_decal)(MopHipDevice *dev, const float *transform, float opacity,
                       int32_t texture_idx);
  void (*remove_decal)(MopHipDevice *dev, int32_t decal_id);
  void (*clear_decals)(MopHipDevice *dev);
  void (*set_volumetric)(MopHipDevice *dev, float density, float r, float g, float b,
                         float anisotropy, int steps);
  void (*set_taa_params)(MopHipDevice *dev, const float *inv_vp_jittered, const float *prev_vp,
                         float jitter_x, float jitter_y, bool first_frame);
  void (*set_ibl_textures)(MopHipDevice *dev, MopHipTexture *irradiance,
                           MopHipTexture *prefiltered, MopHipTexture *brdf_lut);
  void (*draw_skybox)(MopHipDevice *dev, MopHipFramebuffer *fb, MopHipTexture *env_map,
                      const float *inv_vp, const float *cam_pos, float rotation,
                      float intensity);
  void (*draw_overlays)(MopHipDevice *dev, MopHipFramebuffer *fb, const void *prims,
                        uint32_t prim_count, const void *grid_params, int fb_width,
                        int fb_height);
  MopHipShader *(*shader_create)(MopHipDevice *dev, const uint32_t *bytecode, size_t size);
  void (*shader_destroy)(MopHipDevice *dev, MopHipShader *shader);
  bool (*framebuffer_copy_to_texture)(MopHipDevice *dev, MopHipFramebuffer *fb,
                                      MopHipTexture *target);
  bool (*texture_read_rgba8)(MopHipDevice *dev, MopHipTexture *texture, uint8_t *out_buf,
                             size_t buf_size);
  void (*set_gpu_driven_rendering)(MopHipDevice *dev, bool enabled);
} MopHipRhiBackend;

#ifdef __cplusplus
}
#endif

#endif
