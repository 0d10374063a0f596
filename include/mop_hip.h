/*
 * mop_hip.h -- C ABI of the MI355X (gfx950) HIP backend for Master of
 * Puppets' software-rasterizer frame path.
 *
 * The backend plugs in behind MOP's RHI function table.  Every mop_hip_*
 * entry point below replaces one slot of `MopRhiBackend` (src/rhi/rhi.h:132-339)
 * or one of the software calls the viewport core makes directly on the CPU
 * backend's framebuffer (src/core/viewport.c:2433,2699,2729,2953-2959).
 * mop_rhi_backend_hip() returns the filled table; a maintainer adds the
 * MOP_BACKEND_HIP enum value and a case in mop_rhi_get_backend (rhi.c:50-71).
 *
 * Plain C types only: pointers, sizes and POD structs whose layout is
 * identical to the reference's (see the Mirror types section; a test compiles
 * both and compares sizeof/offsetof field by field).
 *
 * Error behaviour follows the RHI: creates return NULL on failure, draw and
 * frame calls are void and log on failure (rhi.h:115-129, cpu_backend.c:584).
 * The library never falls back to a CPU path: without a usable gfx950 device,
 * mop_hip_device_create returns NULL and logs why.
 */
#ifndef MOP_HIP_H
#define MOP_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Mirror types: same field order, types and therefore layout as the
 * reference value types, so a pointer to either may be passed.
 * ------------------------------------------------------------------------ */

typedef struct MopHipVec3 { float x, y, z; } MopHipVec3;        /* types.h:23-25 */
typedef struct MopHipVec4 { float x, y, z, w; } MopHipVec4;     /* types.h:27-29 */
typedef struct MopHipMat4 { float d[16]; } MopHipMat4;          /* types.h:39-41, column-major */
typedef struct MopHipColor { float r, g, b, a; } MopHipColor;   /* types.h:47-49 */

/* 48-byte vertex: types.h:58-63 */
typedef struct MopHipVertex {
  MopHipVec3 position;
  MopHipVec3 normal;
  MopHipColor color;
  float u, v;
} MopHipVertex;

/* light.h:24-28 / 37-62 (72 bytes) */
enum { MOP_HIP_LIGHT_DIRECTIONAL = 0, MOP_HIP_LIGHT_POINT = 1, MOP_HIP_LIGHT_SPOT = 2 };
typedef struct MopHipLight {
  int32_t type;
  MopHipVec3 position;
  MopHipVec3 direction;
  MopHipColor color;
  float intensity;
  float range;
  float spot_inner_cos;
  float spot_outer_cos;
  bool active;
  bool cast_shadows;
  void *viewport;
} MopHipLight;

/* vertex_format.h:25-70 */
enum {
  MOP_HIP_ATTRIB_POSITION = 0, MOP_HIP_ATTRIB_NORMAL = 1, MOP_HIP_ATTRIB_COLOR = 2,
  MOP_HIP_ATTRIB_TEXCOORD0 = 3, MOP_HIP_ATTRIB_JOINTS = 6, MOP_HIP_ATTRIB_WEIGHTS = 7
};
enum {
  MOP_HIP_FORMAT_FLOAT = 0, MOP_HIP_FORMAT_FLOAT2 = 1, MOP_HIP_FORMAT_FLOAT3 = 2,
  MOP_HIP_FORMAT_FLOAT4 = 3, MOP_HIP_FORMAT_UBYTE4 = 4
};
typedef struct MopHipVertexAttrib {
  int32_t semantic;
  int32_t format;
  uint32_t offset;
} MopHipVertexAttrib;
typedef struct MopHipVertexFormat {
  MopHipVertexAttrib attribs[12];
  uint32_t attrib_count;
  uint32_t stride;
} MopHipVertexFormat;

/* postprocess.h:20-39 */
enum {
  MOP_HIP_POST_GAMMA = 1 << 0, MOP_HIP_POST_TONEMAP = 1 << 1,
  MOP_HIP_POST_VIGNETTE = 1 << 2, MOP_HIP_POST_FOG = 1 << 3,
  MOP_HIP_POST_FXAA = 1 << 4
};
typedef struct MopHipFogParams {
  MopHipColor color;
  float near_dist;
  float far_dist;
} MopHipFogParams;

/* types.h:77-96 */
enum { MOP_HIP_BLEND_OPAQUE = 0, MOP_HIP_BLEND_ALPHA = 1, MOP_HIP_BLEND_ADDITIVE = 2,
       MOP_HIP_BLEND_MULTIPLY = 3 };
enum { MOP_HIP_SHADING_FLAT = 0, MOP_HIP_SHADING_SMOOTH = 1 };

typedef struct MopHipDevice MopHipDevice;
typedef struct MopHipBuffer MopHipBuffer;
typedef struct MopHipFramebuffer MopHipFramebuffer;
typedef struct MopHipTexture MopHipTexture;
typedef struct MopHipShader MopHipShader;

typedef struct MopHipBufferDesc { const void *data; size_t size; } MopHipBufferDesc; /* rhi.h:33-36 */
typedef struct MopHipFramebufferDesc { int width, height; } MopHipFramebufferDesc;  /* rhi.h:42-45 */

/* rhi.h:51-110 -- everything one draw needs */
typedef struct MopHipDrawCall {
  MopHipBuffer *vertex_buffer;
  MopHipBuffer *index_buffer;
  uint32_t vertex_count;
  uint32_t index_count;
  uint32_t object_id;
  MopHipMat4 model;
  MopHipMat4 view;
  MopHipMat4 projection;
  MopHipMat4 mvp;
  MopHipColor base_color;
  float opacity;
  MopHipVec3 light_dir;
  float ambient;
  int32_t shading_mode;
  bool wireframe;
  bool depth_test;
  bool depth_write;
  bool backface_cull;
  MopHipTexture *texture;
  MopHipTexture *normal_map;
  MopHipTexture *metallic_roughness_map;
  MopHipTexture *ao_map;
  int32_t blend_mode;
  float metallic;
  float roughness;
  MopHipVec3 emissive;
  const MopHipLight *lights;
  uint32_t light_count;
  MopHipVec3 cam_eye;
  const MopHipVertexFormat *vertex_format;
  float line_width;
  float depth_bias;
  MopHipVec3 aabb_min;
  MopHipVec3 aabb_max;
  bool cast_shadows;
} MopHipDrawCall;

/* ------------------------------------------------------------------------
 * RHI slots (rhi.h:132-339).  Each comment names the slot it fills.
 * ------------------------------------------------------------------------ */

MopHipDevice *mop_hip_device_create(void);                                  /* device_create  rhi.h:135 */
void mop_hip_device_destroy(MopHipDevice *dev);                             /* device_destroy rhi.h:136 */

MopHipBuffer *mop_hip_buffer_create(MopHipDevice *dev, const MopHipBufferDesc *desc); /* rhi.h:139-140; copies desc->data */
void mop_hip_buffer_destroy(MopHipDevice *dev, MopHipBuffer *buf);          /* rhi.h:141 */
void mop_hip_buffer_update(MopHipDevice *dev, MopHipBuffer *buf, const void *data,
                           size_t offset, size_t size);                     /* buffer_update rhi.h:224-225 */
const void *mop_hip_buffer_read(MopHipBuffer *buf);                         /* buffer_read rhi.h:229; host copy */

MopHipFramebuffer *mop_hip_framebuffer_create(MopHipDevice *dev,
                                              const MopHipFramebufferDesc *desc); /* rhi.h:144-145 */
void mop_hip_framebuffer_destroy(MopHipDevice *dev, MopHipFramebuffer *fb); /* rhi.h:146 */
void mop_hip_framebuffer_resize(MopHipDevice *dev, MopHipFramebuffer *fb, int width,
                                int height);                                /* rhi.h:147-148 */

void mop_hip_frame_begin(MopHipDevice *dev, MopHipFramebuffer *fb, MopHipColor clear); /* rhi.h:160-161 */
void mop_hip_frame_end(MopHipDevice *dev, MopHipFramebuffer *fb);          /* rhi.h:162 */
void mop_hip_frame_submit(MopHipDevice *dev, MopHipFramebuffer *fb);       /* rhi.h:163 */
void mop_hip_frame_wait_readback(MopHipDevice *dev, MopHipFramebuffer *fb); /* rhi.h:171 */
void mop_hip_draw(MopHipDevice *dev, MopHipFramebuffer *fb, const MopHipDrawCall *call); /* rhi.h:173-174 */
void mop_hip_draw_instanced(MopHipDevice *dev, MopHipFramebuffer *fb,
                            const MopHipDrawCall *call, const MopHipMat4 *instance_transforms,
                            uint32_t instance_count);                      /* rhi.h:218-221 */

uint32_t mop_hip_pick_read_id(MopHipDevice *dev, MopHipFramebuffer *fb, int x, int y);  /* rhi.h:177-178 */
float mop_hip_pick_read_depth(MopHipDevice *dev, MopHipFramebuffer *fb, int x, int y);  /* rhi.h:179-180 */
const uint8_t *mop_hip_framebuffer_read_color(MopHipDevice *dev, MopHipFramebuffer *fb,
                                              int *out_w, int *out_h);     /* rhi.h:183-185 */
const uint32_t *mop_hip_framebuffer_read_object_id(MopHipDevice *dev, MopHipFramebuffer *fb,
                                                   int *out_w, int *out_h); /* rhi.h:188-191 */
const float *mop_hip_framebuffer_read_depth(MopHipDevice *dev, MopHipFramebuffer *fb,
                                            int *out_w, int *out_h);       /* rhi.h:194-196 */

MopHipTexture *mop_hip_texture_create(MopHipDevice *dev, int w, int h,
                                      const uint8_t *rgba);                /* rhi.h:199-200 */
MopHipTexture *mop_hip_texture_create_hdr(MopHipDevice *dev, int w, int h,
                                          const float *rgba);              /* rhi.h:201-203 */
void mop_hip_texture_destroy(MopHipDevice *dev, MopHipTexture *tex);       /* rhi.h:215 */
bool mop_hip_texture_read_rgba8(MopHipDevice *dev, MopHipTexture *tex, uint8_t *out,
                                size_t size);                              /* rhi.h:319-320 */
float mop_hip_frame_gpu_time_ms(MopHipDevice *dev);                        /* rhi.h:233 */
/* draw_skybox rhi.h:291-293: the environment map as the frame's background,
 * from the inverse view-projection and camera position the core passes
 * (viewport.c:2209-2219); recorded to land before the frame's draws. */
void mop_hip_draw_skybox(MopHipDevice *dev, MopHipFramebuffer *fb, MopHipTexture *env_map,
                         const float *inv_vp, const float *cam_pos, float rotation,
                         float intensity);
void mop_hip_set_exposure(MopHipDevice *dev, float exposure);              /* rhi.h:238 */

/* The filled MopRhiBackend-layout table (name "hip"). */
const void *mop_rhi_backend_hip(void);

/* ------------------------------------------------------------------------
 * Hooks for the viewport core's CPU-gated software calls.  Same arguments as
 * the software function each replaces, plus the device.
 * ------------------------------------------------------------------------ */

/* mop_sw_shadow_render_mesh (rasterizer.c:357-493; called viewport.c:2433).
 * `vertices`/`indices` are the host pointers buffer_read returned (looked up
 * to the resident device buffers) or any host arrays (uploaded).  `size` is
 * the square shadow-map size; `first` != 0 starts a new shadow map cleared to
 * 1.0 (pass_shadow clears before its first mesh, viewport.c:2365-2368). */
void mop_hip_shadow_render_mesh(MopHipDevice *dev, const void *vertices, uint32_t vertex_count,
                                const uint32_t *indices, uint32_t index_count,
                                const MopHipMat4 *model, const MopHipMat4 *light_vp, int size,
                                int first);
/* mop_sw_shadow_set (rasterizer.c:141-146): shading now tests this map. */
void mop_hip_shadow_set(MopHipDevice *dev, int w, int h, const MopHipMat4 *light_vp);
/* mop_sw_shadow_clear (rasterizer.c:148-152). */
void mop_hip_shadow_clear(MopHipDevice *dev);
/* mop_sw_ibl_set (rasterizer.c:167-181; rasterizer.h:212-215): tables are
 * copied to HBM (re-uploaded only when their contents change). */
void mop_hip_ibl_set(MopHipDevice *dev, const float *irradiance, int irr_w, int irr_h,
                     const float *prefiltered, int pf_w, int pf_h, int pf_levels,
                     const float *brdf_lut, int brdf_size, float rotation, float intensity);
/* mop_sw_ibl_clear (rasterizer.c:183). */
void mop_hip_ibl_clear(MopHipDevice *dev);
/* pass_background_hdri_cpu (viewport.c:2128-2180): the core's CPU skybox,
 * with its inputs (vp->env_texture, vp->cam_fov_radians, cam_eye,
 * cam_target, cam_up, env_rotation, env_intensity); the background equals
 * the CPU backend's bit for bit.  A core integrating this backend calls it
 * where it now calls the CPU skybox for the CPU backend. */
void mop_hip_draw_background_hdri(MopHipDevice *dev, MopHipFramebuffer *fb, MopHipTexture *env,
                                  float fov_radians, MopHipVec3 eye, MopHipVec3 target,
                                  MopHipVec3 up, float rotation, float intensity);
/* mop_overlay_builtin_outline (overlay_builtin.c:1309-1400; called
 * viewport.c:2738-2739 and :2906-2907): the object outline painted on the
 * resolved RGBA8 image before the post effects, with the viewport's theme
 * accent and outline opacities and its selection list.  Applied on the
 * device by the frame's tail, so the image read back already carries it;
 * the RGBA8 result equals the overlay's in-place host loop bit for bit.  A
 * core integrating this backend calls it instead of the host overlay. */
void mop_hip_draw_outline(MopHipDevice *dev, MopHipFramebuffer *fb, MopHipColor accent,
                          float alpha_selected, float alpha_unselected,
                          const uint32_t *selected_ids, uint32_t selected_count);
/* mop_sw_hdr_resolve (rasterizer.c:2776-2803). */
void mop_hip_hdr_resolve(MopHipDevice *dev, MopHipFramebuffer *fb, float exposure);
/* mop_postprocess_apply (postprocess.c:24-116): FXAA then fog/tonemap/gamma/vignette. */
void mop_hip_postprocess_apply(MopHipDevice *dev, MopHipFramebuffer *fb, uint32_t effects,
                               const MopHipFogParams *fog);
/* The RGBA32F HDR attachment (MopSwFramebuffer.color_hdr, rasterizer.h:33),
 * read back to host memory owned by the framebuffer. */
const float *mop_hip_framebuffer_read_hdr(MopHipDevice *dev, MopHipFramebuffer *fb, int *out_w,
                                          int *out_h);
/* mop_viewport_read_color's SSAA box downsample (viewport.c:3385-3411) run on
 * the device: the internal image averaged over factor x factor blocks
 * (truncating integer mean per channel), returned at (w / factor) x
 * (h / factor) in host memory owned by the framebuffer.  A core integrating
 * this backend calls it where it now reads the internal image and filters on
 * the host. */
const uint8_t *mop_hip_framebuffer_read_color_downsampled(MopHipDevice *dev,
                                                          MopHipFramebuffer *fb, int factor,
                                                          int *out_w, int *out_h);
/* Asynchronous readback (no counterpart in the reference, whose buffers are
 * host memory; it serves the deferred readback that frame_submit /
 * frame_wait_readback imply, rhi.h:163-171).  Starts copying the current
 * frame's buffers -- mask: 1 RGBA8, 2 depth, 4 object ID, 8 HDR -- to the
 * framebuffer's host buffers on a copy engine once the frame finishes, so
 * the copy overlaps the next frames' kernels; the read_* calls above then
 * wait for it and return the same pointers.  A frame re-run after a scratch
 * overflow, or new contents, makes the read_* calls copy again. */
void mop_hip_framebuffer_readback_async(MopHipDevice *dev, MopHipFramebuffer *fb, uint32_t mask);

/* ------------------------------------------------------------------------
 * Batched execution (no counterpart in the reference, whose frames are
 * synchronous per viewport).  Record frames on many framebuffers with
 * frame_begin/draw (+ mop_hip_frame_request_post), then execute them all in
 * one set of launches; readbacks and mop_hip_frame_wait_readback work per
 * framebuffer as usual.
 * ------------------------------------------------------------------------ */

/* Record the HDR resolve (exposure) and post-process (effects, fog: may be
 * NULL) of the current frame of fb without executing it. */
void mop_hip_frame_request_post(MopHipDevice *dev, MopHipFramebuffer *fb, float exposure,
                                uint32_t effects, const MopHipFogParams *fog);
/* Execute every recorded, not yet executed frame of fbs[0..n) as one batch. */
void mop_hip_frames_execute(MopHipDevice *dev, MopHipFramebuffer *const *fbs, size_t n);
/* ---- skinning and morph targets (device side of a dispatch hook) ----
 * mop_skin_apply (viewport.c:1572-1680) and mop_morph_apply
 * (viewport.c:1763-1805) compute the deformed vertices on the host and upload
 * them with buffer_update.  These entry points take the same inputs -- the
 * mesh's bind-pose snapshot, its vertex format and the bone matrices / morph
 * targets -- and write the deformed vertices straight into the draw's vertex
 * buffer on the device, with the reference's arithmetic; buffer_read returns
 * the deformed bytes afterwards.  The bind pose is copied to the device on
 * the first call for a buffer (and again if the pointer changes).
 * skin: fmt must hold POSITION (float3), JOINTS (ubyte4), WEIGHTS (float4);
 * NORMAL is skinned and renormalized when present.  morph: fmt NULL means
 * MopVertex; targets are target_count x vertex_count x 3 floats.
 * Both return 0, or -1 with mop_hip_last_error() set. */
int mop_hip_skin_apply(MopHipDevice *dev, MopHipBuffer *vb, const void *bind_pose,
                       uint32_t vertex_count, const MopHipVertexFormat *fmt,
                       const MopHipMat4 *bones, uint32_t bone_count);
int mop_hip_morph_apply(MopHipDevice *dev, MopHipBuffer *vb, const void *bind_pose,
                        uint32_t vertex_count, const MopHipVertexFormat *fmt,
                        const float *targets, const float *weights, uint32_t target_count);

/* Frame plans replayed from captured hipGraphs.  mode 1: always, 0: never,
 * 2 (default): for executions of two or more frames (mop_hip_frames_execute
 * batches); a lone frame is device-bound with few launches and runs faster
 * launched directly.  Each execution is planned on the host; a plan
 * identical in every launch and buffer address to the previous one is
 * captured once (two instances with their own pinned upload buffers; one
 * graph per run of same-stream launches, the shadow passes' side stream
 * forked and joined between them) and replayed afterwards, with only the
 * uploaded bytes (draws, lights, views) refreshed.  Stage timers force
 * direct launches.  graph_stats reports replays and captures so far. */
void mop_hip_set_graphs(MopHipDevice *dev, int mode);
void mop_hip_graph_stats(MopHipDevice *dev, uint64_t *replays, uint64_t *captures);

/* Launch-shape tunables (results never depend on them): bbox area up to
 * which one lane walks a screen triangle (default 64), shading tile edge
 * 8/16/32 (default 16), waves working through the large-triangle queue
 * (0 = 8192; the small-triangle pass of batches over 2^27 triangles takes
 * it as its workgroup count).
 * Returns 0, or -1 for an invalid value. */
int mop_hip_set_tuning(MopHipDevice *dev, uint32_t small_area, uint32_t shade_tile,
                       uint32_t raster_grid);
/* Waves per SIMD the shading pass runs at: 4, 5 (default), 6 or 8 (the
 * kernel is compiled for each; fewer waves leave more registers per lane).
 * Results never depend on it.  Returns 0, or -1 for another value. */
int mop_hip_set_shade_waves(MopHipDevice *dev, uint32_t waves);
/* Raster path of the opaque, depth-tested and depth-written draws recorded
 * after the call: 0 (default) the per-pixel (depth, submission index)
 * minimum of the visibility buffer, 1 the tile path the blended draws take
 * (prefix-scan binning to 16x16 tiles in submission order, colour/depth/id
 * tiles in LDS, ballot coverage).  Results are identical; speed differs.
 * No counterpart in the reference (the CPU backend has one raster).
 * Returns 0, or -1 for another value. */
int mop_hip_set_opaque_raster(MopHipDevice *dev, uint32_t tiled);
/* 1 when fb's frame has finished on the device (non-blocking). */
int mop_hip_frame_done(MopHipDevice *dev, MopHipFramebuffer *fb);
/* 1 when the copy mop_hip_framebuffer_readback_async started for fb's
 * current frame has reached host memory, so the read_* calls of what it
 * copied return without waiting (non-blocking); mop_hip_frame_done when no
 * such copy is pending. */
int mop_hip_readback_done(MopHipDevice *dev, MopHipFramebuffer *fb);
/* Wait until everything queued on the device has finished. */
void mop_hip_device_sync(MopHipDevice *dev);

/* Scratch overflow (no counterpart in the reference, whose scratch grows on
 * the host).  Every batch ORs its overflow bits (1 clip queue, 2
 * screen-triangle store, 8 clipped polygon over 12 vertices, 16 tile bins,
 * 32 big-triangle checkpoints) into a per-device word; this returns it after
 * waiting for the queued work, and clears it when reset != 0.  A frame that
 * overflowed is run again with grown stores when it is read back; a caller
 * that never reads a frame back uses this to know that every executed frame
 * was complete. */
uint32_t mop_hip_overflow_flags(MopHipDevice *dev, int reset);
/* Forced capacities of the clip queue (triangles), the checkpoint store and
 * the tile bins (entries) for the next executions (0 = none); dropped for a
 * store once it overflows.  They exercise the overflow recovery. */
int mop_hip_set_scratch_limits(MopHipDevice *dev, uint32_t clip, uint32_t chk, uint32_t bins);
/* The same for every device created afterwards in this process. */
void mop_hip_set_default_scratch_limits(uint32_t clip, uint32_t chk, uint32_t bins);
/* Frames run again after a scratch overflow in this process. */
uint64_t mop_hip_overflow_reruns(void);

/* ------------------------------------------------------------------------
 * Device-resident shadow maps and IBL tables (no counterpart in the
 * reference, whose shadow/IBL state is process-global: rasterizer.c:141-183).
 * They let many scenes keep their lighting state in HBM and let one device
 * render frames of different scenes back to back (batching, serving).
 * ------------------------------------------------------------------------ */

typedef struct MopHipShadowMap MopHipShadowMap;
typedef struct MopHipIbl MopHipIbl;

MopHipShadowMap *mop_hip_shadow_map_create(MopHipDevice *dev, int size);
void mop_hip_shadow_map_destroy(MopHipDevice *dev, MopHipShadowMap *map);
/* Start a caster list for `map`; mop_hip_shadow_draw adds resident meshes
 * (model matrix, light view-projection as pass_shadow computes them,
 * viewport.c:2312-2415); mop_hip_shadow_end renders the map (cleared to 1.0)
 * and makes later draws test it, like mop_sw_shadow_set. */
void mop_hip_shadow_begin(MopHipDevice *dev, MopHipShadowMap *map);
void mop_hip_shadow_draw(MopHipDevice *dev, MopHipBuffer *vb, uint32_t vertex_count,
                         MopHipBuffer *ib, uint32_t index_count, const MopHipMat4 *model,
                         const MopHipMat4 *light_vp);
void mop_hip_shadow_end(MopHipDevice *dev, const MopHipMat4 *light_vp);
/* Make later draws test an already rendered map. */
void mop_hip_shadow_bind(MopHipDevice *dev, const MopHipShadowMap *map,
                         const MopHipMat4 *light_vp);
/* The map's depths (size * size floats) on the host, after the queued work. */
int mop_hip_shadow_map_read(MopHipDevice *dev, const MopHipShadowMap *map, float *out);

/* Allocation-free shadow pass.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Renders one caster
 * mesh (resident buffers vb/ib, model and light view-projection as
 * mop_hip_shadow_draw takes them) into map on the device's stream, clearing
 * the map to depth 1.0 first when clear != 0; call it once per caster, with
 * clear on the first only.  The result equals mop_hip_shadow_begin / draw /
 * end with the same casters.  It allocates nothing: its working memory is a
 * fixed 20 MiB per GPU, held until the process ends, for at most 524,288
 * vertices and 1,048,576 triangles per call (larger meshes are refused:
 * -1 and mop_hip_last_error()).  Every call on one GPU shares that memory,
 * so calls that overlap on one GPU -- from two devices or two streams --
 * race.  Returns 0 when the work is queued. */
int mop_hip_shadow_render_noalloc(MopHipDevice *dev, MopHipShadowMap *map, MopHipBuffer *vb,
                                  uint32_t vertex_count, MopHipBuffer *ib, uint32_t index_count,
                                  const MopHipMat4 *model, const MopHipMat4 *light_vp, int clear);

/* IBL tables computed on the device from an equirect RGBA32F environment
 * (environment.c:115-484: irradiance 64x32, prefiltered 128x64 x 5 levels,
 * BRDF LUT 256^2). */
MopHipIbl *mop_hip_ibl_create(MopHipDevice *dev, const float *env_rgba, int w, int h);
void mop_hip_ibl_destroy(MopHipDevice *dev, MopHipIbl *ibl);
/* Like mop_hip_ibl_set with resident tables. */
void mop_hip_ibl_bind(MopHipDevice *dev, const MopHipIbl *ibl, float rotation, float intensity);
/* Copy the tables to host arrays (64*32*4, 5-level pyramid, 256*256*4 floats). */
int mop_hip_ibl_download(MopHipDevice *dev, const MopHipIbl *ibl, float *irr, float *pf,
                         float *lut);

/* ------------------------------------------------------------------------
 * Diagnostics (no counterpart in the reference).
 * ------------------------------------------------------------------------ */

/* Last error message of this thread ("" when none). */
const char *mop_hip_last_error(void);

/* Per-frame statistics of the last executed frame. */
typedef struct MopHipFrameStats {
  uint32_t triangles_in;     /* triangles submitted (viewport.c:2032 counting) */
  uint32_t subtris;          /* screen triangles after cull/clip/fan */
  uint32_t clipped;          /* triangles that went through the clipper */
  uint32_t big_subtris;      /* screen triangles walked row-parallel */
  uint32_t shadow_tris;      /* shadow-caster triangles */
  float gpu_ms;              /* device time of the frame (hipEvents) */
  uint32_t culled_chunks;    /* chunks of 256 triangles of large draws skipped
                                whole: every triangle trivially rejected */
} MopHipFrameStats;
void mop_hip_get_frame_stats(MopHipDevice *dev, MopHipFrameStats *out);

/* Device-side stage timers (hipEvents on the backend's stream around each
 * stage: vertex, setup, raster, shade, resolve, post, shadow).  Enabling
 * resets the sums; reading synchronises the stream. */
void mop_hip_timers_enable(MopHipDevice *dev, int on);
int mop_hip_timers_read(MopHipDevice *dev, const char *stage, double *total_ms,
                        uint64_t *count);

/* The rate, in G atomics/s, at which the current GPU performs 64-bit atomic
 * mins with no return value -- the order-free raster's visibility update --
 * when every lane of a wave targets its own random address (pattern 0) or
 * the 64 lanes target 512 consecutive bytes (pattern 1).  Times 33.5 M
 * atomics on a 512 MiB buffer it allocates and frees; negative on error. */
double mop_hip_probe_atomic_rate(int pattern);

#ifdef __cplusplus
}
#endif

#endif /* MOP_HIP_H */
