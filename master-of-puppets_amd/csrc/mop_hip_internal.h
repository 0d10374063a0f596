/*
 * mop_hip_internal.h -- device-side data layout shared by the gfx950 kernels
 * (kernels.hip) and the host runtime (runtime.hip).
 *
 * Frame pipeline (one stream, deferred to frame_end):
 *   k_vertex      per draw vertex: clip position (MVP), world position and
 *                 model-space-transformed normal, vertex colour (texture
 *                 modulated) -- cpu_prepare_triangle, cpu_backend.c:224-324
 *   k_classify    per triangle: trivial reject, clip-space backface cull,
 *                 trivial accept -> number of screen triangles, or queue for
 *                 the clipper -- rasterizer.c:2467-2504
 *   k_clip        Sutherland-Hodgman over the 6 planes, fan split, per-fan
 *                 rejects -- rasterizer.c:510-626, 2510-2553
 *   scan          exclusive prefix sum of per-triangle counts: screen
 *                 triangles get indices in submission order, used as the
 *                 depth tie-break sequence number
 *   k_emit        write screen-triangle records
 *   k_raster_*    half-space walk with the reference's incremental edge
 *                 recurrences; 64-bit atomic min of (depth key, sequence)
 *                 into the visibility buffer
 *   k_shade       per pixel: re-derive the winner's edge values, interpolate,
 *                 multi-light Lambert + GGX + PCF + IBL -- rasterizer.c:1926-2140
 *   k_resolve     ACES + exposure -- rasterizer.c:2776-2803
 *   k_post        FXAA then fog / tonemap / gamma / vignette -- postprocess.c
 */
#ifndef MOP_HIP_INTERNAL_H
#define MOP_HIP_INTERNAL_H

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "mop_hip.h"

/* ---- frame plans ---------------------------------------------------------
 * While a plan is active (t_plan set by the runtime), frame launches and
 * uploads are recorded instead of issued: the plan keeps a byte signature of
 * every kernel (function, grid, block, argument bytes) and upload (device
 * destination, size), and the operations themselves.  The runtime then
 * issues the operations directly, or replays a hipGraph captured from an
 * earlier plan with the same signature (upload sources live in a pinned
 * buffer owned by that graph, refilled before each replay). */
struct Plan {
  std::vector<uint8_t> sig;
  /* each op issues on the stream it was recorded for; arg: upload source */
  std::vector<std::function<void(const uint8_t *)>> ops;
  std::vector<hipStream_t> ostream; /* the stream of each op */
  std::vector<uint8_t> marker;      /* 1: a cross-stream event dependency */
  std::vector<uint8_t> blob;        /* host bytes of the plan's uploads */
};
extern thread_local Plan *t_plan;

/* When set (stage timers on, no plan active), the next direct launch is made
 * with hipExtLaunchKernelGGL and these events, which the runtime timestamps
 * at the dispatch's own start and end, then both are cleared. */
extern thread_local hipEvent_t t_kernel_ev[2];

template <typename T> inline void sig_put(std::vector<uint8_t> &s, const T &v) {
  const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
  s.insert(s.end(), p, p + sizeof(T));
}

/* hipLaunchKernelGGL, or its record into the active plan */
template <typename... P, typename... A>
inline void launch(void (*k)(P...), dim3 g, dim3 b, uint32_t shmem, hipStream_t s, A... a) {
  if (Plan *pl = t_plan) {
    sig_put(pl->sig, (uintptr_t)k);
    sig_put(pl->sig, (uintptr_t)s);
    sig_put(pl->sig, g);
    sig_put(pl->sig, b);
    (sig_put(pl->sig, (P)a), ...);
    pl->ops.push_back([=](const uint8_t *) {
      hipLaunchKernelGGL(k, g, b, shmem, s, (P)a...);
    });
    pl->ostream.push_back(s);
    pl->marker.push_back(0);
  } else if (t_kernel_ev[0]) {
    hipEvent_t e0 = t_kernel_ev[0], e1 = t_kernel_ev[1];
    t_kernel_ev[0] = t_kernel_ev[1] = nullptr;
    hipExtLaunchKernelGGL(k, g, b, shmem, s, e0, e1, 0u, (P)a...);
  } else {
    hipLaunchKernelGGL(k, g, b, shmem, s, (P)a...);
  }
}

/* Screen triangles whose clamped bbox has more pixels than this are walked
 * one row per lane instead of one triangle per lane. */
#define MOPH_SMALL_AREA 64
/* Stored vertices per clipped triangle (a triangle cut by 6 planes has at
 * most 9; more is flagged as overflow). */
#define MOPH_MAX_POLY 12

enum {
  MOPH_DRAW_CULL = 1u << 0,
  MOPH_DRAW_FLEX = 1u << 1,
  MOPH_DRAW_DEPTH_TEST = 1u << 2,
  MOPH_DRAW_DEPTH_WRITE = 1u << 3,
  MOPH_DRAW_BLEND_SHIFT = 4, /* 2 bits: MOP_HIP_BLEND_* */
  MOPH_DRAW_SM_SHIFT = 6,    /* 2 bits: MOPH_SM_* */
  MOPH_DRAW_WIRE = 1u << 8,  /* wireframe: three anti-aliased lines per triangle */
};

/* Shading model of a draw, as cpu_backend.c:608-619 and
 * rasterizer.c:2540-2610 / 1158-1243 dispatch it. */
enum {
  MOPH_SM_LIT_SMOOTH = 0,   /* lights, smooth: mop_sw_rasterize_triangle_smooth_ml */
  MOPH_SM_LIT_FLAT = 1,     /* lights, flat: compute_multi_light + filled triangle */
  MOPH_SM_UNLIT_FLAT = 2,   /* no lights, flat: light_dir Lambert + filled triangle */
  MOPH_SM_UNLIT_SMOOTH = 3, /* no lights, smooth: mop_sw_rasterize_triangle_smooth */
};

/* One draw (or one instance of an instanced draw) as the kernels see it. */
struct DrawGPU {
  float mvp[16];
  float model[16];
  const uint8_t *vbuf;
  const uint32_t *ibuf;
  const uint8_t *tex;
  int32_t tex_w, tex_h;
  uint32_t vertex_count;
  uint32_t tri_count; /* index_count / 3 triangles considered */
  uint32_t vtx_base;  /* first slot in the transformed-vertex cache */
  uint32_t tri_base;  /* first slot in the flattened triangle space */
  uint32_t stride;
  int32_t pos_off, pos_fmt, nrm_off, nrm_fmt, col_off, col_fmt, uv_off, uv_fmt;
  uint32_t object_id;
  uint32_t flags;
  float depth_bias;
  float ambient;
  float metallic;
  float roughness;
  float cam_eye[3];
  uint32_t light_off;
  uint32_t light_count;
  uint32_t view; /* index into the batch's views */
  float opacity;
  float light_dir[3]; /* the call's light_dir (unlit models) */
  uint32_t vtx_wg;    /* first workgroup of this draw in the per-vertex grid */
  uint32_t tri_wg;    /* first workgroup of this draw in the per-triangle grid */
  /* order-free path, large draws: per chunk of 256 triangles the object-space
   * box (lo.xyz, non-finite flag) (hi.xyz, 0) of the vertices its valid
   * triangles use, or null (no chunk culling) */
  const float4 *chunk_box;
  /* with chunk_box: per-vertex-workgroup chunk references (k_chunk_bounds) */
  const uint32_t *vref;
};

/* chunks listed per per-vertex workgroup; one used by more is always kept */
#define VREF_CAP 16

/* k_chunk_bounds: the boxes of one (index buffer, vertex buffer) pair */
struct ChunkBoundsArgs {
  const uint8_t *vbuf;
  const uint32_t *ibuf;
  uint32_t tri_count, vertex_count, stride, flex;
  int32_t pos_off, pos_fmt;
  float4 *box;   /* 2 per chunk of 256 triangles */
  uint32_t *vref; /* nvwg counts (zeroed), then nvwg * VREF_CAP chunk indices */
  uint32_t nvwg;  /* per-vertex workgroups: (vertex_count + 255) / 256 */
};

/* Screen-space triangle: exactly the MopSwScreenVertex positions the CPU
 * path hands to mop_sw_rasterize_triangle_smooth_ml. */
struct SubTri {
  float sx[3], sy[3], sz[3], iw[3];
  uint32_t draw;
  uint32_t a[3]; /* attribute slots: < n_verts -> vertex cache, else clip store */
};

/* Counters in device memory. */
enum {
  CNT_CLIP = 0,     /* triangles queued for clipping */
  CNT_BIG = 1,      /* big screen triangles queued */
  CNT_SUBTRI = 2,   /* total screen triangles (written by the scan) */
  CNT_OVERFLOW = 3, /* bit0 clip queue, bit1 subtri store, bit2 shadow big queue,
                       bit3 clipped polygon longer than MOPH_MAX_POLY, bit4 tile bins,
                       bit5 big-triangle checkpoint store */
  CNT_SHADOW_BIG = 4,
  CNT_CHK = 5,    /* order-free path: row checkpoint entries of big triangles */
  CNT_BINS = 6,   /* ordered path: tile-bin entries the batch needed */
  CNT_TICKET = 7, /* k_finish: workgroups done */
  CNT_BIGCHUNK = 8, /* order-free path: 64-row chunks of big screen triangles */
  CNT_TLIST = 9,    /* chunk culling: kept per-triangle workgroups */
  CNT_VLIST = 10,   /* chunk culling: kept per-vertex workgroups */
  CNT_CULLED = 11,  /* chunk culling: chunks of 256 triangles skipped */
  CNT_N = 12
};

/* Overflow bits that make the runtime grow a scratch store and run the
 * frame again (clip queue, screen-triangle store, tile bins, checkpoints). */
#define CNT_OVERFLOW_RERUN 51u

/* One view (framebuffer + the shading state its draws see) of a batch.
 * A batch renders one order-free segment of each of its views in a single
 * set of launches; a single frame is a batch of one view. */
struct ViewGPU {
  int32_t w, h;
  uint32_t stile_base; /* first shading tile of this view in the batch */
  uint32_t tile_base;  /* first 16x16 tile of this view in the batch's tiles */
  unsigned long long *vis;
  float4 *wbuf;
  float *hdr;
  float *depth;
  uint32_t *ids;
  uint32_t *cnt_out; /* device alias of host-pinned counters of this segment */
  /* shadow lookup state (mop_sw_shadow_set) */
  const uint32_t *shadow_map; /* ordered depth keys (key_depth gives the depth) */
  int32_t shadow_w, shadow_h;
  int32_t shadow_on;
  int32_t ibl_on;
  float shadow_vp[16];
  /* IBL state (mop_sw_ibl_set) */
  const float *irr;
  const float *pf;
  const float *lut;
  int32_t irr_w, irr_h, pf_w, pf_h, pf_levels, lut_size;
  float ibl_rot, ibl_int;
  /* 1: this segment opens a cleared frame, and the shading pass writes the
   * clear values (colour, depth 1, id 0) to the pixels no triangle covers;
   * 2: the same but depth and id only (an environment background already
   * wrote the colour) */
  int32_t shade_clear;
  float clear[4];
};

/* Everything the frame kernels read, passed by value. */
struct FrameArgs {
  const DrawGPU *draws;
  const ViewGPU *views;
  uint32_t ndraws;
  uint32_t nviews;
  /* draw-major grids of the per-vertex and per-triangle kernels: workgroup b
   * works on 256 items of the single draw wg_vtx[b] / wg_tri[b]; the maps are
   * written by k_wg_map from the runs (per draw in grid order: its first
   * per-triangle and per-vertex workgroups, the draw) */
  const uint32_t *wg_vtx;
  const uint32_t *wg_tri;
  const uint4 *runs;
  uint32_t n_runs;
  uint32_t n_wg_vtx;
  uint32_t n_wg_tri;
  uint32_t n_verts;
  uint32_t n_tris;
  uint32_t clip_cap;
  uint32_t subtri_cap;
  uint32_t n_tiles;   /* 16x16 tiles over all views (ordered path) */
  uint32_t n_stiles;  /* shade_tile^2 tiles over all views (shading) */
  uint32_t small_area;  /* bbox pixels up to which one lane walks a triangle */
  uint32_t shade_tile;  /* shading tile edge: 8, 16 or 32 */
  uint32_t raster_grid; /* workgroups of the small-triangle raster (0 = auto) */
  uint32_t shade_waves; /* waves per SIMD the shading pass is compiled for: 4, 5, 6 or 8 */
  uint32_t fused;       /* 1: order-free path (no screen-triangle store) */
  float4 *vpos;   /* n_verts clip positions */
  float4 *vscr;   /* n_verts screen positions (sx, sy, sz, 1/w) in the draw's view */
  uint4 *tri_rec; /* order-free path, per triangle: (v0, v1, v2, draw) when drawn
                     whole, (clip slot, 0, 0, draw) when queued for clipping */
  uint32_t *blk_counts; /* order-free path: screen triangles per k_tri workgroup */
  /* order-free path, big screen triangles: the walk state (w0, w1, w2, flat z)
   * at every 16th pixel of every bbox row, so shading replays at most 15
   * steps.  big_chk[q] is the first entry of big-queue entry q, big_off the
   * same indexed by triangle (whole) or n_tris + clip slot * 16 + fan index
   * (clipped); ~0 when the store overflowed. */
  float4 *chk;
  uint32_t chk_cap;
  uint32_t *big_chk;
  uint32_t *big_off;
  /* order-free path: big_map[c] = (big-queue entry, chunk) for every chunk
   * of 64 bbox rows of a big screen triangle with checkpoints.  A chunk
   * holds at least one bbox row and every row at least one checkpoint, so
   * chk_cap entries always suffice. */
  uint2 *big_map;
  /* order-free path, chunk culling (a draw with chunk_box in the batch):
   * the workgroups of the culled draws follow the first tri_direct /
   * vtx_direct ones; tri_vis[b] = 0 when the triangles of per-triangle
   * workgroup b are all trivially rejected; tri_list / vtx_list the kept
   * workgroups of the culled draws as (workgroup, draw) (counts in
   * CNT_TLIST / CNT_VLIST), walked by tri_lgrid / vtx_lgrid workgroups;
   * null without chunk culling */
  uint32_t *tri_vis;
  uint2 *tri_list;
  uint2 *vtx_list;
  uint32_t tri_direct, vtx_direct;
  uint32_t tri_lgrid, vtx_lgrid;
  /* entries of blk_counts the triangle kernel writes (n_wg_tri, or
   * tri_direct + tri_lgrid with lists); the clipped-fan raster's follow */
  uint32_t tri_grid;
  float4 *vattr;  /* n_verts * 3: (n.xyz, wp.x) (wp.yz, c.rg) (c.ba, -, -) */
  float4 *cpos;   /* clip_cap * MOPH_MAX_POLY */
  float4 *cattr;  /* clip_cap * MOPH_MAX_POLY * 3 */
  uint32_t *cnt_poly; /* clip_cap: polygon vertex count per queued triangle */
  uint32_t *tri_counts; /* n_tris (+1): screen-triangle count, scanned in place */
  uint32_t *clip_queue;
  uint32_t *clip_slot; /* n_tris: clip-queue slot of a triangle or ~0 */
  uint32_t *counters;
  SubTri *subtris;
  uint32_t *big_queue;
  const MopHipLight *lights;
  /* the device's overflow words: k_finish ORs the batch's overflow bits into
   * sticky[0] (device memory) and keeps the largest clip-queue, checkpoint
   * and tile-bin needs in sticky[1..3], and stores all four to sticky_out (a
   * device alias of host-mapped memory), so the host can check every
   * executed batch, not only the frames it reads back, and size the next
   * batches */
  uint32_t *sticky;
  uint32_t *sticky_out;
};

/* One view of the batched clear. */
struct ClearViewGPU {
  uint32_t px_base; /* first pixel of the view in the launch */
  uint32_t pad_;
  float clear[4];
  float *hdr;
  float *depth;
  uint32_t *ids;
  unsigned long long *vis;
};
void launch_clear_views(const ClearViewGPU *views, uint32_t n, uint32_t total, hipStream_t s);

/* Per-view arguments of the resolve and post passes. */
struct PostViewGPU {
  int32_t w, h;
  uint32_t tile_base;
  uint32_t effects;
  const float *hdr;
  const float *depth;
  const uint32_t *ids;
  const uint8_t *src;
  uint8_t *dst;
  float exposure;
  float fog_r, fog_g, fog_b, fog_near, fog_far;
  int32_t has_fog;
  int32_t pad_;
  const uint8_t *lut; /* per-channel table of the colour effects, or null */
  /* object outline on the resolved image before the effects
   * (overlay_builtin.c:1309-1400); outline == 0 leaves it off */
  const uint32_t *sel; /* selected object ids, sorted ascending, no repeats */
  uint32_t n_sel;
  uint32_t accent; /* (u8)(accent * 255) per channel, packed r | g << 8 | b << 16 */
  float alpha_sel, alpha_unsel;
  int32_t outline;
  int32_t pad2_;
};

/* One shadow caster mesh (mop_sw_shadow_render_mesh arguments). */
struct ShadowDrawGPU {
  float mvp[16]; /* light_vp * model, mop_mat4_multiply order */
  const uint8_t *vbuf;
  const uint32_t *ibuf;
  uint32_t vertex_count;
  uint32_t tri_count;
  uint32_t tri_base;
  uint32_t vtx_base; /* first slot in the light-space vertex cache */
};

struct ShadowArgsGPU {
  const ShadowDrawGPU *draws;
  uint32_t ndraws;
  uint32_t n_tris;
  int32_t size;
  uint32_t big_cap;
  uint32_t n_verts;
  uint32_t pad_;
  float4 *lclip; /* n_verts light clip positions */
  float4 *lscr;  /* n_verts shadow-map positions (sx, sy, sz, -) */
  uint32_t *keys; /* size*size ordered depth keys */
  uint32_t *big_queue;
  uint32_t *counters;
};

/* Kernel launchers (kernels.hip). */
void launch_wg_map(const FrameArgs &fa, hipStream_t s);
void launch_vertex(const FrameArgs &fa, hipStream_t s);
void launch_chunk_bounds(const ChunkBoundsArgs &a, hipStream_t s);
void launch_cull(const FrameArgs &fa, hipStream_t s);
void cull_list_grids(uint32_t *tri_grid, uint32_t *vtx_grid);
void launch_classify(const FrameArgs &fa, hipStream_t s);
void launch_clip(const FrameArgs &fa, hipStream_t s);
/* Skinning: bind-pose records in, deformed records out (same layout). */
struct SkinArgs {
  const uint8_t *bind;
  uint8_t *out;
  const float *bones; /* bone_count column-major 4x4 */
  uint32_t vertex_count;
  uint32_t stride;
  uint32_t bone_count;
  int32_t pos_off, normal_off, joints_off, weights_off; /* normal_off < 0: none */
  int32_t aligned; /* stride and buffers dword aligned */
};

struct MorphArgs {
  const uint8_t *bind;
  uint8_t *out;
  const float *targets; /* target_count x vertex_count x 3 */
  const float *weights;
  uint32_t vertex_count;
  uint32_t stride;
  uint32_t target_count;
  int32_t pos_off;
  int32_t aligned;
};

void launch_skin(const SkinArgs &a, hipStream_t s);
void launch_morph(const MorphArgs &a, hipStream_t s);

/* workgroups of the order-free clipped-fan raster (k_clip_raster) */
#define CLIP_RASTER_GRID 1024
void launch_fused_setup(const FrameArgs &fa, hipStream_t s);
void launch_raster_big(const FrameArgs &fa, hipStream_t s);
void launch_scan_counts(const FrameArgs &fa, uint32_t *block_sums, hipStream_t s);
void launch_emit(const FrameArgs &fa, hipStream_t s);
void launch_raster(const FrameArgs &fa, hipStream_t s);
void launch_shade(const FrameArgs &fa, hipStream_t s);
void launch_export_counters(const FrameArgs &fa, hipStream_t s);
/* ordered path: bin screen triangles to 16x16 tiles, sort each tile's list
 * by submission index, replay the CPU loop per tile in LDS */
void launch_ordered(const FrameArgs &fa, uint32_t *tile_counts, uint32_t *tile_cursor,
                    uint32_t *bins, uint32_t bin_cap, uint32_t *block_sums, hipStream_t s);
void launch_clear(int npx, const float clear[4], uint8_t *color, float *hdr,
                  float *depth, uint32_t *ids, unsigned long long *vis,
                  hipStream_t s);
void launch_vis_from_depth(int npx, const float *depth,
                           unsigned long long *vis, hipStream_t s);
/* resolve (mode 0) or post (mode 1) over the tiles of n views */
void launch_tail_views(const PostViewGPU *views, uint32_t nviews, uint32_t n_tiles,
                       bool outline, hipStream_t s);
void launch_upload(const void *src, void *dst, size_t n, hipStream_t s);
void launch_copy_u8(size_t n, const uint8_t *src, uint8_t *dst, hipStream_t s);
void launch_shadow_clear(int n, uint32_t *keys, hipStream_t s);
void launch_shadow_raster(const ShadowArgsGPU &sa, hipStream_t s);
void launch_put_shadow_draw(const ShadowDrawGPU &g, ShadowDrawGPU *dst, hipStream_t s);
void launch_downsample(const uint8_t *src, int src_w, uint8_t *dst, int pw, int ph, int sf,
                       hipStream_t s);

/* Environment background of one view (the CPU skybox of viewport.c:2128-2180
 * or the RHI draw_skybox form): written to every pixel's HDR before the
 * frame's draws. */
struct BackgroundGPU {
  const float *env; /* RGBA32F equirect */
  int32_t env_w, env_h;
  int32_t w, h;
  int32_t mode; /* 1: camera basis (CPU path), 2: inverse view-projection */
  float fwd[3], right[3], up[3];
  float half_w, half_h;
  float inv_vp[16];
  float cam[3];
  float rotation, intensity;
  float *hdr;
};
void launch_background(const BackgroundGPU &bg, hipStream_t s);
void launch_shadow_finalize(int n, const uint32_t *keys, float *depth,
                            hipStream_t s);
void launch_fill_u32(size_t n, uint32_t *p, uint32_t v, hipStream_t s);

/* runtime services for the other translation units */
MopHipBuffer *buffer_adopt_device(MopHipDevice *d, void *dev, size_t size);
void *buffer_device_ptr(MopHipBuffer *b);
hipStream_t device_stream(MopHipDevice *d);
void runtime_set_err(const char *fmt, ...);

#endif
