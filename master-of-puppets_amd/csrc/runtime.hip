/*
 * runtime.hip -- host side of the HIP backend: resources, deferred frame
 * recording and execution, readbacks, and the hooks for the viewport core's
 * CPU-gated calls.  All entry points are extern "C" (include/mop_hip.h).
 *
 * Frame model.  frame_begin records the clear, draw/draw_instanced record
 * draws (host copies of the call data; vertex/index data are already
 * resident), frame_end executes the recorded frame on the device stream and
 * resolves HDR -> RGBA8 with the exposure given by set_exposure.  Readbacks
 * synchronise the stream once and copy into pinned host memory owned by the
 * framebuffer (valid until the next frame, resize or destroy, as the RHI
 * requires: viewport.h:157-166).
 *
 * Draws are grouped into segments of consecutive draws of one class.  The
 * class every parity scene uses -- opaque, opacity >= 1, depth test and write
 * on, smooth shading with at least one light, not wireframe -- is order-free:
 * the CPU's strict `z < depth` with first-submitted-wins gives the same image
 * as a lexicographic min over (depth, submission index), which the kernels
 * compute with 64-bit atomics.  Draws of other classes are not rendered by
 * this backend yet; they are reported through mop_hip_last_error and counted
 * in the frame statistics.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "mop_hip_internal.h"

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
/* ------------------------------------------------------------------------ */

static thread_local char t_err[512];

static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
  fprintf(stderr, "[mop-hip] %s\n", t_err);
}

#define HIPCHK(expr)                                                           \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
              __LINE__);                                                       \
    }                                                                          \
  } while (0)

extern "C" const char *mop_hip_last_error(void) { return t_err; }

/* ------------------------------------------------------------------------ */
/* host matrix product in mop_mat4_multiply order (math.c:182-194)           */
/* ------------------------------------------------------------------------ */

static MopHipMat4 mat4_mul(const MopHipMat4 &a, const MopHipMat4 &b) {
  MopHipMat4 r;
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      float sum = 0.0f;
      for (int k = 0; k < 4; k++)
        sum += a.d[k * 4 + row] * b.d[c * 4 + k];
      r.d[c * 4 + row] = sum;
    }
  return r;
}

/* ------------------------------------------------------------------------ */
/* resource types                                                            */
/* ------------------------------------------------------------------------ */

struct MopHipBuffer {
  void *host = nullptr;
  void *dev = nullptr;
  size_t size = 0;
  MopHipDevice *owner = nullptr;
  bool host_stale = false; /* device side rewritten by skinning/morphing */
  /* device copy of the bind pose handed to the skinning/morph entry points */
  void *bind_dev = nullptr;
  const void *bind_src = nullptr;
  size_t bind_size = 0;
  /* contents version, from one process-wide counter: a new or rewritten
   * buffer never carries a version an earlier one had */
  uint64_t version = 0;
};

static std::atomic<uint64_t> g_buf_version{0};
static void buf_touched(MopHipBuffer *b) { b->version = ++g_buf_version; }

struct MopHipTexture {
  int w = 0, h = 0;
  bool hdr = false;
  void *host = nullptr;
  void *dev = nullptr; /* RGBA8 draw textures; RGBA32F environment maps (uploaded on
                          first use as a background) */
};

struct PendingDraw {
  DrawGPU g;              /* filled except vtx/tri bases and light_off */
  MopHipBuffer *vb = nullptr, *ib = nullptr;
  uint32_t light_at;      /* the draw's lights: fb->light_pool[light_at ..] */
  int cls;                /* 0 order-free, 3 ordered, 2 no-op */
  int state;              /* index into the frame's shading-state list */
};

/* Shading state a draw sees (mop_sw_shadow_set / mop_sw_ibl_set state at
 * the time of the draw): device pointers of the bound tables. */
struct ShadeState {
  bool shadow_on = false;
  const uint32_t *shadow_map = nullptr; /* ordered depth keys */
  int shadow_w = 0, shadow_h = 0;
  float shadow_vp[16] = {0};
  bool ibl_on = false;
  const float *irr = nullptr, *pf = nullptr, *lut = nullptr;
  int irr_w = 0, irr_h = 0, pf_w = 0, pf_h = 0, pf_levels = 0, lut_size = 0;
  float ibl_rot = 0.0f, ibl_int = 1.0f;
};

struct MopHipFramebuffer {
  int w = 0, h = 0;
  uint8_t *color[2] = {nullptr, nullptr};
  int cur = 0;
  float *hdr = nullptr, *depth = nullptr;
  uint32_t *ids = nullptr;
  unsigned long long *vis = nullptr;
  float4 *wbuf = nullptr;
  uint8_t *h_color = nullptr;
  float *h_depth = nullptr, *h_hdr = nullptr;
  uint32_t *h_ids = nullptr;
  bool v_color = false, v_depth = false, v_ids = false, v_hdr = false;
  /* recorded frame */
  bool has_clear = false;
  bool shade_clear = false; /* the clear is left to the first segment's shading */
  float clear[4] = {0, 0, 0, 0};
  std::vector<PendingDraw> draws;
  std::vector<MopHipLight> light_pool; /* the recorded draws' lights */
  std::vector<ShadeState> states;
  bool executed = false;
  bool checked = false;
  /* resolve / post requested for this frame (replayed on re-execution) */
  bool resolve_set = false;
  float resolve_exposure = 1.0f;
  uint32_t post_effects = 0;
  bool has_fog = false;
  MopHipFogParams fog = {};
  /* per-segment counters written by the device into pinned host memory */
  uint32_t *h_cnt = nullptr; /* MOPH_MAX_SEGS * CNT_N, host pointer */
  uint32_t *d_cnt = nullptr; /* device alias of h_cnt */
  int nseg = 0;
  hipEvent_t done = nullptr;
  size_t skipped = 0;
  /* asynchronous readback (mop_hip_framebuffer_readback_async): copies of
   * the execution numbered rb_seq into the host buffers, on the device's copy
   * stream, finished when rb_done completes; exec_seq numbers the contents
   * of the device buffers (every execution or re-run of the tail) */
  hipEvent_t rb_done = nullptr;
  bool rb_pending = false;
  uint32_t rb_mask = 0;
  uint64_t rb_seq = 0, exec_seq = 0;
  /* environment background of the recorded frame (mop_hip_draw_background_hdri
   * or the draw_skybox slot) */
  bool has_bg = false;
  BackgroundGPU bg = {};
  /* object outline of the recorded frame (mop_hip_draw_outline) */
  bool outline = false;
  uint32_t outline_accent = 0;
  float outline_alpha_sel = 0.0f, outline_alpha_unsel = 0.0f;
  std::vector<uint32_t> outline_sel; /* sorted, no repeats */
  /* SSAA-downsampled RGBA8 (mop_hip_framebuffer_read_color_downsampled) */
  uint8_t *ds_dev = nullptr, *ds_host = nullptr;
  size_t ds_cap = 0;
  uint64_t ds_seq = ~0ull;
  int ds_factor = 0;
};

#define MOPH_MAX_SEGS 32

/* set when a plan needed a buffer to grow: the frame is planned again
 * without a plan (growing frees memory that recorded operations point at) */
static thread_local bool t_plan_abort = false;
thread_local Plan *t_plan = nullptr;
thread_local hipEvent_t t_kernel_ev[2] = {nullptr, nullptr};

template <class T> struct DevArr {
  T *p = nullptr;
  size_t cap = 0;
  bool ensure(size_t n) {
    if (n <= cap)
      return true;
    if (t_plan) {
      t_plan_abort = true;
      return false;
    }
    if (p)
      (void)hipFree(p);
    size_t nc = n + n / 4 + 256;
    if (hipMalloc((void **)&p, nc * sizeof(T)) != hipSuccess) {
      set_err("hipMalloc of %zu bytes failed", nc * sizeof(T));
      p = nullptr;
      cap = 0;
      return false;
    }
    cap = nc;
    return true;
  }
  void release() {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

/* Chunk boxes of one (index buffer, vertex buffer) pair for the draw
 * parameters they were computed with (k_chunk_bounds); stale once either
 * buffer's version or a parameter changes. */
struct ChunkBoxes {
  uint64_t ib_ver = 0, vb_ver = 0;
  uint32_t tri_count = 0, vertex_count = 0, stride = 0, flex = 0;
  int32_t pos_off = 0, pos_fmt = 0;
  uint64_t batch = 0; /* the batch that last computed them */
  DevArr<float4> box;
  DevArr<uint32_t> vref; /* per-vertex-workgroup chunk references */
};

/* batches with at most this many per-vertex plus per-triangle workgroups
 * upload their workgroup -> draw maps; larger ones build them on the device */
#define WG_MAP_HOST_MAX 32768u

/* draws with at least this many triangles are chunk-culled */
#define CULL_MIN_TRIS 65536u

struct MopHipShadowMap {
  int size = 0;
  DevArr<uint32_t> keys; /* ordered depth keys: the pass's atomic min target and
                            what the PCF lookup reads */
};

struct MopHipIbl {
  DevArr<float> irr, pf, lut;
  int irr_w = 0, irr_h = 0, pf_w = 0, pf_h = 0, pf_levels = 0, lut_size = 0;
  bool has_irr = false, has_pf = false, has_lut = false;
};

struct ShadowDrawHost {
  ShadowDrawGPU g;
};

struct ShadowJob {
  MopHipShadowMap *map;
  std::vector<ShadowDrawHost> casters;
};

/* A captured frame plan: two graph instances, each with its own pinned
 * upload buffer, used alternately so the host refills one while the device
 * may still read the other. */
/* A captured plan: its runs of same-stream operations between cross-stream
 * dependencies are captured as one graph each (graph branches would run in
 * order on one queue); a replay launches each on its own stream and issues
 * the dependencies (event record + wait) between them, so the side stream's
 * shadow passes overlap the main stream's work. */
struct GraphEntry {
  std::vector<uint8_t> sig;
  std::vector<hipGraphExec_t> seg[2]; /* per upload buffer, per segment */
  std::vector<hipStream_t> seg_stream;
  hipGraphExec_t exec[2] = {nullptr, nullptr}; /* unused */
  uint8_t *blob[2] = {nullptr, nullptr};
  uint8_t *dblob[2] = {nullptr, nullptr}; /* device aliases of blob */
  hipEvent_t done[2] = {nullptr, nullptr};
  size_t blob_size = 0;
  int next = 0;
  uint64_t use = 0;
};

/* Optional per-stage device timers: hipEvents recorded on the device
 * stream around each stage while enabled; summed on read. */
enum { TM_VERTEX, TM_SETUP, TM_RASTER, TM_SHADE, TM_RESOLVE, TM_POST, TM_SHADOW, TM_N };
static const char *k_timer_names[TM_N] = {"vertex", "setup", "raster", "shade",
                                          "resolve", "post", "shadow"};

struct StageTimers {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double total_ms[TM_N] = {0};
  uint64_t count[TM_N] = {0};
  size_t next = 0;
};

#define NSTAGE 16
struct StageSlot {
  void *p = nullptr;
  uint8_t *dp = nullptr; /* device alias of p (the upload kernel reads it) */
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool used = false;
};

struct MopHipDevice {
  hipStream_t stream = nullptr;
  /* device-to-host copies of asynchronous readbacks: they overlap the next
   * frames' kernels on `stream` */
  hipStream_t copy = nullptr;
  /* queued shadow passes run on `side`, forked from and joined back into
   * `stream` before the first shading launch that reads the maps */
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool join_pending = false;
  bool shadow_deferred = false; /* direct launches: passes launched at the first join */
  StageTimers tm;
  StageSlot ring[NSTAGE];
  int ring_next = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex mu;
  std::unordered_map<const void *, MopHipBuffer *> host2buf;
  /* frame scratch */
  DevArr<DrawGPU> draws;
  DevArr<MopHipLight> lights;
  DevArr<float4> vpos, vscr, vattr, cpos, cattr;
  DevArr<uint4> tri_rec;
  DevArr<uint32_t> blk_counts;
  DevArr<float4> chk;          /* big-triangle row checkpoints */
  DevArr<uint32_t> big_chk, big_off;
  DevArr<uint2> big_map; /* row chunks of big screen triangles */
  uint32_t chk_cap_hint = 0;
  /* chunk culling of large order-free draws (MOP_HIP_CHUNK_CULL=0 turns it
   * off): boxes per (index buffer, vertex buffer), and the per-frame marks */
  bool chunk_cull = true;
  uint32_t cull_tri_grid = 1024, cull_vtx_grid = 1024; /* list-walk grids */
  std::map<std::pair<const MopHipBuffer *, const MopHipBuffer *>, ChunkBoxes> cboxes;
  uint64_t batch_clock = 0;
  DevArr<uint32_t> tri_vis;
  DevArr<uint32_t> wg_map; /* per-triangle then per-vertex workgroup -> draw */
  DevArr<uint2> tri_list, vtx_list;
  DevArr<uint8_t> arena; /* per-batch upload: counters, draws, lights, views */
  DevArr<float> skin_bones, morph_targets, morph_weights;
  DevArr<uint32_t> cnt_poly, tri_counts, clip_queue, clip_slot, counters, block_sums,
      big_queue;
  DevArr<SubTri> subtris;
  uint32_t clip_cap_hint = 0;
  /* tunables (mop_hip_set_tuning) */
  uint32_t small_area = 64, shade_tile = 16, raster_grid = 0, shade_waves = 5;
  /* mop_hip_set_opaque_raster: order-free draws recorded from now on take the
   * tile path (binned, LDS tiles walked in submission order) */
  uint32_t opaque_tiles = 0;
  DevArr<uint32_t> tile_counts, tile_cursor, bins;
  uint32_t bin_cap_hint = 0;
  /* forced scratch capacities (mop_hip_set_scratch_limits; 0 = none): used
   * for the next executions until an overflow makes the runtime grow the
   * store and run the frame again */
  uint32_t force_clip = 0, force_chk = 0, force_bins = 0;
  /* sticky overflow word: device copy (atomic OR by k_finish) and its
   * host-mapped mirror */
  DevArr<uint32_t> sticky;
  uint32_t *h_sticky = nullptr, *d_sticky_out = nullptr;
  /* device address of `scratch` (mop_hip_shadow_render_noalloc); null when
   * the symbol lookup failed, and the entry point then reports an error */
  uint8_t *noalloc_scratch = nullptr;
  uint32_t *h_counters = nullptr;
  DevArr<ViewGPU> views;
  DevArr<PostViewGPU> pviews;
  DevArr<uint32_t> outline_ids; /* selection lists of the views' outlines */
  DevArr<ClearViewGPU> clear_views;
  DevArr<uint8_t> post_lut; /* 4 x 256 colour-effect tables */
  DevArr<float> ibl_trig;   /* sine/cosine table of the IBL precompute */
  /* shadow: the map the hook API renders into, and the recorded casters */
  MopHipShadowMap *def_shadow = nullptr;
  MopHipShadowMap *rec_shadow = nullptr; /* map the recorded casters target */
  DevArr<uint32_t> shadow_big;
  DevArr<float4> shadow_lclip, shadow_lscr;
  DevArr<uint8_t> shadow_arena; /* counters + caster draws of one pass */
  DevArr<ShadowDrawGPU> shadow_draws_dev;
  std::vector<ShadowDrawHost> shadow_draws;
  std::vector<ShadowJob> shadow_jobs; /* queued shadow passes */
  /* frame plans replayed from hipGraphs (mop_hip_set_graphs) */
  int graphs_mode = 2; /* 0 direct, 1 graphs, 2 graphs for multi-frame executions */
  std::vector<GraphEntry> graphs;
  std::vector<uint8_t> last_sig;
  uint64_t plan_clock = 0;
  uint64_t graph_replays = 0, graph_captures = 0;
  std::vector<MopHipBuffer *> shadow_temp;
  ShadeState cur_state;
  bool state_dirty = true;
  /* IBL tables uploaded through mop_hip_ibl_set */
  MopHipIbl *def_ibl = nullptr;
  uint64_t ibl_hash = 0;
  float exposure = 1.0f;
  MopHipFrameStats stats = {};
  bool timing_pending = false;
};

/* ------------------------------------------------------------------------ */
/* device / buffers / textures                                               */
/* ------------------------------------------------------------------------ */

/* Process-wide forced scratch capacities copied into every device created
 * afterwards (mop_hip_set_default_scratch_limits), and the number of frame
 * re-runs caused by scratch overflow in this process. */
static uint32_t g_force_clip = 0, g_force_chk = 0, g_force_bins = 0;
static std::atomic<uint64_t> g_overflow_reruns{0};

/* Scratch of mop_hip_shadow_render_noalloc, held by every GPU the library
 * runs on until the process ends: [0, 256) counters, [256, 512) the caster
 * record, then the light-space clip and shadow-map positions of up to
 * NOALLOC_MAX_VERTS vertices (16 B each, twice) and the big-triangle queue
 * of up to NOALLOC_MAX_TRIS triangles (4 B each): 20 MiB + 512 B.  In host
 * code this name is a host-side stand-in, not a device address: the device
 * address comes from hipGetSymbolAddress at device creation. */
#define NOALLOC_MAX_VERTS (1u << 19)
#define NOALLOC_MAX_TRIS (1u << 20)
#define NOALLOC_OFF_DRAW 256u
#define NOALLOC_OFF_LCLIP 512u
#define NOALLOC_OFF_LSCR (NOALLOC_OFF_LCLIP + (size_t)NOALLOC_MAX_VERTS * 16u)
#define NOALLOC_OFF_BIG (NOALLOC_OFF_LSCR + (size_t)NOALLOC_MAX_VERTS * 16u)
#define NOALLOC_BYTES (NOALLOC_OFF_BIG + (size_t)NOALLOC_MAX_TRIS * 4u)
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

/* The per-channel colour-effect tables of mop_postprocess_apply without fog
 * or vignette (postprocess.c:45-116): [tonemap | gamma << 1][value].  The
 * sRGB curve calls powf, so the tables are evaluated here with the host C
 * library, the one the reference calls (through a pointer, so nothing is
 * folded at build time with a different rounding). */
static float (*volatile host_powf)(float, float) = powf;

static void post_lut_fill(uint8_t *lut) {
  for (int variant = 0; variant < 4; variant++)
    for (int o = 0; o < 256; o++) {
      float r = (float)o / 255.0f;
      if (variant & 1)
        r = r / (r + 1.0f);
      if (variant & 2) {
        float c = r > 0.0f ? r : 0.0f;
        r = c <= 0.0031308f ? c * 12.92f : 1.055f * host_powf(c, 1.0f / 2.4f) - 0.055f;
      }
      if (r > 1.0f)
        r = 1.0f;
      if (r < 0.0f)
        r = 0.0f;
      lut[variant * 256 + o] = (uint8_t)(r * 255.0f + 0.5f);
    }
}

extern "C" MopHipDevice *mop_hip_device_create(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_err("no HIP device visible");
    return nullptr;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    set_err("hipGetDeviceProperties failed");
    return nullptr;
  }
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_err("device 0 is %s; this library is built for gfx950 only", prop.gcnArchName);
    return nullptr;
  }
  MopHipDevice *d = new MopHipDevice();
  HIPCHK(hipSetDevice(0));
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    set_err("hipStreamCreate failed");
    delete d;
    return nullptr;
  }
  if (hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking) != hipSuccess) {
    set_err("hipStreamCreate failed");
    (void)hipStreamDestroy(d->stream);
    delete d;
    return nullptr;
  }
  if (hipStreamCreateWithFlags(&d->copy, hipStreamNonBlocking) != hipSuccess) {
    set_err("hipStreamCreate failed");
    (void)hipStreamDestroy(d->side);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return nullptr;
  }
  HIPCHK(hipEventCreate(&d->ev0));
  HIPCHK(hipEventCreate(&d->ev1));
  HIPCHK(hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming));
  for (int k = 0; k < NSTAGE; k++)
    HIPCHK(hipEventCreateWithFlags(&d->ring[k].ev, hipEventDisableTiming));
  HIPCHK(hipHostMalloc((void **)&d->h_counters, CNT_N * sizeof(uint32_t), 0));
  if (hipHostMalloc((void **)&d->h_sticky, 64, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&d->d_sticky_out, d->h_sticky, 0) != hipSuccess ||
      !d->sticky.ensure(4)) {
    set_err("overflow word allocation failed");
    mop_hip_device_destroy(d);
    return nullptr;
  }
  memset(d->h_sticky, 0, 4 * sizeof(uint32_t));
  launch_fill_u32(4, d->sticky.p, 0u, d->stream);
  d->force_clip = g_force_clip;
  d->force_chk = g_force_chk;
  d->force_bins = g_force_bins;
  if (const char *e = getenv("MOP_HIP_CHUNK_CULL"))
    d->chunk_cull = atoi(e) != 0;
  cull_list_grids(&d->cull_tri_grid, &d->cull_vtx_grid);
  /* this device's copy of the allocation-free entry point's scratch */
  if (hipGetSymbolAddress((void **)&d->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess ||
      !d->noalloc_scratch) {
    (void)hipGetLastError();
    d->noalloc_scratch = nullptr;
  }
  if (d->post_lut.ensure(4 * 256)) {
    uint8_t lut[4 * 256];
    post_lut_fill(lut);
    HIPCHK(hipMemcpy(d->post_lut.p, lut, sizeof(lut), hipMemcpyHostToDevice));
  }
  return d;
}

extern "C" void mop_hip_device_destroy(MopHipDevice *d) {
  if (!d)
    return;
  (void)hipStreamSynchronize(d->stream);
  d->draws.release();
  d->lights.release();
  d->vpos.release();
  d->vattr.release();
  d->cpos.release();
  d->cattr.release();
  d->cnt_poly.release();
  d->tri_counts.release();
  d->clip_queue.release();
  d->clip_slot.release();
  d->counters.release();
  d->block_sums.release();
  d->big_queue.release();
  d->subtris.release();
  d->shadow_big.release();
  d->shadow_draws_dev.release();
  for (GraphEntry &e : d->graphs)
    for (int k = 0; k < 2; k++) {
      if (e.exec[k])
        (void)hipGraphExecDestroy(e.exec[k]);
      if (e.blob[k])
        (void)hipHostFree(e.blob[k]);
      if (e.done[k])
        (void)hipEventDestroy(e.done[k]);
    }
  d->graphs.clear();
  d->shadow_lclip.release();
  d->shadow_arena.release();
  d->shadow_lscr.release();
  d->vscr.release();
  d->tri_rec.release();
  d->blk_counts.release();
  d->chk.release();
  d->big_chk.release();
  d->big_off.release();
  d->big_map.release();
  for (auto &kv : d->cboxes) {
    kv.second.box.release();
    kv.second.vref.release();
  }
  d->cboxes.clear();
  d->tri_vis.release();
  d->wg_map.release();
  d->tri_list.release();
  d->vtx_list.release();
  d->arena.release();
  d->skin_bones.release();
  d->morph_targets.release();
  d->morph_weights.release();
  d->post_lut.release();
  d->ibl_trig.release();
  mop_hip_shadow_map_destroy(d, d->def_shadow);
  mop_hip_ibl_destroy(d, d->def_ibl);
  for (MopHipBuffer *b : d->shadow_temp) {
    (void)hipFree(b->dev);
    delete b;
  }
  for (int k = 0; k < NSTAGE; k++) {
    if (d->ring[k].p)
      (void)hipHostFree(d->ring[k].p);
    (void)hipEventDestroy(d->ring[k].ev);
  }
  d->views.release();
  d->pviews.release();
  d->outline_ids.release();
  d->clear_views.release();
  d->tile_counts.release();
  d->tile_cursor.release();
  d->bins.release();
  if (d->h_counters)
    (void)hipHostFree(d->h_counters);
  if (d->h_sticky)
    (void)hipHostFree(d->h_sticky);
  d->sticky.release();
  (void)hipEventDestroy(d->ev0);
  (void)hipEventDestroy(d->ev1);
  (void)hipEventDestroy(d->ev_fork);
  (void)hipEventDestroy(d->ev_join);
  (void)hipStreamDestroy(d->side);
  if (d->copy) {
    (void)hipStreamSynchronize(d->copy);
    (void)hipStreamDestroy(d->copy);
  }
  (void)hipStreamDestroy(d->stream);
  delete d;
}

/* Pinned staging ring: each upload takes the next slot, waiting only for
 * the copy that used the slot NSTAGE uploads ago. */
static uint8_t *stage_acquire(MopHipDevice *d, size_t bytes, int *slot) {
  if (Plan *pl = t_plan) {
    /* staged into the plan's upload blob (256-byte aligned regions) */
    size_t off = (pl->blob.size() + 255) & ~(size_t)255;
    pl->blob.resize(off + bytes);
    *slot = -1;
    return pl->blob.data() + off;
  }
  int k = d->ring_next;
  d->ring_next = (k + 1) % NSTAGE;
  StageSlot &sl = d->ring[k];
  if (sl.used)
    HIPCHK(hipEventSynchronize(sl.ev));
  sl.used = false;
  if (bytes > sl.cap) {
    if (sl.p)
      (void)hipHostFree(sl.p);
    size_t nc = bytes + bytes / 2 + 4096;
    if (hipHostMalloc(&sl.p, nc, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&sl.dp, sl.p, 0) != hipSuccess) {
      if (sl.p)
        (void)hipHostFree(sl.p);
      sl.p = nullptr;
      sl.dp = nullptr;
      sl.cap = 0;
      set_err("hipHostMalloc(%zu) failed", nc);
      return nullptr;
    }
    sl.cap = nc;
  }
  *slot = k;
  return (uint8_t *)sl.p;
}

static void stage_release(MopHipDevice *d, int k) {
  if (k < 0)
    return;
  HIPCHK(hipEventRecord(d->ring[k].ev, d->stream));
  d->ring[k].used = true;
}

/* Host-to-device copy of staged bytes (from stage_acquire), or its record
 * into the active plan (source = an offset into the plan's upload blob). */
static void upload(MopHipDevice *d, void *dst, const uint8_t *stg, size_t n) {
  if (Plan *pl = t_plan) {
    size_t off = (size_t)(stg - pl->blob.data());
    hipStream_t s = d->stream; /* the main or the side stream */
    sig_put(pl->sig, (uintptr_t)dst);
    sig_put(pl->sig, n);
    sig_put(pl->sig, (uintptr_t)s);
    pl->ops.push_back([=](const uint8_t *base) { launch_upload(base + off, dst, n, s); });
    pl->ostream.push_back(s);
    pl->marker.push_back(0);
    return;
  }
  for (int k = 0; k < NSTAGE; k++) {
    const uint8_t *p = (const uint8_t *)d->ring[k].p;
    if (p && stg >= p && stg + n <= p + d->ring[k].cap) {
      launch_upload(d->ring[k].dp + (stg - p), dst, n, d->stream);
      return;
    }
  }
  set_err("upload source is not a staging slot");
}

static void tm_collect(MopHipDevice *d);

static hipEvent_t tm_event(MopHipDevice *d) {
  StageTimers &t = d->tm;
  if (t.next == t.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess)
      return nullptr;
    t.pool.push_back(e);
  }
  return t.pool[t.next++];
}

/* Stage timer: events recorded on the stream around the stage's launches,
 * or -- for a stage of one kernel launched directly (one_kernel) -- the
 * timestamps of that kernel's own dispatch (hipExtLaunchKernelGGL), which
 * leave out the command processor's time between an event and a dispatch. */
struct TmScope {
  MopHipDevice *d;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  bool exact = false;
  TmScope(MopHipDevice *dev, int which, bool one_kernel = false) : d(dev), id(which) {
    if (!d->tm.on || !(a = tm_event(d)))
      return;
    if (one_kernel && !t_plan && (b = tm_event(d))) {
      exact = true;
      t_kernel_ev[0] = a;
      t_kernel_ev[1] = b;
    } else {
      (void)hipEventRecord(a, d->stream);
    }
  }
  ~TmScope() {
    if (!a)
      return;
    if (exact) {
      if (t_kernel_ev[0]) { /* nothing was launched: an empty interval */
        t_kernel_ev[0] = t_kernel_ev[1] = nullptr;
        (void)hipEventRecord(a, d->stream);
        (void)hipEventRecord(b, d->stream);
      }
    } else {
      b = tm_event(d);
      if (!b)
        return;
      (void)hipEventRecord(b, d->stream);
    }
    d->tm.pending.push_back({id, {a, b}});
    if (d->tm.pending.size() >= 1024)
      tm_collect(d);
  }
};

static void tm_collect(MopHipDevice *d) {
  StageTimers &t = d->tm;
  if (t.pending.empty())
    return;
  (void)hipStreamSynchronize(d->stream);
  for (auto &p : t.pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      t.total_ms[p.first] += ms;
      t.count[p.first]++;
    }
  }
  t.pending.clear();
  t.next = 0;
}

/* Enable/disable and reset the stage timers. */
extern "C" void mop_hip_timers_enable(MopHipDevice *d, int on) {
  if (!d)
    return;
  tm_collect(d);
  d->tm.on = on != 0;
  for (int i = 0; i < TM_N; i++) {
    d->tm.total_ms[i] = 0.0;
    d->tm.count[i] = 0;
  }
}

/* Total milliseconds and number of timed instances of stage `name`
 * (vertex, setup, raster, shade, resolve, post, shadow); -1 if unknown. */
extern "C" int mop_hip_timers_read(MopHipDevice *d, const char *name, double *total_ms,
                                   uint64_t *count) {
  if (!d || !name)
    return -1;
  tm_collect(d);
  for (int i = 0; i < TM_N; i++)
    if (strcmp(name, k_timer_names[i]) == 0) {
      if (total_ms)
        *total_ms = d->tm.total_ms[i];
      if (count)
        *count = d->tm.count[i];
      return 0;
    }
  return -1;
}

extern "C" MopHipBuffer *mop_hip_buffer_create(MopHipDevice *d, const MopHipBufferDesc *desc) {
  if (!d || !desc)
    return nullptr;
  MopHipBuffer *b = new MopHipBuffer();
  b->size = desc->size;
  b->host = malloc(desc->size ? desc->size : 1);
  if (!b->host) {
    delete b;
    return nullptr;
  }
  if (desc->data && desc->size)
    memcpy(b->host, desc->data, desc->size);
  if (hipMalloc(&b->dev, desc->size ? desc->size : 4) != hipSuccess) {
    set_err("hipMalloc(%zu) for buffer failed", desc->size);
    free(b->host);
    delete b;
    return nullptr;
  }
  if (desc->size)
    HIPCHK(hipMemcpy(b->dev, b->host, desc->size, hipMemcpyHostToDevice));
  b->owner = d;
  buf_touched(b);
  std::lock_guard<std::mutex> lk(d->mu);
  d->host2buf[b->host] = b;
  return b;
}

extern "C" void mop_hip_buffer_destroy(MopHipDevice *d, MopHipBuffer *b) {
  if (!b)
    return;
  if (d) {
    (void)hipStreamSynchronize(d->stream);
    std::lock_guard<std::mutex> lk(d->mu);
    d->host2buf.erase(b->host);
    for (auto it = d->cboxes.begin(); it != d->cboxes.end();) {
      if (it->first.first == b || it->first.second == b) {
        it->second.box.release();
        it->second.vref.release();
        it = d->cboxes.erase(it);
      } else {
        ++it;
      }
    }
  }
  (void)hipFree(b->dev);
  if (b->bind_dev)
    (void)hipFree(b->bind_dev);
  free(b->host);
  delete b;
}

extern "C" void mop_hip_buffer_update(MopHipDevice *d, MopHipBuffer *b, const void *data,
                                      size_t offset, size_t size) {
  if (!b || !data || offset + size > b->size)
    return;
  if (d)
    (void)hipStreamSynchronize(d->stream);
  memcpy((uint8_t *)b->host + offset, data, size);
  HIPCHK(hipMemcpy((uint8_t *)b->dev + offset, (uint8_t *)b->host + offset, size,
                   hipMemcpyHostToDevice));
  buf_touched(b);
}

/* A buffer whose contents the device wrote (a mesh decoded in HBM): the
 * host copy is fetched on the first buffer_read. */
MopHipBuffer *buffer_adopt_device(MopHipDevice *d, void *dev, size_t size) {
  MopHipBuffer *b = new MopHipBuffer();
  b->size = size;
  b->host = malloc(size ? size : 1);
  if (!b->host) {
    delete b;
    return nullptr;
  }
  b->dev = dev;
  b->owner = d;
  b->host_stale = true;
  buf_touched(b);
  std::lock_guard<std::mutex> lk(d->mu);
  d->host2buf[b->host] = b;
  return b;
}

void *buffer_device_ptr(MopHipBuffer *b) { return b ? b->dev : nullptr; }
hipStream_t device_stream(MopHipDevice *d) { return d->stream; }

void runtime_set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
  fprintf(stderr, "[mop-hip] %s\n", t_err);
}

/* Host-visible bytes (the shadow pass and skinning read them,
 * viewport.c:2426-2429, 1539); refreshed from the device after skinning. */
extern "C" const void *mop_hip_buffer_read(MopHipBuffer *b) {
  if (!b)
    return nullptr;
  if (b->host_stale && b->owner) {
    HIPCHK(hipStreamSynchronize(b->owner->stream));
    HIPCHK(hipMemcpy(b->host, b->dev, b->size, hipMemcpyDeviceToHost));
    b->host_stale = false;
  }
  return b->host;
}

/* ------------------------------------------------------------------------ */
/* skinning and morph targets                                                */
/* ------------------------------------------------------------------------ */

static const MopHipVertexAttrib *find_attr(const MopHipVertexFormat *f, int sem);

/* the bind pose of vb on the device (uploaded once per bind-pose pointer) */
static bool bind_pose_on_device(MopHipDevice *d, MopHipBuffer *vb, const void *bind,
                                size_t bytes) {
  if (vb->bind_dev && vb->bind_src == bind && vb->bind_size == bytes)
    return true;
  HIPCHK(hipStreamSynchronize(d->stream));
  if (vb->bind_dev)
    (void)hipFree(vb->bind_dev);
  vb->bind_dev = nullptr;
  if (hipMalloc(&vb->bind_dev, bytes ? bytes : 4) != hipSuccess) {
    set_err("hipMalloc(%zu) for a bind pose failed", bytes);
    return false;
  }
  HIPCHK(hipMemcpy(vb->bind_dev, bind, bytes, hipMemcpyHostToDevice));
  vb->bind_src = bind;
  vb->bind_size = bytes;
  return true;
}

extern "C" int mop_hip_skin_apply(MopHipDevice *d, MopHipBuffer *vb, const void *bind_pose,
                                  uint32_t vertex_count, const MopHipVertexFormat *fmt,
                                  const MopHipMat4 *bones, uint32_t bone_count) {
  if (!d || !vb || !bind_pose || !fmt || !bones || bone_count == 0) {
    set_err("skin_apply: missing argument");
    return -1;
  }
  const MopHipVertexAttrib *pa = find_attr(fmt, MOP_HIP_ATTRIB_POSITION);
  const MopHipVertexAttrib *na = find_attr(fmt, MOP_HIP_ATTRIB_NORMAL);
  const MopHipVertexAttrib *ja = find_attr(fmt, MOP_HIP_ATTRIB_JOINTS);
  const MopHipVertexAttrib *wa = find_attr(fmt, MOP_HIP_ATTRIB_WEIGHTS);
  if (!pa || !ja || !wa) {
    set_err("skin_apply: format lacks POSITION, JOINTS or WEIGHTS");
    return -1;
  }
  size_t stride = fmt->stride, bytes = (size_t)vertex_count * stride;
  if (stride == 0 || bytes > vb->size || pa->offset + 12 > stride || ja->offset + 4 > stride ||
      wa->offset + 16 > stride || (na && na->offset + 12 > stride)) {
    set_err("skin_apply: vertex format does not fit the buffer");
    return -1;
  }
  if (!bind_pose_on_device(d, vb, bind_pose, bytes) || !d->skin_bones.ensure((size_t)bone_count * 16))
    return -1;
  HIPCHK(hipMemcpyAsync(d->skin_bones.p, bones, (size_t)bone_count * sizeof(MopHipMat4),
                        hipMemcpyHostToDevice, d->stream));
  SkinArgs a;
  a.bind = (const uint8_t *)vb->bind_dev;
  a.out = (uint8_t *)vb->dev;
  a.bones = d->skin_bones.p;
  a.vertex_count = vertex_count;
  a.stride = (uint32_t)stride;
  a.bone_count = bone_count;
  a.pos_off = (int32_t)pa->offset;
  a.normal_off = na ? (int32_t)na->offset : -1;
  a.joints_off = (int32_t)ja->offset;
  a.weights_off = (int32_t)wa->offset;
  a.aligned = (stride % 4) == 0;
  launch_skin(a, d->stream);
  vb->host_stale = true;
  buf_touched(vb);
  return 0;
}

extern "C" int mop_hip_morph_apply(MopHipDevice *d, MopHipBuffer *vb, const void *bind_pose,
                                   uint32_t vertex_count, const MopHipVertexFormat *fmt,
                                   const float *targets, const float *weights,
                                   uint32_t target_count) {
  if (!d || !vb || !bind_pose || !targets || !weights || target_count == 0) {
    set_err("morph_apply: missing argument");
    return -1;
  }
  size_t stride = fmt ? fmt->stride : sizeof(MopHipVertex);
  int32_t pos_off = 0; /* position is at offset 0 in MopVertex */
  if (fmt) {
    const MopHipVertexAttrib *pa = find_attr(fmt, MOP_HIP_ATTRIB_POSITION);
    if (!pa) {
      set_err("morph_apply: format lacks POSITION");
      return -1;
    }
    pos_off = (int32_t)pa->offset;
  }
  size_t bytes = (size_t)vertex_count * stride;
  size_t tcount = (size_t)target_count * vertex_count * 3;
  if (stride == 0 || bytes > vb->size || (size_t)pos_off + 12 > stride) {
    set_err("morph_apply: vertex format does not fit the buffer");
    return -1;
  }
  if (!bind_pose_on_device(d, vb, bind_pose, bytes) || !d->morph_targets.ensure(tcount) ||
      !d->morph_weights.ensure(target_count))
    return -1;
  HIPCHK(hipMemcpyAsync(d->morph_targets.p, targets, tcount * sizeof(float),
                        hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipMemcpyAsync(d->morph_weights.p, weights, (size_t)target_count * sizeof(float),
                        hipMemcpyHostToDevice, d->stream));
  MorphArgs a;
  a.bind = (const uint8_t *)vb->bind_dev;
  a.out = (uint8_t *)vb->dev;
  a.targets = d->morph_targets.p;
  a.weights = d->morph_weights.p;
  a.vertex_count = vertex_count;
  a.stride = (uint32_t)stride;
  a.target_count = target_count;
  a.pos_off = pos_off;
  a.aligned = (stride % 4) == 0;
  launch_morph(a, d->stream);
  vb->host_stale = true;
  buf_touched(vb);
  return 0;
}

extern "C" MopHipTexture *mop_hip_texture_create(MopHipDevice *d, int w, int h,
                                                 const uint8_t *rgba) {
  if (!d || w <= 0 || h <= 0 || !rgba)
    return nullptr;
  MopHipTexture *t = new MopHipTexture();
  t->w = w;
  t->h = h;
  size_t n = (size_t)w * h * 4;
  t->host = malloc(n);
  memcpy(t->host, rgba, n);
  if (hipMalloc(&t->dev, n) != hipSuccess) {
    free(t->host);
    delete t;
    return nullptr;
  }
  HIPCHK(hipMemcpy(t->dev, rgba, n, hipMemcpyHostToDevice));
  return t;
}

extern "C" MopHipTexture *mop_hip_texture_create_hdr(MopHipDevice *d, int w, int h,
                                                     const float *rgba) {
  if (!d || w <= 0 || h <= 0 || !rgba)
    return nullptr;
  MopHipTexture *t = new MopHipTexture();
  t->w = w;
  t->h = h;
  t->hdr = true;
  size_t n = (size_t)w * h * 4 * sizeof(float);
  t->host = malloc(n);
  memcpy(t->host, rgba, n);
  return t;
}

/* Device copy of an RGBA32F texture (environment map), made on first use. */
static const float *texture_device_f32(MopHipDevice *d, MopHipTexture *t) {
  if (!t || !t->hdr || !t->host)
    return nullptr;
  if (!t->dev) {
    size_t n = (size_t)t->w * t->h * 4 * sizeof(float);
    if (hipMalloc(&t->dev, n) != hipSuccess) {
      t->dev = nullptr;
      set_err("environment map upload failed");
      return nullptr;
    }
    HIPCHK(hipMemcpy(t->dev, t->host, n, hipMemcpyHostToDevice));
  }
  return (const float *)t->dev;
}

/* pass_background_hdri_cpu (viewport.c:2128-2180) for this backend: the
 * environment as the frame's background, from the camera the CPU path uses
 * (vertical fov in radians, eye, target, up).  The camera basis and
 * tan(fov/2) are formed here in the CPU path's order (math.c:30-50); the
 * device maps each pixel's ray as it does, so the background is the CPU
 * backend's bit for bit.  Recorded like a draw: it lands after the frame's
 * clear and before its draws. */
extern "C" void mop_hip_draw_background_hdri(MopHipDevice *d, MopHipFramebuffer *fb,
                                             MopHipTexture *env, float fov_radians,
                                             MopHipVec3 eye, MopHipVec3 target, MopHipVec3 up,
                                             float rotation, float intensity) {
  if (!d || !fb)
    return;
  const float *ed = texture_device_f32(d, env);
  if (!ed)
    return;
  auto sub = [](MopHipVec3 a, MopHipVec3 b) { return MopHipVec3{a.x - b.x, a.y - b.y, a.z - b.z}; };
  auto cross = [](MopHipVec3 a, MopHipVec3 b) {
    return MopHipVec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  };
  auto norm = [](MopHipVec3 v) {
    float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (len < 1e-8f)
      return MopHipVec3{0.0f, 0.0f, 0.0f};
    float inv = 1.0f / len;
    return MopHipVec3{v.x * inv, v.y * inv, v.z * inv};
  };
  MopHipVec3 fwd = norm(sub(target, eye));
  MopHipVec3 right = norm(cross(fwd, up));
  MopHipVec3 upv = cross(right, fwd);
  float aspect = (float)fb->w / (float)fb->h;
  float half_h = tanf(fov_radians * 0.5f);
  float half_w = half_h * aspect;
  BackgroundGPU &bg = fb->bg;
  memset(&bg, 0, sizeof(bg));
  bg.env = ed;
  bg.env_w = env->w;
  bg.env_h = env->h;
  bg.mode = 1;
  bg.fwd[0] = fwd.x; bg.fwd[1] = fwd.y; bg.fwd[2] = fwd.z;
  bg.right[0] = right.x; bg.right[1] = right.y; bg.right[2] = right.z;
  bg.up[0] = upv.x; bg.up[1] = upv.y; bg.up[2] = upv.z;
  bg.half_w = half_w;
  bg.half_h = half_h;
  bg.rotation = rotation;
  bg.intensity = intensity;
  fb->has_bg = true;
  fb->executed = false;
}

/* The RHI's draw_skybox slot (rhi.h:291-293), which the core calls for
 * backends other than the CPU one (viewport.c:2209-2219): the background
 * from the inverse view-projection and camera position. */
extern "C" void mop_hip_draw_skybox(MopHipDevice *d, MopHipFramebuffer *fb, MopHipTexture *env,
                                    const float *inv_vp, const float *cam_pos, float rotation,
                                    float intensity) {
  if (!d || !fb || !inv_vp || !cam_pos)
    return;
  const float *ed = texture_device_f32(d, env);
  if (!ed)
    return;
  BackgroundGPU &bg = fb->bg;
  memset(&bg, 0, sizeof(bg));
  bg.env = ed;
  bg.env_w = env->w;
  bg.env_h = env->h;
  bg.mode = 2;
  memcpy(bg.inv_vp, inv_vp, sizeof(bg.inv_vp));
  bg.cam[0] = cam_pos[0];
  bg.cam[1] = cam_pos[1];
  bg.cam[2] = cam_pos[2];
  bg.rotation = rotation;
  bg.intensity = intensity;
  fb->has_bg = true;
  fb->executed = false;
}

extern "C" void mop_hip_texture_destroy(MopHipDevice *d, MopHipTexture *t) {
  if (!t)
    return;
  if (d)
    (void)hipStreamSynchronize(d->stream);
  if (t->dev)
    (void)hipFree(t->dev);
  free(t->host);
  delete t;
}

extern "C" bool mop_hip_texture_read_rgba8(MopHipDevice *d, MopHipTexture *t, uint8_t *out,
                                           size_t size) {
  (void)d;
  if (!t || t->hdr || !out || size < (size_t)t->w * t->h * 4)
    return false;
  memcpy(out, t->host, (size_t)t->w * t->h * 4);
  return true;
}

/* ------------------------------------------------------------------------ */
/* framebuffers                                                              */
/* ------------------------------------------------------------------------ */

static void fb_free(MopHipFramebuffer *fb) {
  /* RGBA8, depth and id share one block on each side (color[0] / h_color
   * head it) so a readback of the three is one copy */
  if (fb->color[0]) (void)hipFree(fb->color[0]);
  if (fb->hdr) (void)hipFree(fb->hdr);
  if (fb->vis) (void)hipFree(fb->vis);
  if (fb->wbuf) (void)hipFree(fb->wbuf);
  if (fb->h_color) (void)hipHostFree(fb->h_color);
  if (fb->h_hdr) (void)hipHostFree(fb->h_hdr);
  if (fb->ds_dev) (void)hipFree(fb->ds_dev);
  if (fb->ds_host) (void)hipHostFree(fb->ds_host);
  fb->ds_dev = fb->ds_host = nullptr;
  fb->ds_cap = 0;
  fb->ds_seq = ~0ull;
  if (fb->h_cnt) (void)hipHostFree(fb->h_cnt);
  if (fb->done) (void)hipEventDestroy(fb->done);
  if (fb->rb_done) (void)hipEventDestroy(fb->rb_done);
  fb->rb_done = nullptr;
  fb->rb_pending = false;
  fb->h_cnt = fb->d_cnt = nullptr;
  fb->done = nullptr;
  fb->color[0] = fb->color[1] = nullptr;
  fb->hdr = fb->depth = nullptr;
  fb->ids = nullptr;
  fb->vis = nullptr;
  fb->wbuf = nullptr;
  fb->h_color = nullptr;
  fb->h_depth = fb->h_hdr = nullptr;
  fb->h_ids = nullptr;
}

static bool fb_alloc(MopHipDevice *d, MopHipFramebuffer *fb, int w, int h) {
  if (w <= 0 || h <= 0 || w > 16384 || h > 16384) {
    set_err("framebuffer dimensions out of range: %dx%d", w, h);
    return false;
  }
  size_t n = (size_t)w * h;
  fb->w = w;
  fb->h = h;
  /* [RGBA8 | depth | id], 12 B/px, on the device and in pinned host memory */
  bool ok = hipMalloc(&fb->color[0], n * 12) == hipSuccess &&
            hipMalloc(&fb->hdr, n * 16) == hipSuccess &&
            hipMalloc(&fb->vis, n * 8) == hipSuccess &&
            hipMalloc(&fb->wbuf, n * 16) == hipSuccess &&
            hipHostMalloc((void **)&fb->h_color, n * 12, 0) == hipSuccess &&
            hipHostMalloc((void **)&fb->h_hdr, n * 16, 0) == hipSuccess &&
            hipHostMalloc((void **)&fb->h_cnt, MOPH_MAX_SEGS * CNT_N * 4,
                          hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess &&
            hipHostGetDevicePointer((void **)&fb->d_cnt, fb->h_cnt, 0) == hipSuccess &&
            hipEventCreateWithFlags(&fb->done, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&fb->rb_done, hipEventDisableTiming) == hipSuccess;
  if (ok) {
    fb->depth = (float *)(fb->color[0] + n * 4);
    fb->ids = (uint32_t *)(fb->color[0] + n * 8);
    fb->h_depth = (float *)(fb->h_color + n * 4);
    fb->h_ids = (uint32_t *)(fb->h_color + n * 8);
  }
  if (!ok) {
    set_err("framebuffer allocation %dx%d failed", w, h);
    fb_free(fb);
    return false;
  }
  /* calloc'd like mop_sw_framebuffer_alloc: colour/HDR/ids zero */
  float zero[4] = {0, 0, 0, 0};
  launch_clear((int)n, zero, fb->color[0], fb->hdr, fb->depth, fb->ids, fb->vis, d->stream);
  HIPCHK(hipStreamSynchronize(d->stream));
  fb->cur = 0;
  fb->v_color = fb->v_depth = fb->v_ids = fb->v_hdr = false;
  return true;
}

extern "C" MopHipFramebuffer *mop_hip_framebuffer_create(MopHipDevice *d,
                                                         const MopHipFramebufferDesc *desc) {
  if (!d || !desc)
    return nullptr;
  MopHipFramebuffer *fb = new MopHipFramebuffer();
  if (!fb_alloc(d, fb, desc->width, desc->height)) {
    delete fb;
    return nullptr;
  }
  return fb;
}

extern "C" void mop_hip_framebuffer_destroy(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!fb)
    return;
  if (d) {
    (void)hipStreamSynchronize(d->stream);
    (void)hipStreamSynchronize(d->copy);
  }
  fb_free(fb);
  delete fb;
}

extern "C" void mop_hip_framebuffer_resize(MopHipDevice *d, MopHipFramebuffer *fb, int w,
                                           int h) {
  if (!d || !fb)
    return;
  (void)hipStreamSynchronize(d->stream);
  (void)hipStreamSynchronize(d->copy);
  fb_free(fb);
  fb->draws.clear();
  fb->light_pool.clear();
  fb->states.clear();
  fb->executed = false;
  fb->checked = false;
  fb_alloc(d, fb, w, h);
}

/* ------------------------------------------------------------------------ */
/* frame recording                                                           */
/* ------------------------------------------------------------------------ */

static void invalidate_host(MopHipFramebuffer *fb) {
  fb->v_color = fb->v_depth = fb->v_ids = fb->v_hdr = false;
}

/* Before new contents go into fb's device buffers: the kernels wait (on the
 * device) for an asynchronous readback still copying the old ones, and that
 * readback no longer counts as the buffers' contents. */
static void fb_new_contents(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (fb->rb_pending)
    HIPCHK(hipStreamWaitEvent(d->stream, fb->rb_done, 0));
  fb->exec_seq++;
  invalidate_host(fb);
}

extern "C" void mop_hip_frame_begin(MopHipDevice *d, MopHipFramebuffer *fb, MopHipColor c) {
  if (!d || !fb)
    return;
  fb->has_clear = true;
  fb->clear[0] = c.r;
  fb->clear[1] = c.g;
  fb->clear[2] = c.b;
  fb->clear[3] = c.a;
  fb->draws.clear();
  fb->light_pool.clear();
  fb->states.clear();
  fb->executed = false;
  fb->checked = false;
  fb->resolve_set = false;
  fb->post_effects = 0;
  fb->has_fog = false;
  fb->has_bg = false;
  fb->outline = false;
  d->state_dirty = true;
  invalidate_host(fb);
}

static int current_state(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (d->state_dirty || fb->states.empty()) {
    fb->states.push_back(d->cur_state);
    d->state_dirty = false;
  }
  return (int)fb->states.size() - 1;
}

static const MopHipVertexAttrib *find_attr(const MopHipVertexFormat *f, int sem) {
  for (uint32_t i = 0; i < f->attrib_count && i < 12; i++)
    if (f->attribs[i].semantic == sem)
      return &f->attribs[i];
  return nullptr;
}

static void record_draw(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c,
                        const MopHipMat4 &model, const MopHipMat4 &mvp) {
  if (!c->vertex_buffer || !c->index_buffer)
    return;
  PendingDraw pd;
  memset(&pd.g, 0, sizeof(pd.g));
  DrawGPU &g = pd.g;
  memcpy(g.mvp, mvp.d, sizeof(g.mvp));
  memcpy(g.model, model.d, sizeof(g.model));
  g.vbuf = (const uint8_t *)c->vertex_buffer->dev;
  g.ibuf = (const uint32_t *)c->index_buffer->dev;
  pd.vb = c->vertex_buffer;
  pd.ib = c->index_buffer;
  if (c->texture && !c->texture->hdr && c->texture->w >= 1 && c->texture->h >= 1) {
    g.tex = (const uint8_t *)c->texture->dev;
    g.tex_w = c->texture->w;
    g.tex_h = c->texture->h;
  }
  g.vertex_count = c->vertex_count;
  /* triangles i with i+2 < index_count, i += 3 (cpu_backend.c:589) */
  g.tri_count = c->index_count / 3;
  size_t ib_elems = c->index_buffer->size / 4;
  if ((size_t)g.tri_count * 3 > ib_elems)
    g.tri_count = (uint32_t)(ib_elems / 3);
  g.stride = sizeof(MopHipVertex);
  g.nrm_off = g.col_off = g.uv_off = -1;
  if (c->vertex_format) {
    const MopHipVertexFormat *f = c->vertex_format;
    const MopHipVertexAttrib *pa = find_attr(f, MOP_HIP_ATTRIB_POSITION);
    if (!pa) {
      set_err("draw with a vertex format lacking POSITION skipped");
      return;
    }
    g.flags |= MOPH_DRAW_FLEX;
    g.stride = f->stride;
    g.pos_off = (int32_t)pa->offset;
    g.pos_fmt = pa->format;
    const MopHipVertexAttrib *na = find_attr(f, MOP_HIP_ATTRIB_NORMAL);
    const MopHipVertexAttrib *ca = find_attr(f, MOP_HIP_ATTRIB_COLOR);
    const MopHipVertexAttrib *ua = find_attr(f, MOP_HIP_ATTRIB_TEXCOORD0);
    if (na) { g.nrm_off = (int32_t)na->offset; g.nrm_fmt = na->format; }
    if (ca) { g.col_off = (int32_t)ca->offset; g.col_fmt = ca->format; }
    if (ua) { g.uv_off = (int32_t)ua->offset; g.uv_fmt = ua->format; }
  }
  /* vertices the device may touch: the draw's vertex_count, clamped to the buffer */
  size_t vmax = c->vertex_buffer->size / (g.stride ? g.stride : 1);
  if (g.vertex_count > vmax)
    g.vertex_count = (uint32_t)vmax;
  g.object_id = c->object_id;
  if (c->backface_cull)
    g.flags |= MOPH_DRAW_CULL;
  g.depth_bias = c->depth_bias;
  g.ambient = c->ambient;
  g.metallic = c->metallic;
  g.roughness = c->roughness;
  g.cam_eye[0] = c->cam_eye.x;
  g.cam_eye[1] = c->cam_eye.y;
  g.cam_eye[2] = c->cam_eye.z;
  /* lights go to the frame's pool; a draw whose lights equal the previous
   * draw's reuses its entries (the core passes the viewport's array to every
   * draw of a frame) */
  pd.light_at = (uint32_t)fb->light_pool.size();
  if (c->lights && c->light_count) {
    size_t n = c->light_count;
    if (!fb->draws.empty() && fb->draws.back().g.light_count == n &&
        memcmp(fb->light_pool.data() + fb->draws.back().light_at, c->lights,
               n * sizeof(MopHipLight)) == 0)
      pd.light_at = fb->draws.back().light_at;
    else
      fb->light_pool.insert(fb->light_pool.end(), c->lights, c->lights + n);
  }
  g.light_count = c->lights ? c->light_count : 0;
  if (c->depth_test)
    g.flags |= MOPH_DRAW_DEPTH_TEST;
  if (c->depth_write)
    g.flags |= MOPH_DRAW_DEPTH_WRITE;
  g.flags |= ((uint32_t)c->blend_mode & 3u) << MOPH_DRAW_BLEND_SHIFT;
  g.opacity = c->opacity;
  g.light_dir[0] = c->light_dir.x;
  g.light_dir[1] = c->light_dir.y;
  g.light_dir[2] = c->light_dir.z;
  /* shading model: lit draws go through mop_sw_rasterize_triangle_full, the
   * rest through mop_sw_rasterize_triangle (cpu_backend.c:608-619) */
  bool lit = c->lights && c->light_count > 0;
  bool smooth = c->shading_mode == MOP_HIP_SHADING_SMOOTH;
  int sm = lit ? (smooth ? MOPH_SM_LIT_SMOOTH : MOPH_SM_LIT_FLAT)
               : (smooth ? MOPH_SM_UNLIT_SMOOTH : MOPH_SM_UNLIT_FLAT);
  g.flags |= (uint32_t)sm << MOPH_DRAW_SM_SHIFT;
  if (c->wireframe)
    g.flags |= MOPH_DRAW_WIRE;
  /* Order-free: opaque writes under a strict depth test, so the result is
   * the (depth, submission) minimum.  The unlit smooth model blends at
   * silhouettes against the depth already drawn, so it is always ordered. */
  bool order_free = !c->wireframe && sm != MOPH_SM_UNLIT_SMOOTH && c->depth_test &&
                    c->depth_write && c->blend_mode == MOP_HIP_BLEND_OPAQUE &&
                    c->opacity >= 1.0f;
  /* An alpha/additive draw at opacity <= 0 changes no colour: every fragment
   * blends as dst*1 + src*0 and writes depth/id only above alpha 0.5
   * (rasterizer.c:1080-1083, 1453-1470, 2109-2140).  The core draws its
   * hidden grid that way; it is dropped. */
  bool invisible = (c->blend_mode == MOP_HIP_BLEND_ALPHA ||
                    c->blend_mode == MOP_HIP_BLEND_ADDITIVE) && !(c->opacity > 0.0f);
  /* 0 order-free, 3 ordered (tile path), 2 no-op */
  if (c->wireframe)
    invisible = false; /* lines ignore blend mode and opacity */
  pd.cls = order_free && !d->opaque_tiles ? 0 : (invisible ? 2 : 3);
  pd.state = current_state(d, fb);
  fb->draws.push_back(std::move(pd));
  fb->executed = false;
}

extern "C" void mop_hip_draw(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c) {
  if (!d || !fb || !c)
    return;
  record_draw(d, fb, c, c->model, c->mvp);
}

extern "C" void mop_hip_draw_instanced(MopHipDevice *d, MopHipFramebuffer *fb,
                                       const MopHipDrawCall *c, const MopHipMat4 *xf,
                                       uint32_t count) {
  if (!d || !fb || !c || !xf || count == 0)
    return;
  /* per instance: model = transform, mvp = P * (V * M) (cpu_backend.c:638-657) */
  for (uint32_t i = 0; i < count; i++) {
    MopHipMat4 vm = mat4_mul(c->view, xf[i]);
    MopHipMat4 mvp = mat4_mul(c->projection, vm);
    record_draw(d, fb, c, xf[i], mvp);
  }
}

/* ------------------------------------------------------------------------ */
/* execution                                                                 */
/* ------------------------------------------------------------------------ */

/* The largest need of scratch store k (1 clip queue, 2 checkpoints, 3 tile
 * bins) any finished batch reported, with headroom: a batch that overflowed
 * a store sizes the following ones even when its own frames are never read
 * back (and so never run again).  The checkpoint store gets half again (its
 * need follows how close big triangles come to the camera, which varies
 * most between batches of one stream), the others a quarter.  0 before any
 * report. */
static uint32_t grown_need(const MopHipDevice *d, int k) {
  uint32_t v = __atomic_load_n(d->h_sticky + k, __ATOMIC_ACQUIRE);
  if (!v)
    return 0;
  uint64_t g = (uint64_t)v + (k == 2 ? v / 2 : v / 4) + 1024;
  return g > 0xffffff00ull ? 0xffffff00u : (uint32_t)g;
}

static bool ensure_frame_scratch(MopHipDevice *d, uint32_t n_verts, uint32_t n_tris,
                                 uint32_t ndraws, uint32_t nlights, uint32_t nviews,
                                 bool fused, uint32_t &clip_cap, uint32_t &subtri_cap) {
  clip_cap = n_tris / 16;
  if (clip_cap < 4096)
    clip_cap = 4096;
  if (d->clip_cap_hint > clip_cap)
    clip_cap = d->clip_cap_hint;
  if (d->force_clip)
    clip_cap = d->force_clip;
  else if (grown_need(d, 1) > clip_cap)
    clip_cap = grown_need(d, 1);
  uint64_t sc = (uint64_t)n_tris + 10ull * clip_cap + 1;
  if (sc > 0xffffff00ull)
    sc = 0xffffff00ull;
  subtri_cap = (uint32_t)sc;
  (void)ndraws;
  (void)nlights;
  (void)nviews;
  bool ok = d->vpos.ensure(n_verts ? n_verts : 1) &&
            d->vattr.ensure((size_t)(n_verts ? n_verts : 1) * 3) &&
            d->cpos.ensure((size_t)clip_cap * MOPH_MAX_POLY) &&
            d->cattr.ensure((size_t)clip_cap * MOPH_MAX_POLY * 3) &&
            d->cnt_poly.ensure(clip_cap) && d->clip_queue.ensure(clip_cap) &&
            d->counters.ensure(CNT_N) && d->block_sums.ensure(n_tris / (256 * 16) + 2) &&
            d->big_queue.ensure(subtri_cap);
  /* the screen-triangle store and the per-triangle scan belong to the
   * ordered path (and batches too large for the fused keys) only */
  if (ok && !fused)
    ok = d->tri_counts.ensure((size_t)n_tris + 1) && d->clip_slot.ensure(n_tris ? n_tris : 1) &&
         d->subtris.ensure(subtri_cap);
  return ok;
}

static uint32_t view_tiles(int w, int h) {
  return (uint32_t)(((w + 15) / 16) * ((h + 15) / 16));
}

static void fill_view(MopHipFramebuffer *fb, const ShadeState &st, int seg, ViewGPU &v) {
  memset(&v, 0, sizeof(v));
  v.w = fb->w;
  v.h = fb->h;
  v.vis = fb->vis;
  v.wbuf = fb->wbuf;
  v.hdr = fb->hdr;
  v.depth = fb->depth;
  v.ids = fb->ids;
  v.cnt_out = fb->d_cnt + (size_t)(seg < MOPH_MAX_SEGS ? seg : MOPH_MAX_SEGS - 1) * CNT_N;
  v.shadow_on = (st.shadow_on && st.shadow_map) ? 1 : 0;
  v.shadow_map = st.shadow_map;
  v.shadow_w = st.shadow_w;
  v.shadow_h = st.shadow_h;
  memcpy(v.shadow_vp, st.shadow_vp, sizeof(v.shadow_vp));
  v.ibl_on = st.ibl_on ? 1 : 0;
  v.irr = st.irr;
  v.pf = st.pf;
  v.lut = st.lut;
  v.irr_w = st.irr_w;
  v.irr_h = st.irr_h;
  v.pf_w = st.pf_w;
  v.pf_h = st.pf_h;
  v.pf_levels = st.pf_levels;
  v.lut_size = st.lut_size;
  v.ibl_rot = st.ibl_rot;
  v.ibl_int = st.ibl_int;
  if (seg == 0 && fb->shade_clear) {
    v.shade_clear = fb->has_bg ? 2 : 1;
    memcpy(v.clear, fb->clear, sizeof(v.clear));
  }
}

struct SegJob {
  MopHipFramebuffer *fb;
  size_t b, e;
  int seg;
};

/* One order-free segment of each job's framebuffer, all in one set of
 * launches (views are independent: own visibility buffer, own state). */
static void join_shadows(MopHipDevice *d);

/* The chunk boxes a large draw is culled with (*out), computed by
 * k_chunk_bounds on the main stream when missing or stale.  *out stays null
 * when the pair's boxes were computed in this batch for other draw
 * parameters.  Under a plan a computation aborts the plan instead (the
 * frame is launched again directly). */
static bool draw_chunk_boxes(MopHipDevice *d, const PendingDraw &pd, const DrawGPU &g,
                             uint64_t batch, const float4 **out, const uint32_t **vref) {
  *out = nullptr;
  *vref = nullptr;
  ChunkBoxes *cb;
  {
    std::lock_guard<std::mutex> lk(d->mu);
    cb = &d->cboxes[std::make_pair((const MopHipBuffer *)pd.ib, (const MopHipBuffer *)pd.vb)];
  }
  uint32_t flex = (g.flags & MOPH_DRAW_FLEX) ? 1u : 0u;
  if (cb->box.p && cb->ib_ver == pd.ib->version && cb->vb_ver == pd.vb->version &&
      cb->tri_count == g.tri_count && cb->vertex_count == g.vertex_count &&
      cb->stride == g.stride && cb->flex == flex &&
      (!flex || (cb->pos_off == g.pos_off && cb->pos_fmt == g.pos_fmt))) {
    *out = cb->box.p;
    *vref = cb->vref.p;
    return true;
  }
  if (cb->batch == batch)
    return true;
  if (t_plan) {
    t_plan_abort = true;
    return false;
  }
  uint32_t nch = (g.tri_count + 255) / 256, nvwg = (g.vertex_count + 255) / 256;
  if (!cb->box.ensure((size_t)nch * 2) || !cb->vref.ensure((size_t)nvwg * (1 + VREF_CAP) + 1))
    return false;
  cb->ib_ver = pd.ib->version;
  cb->vb_ver = pd.vb->version;
  cb->tri_count = g.tri_count;
  cb->vertex_count = g.vertex_count;
  cb->stride = g.stride;
  cb->flex = flex;
  cb->pos_off = g.pos_off;
  cb->pos_fmt = g.pos_fmt;
  cb->batch = batch;
  ChunkBoundsArgs a;
  a.vbuf = g.vbuf;
  a.ibuf = g.ibuf;
  a.tri_count = g.tri_count;
  a.vertex_count = g.vertex_count;
  a.stride = g.stride;
  a.flex = flex;
  a.pos_off = g.pos_off;
  a.pos_fmt = g.pos_fmt;
  a.box = cb->box.p;
  a.vref = cb->vref.p;
  a.nvwg = nvwg;
  launch_fill_u32(nvwg, cb->vref.p, 0u, d->stream);
  launch_chunk_bounds(a, d->stream);
  *out = cb->box.p;
  *vref = cb->vref.p;
  return true;
}

static bool run_batch(MopHipDevice *d, const std::vector<SegJob> &jobs, bool ordered) {
  std::vector<DrawGPU> gd;
  std::vector<MopHipLight> lights;
  std::vector<ViewGPU> views;
  std::vector<uint4> runs; /* per draw in grid order: first workgroups, draw */
  uint32_t n_wg_vtx = 0, n_wg_tri = 0;
  std::vector<std::pair<size_t, const PendingDraw *>> cull_cand;
  uint32_t n_verts = 0, n_tris = 0, tiles = 0, stiles = 0;
  int T = (int)d->shade_tile;
  for (size_t j = 0; j < jobs.size(); j++) {
    MopHipFramebuffer *fb = jobs[j].fb;
    ViewGPU v;
    fill_view(fb, fb->states[fb->draws[jobs[j].b].state], jobs[j].seg, v);
    v.tile_base = tiles;
    tiles += view_tiles(fb->w, fb->h);
    v.stile_base = stiles;
    stiles += (uint32_t)(((fb->w + T - 1) / T) * ((fb->h + T - 1) / T));
    views.push_back(v);
    for (size_t i = jobs[j].b; i < jobs[j].e; i++) {
      PendingDraw &pd = fb->draws[i];
      DrawGPU g = pd.g;
      g.view = (uint32_t)j;
      g.vtx_base = n_verts;
      g.tri_base = n_tris;
      const MopHipLight *pl = fb->light_pool.data() + pd.light_at;
      if (i > jobs[j].b && fb->draws[i - 1].light_at == pd.light_at &&
          fb->draws[i - 1].g.light_count == g.light_count) {
        g.light_off = gd.back().light_off; /* same pool entries as the previous draw */
      } else if (!gd.empty() && gd.back().light_count == g.light_count &&
                 (g.light_count == 0 ||
                  memcmp(&lights[gd.back().light_off], pl,
                         g.light_count * sizeof(MopHipLight)) == 0)) {
        g.light_off = gd.back().light_off;
      } else {
        g.light_off = (uint32_t)lights.size();
        lights.insert(lights.end(), pl, pl + g.light_count);
      }
      if ((uint64_t)n_verts + g.vertex_count > 0xffffffffull ||
          (uint64_t)n_tris + g.tri_count > 0x7fffffffull) {
        set_err("batch exceeds 2^31 triangles; remaining draws skipped");
        break;
      }
      n_verts += g.vertex_count;
      n_tris += g.tri_count;
      g.chunk_box = nullptr;
      if (!ordered && d->chunk_cull && g.tri_count >= CULL_MIN_TRIS && pd.ib && pd.vb &&
          !pd.vb->bind_dev)
        cull_cand.push_back(std::make_pair(gd.size(), &pd));
      gd.push_back(g);
    }
  }
  /* draw-major workgroup grids; the workgroups of the chunk-culling
   * candidates come after all others */
  std::vector<char> cand(gd.size(), 0);
  for (const auto &cc : cull_cand)
    cand[cc.first] = 1;
  uint32_t direct_vtx = 0, direct_tri = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (size_t k = 0; k < gd.size(); k++) {
      if (cand[k] != pass)
        continue;
      DrawGPU &g = gd[k];
      g.vtx_wg = n_wg_vtx;
      g.tri_wg = n_wg_tri;
      runs.push_back(make_uint4(n_wg_tri, n_wg_vtx, (uint32_t)k, 0u));
      n_wg_vtx += (g.vertex_count + 255) / 256;
      n_wg_tri += (g.tri_count + 255) / 256;
    }
    if (pass == 0) {
      direct_vtx = n_wg_vtx;
      direct_tri = n_wg_tri;
    }
  }
  uint32_t clip_cap, subtri_cap;
  /* order-free batches use the fused path while (triangle << 4 | fan index)
   * fits the 32-bit sequence half of the visibility key */
  bool fused = !ordered && n_tris < (1u << 27);
  if (!ensure_frame_scratch(d, n_verts, n_tris, (uint32_t)gd.size(), (uint32_t)lights.size(),
                            (uint32_t)views.size(), fused, clip_cap, subtri_cap))
    return false;
  /* chunk culling of the large draws of an order-free fused batch; the
   * per-frame marks start all zero (k_vertex clears what k_cull set) */
  bool any_cull = false;
  if (fused && !cull_cand.empty()) {
    uint64_t batch = ++d->batch_clock;
    for (const auto &cc : cull_cand) {
      const float4 *box;
      const uint32_t *vref;
      if (!draw_chunk_boxes(d, *cc.second, gd[cc.first], batch, &box, &vref))
        return false;
      gd[cc.first].chunk_box = box;
      gd[cc.first].vref = vref;
      any_cull = any_cull || box;
    }
    if (any_cull && (!d->tri_vis.ensure((size_t)n_wg_tri + 1) ||
                     !d->tri_list.ensure((size_t)n_wg_tri + 1) ||
                     !d->vtx_list.ensure((size_t)n_wg_vtx + 1)))
      return false;
  }
  /* one upload per batch: [counters (zeroed) | draws | lights | views], each
   * 256-byte aligned */
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t oc = 0, od = al(CNT_N * sizeof(uint32_t));
  size_t bd = gd.size() * sizeof(DrawGPU), bl = lights.size() * sizeof(MopHipLight),
         bv = views.size() * sizeof(ViewGPU);
  /* the workgroup -> draw maps: small ones written here and uploaded with
   * the tables, large ones (a 12-view serving batch has 0.6 M entries)
   * written on the device from the runs (k_wg_map, one launch more) */
  bool host_maps = (size_t)n_wg_vtx + n_wg_tri <= WG_MAP_HOST_MAX;
  size_t br = host_maps ? ((size_t)n_wg_tri + n_wg_vtx) * sizeof(uint32_t)
                        : runs.size() * sizeof(uint4);
  size_t ol = od + al(bd), ov = ol + al(bl), orun = ov + al(bv), total = orun + br;
  if (!d->arena.ensure(total))
    return false;
  int slot;
  uint8_t *stg = stage_acquire(d, total, &slot);
  if (!stg)
    return false;
  memset(stg + oc, 0, CNT_N * sizeof(uint32_t));
  memcpy(stg + od, gd.data(), bd);
  if (bl)
    memcpy(stg + ol, lights.data(), bl);
  memcpy(stg + ov, views.data(), bv);
  if (br && host_maps) {
    uint32_t *m = (uint32_t *)(stg + orun);
    for (const uint4 &r : runs) {
      const DrawGPU &g = gd[r.z];
      std::fill(m + r.x, m + r.x + (g.tri_count + 255) / 256, r.z);
      std::fill(m + n_wg_tri + r.y, m + n_wg_tri + r.y + (g.vertex_count + 255) / 256, r.z);
    }
  } else if (br) {
    memcpy(stg + orun, runs.data(), br);
  }
  upload(d, d->arena.p, stg, total);
  stage_release(d, slot);

  FrameArgs fa;
  memset(&fa, 0, sizeof(fa));
  fa.draws = (const DrawGPU *)(d->arena.p + od);
  fa.views = (const ViewGPU *)(d->arena.p + ov);
  fa.ndraws = (uint32_t)gd.size();
  fa.nviews = (uint32_t)views.size();
  if (host_maps) {
    fa.wg_tri = (const uint32_t *)(d->arena.p + orun);
    fa.wg_vtx = fa.wg_tri + n_wg_tri;
  } else {
    if (!d->wg_map.ensure((size_t)n_wg_vtx + n_wg_tri + 1))
      return false;
    fa.runs = (const uint4 *)(d->arena.p + orun);
    fa.n_runs = (uint32_t)runs.size();
    fa.wg_tri = d->wg_map.p;
    fa.wg_vtx = d->wg_map.p + n_wg_tri;
  }
  fa.n_wg_vtx = n_wg_vtx;
  fa.n_wg_tri = n_wg_tri;
  fa.n_verts = n_verts;
  fa.n_tris = n_tris;
  fa.clip_cap = clip_cap;
  fa.subtri_cap = subtri_cap;
  fa.n_tiles = tiles;
  fa.n_stiles = stiles;
  fa.small_area = d->small_area;
  fa.shade_tile = d->shade_tile;
  fa.shade_waves = d->shade_waves;
  fa.raster_grid = d->raster_grid;
  fa.fused = fused ? 1u : 0u;
  if (fa.fused) {
    uint32_t chk_cap = d->chk_cap_hint > (1u << 20) ? d->chk_cap_hint : (1u << 20);
    if (grown_need(d, 2) > chk_cap)
      chk_cap = grown_need(d, 2);
    if (d->force_chk)
      chk_cap = d->force_chk;
    if (!d->vscr.ensure(n_verts ? n_verts : 1) || !d->tri_rec.ensure(n_tris ? n_tris : 1) ||
        !d->blk_counts.ensure((size_t)n_wg_tri + 1 + CLIP_RASTER_GRID) ||
        !d->chk.ensure(chk_cap) || !d->big_chk.ensure(subtri_cap) ||
        !d->big_off.ensure((size_t)n_tris + (size_t)clip_cap * 16 + 1) ||
        !d->big_map.ensure(chk_cap))
      return false;
    fa.vscr = d->vscr.p;
    fa.tri_rec = d->tri_rec.p;
    fa.blk_counts = d->blk_counts.p;
    fa.chk = d->chk.p;
    fa.chk_cap = chk_cap;
    fa.big_chk = d->big_chk.p;
    fa.big_off = d->big_off.p;
    fa.big_map = d->big_map.p;
    fa.tri_grid = fa.n_wg_tri;
    if (any_cull) {
      /* the kept workgroups are walked by fixed grids */
      fa.tri_vis = d->tri_vis.p;
      fa.tri_list = d->tri_list.p;
      fa.vtx_list = d->vtx_list.p;
      fa.tri_direct = direct_tri;
      fa.vtx_direct = direct_vtx;
      fa.tri_lgrid = std::min<uint32_t>(fa.n_wg_tri - direct_tri, d->cull_tri_grid);
      fa.vtx_lgrid = std::min<uint32_t>(fa.n_wg_vtx - direct_vtx, d->cull_vtx_grid);
      fa.tri_grid = fa.tri_direct + fa.tri_lgrid;
    }
  }
  fa.vpos = d->vpos.p;
  fa.vattr = d->vattr.p;
  fa.cpos = d->cpos.p;
  fa.cattr = d->cattr.p;
  fa.cnt_poly = d->cnt_poly.p;
  fa.tri_counts = d->tri_counts.p;
  fa.clip_queue = d->clip_queue.p;
  fa.clip_slot = d->clip_slot.p;
  fa.counters = (uint32_t *)(d->arena.p + oc);
  fa.subtris = d->subtris.p;
  fa.big_queue = d->big_queue.p;
  fa.lights = (const MopHipLight *)(d->arena.p + ol);
  fa.sticky = d->sticky.p;
  fa.sticky_out = d->d_sticky_out;

  if (!ordered)
    for (const SegJob &j : jobs)
      if (j.seg > 0 || !j.fb->has_clear)
        launch_vis_from_depth(j.fb->w * j.fb->h, j.fb->depth, j.fb->vis, d->stream);
  uint32_t bin_cap = 0;
  if (ordered) {
    uint64_t bc = 2ull * subtri_cap + 65536;
    if (d->bin_cap_hint > bc)
      bc = d->bin_cap_hint;
    if (grown_need(d, 3) > bc)
      bc = grown_need(d, 3);
    if (d->force_bins)
      bc = d->force_bins;
    if (bc > 0xffffff00ull)
      bc = 0xffffff00ull;
    bin_cap = (uint32_t)bc;
    if (!d->tile_counts.ensure((size_t)tiles + 1) || !d->tile_cursor.ensure(tiles ? tiles : 1) ||
        !d->bins.ensure(bin_cap) || !d->block_sums.ensure(tiles / (256 * 16) + 2))
      return false;
  }
  if (fa.ndraws)
    launch_wg_map(fa, d->stream);
  if (fa.ndraws) {
    {
      TmScope t(d, TM_VERTEX);
      if (fa.tri_list)
        launch_cull(fa, d->stream);
      launch_vertex(fa, d->stream);
    }
    if (fa.fused) {
      {
        TmScope t(d, TM_SETUP);
        launch_fused_setup(fa, d->stream);
      }
      {
        TmScope t(d, TM_RASTER);
        launch_raster_big(fa, d->stream);
      }
      join_shadows(d); /* outside the stage timer: the wait is not shading */
      {
        TmScope t(d, TM_SHADE, true);
        launch_shade(fa, d->stream);
      }
      launch_export_counters(fa, d->stream);
      return true;
    }
    {
      TmScope t(d, TM_SETUP);
      launch_classify(fa, d->stream);
      launch_clip(fa, d->stream);
      launch_scan_counts(fa, d->block_sums.p, d->stream);
      launch_emit(fa, d->stream);
    }
    if (ordered) {
      join_shadows(d);
      TmScope t(d, TM_SHADE);
      launch_ordered(fa, d->tile_counts.p, d->tile_cursor.p, d->bins.p, bin_cap,
                     d->block_sums.p, d->stream);
    } else {
      {
        TmScope t(d, TM_RASTER);
        launch_raster(fa, d->stream);
      }
      join_shadows(d); /* outside the stage timer: the wait is not shading */
      {
        TmScope t(d, TM_SHADE);
        launch_shade(fa, d->stream);
      }
    }
  } else if (!ordered) {
    join_shadows(d);
    launch_shade(fa, d->stream); /* clears and leaves the visibility buffers clear */
  }
  launch_export_counters(fa, d->stream);
  return true;
}

/* The frame tail of every executed frame in one launch: HDR resolve with the
 * frame's exposure, then FXAA and the per-pixel effects when a post request
 * was recorded (mop_hip_frame_request_post / the postprocess hook).  The
 * final RGBA8 lands in color[0]. */
static void run_tail(MopHipDevice *d, const std::vector<MopHipFramebuffer *> &fbs) {
  if (fbs.empty())
    return;
  std::vector<PostViewGPU> pv(fbs.size());
  size_t n_sel = 0;
  for (MopHipFramebuffer *fb : fbs)
    if (fb->outline)
      n_sel += fb->outline_sel.size();
  if (n_sel && !d->outline_ids.ensure(n_sel))
    return;
  size_t sel_off = 0;
  uint32_t tiles = 0;
  for (size_t i = 0; i < fbs.size(); i++) {
    MopHipFramebuffer *fb = fbs[i];
    PostViewGPU &v = pv[i];
    memset(&v, 0, sizeof(v));
    v.w = fb->w;
    v.h = fb->h;
    v.tile_base = tiles;
    tiles += view_tiles(fb->w, fb->h);
    v.hdr = fb->hdr;
    v.ids = fb->ids;
    v.depth = fb->depth;
    v.dst = fb->color[0];
    v.exposure = fb->resolve_set ? fb->resolve_exposure : d->exposure;
    v.effects = fb->post_effects;
    if (fb->has_fog) {
      v.has_fog = 1;
      v.fog_r = fb->fog.color.r;
      v.fog_g = fb->fog.color.g;
      v.fog_b = fb->fog.color.b;
      v.fog_near = fb->fog.near_dist;
      v.fog_far = fb->fog.far_dist;
    }
    /* without fog or vignette the colour effects map each channel value on
     * its own: a 256-entry table per (tonemap, gamma) */
    bool fog_on = (fb->post_effects & MOP_HIP_POST_FOG) && fb->has_fog;
    if (fb->post_effects && !fog_on && !(fb->post_effects & MOP_HIP_POST_VIGNETTE) &&
        d->post_lut.p) {
      int variant = ((fb->post_effects & MOP_HIP_POST_TONEMAP) ? 1 : 0) |
                    ((fb->post_effects & MOP_HIP_POST_GAMMA) ? 2 : 0);
      v.lut = d->post_lut.p + variant * 256;
    }
    if (fb->outline) {
      v.outline = 1;
      v.sel = d->outline_ids.p + sel_off;
      v.n_sel = (uint32_t)fb->outline_sel.size();
      sel_off += fb->outline_sel.size();
      v.accent = fb->outline_accent;
      v.alpha_sel = fb->outline_alpha_sel;
      v.alpha_unsel = fb->outline_alpha_unsel;
    }
    fb->cur = 0;
    fb->v_color = false;
  }
  if (!d->pviews.ensure(pv.size()))
    return;
  /* views and selection lists staged together (the second copy's source
   * 16-byte aligned, as uploads read 16 bytes per lane) */
  size_t pv_bytes = pv.size() * sizeof(PostViewGPU);
  size_t pv_pad = (pv_bytes + 15) & ~(size_t)15;
  int slot;
  uint8_t *stg = stage_acquire(d, pv_pad + n_sel * 4, &slot);
  if (!stg)
    return;
  memcpy(stg, pv.data(), pv_bytes);
  for (MopHipFramebuffer *fb : fbs)
    if (fb->outline && !fb->outline_sel.empty()) {
      memcpy(stg + pv_pad, fb->outline_sel.data(), fb->outline_sel.size() * 4);
      pv_pad += fb->outline_sel.size() * 4;
    }
  upload(d, d->pviews.p, stg, pv_bytes);
  if (n_sel)
    upload(d, d->outline_ids.p, stg + ((pv_bytes + 15) & ~(size_t)15), n_sel * 4);
  stage_release(d, slot);
  TmScope t(d, TM_POST);
  bool any_outline = false;
  for (MopHipFramebuffer *fb : fbs)
    any_outline |= fb->outline;
  launch_tail_views(d->pviews.p, (uint32_t)pv.size(), tiles, any_outline, d->stream);
}

static void free_graph_entry(GraphEntry &e) {
  for (int k = 0; k < 2; k++) {
    if (e.done[k])
      (void)hipEventSynchronize(e.done[k]);
    for (hipGraphExec_t x : e.seg[k])
      if (x)
        (void)hipGraphExecDestroy(x);
    e.seg[k].clear();
    if (e.exec[k])
      (void)hipGraphExecDestroy(e.exec[k]);
    if (e.blob[k])
      (void)hipHostFree(e.blob[k]);
    if (e.done[k])
      (void)hipEventDestroy(e.done[k]);
    e.exec[k] = nullptr;
    e.blob[k] = nullptr;
    e.dblob[k] = nullptr;
    e.done[k] = nullptr;
  }
}

/* Issue a plan's operations on the stream, uploads from a staging slot. */
static void issue_direct(MopHipDevice *d, const Plan &pl) {
  const uint8_t *base = nullptr; /* device alias of the staged blob */
  int slot = -1;
  if (!pl.blob.empty()) {
    uint8_t *stg = stage_acquire(d, pl.blob.size(), &slot);
    if (!stg)
      return;
    memcpy(stg, pl.blob.data(), pl.blob.size());
    base = d->ring[slot].dp;
  }
  for (const auto &op : pl.ops)
    op(base);
  stage_release(d, slot);
}

/* Capture the plan twice (one set of graphs per upload buffer). */
static bool capture_entry(MopHipDevice *d, const Plan &pl, GraphEntry &e) {
  e.sig = pl.sig;
  e.blob_size = pl.blob.size();
  e.seg_stream.clear();
  for (size_t i = 0; i < pl.ops.size(); i++)
    if (!pl.marker[i] && (i == 0 || pl.marker[i - 1] || pl.ostream[i] != pl.ostream[i - 1]))
      e.seg_stream.push_back(pl.ostream[i]);
  for (int k = 0; k < 2; k++) {
    if (hipHostMalloc((void **)&e.blob[k], e.blob_size + 64, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&e.dblob[k], e.blob[k], 0) != hipSuccess ||
        hipEventCreateWithFlags(&e.done[k], hipEventDisableTiming) != hipSuccess) {
      free_graph_entry(e);
      return false;
    }
    if (e.blob_size)
      memcpy(e.blob[k], pl.blob.data(), e.blob_size);
    size_t i = 0;
    while (i < pl.ops.size()) {
      if (pl.marker[i]) {
        i++;
        continue;
      }
      hipStream_t st = pl.ostream[i];
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        free_graph_entry(e);
        return false;
      }
      for (; i < pl.ops.size() && !pl.marker[i] && pl.ostream[i] == st; i++)
        pl.ops[i](e.dblob[k]);
      hipGraph_t g = nullptr;
      hipGraphExec_t x = nullptr;
      hipError_t er = hipStreamEndCapture(st, &g);
      if (er != hipSuccess || !g || hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g)
          (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        free_graph_entry(e);
        return false;
      }
      (void)hipGraphDestroy(g);
      e.seg[k].push_back(x);
    }
  }
  return true;
}

static void replay_entry(MopHipDevice *d, GraphEntry &e, const Plan &pl) {
  int k = e.next;
  e.next ^= 1;
  e.use = ++d->plan_clock;
  HIPCHK(hipEventSynchronize(e.done[k])); /* its previous replay has read the buffer */
  if (e.blob_size)
    memcpy(e.blob[k], pl.blob.data(), e.blob_size);
  size_t j = 0;
  for (size_t i = 0; i < pl.ops.size(); i++) {
    if (pl.marker[i]) {
      pl.ops[i](nullptr); /* event record on one stream, wait on the other */
    } else if (i == 0 || pl.marker[i - 1] || pl.ostream[i] != pl.ostream[i - 1]) {
      HIPCHK(hipGraphLaunch(e.seg[k][j], e.seg_stream[j]));
      j++;
    }
  }
  HIPCHK(hipEventRecord(e.done[k], d->stream));
  d->graph_replays++;
}

/* A plan seen for the second time in a row is captured; a captured plan is
 * replayed; anything else is issued directly. */
static void issue_plan(MopHipDevice *d, const Plan &pl) {
  for (GraphEntry &e : d->graphs)
    if (e.sig == pl.sig) {
      replay_entry(d, e, pl);
      return;
    }
  if (!pl.ops.empty() && pl.sig == d->last_sig) {
    GraphEntry e;
    if (capture_entry(d, pl, e)) {
      if (d->graphs.size() >= 8) {
        size_t lru = 0;
        for (size_t i = 1; i < d->graphs.size(); i++)
          if (d->graphs[i].use < d->graphs[lru].use)
            lru = i;
        free_graph_entry(d->graphs[lru]);
        d->graphs.erase(d->graphs.begin() + (long)lru);
      }
      d->graphs.push_back(std::move(e));
      d->graph_captures++;
      replay_entry(d, d->graphs.back(), pl);
      return;
    }
  }
  d->last_sig = pl.sig;
  issue_direct(d, pl);
}

static void frame_body(MopHipDevice *d, const std::vector<MopHipFramebuffer *> &todo);
static void run_shadow_job(MopHipDevice *d, MopHipShadowMap *map,
                           const std::vector<ShadowDrawHost> &casters);

static void execute_frames(MopHipDevice *d, MopHipFramebuffer *const *fbs, size_t n) {
  std::vector<MopHipFramebuffer *> todo;
  for (size_t i = 0; i < n; i++)
    if (fbs[i] && !fbs[i]->executed)
      todo.push_back(fbs[i]);
  if (todo.empty())
    return;
  for (MopHipFramebuffer *fb : todo)
    fb_new_contents(d, fb);
  HIPCHK(hipEventRecord(d->ev0, d->stream));
  bool planned = false;
  /* a lone frame is device-bound with few launches: direct launches run it
   * faster than graph replays on this runtime; a batch of frames gains
   * from the replay */
  bool use_graphs = d->graphs_mode == 1 || (d->graphs_mode == 2 && todo.size() >= 2);
  if (use_graphs && !d->tm.on) {
    Plan pl;
    t_plan = &pl;
    t_plan_abort = false;
    frame_body(d, todo);
    t_plan = nullptr;
    if (!t_plan_abort) {
      issue_plan(d, pl);
      planned = true;
    }
    t_plan_abort = false;
  }
  if (!planned)
    frame_body(d, todo); /* direct (timers on, graphs off, or a buffer had to grow) */
  d->shadow_jobs.clear();
  for (MopHipFramebuffer *fb : todo) {
    HIPCHK(hipEventRecord(fb->done, d->stream));
    fb->executed = true;
    fb->checked = false;
    invalidate_host(fb);
  }
  HIPCHK(hipEventRecord(d->ev1, d->stream));
  d->timing_pending = true;
}

/* Fork the side stream from the main one (record, wait) or join it back;
 * recorded into the active plan like any launch (a captured graph gets the
 * two branches). */
static void stream_dep(MopHipDevice *d, hipStream_t from, hipStream_t to, hipEvent_t ev) {
  auto dep = [from, to, ev]() {
    HIPCHK(hipEventRecord(ev, from));
    HIPCHK(hipStreamWaitEvent(to, ev, 0));
  };
  if (Plan *pl = t_plan) {
    sig_put(pl->sig, (uintptr_t)ev);
    pl->ops.push_back([dep](const uint8_t *) { dep(); });
    pl->ostream.push_back(to);
    pl->marker.push_back(1);
    return;
  }
  dep();
}

/* Before a launch that reads shadow maps: wait for the side stream's passes
 * (launching them first when they were deferred). */
static void join_shadows(MopHipDevice *d) {
  if (d->shadow_deferred) {
    HIPCHK(hipStreamWaitEvent(d->side, d->ev_fork, 0));
    hipStream_t main = d->stream;
    d->stream = d->side;
    for (const ShadowJob &j : d->shadow_jobs)
      run_shadow_job(d, j.map, j.casters);
    d->stream = main;
    d->shadow_deferred = false;
    d->join_pending = true;
  }
  if (!d->join_pending)
    return;
  stream_dep(d, d->side, d->stream, d->ev_join);
  d->join_pending = false;
}

/* Everything a frame execution launches, in order: queued shadow passes,
 * clears, the segments of every frame (order-free and ordered batches in
 * rounds), the tail.  Runs under a plan or directly. */
static void frame_body(MopHipDevice *d, const std::vector<MopHipFramebuffer *> &todo) {
  if (!d->shadow_jobs.empty()) {
    /* the shadow passes overlap the vertex, triangle and raster kernels of
     * the views: they run on the side stream (every launch, upload and
     * staging event of run_shadow_job goes to d->stream, swapped here) */
    if (t_plan) {
      stream_dep(d, d->stream, d->side, d->ev_fork);
      hipStream_t main = d->stream;
      d->stream = d->side;
      for (const ShadowJob &j : d->shadow_jobs)
        run_shadow_job(d, j.map, j.casters);
      d->stream = main;
      d->join_pending = true;
    } else {
      /* launched directly, the side stream forks here but its passes are
       * launched at the first join, after the views' geometry kernels: the
       * main stream's work then starts without waiting for the host to
       * launch the passes */
      HIPCHK(hipEventRecord(d->ev_fork, d->stream));
      d->shadow_deferred = true;
    }
  }
  std::vector<std::vector<std::pair<size_t, size_t>>> segs(todo.size());
  size_t rounds = 0;
  for (size_t i = 0; i < todo.size(); i++) {
    MopHipFramebuffer *fb = todo[i];
    fb->skipped = 0;
    size_t a = 0, m = fb->draws.size();
    while (a < m) {
      size_t b = a + 1;
      while (b < m && fb->draws[b].cls == fb->draws[a].cls &&
             fb->draws[b].state == fb->draws[a].state)
        b++;
      if (fb->draws[a].cls == 0 || fb->draws[a].cls == 3)
        segs[i].push_back({a, b});
      else if (fb->draws[a].cls == 1)
        fb->skipped += b - a;
      a = b;
    }
    fb->nseg = (int)segs[i].size();
    if (fb->nseg > MOPH_MAX_SEGS)
      set_err("frame has %d draw segments; counters of segments past %d are merged",
              fb->nseg, MOPH_MAX_SEGS);
    if (segs[i].size() > rounds)
      rounds = segs[i].size();
    if (fb->skipped)
      set_err("%zu draw(s) of a class this backend does not render yet were skipped",
              fb->skipped);
  }
  /* the clears of every frame in one launch, except for frames whose first
   * segment is order-free: that segment's shading pass visits every pixel,
   * writes the clear values where nothing was drawn and leaves the
   * visibility buffer cleared (it is cleared when allocated and after every
   * order-free shading pass, so it is clear at the start of every frame) */
  std::vector<ClearViewGPU> cv;
  uint32_t cpx = 0;
  for (size_t i = 0; i < todo.size(); i++) {
    MopHipFramebuffer *fb = todo[i];
    fb->shade_clear = fb->has_clear && !segs[i].empty() &&
                      fb->draws[segs[i][0].first].cls == 0;
    if (fb->has_clear && !fb->shade_clear) {
      ClearViewGPU c;
      memset(&c, 0, sizeof(c));
      c.px_base = cpx;
      cpx += (uint32_t)(fb->w * fb->h);
      memcpy(c.clear, fb->clear, sizeof(c.clear));
      c.hdr = fb->hdr;
      c.depth = fb->depth;
      c.ids = fb->ids;
      c.vis = fb->vis;
      cv.push_back(c);
    }
  }
  if (!cv.empty() && d->clear_views.ensure(cv.size())) {
    int slot;
    uint8_t *stg = stage_acquire(d, cv.size() * sizeof(ClearViewGPU), &slot);
    if (stg) {
      memcpy(stg, cv.data(), cv.size() * sizeof(ClearViewGPU));
      upload(d, d->clear_views.p, stg, cv.size() * sizeof(ClearViewGPU));
      stage_release(d, slot);
      launch_clear_views(d->clear_views.p, (uint32_t)cv.size(), cpx, d->stream);
    }
  }
  /* environment backgrounds, after the clears and before any draw (the
   * core's background pass follows frame_begin, viewport.c:3211-3219) */
  for (MopHipFramebuffer *fb : todo)
    if (fb->has_bg) {
      BackgroundGPU bg = fb->bg;
      bg.hdr = fb->hdr;
      bg.w = fb->w;
      bg.h = fb->h;
      launch_background(bg, d->stream);
    }
  for (size_t r = 0; r < rounds; r++) {
    std::vector<SegJob> jobs[2];
    for (size_t i = 0; i < todo.size(); i++)
      if (r < segs[i].size()) {
        bool ord = todo[i]->draws[segs[i][r].first].cls == 3;
        jobs[ord ? 1 : 0].push_back({todo[i], segs[i][r].first, segs[i][r].second, (int)r});
      }
    if (!jobs[0].empty())
      run_batch(d, jobs[0], false);
    if (!jobs[1].empty())
      run_batch(d, jobs[1], true);
  }
  join_shadows(d);
  run_tail(d, todo);
}

/* Wait for fb's frame, re-run it when a scratch capacity overflowed, and
 * take its statistics. */
static void check_frame(MopHipDevice *d, MopHipFramebuffer *fb) {
  for (int attempt = 0; attempt < 4; attempt++) {
    if (!fb->executed)
      execute_frames(d, &fb, 1);
    if (fb->checked)
      return;
    HIPCHK(hipEventSynchronize(fb->done));
    uint32_t ovf = 0, need_clip = 0, need_chk = 0, need_bins = 0;
    MopHipFrameStats st = {};
    int ns = fb->nseg < MOPH_MAX_SEGS ? fb->nseg : MOPH_MAX_SEGS;
    for (int k = 0; k < ns; k++) {
      const uint32_t *c = fb->h_cnt + (size_t)k * CNT_N;
      ovf |= c[CNT_OVERFLOW];
      if (c[CNT_CLIP] > need_clip)
        need_clip = c[CNT_CLIP];
      if (c[CNT_CHK] > need_chk)
        need_chk = c[CNT_CHK];
      if (c[CNT_BINS] > need_bins)
        need_bins = c[CNT_BINS];
      st.subtris += c[CNT_SUBTRI];
      st.clipped += c[CNT_CLIP];
      st.big_subtris += c[CNT_BIG];
      st.culled_chunks += c[CNT_CULLED];
    }
    if (ovf & 8u)
      set_err("a clipped polygon exceeded %d vertices and was dropped", MOPH_MAX_POLY);
    if (ovf & CNT_OVERFLOW_RERUN) {
      /* clip queue, screen-triangle store, tile bins or checkpoint store too
       * small: grow to what the batch needed (the counters kept counting past
       * the capacity), drop any forced capacity, re-run */
      if (ovf & 3u) {
        d->clip_cap_hint = need_clip + need_clip / 4 + 1024;
        d->force_clip = 0;
      }
      if (ovf & 32u) {
        uint64_t c = (uint64_t)need_chk + need_chk / 4 + 65536;
        d->chk_cap_hint = c > 0xffffff00ull ? 0xffffff00u : (uint32_t)c;
        d->force_chk = 0;
      }
      if (ovf & 16u) {
        uint64_t c = (uint64_t)need_bins + need_bins / 4 + 65536;
        if (c < (uint64_t)d->bins.cap * 2)
          c = (uint64_t)d->bins.cap * 2;
        d->bin_cap_hint = c > 0xffffff00ull ? 0xffffff00u : (uint32_t)c;
        d->force_bins = 0;
      }
      g_overflow_reruns++;
      fb->executed = false;
      continue;
    }
    fb->checked = true;
    st.triangles_in = 0;
    for (const PendingDraw &pd : fb->draws)
      if (pd.cls == 0 || pd.cls == 3)
        st.triangles_in += pd.g.tri_count;
    st.shadow_tris = d->stats.shadow_tris;
    st.gpu_ms = d->stats.gpu_ms;
    d->stats = st;
    return;
  }
  set_err("frame scratch overflow persisted; frame left incomplete");
  fb->checked = true;
}

extern "C" void mop_hip_frame_end(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!d || !fb)
    return;
  execute_frames(d, &fb, 1);
}

extern "C" void mop_hip_frame_submit(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!d || !fb)
    return;
  execute_frames(d, &fb, 1);
}

extern "C" void mop_hip_frame_wait_readback(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!d || !fb)
    return;
  check_frame(d, fb);
}

/* Execute the recorded frames of many framebuffers as one batch (the
 * batched form: one set of launches for all of them). */
extern "C" void mop_hip_frames_execute(MopHipDevice *d, MopHipFramebuffer *const *fbs,
                                       size_t n) {
  if (!d || !fbs)
    return;
  execute_frames(d, fbs, n);
}

/* Launch-shape tunables of the frame kernels: bbox area up to which one lane
 * walks a screen triangle, shading tile edge (8, 16 or 32 pixels) and the
 * number of waves of the large-triangle raster (0 = 8192).  Results do not
 * depend on them; speed does. */
extern "C" int mop_hip_set_shade_waves(MopHipDevice *d, uint32_t waves) {
  if (!d || (waves != 4 && waves != 5 && waves != 6 && waves != 8))
    return -1;
  d->shade_waves = waves;
  return 0;
}

extern "C" int mop_hip_set_tuning(MopHipDevice *d, uint32_t small_area, uint32_t shade_tile,
                                  uint32_t raster_grid) {
  if (!d || (shade_tile != 8 && shade_tile != 16 && shade_tile != 32) || small_area < 1)
    return -1;
  d->small_area = small_area;
  d->shade_tile = shade_tile;
  d->raster_grid = raster_grid;
  return 0;
}

/* Raster path of the opaque depth-tested draws recorded from now on: 0 (the
 * default) the visibility buffer's (depth, submission) minimum, 1 the tile
 * path of the blended draws -- screen triangles binned to 16x16 tiles by
 * prefix scan in submission order, each tile's colour, depth and id held in
 * LDS and its list walked in order with ballot coverage skips.  Both give
 * the CPU backend's frame; the first is faster on the workloads here. */
extern "C" int mop_hip_set_opaque_raster(MopHipDevice *d, uint32_t tiled) {
  if (!d || tiled > 1)
    return -1;
  d->opaque_tiles = tiled;
  return 0;
}

/* 1 when fb's frame has executed on the device (non-blocking), else 0. */
extern "C" int mop_hip_frame_done(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!d || !fb || !fb->executed)
    return 0;
  return hipEventQuery(fb->done) == hipSuccess ? 1 : 0;
}

/* 1 when fb's frame has executed and the copy a
 * mop_hip_framebuffer_readback_async call started for it has reached host
 * memory; with no such copy pending, mop_hip_frame_done. */
extern "C" int mop_hip_readback_done(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!d || !fb || !fb->executed)
    return 0;
  if (fb->rb_pending && fb->rb_seq == fb->exec_seq)
    return hipEventQuery(fb->rb_done) == hipSuccess ? 1 : 0;
  return hipEventQuery(fb->done) == hipSuccess ? 1 : 0;
}

/* Wait for all work queued on the device. */
extern "C" void mop_hip_device_sync(MopHipDevice *d) {
  if (d)
    HIPCHK(hipStreamSynchronize(d->stream));
}

/* OR of the overflow bits of every batch executed on the device since the
 * last reset (waits for the queued work); reset != 0 clears it. */
extern "C" uint32_t mop_hip_overflow_flags(MopHipDevice *d, int reset) {
  if (!d)
    return 0;
  HIPCHK(hipStreamSynchronize(d->side));
  HIPCHK(hipStreamSynchronize(d->stream));
  uint32_t v = __atomic_load_n(d->h_sticky, __ATOMIC_ACQUIRE);
  if (reset) { /* the flags only; the needs keep sizing later batches */
    launch_fill_u32(1, d->sticky.p, 0u, d->stream);
    HIPCHK(hipStreamSynchronize(d->stream));
    __atomic_store_n(d->h_sticky, 0u, __ATOMIC_RELEASE);
  }
  return v;
}

/* Forced scratch capacities for the next executions on d (0 = none): the
 * clip queue (triangles), the big-triangle checkpoint store (entries) and
 * the ordered path's tile bins (entries).  A frame that overflows one is
 * grown and run again at readback, as any overflow is; the forced value is
 * then dropped.  Exists to exercise that recovery. */
extern "C" int mop_hip_set_scratch_limits(MopHipDevice *d, uint32_t clip, uint32_t chk,
                                          uint32_t bins) {
  if (!d)
    return -1;
  d->force_clip = clip;
  d->force_chk = chk;
  d->force_bins = bins;
  return 0;
}

/* The same for every device created afterwards in this process (a device
 * created by a caller that does not expose it, such as the viewport core). */
extern "C" void mop_hip_set_default_scratch_limits(uint32_t clip, uint32_t chk, uint32_t bins) {
  g_force_clip = clip;
  g_force_chk = chk;
  g_force_bins = bins;
}

/* Frames run again after a scratch overflow, in this process. */
extern "C" uint64_t mop_hip_overflow_reruns(void) { return g_overflow_reruns.load(); }

extern "C" void mop_hip_set_exposure(MopHipDevice *d, float e) {
  if (d)
    d->exposure = e;
}

extern "C" float mop_hip_frame_gpu_time_ms(MopHipDevice *d) {
  if (!d)
    return 0.0f;
  if (d->timing_pending) {
    HIPCHK(hipEventSynchronize(d->ev1));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, d->ev0, d->ev1) == hipSuccess)
      d->stats.gpu_ms = ms;
    d->timing_pending = false;
  }
  return d->stats.gpu_ms;
}

extern "C" void mop_hip_get_frame_stats(MopHipDevice *d, MopHipFrameStats *out) {
  if (!d || !out)
    return;
  mop_hip_frame_gpu_time_ms(d);
  *out = d->stats;
}

/* ------------------------------------------------------------------------ */
/* readbacks                                                                 */
/* ------------------------------------------------------------------------ */

static void ensure_ready(MopHipDevice *d, MopHipFramebuffer *fb) { check_frame(d, fb); }

/* True when an asynchronous readback of the current contents copied `bit`:
 * waits for that copy. */
static bool take_async(MopHipFramebuffer *fb, uint32_t bit) {
  if (!fb->rb_pending || !(fb->rb_mask & bit) || fb->rb_seq != fb->exec_seq)
    return false;
  HIPCHK(hipEventSynchronize(fb->rb_done));
  return true;
}

/* Start copying the current frame's buffers (mask: 1 RGBA8, 2 depth, 4
 * object ID, 8 HDR) into the framebuffer's host buffers on the copy stream,
 * after the frame finishes; the read_* calls then wait for that copy
 * instead of issuing their own.  Executes a recorded frame first. */
extern "C" void mop_hip_framebuffer_readback_async(MopHipDevice *d, MopHipFramebuffer *fb,
                                                   uint32_t mask) {
  if (!d || !fb || !(mask & 15u))
    return;
  if (!fb->executed)
    execute_frames(d, &fb, 1);
  uint32_t requested = mask & 15u;
  size_t n = (size_t)fb->w * fb->h;
  HIPCHK(hipStreamWaitEvent(d->copy, fb->done, 0));
  if ((mask & 7u) == 7u && fb->cur == 0) {
    /* RGBA8, depth and id are one block: one copy */
    HIPCHK(hipMemcpyAsync(fb->h_color, fb->color[0], n * 12, hipMemcpyDeviceToHost, d->copy));
    mask &= ~7u;
  }
  if (mask & 1u)
    HIPCHK(hipMemcpyAsync(fb->h_color, fb->color[fb->cur], n * 4, hipMemcpyDeviceToHost,
                          d->copy));
  if (mask & 2u)
    HIPCHK(hipMemcpyAsync(fb->h_depth, fb->depth, n * 4, hipMemcpyDeviceToHost, d->copy));
  if (mask & 4u)
    HIPCHK(hipMemcpyAsync(fb->h_ids, fb->ids, n * 4, hipMemcpyDeviceToHost, d->copy));
  if (mask & 8u)
    HIPCHK(hipMemcpyAsync(fb->h_hdr, fb->hdr, n * 16, hipMemcpyDeviceToHost, d->copy));
  HIPCHK(hipEventRecord(fb->rb_done, d->copy));
  fb->rb_pending = true;
  fb->rb_mask = requested;
  fb->rb_seq = fb->exec_seq;
}

extern "C" const uint8_t *mop_hip_framebuffer_read_color(MopHipDevice *d, MopHipFramebuffer *fb,
                                                         int *ow, int *oh) {
  if (!d || !fb)
    return nullptr;
  ensure_ready(d, fb);
  if (ow) *ow = fb->w;
  if (oh) *oh = fb->h;
  if (!fb->v_color) {
    if (!take_async(fb, 1u)) {
      HIPCHK(hipMemcpyAsync(fb->h_color, fb->color[fb->cur], (size_t)fb->w * fb->h * 4, hipMemcpyDeviceToHost,
                            d->stream));
      HIPCHK(hipStreamSynchronize(d->stream));
    }
    fb->v_color = true;
  }
  return fb->h_color;
}

extern "C" const uint32_t *mop_hip_framebuffer_read_object_id(MopHipDevice *d,
                                                              MopHipFramebuffer *fb, int *ow,
                                                              int *oh) {
  if (!d || !fb)
    return nullptr;
  ensure_ready(d, fb);
  if (ow) *ow = fb->w;
  if (oh) *oh = fb->h;
  if (!fb->v_ids) {
    if (!take_async(fb, 4u)) {
      HIPCHK(hipMemcpyAsync(fb->h_ids, fb->ids, (size_t)fb->w * fb->h * 4, hipMemcpyDeviceToHost,
                            d->stream));
      HIPCHK(hipStreamSynchronize(d->stream));
    }
    fb->v_ids = true;
  }
  return fb->h_ids;
}

extern "C" const float *mop_hip_framebuffer_read_depth(MopHipDevice *d, MopHipFramebuffer *fb,
                                                       int *ow, int *oh) {
  if (!d || !fb)
    return nullptr;
  ensure_ready(d, fb);
  if (ow) *ow = fb->w;
  if (oh) *oh = fb->h;
  if (!fb->v_depth) {
    if (!take_async(fb, 2u)) {
      HIPCHK(hipMemcpyAsync(fb->h_depth, fb->depth, (size_t)fb->w * fb->h * 4, hipMemcpyDeviceToHost,
                            d->stream));
      HIPCHK(hipStreamSynchronize(d->stream));
    }
    fb->v_depth = true;
  }
  return fb->h_depth;
}

extern "C" const float *mop_hip_framebuffer_read_hdr(MopHipDevice *d, MopHipFramebuffer *fb,
                                                     int *ow, int *oh) {
  if (!d || !fb)
    return nullptr;
  ensure_ready(d, fb);
  if (ow) *ow = fb->w;
  if (oh) *oh = fb->h;
  if (!fb->v_hdr) {
    if (!take_async(fb, 8u)) {
      HIPCHK(hipMemcpyAsync(fb->h_hdr, fb->hdr, (size_t)fb->w * fb->h * 16, hipMemcpyDeviceToHost,
                            d->stream));
      HIPCHK(hipStreamSynchronize(d->stream));
    }
    fb->v_hdr = true;
  }
  return fb->h_hdr;
}

/* The RGBA8 image box-filtered by an integer factor on the device
 * (viewport.c:3385-3411 does it on the host after framebuffer_read_color):
 * the internal w x h image to (w / factor) x (h / factor), each channel the
 * truncated integer mean of its factor x factor block.  The pointer is host
 * memory owned by the framebuffer, valid until the next frame. */
extern "C" const uint8_t *mop_hip_framebuffer_read_color_downsampled(MopHipDevice *d,
                                                                     MopHipFramebuffer *fb,
                                                                     int factor, int *ow,
                                                                     int *oh) {
  if (!d || !fb || factor < 1 || factor > 16)
    return nullptr;
  ensure_ready(d, fb);
  int pw = fb->w / factor, ph = fb->h / factor;
  if (ow) *ow = pw;
  if (oh) *oh = ph;
  if (pw < 1 || ph < 1)
    return nullptr;
  if (fb->ds_seq == fb->exec_seq && fb->ds_factor == factor)
    return fb->ds_host;
  size_t bytes = (size_t)pw * ph * 4;
  if (bytes > fb->ds_cap) {
    if (fb->ds_dev) (void)hipFree(fb->ds_dev);
    if (fb->ds_host) (void)hipHostFree(fb->ds_host);
    fb->ds_dev = fb->ds_host = nullptr;
    fb->ds_cap = 0;
    if (hipMalloc(&fb->ds_dev, bytes) != hipSuccess ||
        hipHostMalloc((void **)&fb->ds_host, bytes, 0) != hipSuccess) {
      set_err("downsample buffer allocation failed");
      return nullptr;
    }
    fb->ds_cap = bytes;
  }
  launch_downsample(fb->color[fb->cur], fb->w, fb->ds_dev, pw, ph, factor, d->stream);
  HIPCHK(hipMemcpyAsync(fb->ds_host, fb->ds_dev, bytes, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  fb->ds_seq = fb->exec_seq;
  fb->ds_factor = factor;
  return fb->ds_host;
}

extern "C" uint32_t mop_hip_pick_read_id(MopHipDevice *d, MopHipFramebuffer *fb, int x, int y) {
  if (!d || !fb || x < 0 || x >= fb->w || y < 0 || y >= fb->h)
    return 0;
  const uint32_t *ids = mop_hip_framebuffer_read_object_id(d, fb, nullptr, nullptr);
  return ids ? ids[(size_t)y * fb->w + x] : 0;
}

extern "C" float mop_hip_pick_read_depth(MopHipDevice *d, MopHipFramebuffer *fb, int x, int y) {
  if (!d || !fb || x < 0 || x >= fb->w || y < 0 || y >= fb->h)
    return 1.0f;
  const float *dp = mop_hip_framebuffer_read_depth(d, fb, nullptr, nullptr);
  return dp ? dp[(size_t)y * fb->w + x] : 1.0f;
}

/* ------------------------------------------------------------------------ */
/* hooks: shadow, IBL, resolve, post                                         */
/* ------------------------------------------------------------------------ */

static MopHipBuffer *lookup_or_upload(MopHipDevice *d, const void *host, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(d->mu);
    auto it = d->host2buf.find(host);
    if (it != d->host2buf.end() && it->second->size >= bytes)
      return it->second;
  }
  MopHipBuffer *b = new MopHipBuffer();
  b->size = bytes;
  if (hipMalloc(&b->dev, bytes ? bytes : 4) != hipSuccess) {
    delete b;
    return nullptr;
  }
  HIPCHK(hipMemcpy(b->dev, host, bytes, hipMemcpyHostToDevice));
  buf_touched(b);
  d->shadow_temp.push_back(b);
  return b;
}

extern "C" MopHipShadowMap *mop_hip_shadow_map_create(MopHipDevice *d, int size) {
  if (!d || size <= 0 || size > 16384) {
    set_err("shadow map size %d out of range", size);
    return nullptr;
  }
  MopHipShadowMap *m = new MopHipShadowMap();
  m->size = size;
  size_t n = (size_t)size * size;
  if (!m->keys.ensure(n)) {
    delete m;
    return nullptr;
  }
  return m;
}

extern "C" void mop_hip_shadow_map_destroy(MopHipDevice *d, MopHipShadowMap *m) {
  if (!m)
    return;
  if (d) {
    /* drop queued passes into this map; their frames would read it anyway */
    for (size_t i = 0; i < d->shadow_jobs.size();)
      if (d->shadow_jobs[i].map == m)
        d->shadow_jobs.erase(d->shadow_jobs.begin() + (long)i);
      else
        i++;
    (void)hipStreamSynchronize(d->stream);
    /* graphs captured with this map's pointers must not be replayed */
    for (GraphEntry &e : d->graphs)
      free_graph_entry(e);
    d->graphs.clear();
    d->last_sig.clear();
  }
  m->keys.release();
  delete m;
}

/* Frame plans replayed from captured hipGraphs: 1 always, 0 never (every
 * launch issued one by one), 2 (default) for executions of two or more
 * frames. */
extern "C" void mop_hip_set_graphs(MopHipDevice *d, int mode) {
  if (!d)
    return;
  d->graphs_mode = (mode == 0 || mode == 1) ? mode : 2;
  if (d->graphs_mode == 0) {
    for (GraphEntry &e : d->graphs)
      free_graph_entry(e);
    d->graphs.clear();
    d->last_sig.clear();
  }
}

extern "C" void mop_hip_graph_stats(MopHipDevice *d, uint64_t *replays, uint64_t *captures) {
  if (!d)
    return;
  if (replays)
    *replays = d->graph_replays;
  if (captures)
    *captures = d->graph_captures;
}

/* Begin a caster list for `map` (cleared to depth 1.0 when rendered). */
static void flush_shadow_jobs(MopHipDevice *d);

static void shadow_begin(MopHipDevice *d, MopHipShadowMap *map) {
  d->shadow_draws.clear();
  if (!d->shadow_temp.empty()) {
    flush_shadow_jobs(d); /* queued passes still read the temporaries */
    (void)hipStreamSynchronize(d->stream);
    for (MopHipBuffer *b : d->shadow_temp) {
      (void)hipFree(b->dev);
      delete b;
    }
    d->shadow_temp.clear();
  }
  d->rec_shadow = map;
}

static void shadow_add(MopHipDevice *d, const void *vdev, uint32_t vertex_count,
                       const void *idev, uint32_t index_count, const MopHipMat4 &model,
                       const MopHipMat4 &light_vp) {
  ShadowDrawHost sh;
  memset(&sh, 0, sizeof(sh));
  /* mvp = light_vp * model (rasterizer.c:361) */
  MopHipMat4 mvp = mat4_mul(light_vp, model);
  memcpy(sh.g.mvp, mvp.d, sizeof(sh.g.mvp));
  sh.g.vbuf = (const uint8_t *)vdev;
  sh.g.ibuf = (const uint32_t *)idev;
  sh.g.vertex_count = vertex_count;
  sh.g.tri_count = index_count / 3;
  d->shadow_draws.push_back(sh);
}

/* Rasterize a caster list into its map (ordered depth keys).  Shadow passes
 * are queued at shadow_end / shadow_set and issued at the start of the next
 * frame execution (in the same plan), before any draw that reads them. */
static void run_shadow_job(MopHipDevice *d, MopHipShadowMap *map,
                           const std::vector<ShadowDrawHost> &casters) {
  if (!map)
    return;
  TmScope t(d, TM_SHADOW);
  int size = map->size;
  size_t npx = (size_t)size * size;
  launch_shadow_clear((int)npx, map->keys.p, d->stream);
  uint32_t n_tris = 0, n_verts = 0;
  std::vector<ShadowDrawGPU> gd;
  for (const ShadowDrawHost &sh : casters) {
    ShadowDrawGPU g = sh.g;
    g.tri_base = n_tris;
    g.vtx_base = n_verts;
    n_tris += g.tri_count;
    n_verts += g.vertex_count;
    gd.push_back(g);
  }
  if (!gd.empty() && n_tris) {
    /* the big-triangle queue holds every caster in the worst case */
    /* one upload: [counters (zeroed) | caster draws] */
    size_t od = 256, bytes = gd.size() * sizeof(ShadowDrawGPU);
    if (!d->shadow_big.ensure(n_tris) || !d->shadow_arena.ensure(od + bytes) ||
        !d->shadow_lclip.ensure(n_verts ? n_verts : 1) ||
        !d->shadow_lscr.ensure(n_verts ? n_verts : 1))
      return;
    int slot;
    uint8_t *stg = stage_acquire(d, od + bytes, &slot);
    if (!stg)
      return;
    memset(stg, 0, CNT_N * sizeof(uint32_t));
    memcpy(stg + od, gd.data(), bytes);
    upload(d, d->shadow_arena.p, stg, od + bytes);
    stage_release(d, slot);
    ShadowArgsGPU sa;
    memset(&sa, 0, sizeof(sa)); /* padding bytes enter plan signatures */
    sa.draws = (const ShadowDrawGPU *)(d->shadow_arena.p + od);
    sa.ndraws = (uint32_t)gd.size();
    sa.n_tris = n_tris;
    sa.size = size;
    sa.big_cap = (uint32_t)d->shadow_big.cap;
    sa.keys = map->keys.p;
    sa.big_queue = d->shadow_big.p;
    sa.counters = (uint32_t *)d->shadow_arena.p;
    sa.n_verts = n_verts;
    sa.pad_ = 0;
    sa.lclip = d->shadow_lclip.p;
    sa.lscr = d->shadow_lscr.p;
    launch_shadow_raster(sa, d->stream);
    d->stats.shadow_tris = n_tris;
  }

}

static void queue_shadow(MopHipDevice *d) {
  d->shadow_jobs.push_back({d->rec_shadow, d->shadow_draws});
}

/* issue the queued shadow passes now (outside a frame) */
static void flush_shadow_jobs(MopHipDevice *d) {
  if (d->shadow_jobs.empty())
    return;
  for (const ShadowJob &j : d->shadow_jobs)
    run_shadow_job(d, j.map, j.casters);
  d->shadow_jobs.clear();
}

static void shadow_enable(MopHipDevice *d, const MopHipShadowMap *map, const MopHipMat4 &lvp) {
  d->cur_state.shadow_on = true;
  d->cur_state.shadow_map = map->keys.p;
  d->cur_state.shadow_w = map->size;
  d->cur_state.shadow_h = map->size;
  memcpy(d->cur_state.shadow_vp, lvp.d, sizeof(d->cur_state.shadow_vp));
  d->state_dirty = true;
}

extern "C" void mop_hip_shadow_render_mesh(MopHipDevice *d, const void *vertices,
                                           uint32_t vertex_count, const uint32_t *indices,
                                           uint32_t index_count, const MopHipMat4 *model,
                                           const MopHipMat4 *light_vp, int size, int first) {
  if (!d || !vertices || !indices || !model || !light_vp || size <= 0)
    return;
  if (first) {
    if (!d->def_shadow || d->def_shadow->size != size) {
      mop_hip_shadow_map_destroy(d, d->def_shadow);
      d->def_shadow = mop_hip_shadow_map_create(d, size);
      if (!d->def_shadow)
        return;
    }
    shadow_begin(d, d->def_shadow);
  }
  MopHipBuffer *vb = lookup_or_upload(d, vertices, (size_t)vertex_count * sizeof(MopHipVertex));
  MopHipBuffer *ib = lookup_or_upload(d, indices, (size_t)index_count * sizeof(uint32_t));
  if (!vb || !ib)
    return;
  shadow_add(d, vb->dev, vertex_count, ib->dev, index_count, *model, *light_vp);
}

extern "C" void mop_hip_shadow_set(MopHipDevice *d, int w, int h, const MopHipMat4 *light_vp) {
  if (!d || !light_vp || !d->rec_shadow)
    return;
  if (w != d->rec_shadow->size || h != d->rec_shadow->size) {
    set_err("shadow map %dx%d does not match the rendered %d^2 map", w, h,
            d->rec_shadow->size);
    return;
  }
  queue_shadow(d);
  shadow_enable(d, d->rec_shadow, *light_vp);
}

extern "C" void mop_hip_shadow_clear(MopHipDevice *d) {
  if (!d)
    return;
  d->cur_state.shadow_on = false;
  d->cur_state.shadow_map = nullptr;
  d->state_dirty = true;
}

/* Resident-scene variant: casters are device buffers already. */
extern "C" void mop_hip_shadow_begin(MopHipDevice *d, MopHipShadowMap *map) {
  if (d && map)
    shadow_begin(d, map);
}

extern "C" void mop_hip_shadow_draw(MopHipDevice *d, MopHipBuffer *vb, uint32_t vertex_count,
                                    MopHipBuffer *ib, uint32_t index_count,
                                    const MopHipMat4 *model, const MopHipMat4 *light_vp) {
  if (!d || !vb || !ib || !model || !light_vp || !d->rec_shadow)
    return;
  /* the kernels read every vertex and index counted here: clamp to the buffers */
  if ((size_t)vertex_count * sizeof(MopHipVertex) > vb->size)
    vertex_count = (uint32_t)(vb->size / sizeof(MopHipVertex));
  if ((size_t)index_count * sizeof(uint32_t) > ib->size)
    index_count = (uint32_t)(ib->size / sizeof(uint32_t));
  shadow_add(d, vb->dev, vertex_count, ib->dev, index_count, *model, *light_vp);
}

extern "C" void mop_hip_shadow_end(MopHipDevice *d, const MopHipMat4 *light_vp) {
  if (!d || !light_vp || !d->rec_shadow)
    return;
  queue_shadow(d);
  shadow_enable(d, d->rec_shadow, *light_vp);
}

/* Render one caster mesh into map on the device's stream with no
 * allocation: the light-space vertex cache, the big-triangle queue, the
 * counters and the caster record live in the file-scope `scratch` array
 * (through its device address noalloc_scratch), the record is written by a
 * kernel from its arguments and the counters and the map are cleared by fill
 * kernels.  The same kernels as mop_hip_shadow_begin/draw/end, so the map
 * is identical; the depth minimum does not depend on caster order. */
extern "C" int mop_hip_shadow_render_noalloc(MopHipDevice *d, MopHipShadowMap *map,
                                             MopHipBuffer *vb, uint32_t vertex_count,
                                             MopHipBuffer *ib, uint32_t index_count,
                                             const MopHipMat4 *model,
                                             const MopHipMat4 *light_vp, int clear) {
  if (!d || !map || !vb || !ib || !model || !light_vp) {
    set_err("shadow_render_noalloc: null argument");
    return -1;
  }
  if (!d->noalloc_scratch) {
    set_err("shadow_render_noalloc: scratch symbol unavailable on this device");
    return -1;
  }
  if ((size_t)vertex_count * sizeof(MopHipVertex) > vb->size ||
      (size_t)index_count * sizeof(uint32_t) > ib->size) {
    set_err("shadow_render_noalloc: counts exceed the buffers");
    return -1;
  }
  uint32_t n_tris = index_count / 3;
  if (vertex_count > NOALLOC_MAX_VERTS || n_tris > NOALLOC_MAX_TRIS) {
    set_err("shadow_render_noalloc: %u vertices / %u triangles exceed the scratch "
            "maximum of %u / %u", vertex_count, n_tris, NOALLOC_MAX_VERTS, NOALLOC_MAX_TRIS);
    return -1;
  }
  uint8_t *base = d->noalloc_scratch;
  int size = map->size;
  if (clear)
    launch_shadow_clear(size * size, map->keys.p, d->stream);
  if (n_tris == 0 || vertex_count == 0)
    return 0;
  ShadowDrawGPU g;
  memset(&g, 0, sizeof(g));
  MopHipMat4 mvp = mat4_mul(*light_vp, *model); /* rasterizer.c:361 */
  memcpy(g.mvp, mvp.d, sizeof(g.mvp));
  g.vbuf = (const uint8_t *)vb->dev;
  g.ibuf = (const uint32_t *)ib->dev;
  g.vertex_count = vertex_count;
  g.tri_count = n_tris;
  g.tri_base = 0;
  g.vtx_base = 0;
  ShadowDrawGPU *draw = (ShadowDrawGPU *)(base + NOALLOC_OFF_DRAW);
  uint32_t *counters = (uint32_t *)base;
  launch_fill_u32(CNT_N, counters, 0u, d->stream);
  launch_put_shadow_draw(g, draw, d->stream);
  ShadowArgsGPU sa;
  memset(&sa, 0, sizeof(sa));
  sa.draws = draw;
  sa.ndraws = 1;
  sa.n_tris = n_tris;
  sa.size = size;
  sa.big_cap = NOALLOC_MAX_TRIS;
  sa.n_verts = vertex_count;
  sa.lclip = (float4 *)(base + NOALLOC_OFF_LCLIP);
  sa.lscr = (float4 *)(base + NOALLOC_OFF_LSCR);
  sa.keys = map->keys.p;
  sa.big_queue = (uint32_t *)(base + NOALLOC_OFF_BIG);
  sa.counters = counters;
  launch_shadow_raster(sa, d->stream);
  return 0;
}

/* The map's depths (the ordered keys decoded) into a host array of size^2
 * floats, after the queued work (queued shadow passes included). */
extern "C" int mop_hip_shadow_map_read(MopHipDevice *d, const MopHipShadowMap *map, float *out) {
  if (!d || !map || !out)
    return -1;
  flush_shadow_jobs(d);
  size_t n = (size_t)map->size * map->size;
  std::vector<uint32_t> keys(n);
  if (hipMemcpyAsync(keys.data(), map->keys.p, n * 4, hipMemcpyDeviceToHost, d->stream) !=
          hipSuccess ||
      hipStreamSynchronize(d->stream) != hipSuccess) {
    set_err("shadow map read failed");
    return -1;
  }
  for (size_t i = 0; i < n; i++) { /* inverse of the ordered key map */
    uint32_t k = keys[i];
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    memcpy(&out[i], &u, 4);
  }
  return 0;
}

extern "C" void mop_hip_shadow_bind(MopHipDevice *d, const MopHipShadowMap *map,
                                    const MopHipMat4 *light_vp) {
  if (!d || !map || !light_vp)
    return;
  shadow_enable(d, map, *light_vp);
}

static uint64_t hash_bytes(const void *p, size_t n, uint64_t h) {
  const uint8_t *b = (const uint8_t *)p;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t v;
    memcpy(&v, b + i, 8);
    h ^= v;
    h *= 0x100000001b3ull;
    h ^= h >> 29;
  }
  for (; i < n; i++) {
    h ^= b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

static size_t pf_floats(int pf_w, int pf_h, int pf_levels) {
  size_t n = 0;
  for (int l = 0; l < pf_levels; l++) {
    int lw = pf_w >> l, lh = pf_h >> l;
    if (lw < 1) lw = 1;
    if (lh < 1) lh = 1;
    n += (size_t)lw * lh * 4;
  }
  return n;
}

static void ibl_enable(MopHipDevice *d, const MopHipIbl *ibl, float rotation, float intensity) {
  ShadeState &st = d->cur_state;
  st.irr = ibl->has_irr ? ibl->irr.p : nullptr;
  st.pf = ibl->has_pf ? ibl->pf.p : nullptr;
  st.lut = ibl->has_lut ? ibl->lut.p : nullptr;
  st.irr_w = ibl->irr_w;
  st.irr_h = ibl->irr_h;
  st.pf_w = ibl->pf_w;
  st.pf_h = ibl->pf_h;
  st.pf_levels = ibl->pf_levels;
  st.lut_size = ibl->lut_size;
  st.ibl_rot = rotation;
  st.ibl_int = intensity;
  st.ibl_on = st.irr || (st.pf && st.lut);
  d->state_dirty = true;
}

extern "C" void mop_hip_ibl_set(MopHipDevice *d, const float *irradiance, int irr_w, int irr_h,
                                const float *prefiltered, int pf_w, int pf_h, int pf_levels,
                                const float *brdf_lut, int brdf_size, float rotation,
                                float intensity) {
  if (!d)
    return;
  size_t n_irr = irradiance ? (size_t)irr_w * irr_h * 4 : 0;
  size_t n_pf = prefiltered ? pf_floats(pf_w, pf_h, pf_levels) : 0;
  size_t n_lut = brdf_lut ? (size_t)brdf_size * brdf_size * 4 : 0;
  uint64_t h = 1469598103934665603ull;
  int dims[6] = {irr_w, irr_h, pf_w, pf_h, pf_levels, brdf_size};
  h = hash_bytes(dims, sizeof(dims), h);
  if (n_irr) h = hash_bytes(irradiance, n_irr * 4, h);
  if (n_pf) h = hash_bytes(prefiltered, n_pf * 4, h);
  if (n_lut) h = hash_bytes(brdf_lut, n_lut * 4, h);
  if (!d->def_ibl)
    d->def_ibl = new MopHipIbl();
  MopHipIbl *ibl = d->def_ibl;
  if (h != d->ibl_hash) {
    (void)hipStreamSynchronize(d->stream);
    ibl->has_irr = n_irr && ibl->irr.ensure(n_irr);
    ibl->has_pf = n_pf && ibl->pf.ensure(n_pf);
    ibl->has_lut = n_lut && ibl->lut.ensure(n_lut);
    if (ibl->has_irr)
      HIPCHK(hipMemcpy(ibl->irr.p, irradiance, n_irr * 4, hipMemcpyHostToDevice));
    if (ibl->has_pf)
      HIPCHK(hipMemcpy(ibl->pf.p, prefiltered, n_pf * 4, hipMemcpyHostToDevice));
    if (ibl->has_lut)
      HIPCHK(hipMemcpy(ibl->lut.p, brdf_lut, n_lut * 4, hipMemcpyHostToDevice));
    d->ibl_hash = h;
  }
  ibl->irr_w = irr_w;
  ibl->irr_h = irr_h;
  ibl->pf_w = pf_w;
  ibl->pf_h = pf_h;
  ibl->pf_levels = pf_levels;
  ibl->lut_size = brdf_size;
  ibl_enable(d, ibl, rotation, intensity);
}

extern "C" void mop_hip_ibl_clear(MopHipDevice *d) {
  if (!d)
    return;
  d->cur_state.ibl_on = false;
  d->cur_state.irr = d->cur_state.pf = d->cur_state.lut = nullptr;
  d->state_dirty = true;
}

void launch_ibl_precompute(const float *env, int ew, int eh, float *irr, float *pf, float *lut,
                           const float *trig, hipStream_t s);
size_t ibl_pf_texels(void);
size_t ibl_trig_floats(void);
void ibl_trig_fill(float *tab);

extern "C" MopHipIbl *mop_hip_ibl_create(MopHipDevice *d, const float *env_rgba, int w, int h) {
  if (!d || !env_rgba || w < 2 || h < 2)
    return nullptr;
  MopHipIbl *ibl = new MopHipIbl();
  ibl->irr_w = 64;
  ibl->irr_h = 32;
  ibl->pf_w = 128;
  ibl->pf_h = 64;
  ibl->pf_levels = 5;
  ibl->lut_size = 256;
  size_t n_env = (size_t)w * h * 4;
  DevArr<float> env;
  bool ok = env.ensure(n_env) && ibl->irr.ensure(64 * 32 * 4) &&
            ibl->pf.ensure(ibl_pf_texels() * 4) && ibl->lut.ensure(256 * 256 * 4);
  if (ok && !d->ibl_trig.p) {
    /* the precompute's sine/cosine arguments, evaluated once per device */
    std::vector<float> tab(ibl_trig_floats());
    ibl_trig_fill(tab.data());
    ok = d->ibl_trig.ensure(tab.size()) &&
         hipMemcpy(d->ibl_trig.p, tab.data(), tab.size() * sizeof(float),
                   hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    env.release();
    mop_hip_ibl_destroy(d, ibl);
    return nullptr;
  }
  HIPCHK(hipMemcpy(env.p, env_rgba, n_env * sizeof(float), hipMemcpyHostToDevice));
  launch_ibl_precompute(env.p, w, h, ibl->irr.p, ibl->pf.p, ibl->lut.p, d->ibl_trig.p,
                        d->stream);
  HIPCHK(hipStreamSynchronize(d->stream));
  env.release();
  ibl->has_irr = ibl->has_pf = ibl->has_lut = true;
  return ibl;
}

extern "C" void mop_hip_ibl_destroy(MopHipDevice *d, MopHipIbl *ibl) {
  if (!ibl)
    return;
  if (d)
    (void)hipStreamSynchronize(d->stream);
  ibl->irr.release();
  ibl->pf.release();
  ibl->lut.release();
  delete ibl;
}

extern "C" void mop_hip_ibl_bind(MopHipDevice *d, const MopHipIbl *ibl, float rotation,
                                 float intensity) {
  if (!d || !ibl)
    return;
  ibl_enable(d, ibl, rotation, intensity);
}

extern "C" int mop_hip_ibl_download(MopHipDevice *d, const MopHipIbl *ibl, float *irr,
                                    float *pf, float *lut) {
  if (!d || !ibl)
    return -1;
  (void)hipStreamSynchronize(d->stream);
  if (irr)
    HIPCHK(hipMemcpy(irr, ibl->irr.p, (size_t)ibl->irr_w * ibl->irr_h * 16, hipMemcpyDeviceToHost));
  if (pf)
    HIPCHK(hipMemcpy(pf, ibl->pf.p, pf_floats(ibl->pf_w, ibl->pf_h, ibl->pf_levels) * 4,
                     hipMemcpyDeviceToHost));
  if (lut)
    HIPCHK(hipMemcpy(lut, ibl->lut.p, (size_t)ibl->lut_size * ibl->lut_size * 16,
                     hipMemcpyDeviceToHost));
  return 0;
}

extern "C" void mop_hip_hdr_resolve(MopHipDevice *d, MopHipFramebuffer *fb, float exposure) {
  if (!d || !fb)
    return;
  fb->resolve_set = true;
  fb->resolve_exposure = exposure;
  if (!fb->executed) {
    execute_frames(d, &fb, 1);
    return;
  }
  /* the tail recomputes from the HDR buffer, so re-running it is exact */
  fb_new_contents(d, fb);
  std::vector<MopHipFramebuffer *> one(1, fb);
  run_tail(d, one);
  HIPCHK(hipEventRecord(fb->done, d->stream));
}

/* Record the resolve and post-process requests of the current frame without
 * executing anything (frames executed later by mop_hip_frames_execute). */
extern "C" void mop_hip_frame_request_post(MopHipDevice *d, MopHipFramebuffer *fb,
                                           float exposure, uint32_t effects,
                                           const MopHipFogParams *fog) {
  if (!d || !fb)
    return;
  fb->resolve_set = true;
  fb->resolve_exposure = exposure;
  fb->post_effects = effects;
  fb->has_fog = fog != nullptr;
  if (fog)
    fb->fog = *fog;
}

extern "C" void mop_hip_draw_outline(MopHipDevice *d, MopHipFramebuffer *fb, MopHipColor accent,
                                     float alpha_selected, float alpha_unselected,
                                     const uint32_t *selected_ids, uint32_t selected_count) {
  if (!d || !fb || (selected_count && !selected_ids))
    return;
  fb->outline = true;
  /* the accent as the overlay converts it (overlay_builtin.c:1330-1333) */
  uint32_t ar = (uint8_t)(accent.r * 255.0f), ag = (uint8_t)(accent.g * 255.0f),
           ab = (uint8_t)(accent.b * 255.0f);
  fb->outline_accent = ar | (ag << 8) | (ab << 16);
  fb->outline_alpha_sel = alpha_selected;
  fb->outline_alpha_unsel = alpha_unselected;
  fb->outline_sel.assign(selected_ids, selected_ids + selected_count);
  std::sort(fb->outline_sel.begin(), fb->outline_sel.end());
  fb->outline_sel.erase(std::unique(fb->outline_sel.begin(), fb->outline_sel.end()),
                        fb->outline_sel.end());
  if (!fb->executed)
    return; /* applied by the frame's tail */
  /* the tail recomputes from the HDR buffer: resolve, outline, effects */
  fb_new_contents(d, fb);
  std::vector<MopHipFramebuffer *> one(1, fb);
  run_tail(d, one);
  HIPCHK(hipEventRecord(fb->done, d->stream));
}

extern "C" void mop_hip_postprocess_apply(MopHipDevice *d, MopHipFramebuffer *fb,
                                          uint32_t effects, const MopHipFogParams *fog) {
  if (!d || !fb || effects == 0)
    return;
  fb->post_effects = effects;
  fb->has_fog = fog != nullptr;
  if (fog)
    fb->fog = *fog;
  if (!fb->executed) {
    execute_frames(d, &fb, 1);
    return;
  }
  /* resolve + post from the HDR buffer: the same as post on the resolved image */
  fb_new_contents(d, fb);
  std::vector<MopHipFramebuffer *> one(1, fb);
  run_tail(d, one);
  HIPCHK(hipEventRecord(fb->done, d->stream));
}
