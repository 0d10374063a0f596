/*
 * kernels.hip -- gfx950 kernels of the software-rasterizer frame path.
 *
 * Every arithmetic step that feeds coverage, depth or the object-ID buffer
 * follows the reference expression by expression (see devmath.h), because
 * those buffers must come out bit-identical to the CPU backend.  The layout
 * around that arithmetic is MI355X-first: one lane per vertex / triangle /
 * pixel, a visibility buffer updated with 64-bit atomic min in L2 instead of
 * ordered per-tile loops, and row-per-lane walks for large triangles.
 */
#include "devmath.h"
#include "cull.h"
#include "libmf.h"
#include "mop_hip_internal.h"
#include "shade_key.h"

#define WG 256

/* --------------------------------------------------------------------- */
/* draw lookup: draws are sorted by vtx_base / tri_base                   */
/* --------------------------------------------------------------------- */

DEV uint32_t find_draw_by_vtx(const DrawGPU *d, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (d[mid].vtx_base <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

DEV uint32_t find_draw_by_tri(const DrawGPU *d, uint32_t n, uint32_t t) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (d[mid].tri_base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* The same searches for one lane of a wave: the wave's first active lane
 * searches with wave-uniform (scalar) loads, and a lane whose item lies
 * outside that draw searches on its own.  Consecutive lanes nearly always
 * share a draw. */
DEV uint32_t find_draw_by_vtx_wave(const DrawGPU *d, uint32_t n, uint32_t v) {
  uint32_t v0 = __builtin_amdgcn_readfirstlane(v);
  uint32_t i = find_draw_by_vtx(d, n, v0);
  uint32_t hi = (i + 1 < n) ? d[i + 1].vtx_base : 0xffffffffu;
  if (v >= d[i].vtx_base && v < hi)
    return i;
  return find_draw_by_vtx(d, n, v);
}

DEV uint32_t find_draw_by_tri_wave(const DrawGPU *d, uint32_t n, uint32_t t) {
  uint32_t t0 = __builtin_amdgcn_readfirstlane(t);
  uint32_t i = find_draw_by_tri(d, n, t0);
  uint32_t hi = (i + 1 < n) ? d[i + 1].tri_base : 0xffffffffu;
  if (t >= d[i].tri_base && t < hi)
    return i;
  return find_draw_by_tri(d, n, t);
}

/* --------------------------------------------------------------------- */
/* k_vertex: cpu_prepare_triangle (cpu_backend.c:224-324) and the flexible */
/* format variant (cpu_backend.c:330-526), once per vertex.               */
/* --------------------------------------------------------------------- */

DEV void read_f3(const uint8_t *base, int off, int fmt, float o[3]) {
  const float *f = (const float *)(base + off);
  switch (fmt) {
  case MOP_HIP_FORMAT_FLOAT3:
  case MOP_HIP_FORMAT_FLOAT4:
    o[0] = f[0]; o[1] = f[1]; o[2] = f[2];
    break;
  case MOP_HIP_FORMAT_FLOAT2:
    o[0] = f[0]; o[1] = f[1]; o[2] = 0.0f;
    break;
  case MOP_HIP_FORMAT_FLOAT:
    o[0] = f[0]; o[1] = 0.0f; o[2] = 0.0f;
    break;
  default:
    o[0] = o[1] = o[2] = 0.0f;
  }
}

DEV void read_f4(const uint8_t *base, int off, int fmt, float o[4]) {
  const float *f = (const float *)(base + off);
  switch (fmt) {
  case MOP_HIP_FORMAT_FLOAT4:
    o[0] = f[0]; o[1] = f[1]; o[2] = f[2]; o[3] = f[3];
    break;
  case MOP_HIP_FORMAT_FLOAT3:
    o[0] = f[0]; o[1] = f[1]; o[2] = f[2]; o[3] = 1.0f;
    break;
  case MOP_HIP_FORMAT_FLOAT2:
    o[0] = f[0]; o[1] = f[1]; o[2] = 0.0f; o[3] = 1.0f;
    break;
  case MOP_HIP_FORMAT_FLOAT:
    o[0] = f[0]; o[1] = 0.0f; o[2] = 0.0f; o[3] = 1.0f;
    break;
  default:
    o[0] = o[1] = o[2] = 0.0f; o[3] = 1.0f;
  }
}

DEV void read_f2(const uint8_t *base, int off, int fmt, float o[2]) {
  const float *f = (const float *)(base + off);
  switch (fmt) {
  case MOP_HIP_FORMAT_FLOAT2:
  case MOP_HIP_FORMAT_FLOAT3:
  case MOP_HIP_FORMAT_FLOAT4:
    o[0] = f[0]; o[1] = f[1];
    break;
  case MOP_HIP_FORMAT_FLOAT:
    o[0] = f[0]; o[1] = 0.0f;
    break;
  default:
    o[0] = o[1] = 0.0f;
  }
}

/* The workgroup -> draw maps of the batch: entry i < n_wg_tri of the
 * per-triangle grid, then the per-vertex grid; the draw is the last run
 * starting at or before the workgroup (binary search over the runs, whose
 * starts do not decrease; a draw with no workgroups shares its start with
 * the next run). */
__global__ __launch_bounds__(WG) void k_wg_map(FrameArgs fa) {
  uint32_t i = blockIdx.x * WG + threadIdx.x;
  if (i >= fa.n_wg_tri + fa.n_wg_vtx)
    return;
  bool tri = i < fa.n_wg_tri;
  uint32_t b = tri ? i : i - fa.n_wg_tri;
  uint32_t lo = 0, hi = fa.n_runs; /* runs[lo].start <= b < runs[hi].start */
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    uint4 r = fa.runs[mid];
    if ((tri ? r.x : r.y) <= b)
      lo = mid;
    else
      hi = mid;
  }
  uint32_t *map = (uint32_t *)(tri ? fa.wg_tri : fa.wg_vtx);
  map[b] = fa.runs[lo].z;
}

void launch_wg_map(const FrameArgs &fa, hipStream_t s) {
  uint32_t n = fa.n_wg_tri + fa.n_wg_vtx;
  if (n && fa.n_runs)
    launch(k_wg_map, dim3((n + WG - 1) / WG), dim3(WG), 0, s, fa);
}

/* the 256 vertices of per-vertex workgroup b (draw-major: the draw is
 * uniform over the workgroup, scalar loads) */
DEV void vertex_work(const FrameArgs &fa, const DrawGPU *draws, uint32_t b, uint32_t di) {
  const DrawGPU &d = draws[di];
  uint32_t i = (b - d.vtx_wg) * WG + threadIdx.x;
  if (i >= d.vertex_count)
    return;
  uint32_t v = d.vtx_base + i;
  const uint8_t *vb = d.vbuf + (size_t)i * d.stride;
  float p[3], n[3], c[4], uv[2];
  bool has_n = true;
  if (!(d.flags & MOPH_DRAW_FLEX)) {
    const float *f = (const float *)vb;
    p[0] = f[0]; p[1] = f[1]; p[2] = f[2];
    n[0] = f[3]; n[1] = f[4]; n[2] = f[5];
    c[0] = f[6]; c[1] = f[7]; c[2] = f[8]; c[3] = f[9];
    uv[0] = f[10]; uv[1] = f[11];
  } else {
    read_f3(vb, d.pos_off, d.pos_fmt, p);
    if (d.nrm_off >= 0)
      read_f3(vb, d.nrm_off, d.nrm_fmt, n);
    else
      has_n = false;
    if (d.col_off >= 0)
      read_f4(vb, d.col_off, d.col_fmt, c);
    else
      c[0] = c[1] = c[2] = c[3] = 1.0f;
    if (d.uv_off >= 0)
      read_f2(vb, d.uv_off, d.uv_fmt, uv);
    else
      uv[0] = uv[1] = 0.0f;
  }
  float4 clip = mat4_mul_vec4(d.mvp, p[0], p[1], p[2], 1.0f);
  if (d.depth_bias != 0.0f)
    clip.z += d.depth_bias * clip.w;
  float4 wp = mat4_mul_vec4(d.model, p[0], p[1], p[2], 1.0f);
  float nx, ny, nz;
  if (has_n) {
    float4 tn = mat4_mul_vec4(d.model, n[0], n[1], n[2], 0.0f);
    nx = tn.x; ny = tn.y; nz = tn.z;
  } else {
    nx = 0.0f; ny = 1.0f; nz = 0.0f;
  }
  if (d.tex) {
    /* nearest texel at the vertex UV modulates the colour (cpu_backend.c:277-303) */
    float tu = uv[0] - floorf(uv[0]);
    float tv = uv[1] - floorf(uv[1]);
    int tx = cvt_i32(tu * (float)(d.tex_w - 1) + 0.5f);
    int ty = cvt_i32(tv * (float)(d.tex_h - 1) + 0.5f);
    if (tx < 0) tx = 0;
    if (tx >= d.tex_w) tx = d.tex_w - 1;
    if (ty < 0) ty = 0;
    if (ty >= d.tex_h) ty = d.tex_h - 1;
    const uint8_t *t = d.tex + ((size_t)ty * (size_t)d.tex_w + (size_t)tx) * 4;
    c[0] *= (float)t[0] / 255.0f;
    c[1] *= (float)t[1] / 255.0f;
    c[2] *= (float)t[2] / 255.0f;
    c[3] *= (float)t[3] / 255.0f;
  }
  fa.vpos[v] = clip;
  if (fa.vscr) {
    /* the fan-vertex screen transform (rasterizer.c:2520-2532); an unclipped
     * triangle's fan vertices are its own vertices */
    const ViewGPU &vw = fa.views[d.view];
    float4 sc;
    screen_xform(clip, (float)vw.w * 0.5f, (float)vw.h * 0.5f, sc.x, sc.y, sc.z, sc.w);
    fa.vscr[v] = sc;
  }
  float4 *a = fa.vattr + (size_t)v * 3;
  a[0] = make_float4(nx, ny, nz, wp.x);
  a[1] = make_float4(wp.y, wp.z, c[0], c[1]);
  a[2] = make_float4(c[2], c[3], 0.0f, 0.0f);
}

__global__ __launch_bounds__(WG) void k_vertex(FrameArgs fa) {
  vertex_work(fa, fa.draws, blockIdx.x, fa.wg_vtx[blockIdx.x]);
}

/* Chunk-culled batches: the leading vtx_direct per-vertex workgroups (draws
 * not culled by chunks) one per workgroup as in k_vertex, then the ones
 * k_cull_vertices kept, walked by the last vtx_lgrid workgroups (a skipped
 * workgroup costs no launch).  The tables are kernel arguments declared
 * unaliased, so the vertex stores of one pass of the walk cannot clobber
 * them and the draw's fields stay scalar loads; the walk's entries are
 * (workgroup, draw) pairs, the next one loaded while the current one is
 * worked on. */
__global__ __launch_bounds__(WG) void k_vertex_list(FrameArgs fa, const DrawGPU *__restrict__ draws,
                                                    const uint32_t *__restrict__ wg_vtx,
                                                    const uint2 *__restrict__ list) {
  if (blockIdx.x < fa.vtx_direct) {
    vertex_work(fa, draws, blockIdx.x, wg_vtx[blockIdx.x]);
    return;
  }
  uint32_t n = fa.counters[CNT_VLIST], e = blockIdx.x - fa.vtx_direct;
  uint32_t stride = gridDim.x - fa.vtx_direct;
  uint2 cur = e < n ? list[e] : make_uint2(0u, 0u);
  for (; e < n; e += stride) {
    uint2 nxt = e + stride < n ? list[e + stride] : make_uint2(0u, 0u);
    vertex_work(fa, draws, cur.x, cur.y);
    cur = nxt;
  }
}

void launch_vertex(const FrameArgs &fa, hipStream_t s) {
  if (!fa.n_wg_vtx)
    return;
  if (fa.vtx_list)
    launch(k_vertex_list, dim3(fa.vtx_direct + fa.vtx_lgrid), dim3(WG), 0, s, fa, fa.draws,
           fa.wg_vtx, (const uint2 *)fa.vtx_list);
  else
    launch(k_vertex, dim3(fa.n_wg_vtx), dim3(WG), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* Chunk culling (order-free path, large draws).  A chunk of 256           */
/* triangles whose object-space box fails one trivial-reject plane with a */
/* rounding margin (box_rejected, cull.h) holds only triangles the raster */
/* rejects (rasterizer.c:2469-2476): they add no coverage, no clip entry  */
/* and no count, so k_tri skips the chunk, and k_vertex transforms only   */
/* the vertex workgroups the kept chunks use.                             */
/* --------------------------------------------------------------------- */

/* Per chunk (one workgroup): the box of the positions k_vertex reads for
 * the vertices of the chunk's valid triangles (indices below vertex_count),
 * a flag for a non-finite coordinate (such a chunk is never culled), and
 * the chunk's entry in the reference list of every per-vertex workgroup it
 * uses: vref[w] counts the chunks using workgroup w, the first VREF_CAP of
 * them listed at vref[nvwg + w * VREF_CAP ..].  The distinct workgroups of
 * a chunk are found with an LDS hash set (at most 768 of 1024 slots). */
__global__ __launch_bounds__(WG) void k_chunk_bounds(ChunkBoundsArgs a) {
  __shared__ float red[WG / 64][7];
  __shared__ uint32_t set[1024];
  for (uint32_t q = threadIdx.x; q < 1024; q += WG)
    set[q] = ~0u;
  __syncthreads();
  uint32_t k = blockIdx.x * WG + threadIdx.x;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float bad = 0.0f;
  if (k < a.tri_count) {
    for (int j = 0; j < 3; j++) {
      uint32_t i = a.ibuf[3 * k + j];
      if (i >= a.vertex_count)
        continue;
      const uint8_t *vb = a.vbuf + (size_t)i * a.stride;
      float p[3];
      if (!a.flex) {
        const float *f = (const float *)vb;
        p[0] = f[0]; p[1] = f[1]; p[2] = f[2];
      } else {
        read_f3(vb, a.pos_off, a.pos_fmt, p);
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (!(fabsf(p[c]) <= 3.402823466e38f))
          bad = 1.0f;
        lo[c] = fminf(lo[c], p[c]);
        hi[c] = fmaxf(hi[c], p[c]);
      }
      uint32_t w = i >> 8;
      for (uint32_t h = (w * 2654435761u) >> 22;; h = (h + 1) & 1023u) {
        uint32_t old = atomicCAS(&set[h], ~0u, w);
        if (old == ~0u) {
          uint32_t slot = atomicAdd(&a.vref[w], 1u);
          if (slot < VREF_CAP)
            a.vref[a.nvwg + (size_t)w * VREF_CAP + slot] = blockIdx.x;
          break;
        }
        if (old == w)
          break;
      }
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      lo[c] = fminf(lo[c], __shfl_xor(lo[c], o));
      hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], o));
    }
    bad = fmaxf(bad, __shfl_xor(bad, o));
  }
  if ((threadIdx.x & 63u) == 0) {
    float *r = red[threadIdx.x >> 6];
    r[0] = lo[0]; r[1] = lo[1]; r[2] = lo[2];
    r[3] = hi[0]; r[4] = hi[1]; r[5] = hi[2];
    r[6] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < WG / 64; w++) {
      for (int c = 0; c < 3; c++) {
        lo[c] = fminf(lo[c], red[w][c]);
        hi[c] = fmaxf(hi[c], red[w][3 + c]);
      }
      bad = fmaxf(bad, red[w][6]);
    }
    a.box[2 * blockIdx.x] = make_float4(lo[0], lo[1], lo[2], bad);
    a.box[2 * blockIdx.x + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
  }
}

void launch_chunk_bounds(const ChunkBoundsArgs &a, hipStream_t s) {
  uint32_t nb = (a.tri_count + WG - 1) / WG;
  if (nb)
    launch(k_chunk_bounds, dim3(nb), dim3(WG), 0, s, a);
}

#define CULL_WG 1024

/* list[base + rank] = item for the lanes of the workgroup with keep, base
 * from one atomic per workgroup: atomics on one address are performed one
 * after another at the memory side, so one per wave (6.4 k for a 12-view
 * serving batch) took 0.12 ms.  The order of the list is immaterial to its
 * consumers. */
DEV void block_append(uint2 *list, uint32_t *count, bool keep, uint2 item) {
  __shared__ uint32_t wave_base[CULL_WG / 64];
  __shared__ uint32_t base;
  unsigned long long m = __ballot(keep);
  uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0)
    wave_base[w] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (int i = 0; i < CULL_WG / 64; i++) {
      uint32_t c = wave_base[i];
      wave_base[i] = sum;
      sum += c;
    }
    base = sum ? atomicAdd(count, sum) : 0u;
  }
  __syncthreads();
  if (keep) {
    uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    list[base + wave_base[w] + rank] = item;
  }
}

/* One lane per per-triangle workgroup b of the culled draws (b >=
 * tri_direct): kept unless the chunk's box fails a plane; tri_vis[b] and
 * the list of kept workgroups k_tri_list walks. */
__global__ __launch_bounds__(CULL_WG) void k_cull_chunks(FrameArgs fa) {
  uint32_t b = fa.tri_direct + blockIdx.x * CULL_WG + threadIdx.x;
  bool keep = false;
  uint32_t di = 0;
  if (b < fa.n_wg_tri) {
    di = fa.wg_tri[b];
    const DrawGPU &d = fa.draws[di];
    keep = true;
    if (d.chunk_box) {
      uint32_t c = b - d.tri_wg;
      float4 lo = d.chunk_box[2 * c], hi = d.chunk_box[2 * c + 1];
      keep = !(lo.w == 0.0f && box_rejected(d.mvp, d.depth_bias, lo, hi));
    }
    fa.tri_vis[b] = keep ? 1u : 0u;
  }
  block_append(fa.tri_list, &fa.counters[CNT_TLIST], keep, make_uint2(b, di));
}

/* One lane per per-vertex workgroup b of the culled draws (b >= vtx_direct):
 * kept when a kept chunk uses it or more than VREF_CAP chunks do. */
__global__ __launch_bounds__(CULL_WG) void k_cull_vertices(FrameArgs fa) {
  uint32_t b = fa.vtx_direct + blockIdx.x * CULL_WG + threadIdx.x;
  bool need = false;
  uint32_t di = 0;
  if (b < fa.n_wg_vtx) {
    di = fa.wg_vtx[b];
    const DrawGPU &d = fa.draws[di];
    need = true;
    if (d.chunk_box) {
      uint32_t w = b - d.vtx_wg, nvwg = (d.vertex_count + WG - 1) / WG;
      uint32_t n = d.vref[w];
      if (n <= VREF_CAP) {
        const uint32_t *l = d.vref + nvwg + (size_t)w * VREF_CAP;
        need = false;
        for (uint32_t q = 0; q < n; q++)
          need = need || fa.tri_vis[d.tri_wg + l[q]] != 0;
      }
    }
  }
  block_append(fa.vtx_list, &fa.counters[CNT_VLIST], need, make_uint2(b, di));
}

void launch_cull(const FrameArgs &fa, hipStream_t s) {
  if (fa.n_wg_tri > fa.tri_direct)
    launch(k_cull_chunks, dim3((fa.n_wg_tri - fa.tri_direct + CULL_WG - 1) / CULL_WG),
           dim3(CULL_WG), 0, s, fa);
  if (fa.n_wg_vtx > fa.vtx_direct)
    launch(k_cull_vertices, dim3((fa.n_wg_vtx - fa.vtx_direct + CULL_WG - 1) / CULL_WG),
           dim3(CULL_WG), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* classification and fan checks (rasterizer.c:2458-2553)                */
/* --------------------------------------------------------------------- */

DEV bool trivially_rejected(float4 a, float4 b, float4 c) {
  return (a.x < -a.w && b.x < -b.w && c.x < -c.w) ||
         (a.x > a.w && b.x > b.w && c.x > c.w) ||
         (a.y < -a.w && b.y < -b.w && c.y < -c.w) ||
         (a.y > a.w && b.y > b.w && c.y > c.w) ||
         (a.z < -a.w && b.z < -b.w && c.z < -c.w) ||
         (a.z > a.w && b.z > b.w && c.z > c.w);
}

DEV bool inside_all(float4 a) {
  return a.x >= -a.w && a.x <= a.w && a.y >= -a.w && a.y <= a.w &&
         a.z >= -a.w && a.z <= a.w;
}

/* One fan triangle survives: |w| guard, screen-space cull, bbox, |area|. */
/* A fan triangle that reaches the raster: |w| >= 1e-7, not culled by screen
 * area and (filled draws only) with a non-empty clamped bbox and non-zero
 * area.  Wireframe triangles draw their edges whatever their area
 * (rasterizer.c:2515-2537, 2598-2606). */
DEV bool fan_ok(float4 p0, float4 p1, float4 p2, bool cull, bool wire, int W, int H) {
  if (fabsf(p0.w) < 1e-7f || fabsf(p1.w) < 1e-7f || fabsf(p2.w) < 1e-7f)
    return false;
  float half_w = (float)W * 0.5f, half_h = (float)H * 0.5f;
  float sx[3], sy[3], sz[3], iw[3];
  screen_xform(p0, half_w, half_h, sx[0], sy[0], sz[0], iw[0]);
  screen_xform(p1, half_w, half_h, sx[1], sy[1], sz[1], iw[1]);
  screen_xform(p2, half_w, half_h, sx[2], sy[2], sz[2], iw[2]);
  float signed_area =
      (sx[1] - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (sy[1] - sy[0]);
  if (cull && signed_area >= 0.0f)
    return false;
  if (wire)
    return true;
  EdgeSetup es;
  return edge_setup(sx, sy, W, H, es);
}

__global__ __launch_bounds__(WG) void k_classify(FrameArgs fa) {
  uint32_t t = blockIdx.x * WG + threadIdx.x;
  if (t >= fa.n_tris)
    return;
  uint32_t di = find_draw_by_tri(fa.draws, fa.ndraws, t);
  const DrawGPU &d = fa.draws[di];
  uint32_t k = t - d.tri_base;
  uint32_t i0 = d.ibuf[3 * k], i1 = d.ibuf[3 * k + 1], i2 = d.ibuf[3 * k + 2];
  uint32_t cnt = 0, slot = 0xffffffffu;
  if (i0 < d.vertex_count && i1 < d.vertex_count && i2 < d.vertex_count) {
    float4 a = fa.vpos[d.vtx_base + i0];
    float4 b = fa.vpos[d.vtx_base + i1];
    float4 c = fa.vpos[d.vtx_base + i2];
    bool cull = (d.flags & MOPH_DRAW_CULL) != 0;
    bool alive = !trivially_rejected(a, b, c);
    if (alive && cull && a.w > 0.0f && b.w > 0.0f && c.w > 0.0f) {
      float ex = b.x * a.w - a.x * b.w;
      float ey = b.y * a.w - a.y * b.w;
      float fx = c.x * a.w - a.x * c.w;
      float fy = c.y * a.w - a.y * c.w;
      if (ex * fy - ey * fx <= 0.0f)
        alive = false;
    }
    if (alive) {
      if (a.w > 0.0f && b.w > 0.0f && c.w > 0.0f && inside_all(a) &&
          inside_all(b) && inside_all(c)) {
        const ViewGPU &vw = fa.views[d.view];
        cnt = fan_ok(a, b, c, cull, (d.flags & MOPH_DRAW_WIRE) != 0, vw.w, vw.h) ? 1u : 0u;
      } else {
        uint32_t q = atomicAdd(&fa.counters[CNT_CLIP], 1u);
        if (q < fa.clip_cap) {
          fa.clip_queue[q] = t;
          slot = q;
        } else {
          atomicOr(&fa.counters[CNT_OVERFLOW], 1u);
        }
      }
    }
  }
  fa.tri_counts[t] = cnt;
  fa.clip_slot[t] = slot;
}

void launch_classify(const FrameArgs &fa, hipStream_t s) {
  if (!fa.n_tris)
    return;
  launch(k_classify, dim3((fa.n_tris + WG - 1) / WG), dim3(WG), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* k_clip: mop_sw_clip_polygon (rasterizer.c:510-626) for queued triangles */
/* Attributes carried: position, normal, world position, colour -- the    */
/* ones the smooth multi-light path reads.                               */
/* --------------------------------------------------------------------- */

struct CV {
  float p[4];
  float a[10]; /* n.xyz, wp.xyz, c.rgba */
};

DEV float clip_dist(int plane, const float *p) {
  /* FRUSTUM_PLANES (rasterizer.c:510-517), all four products kept */
  const float N[6][4] = {{1.0f, 0.0f, 0.0f, 1.0f},  {-1.0f, 0.0f, 0.0f, 1.0f},
                         {0.0f, 1.0f, 0.0f, 1.0f},  {0.0f, -1.0f, 0.0f, 1.0f},
                         {0.0f, 0.0f, 1.0f, 1.0f},  {0.0f, 0.0f, -1.0f, 1.0f}};
  return N[plane][0] * p[0] + N[plane][1] * p[1] + N[plane][2] * p[2] +
         N[plane][3] * p[3];
}

DEV void lerp_cv(const CV &a, const CV &b, float t, CV &r) {
#pragma unroll
  for (int k = 0; k < 4; k++)
    r.p[k] = a.p[k] + t * (b.p[k] - a.p[k]);
#pragma unroll
  for (int k = 0; k < 10; k++)
    r.a[k] = a.a[k] + t * (b.a[k] - a.a[k]);
}

/* One Sutherland-Hodgman plane (clip_against_plane, rasterizer.c:553-593).
 * A triangle clipped by the six planes never exceeds 9 vertices, so buffers
 * of CLIP_CAP entries hold every polygon the reference's 24-entry buffers do;
 * a longer one (impossible) is reported as n > MOPH_MAX_POLY. */
#define CLIP_CAP 12

DEV int clip_plane(const CV *in, int n, CV *out, int plane) {
  if (n < 1)
    return 0;
  int oc = 0;
  CV prev = in[n - 1];
  float prev_d = clip_dist(plane, prev.p);
  for (int i = 0; i < n; i++) {
    CV curr = in[i];
    float curr_d = clip_dist(plane, curr.p);
    if (curr_d >= 0.0f) {
      if (prev_d < 0.0f) {
        float t = prev_d / (prev_d - curr_d);
        if (oc < CLIP_CAP)
          lerp_cv(prev, curr, t, out[oc]);
        oc++;
      }
      if (oc < CLIP_CAP)
        out[oc] = curr;
      oc++;
    } else if (prev_d >= 0.0f) {
      float t = prev_d / (prev_d - curr_d);
      if (oc < CLIP_CAP)
        lerp_cv(prev, curr, t, out[oc]);
      oc++;
    }
    prev = curr;
    prev_d = curr_d;
  }
  return oc;
}

DEV void load_cv(const FrameArgs &fa, uint32_t v, CV &o) {
  float4 p = fa.vpos[v];
  const float4 *a = fa.vattr + (size_t)v * 3;
  float4 a0 = a[0], a1 = a[1], a2 = a[2];
  o.p[0] = p.x; o.p[1] = p.y; o.p[2] = p.z; o.p[3] = p.w;
  o.a[0] = a0.x; o.a[1] = a0.y; o.a[2] = a0.z;
  o.a[3] = a0.w; o.a[4] = a1.x; o.a[5] = a1.y;
  o.a[6] = a1.z; o.a[7] = a1.w; o.a[8] = a2.x; o.a[9] = a2.y;
}

/* The two polygon buffers of each lane live in LDS (1344 B per lane), not in
 * per-lane scratch.  A clip is a long serial chain (six planes over up to
 * nine vertices), so workgroups are small: the few clipped triangles of a
 * frame spread over many CUs instead of sharing a wave. */
#define CLIP_WG 16

__global__ __launch_bounds__(CLIP_WG) void k_clip(FrameArgs fa) {
  __shared__ CV pool[CLIP_WG][2][CLIP_CAP];
  uint32_t nq = fa.counters[CNT_CLIP];
  if (nq > fa.clip_cap)
    nq = fa.clip_cap;
  for (uint32_t q = blockIdx.x * CLIP_WG + threadIdx.x; q < nq; q += gridDim.x * CLIP_WG) {
    uint32_t t = fa.clip_queue[q];
    uint32_t di = fa.fused ? fa.tri_rec[t].w : find_draw_by_tri(fa.draws, fa.ndraws, t);
    const DrawGPU &d = fa.draws[di];
    uint32_t k = t - d.tri_base;
    CV *bufA = pool[threadIdx.x][0], *bufB = pool[threadIdx.x][1];
    load_cv(fa, d.vtx_base + d.ibuf[3 * k], bufA[0]);
    load_cv(fa, d.vtx_base + d.ibuf[3 * k + 1], bufA[1]);
    load_cv(fa, d.vtx_base + d.ibuf[3 * k + 2], bufA[2]);
    int n = 3;
    CV *src = bufA, *dst = bufB;
    for (int pl = 0; pl < 6 && n > 0; pl++) {
      n = clip_plane(src, n, dst, pl);
      if (n > CLIP_CAP)
        break;
      CV *tmp = src;
      src = dst;
      dst = tmp;
    }
    uint32_t cnt = 0;
    if (n > MOPH_MAX_POLY) {
      atomicOr(&fa.counters[CNT_OVERFLOW], 8u);
      n = 0;
    }
    if (n >= 3) {
      bool cull = (d.flags & MOPH_DRAW_CULL) != 0;
      const ViewGPU &vw = fa.views[d.view];
      for (int i = 1; i < n - 1; i++) {
        float4 p0 = make_float4(src[0].p[0], src[0].p[1], src[0].p[2], src[0].p[3]);
        float4 p1 = make_float4(src[i].p[0], src[i].p[1], src[i].p[2], src[i].p[3]);
        float4 p2 = make_float4(src[i + 1].p[0], src[i + 1].p[1], src[i + 1].p[2],
                                src[i + 1].p[3]);
        if (fan_ok(p0, p1, p2, cull, (d.flags & MOPH_DRAW_WIRE) != 0, vw.w, vw.h))
          cnt++;
      }
      size_t base = (size_t)q * MOPH_MAX_POLY;
      for (int i = 0; i < n; i++) {
        fa.cpos[base + i] = make_float4(src[i].p[0], src[i].p[1], src[i].p[2], src[i].p[3]);
        float4 *ca = fa.cattr + (base + i) * 3;
        ca[0] = make_float4(src[i].a[0], src[i].a[1], src[i].a[2], src[i].a[3]);
        ca[1] = make_float4(src[i].a[4], src[i].a[5], src[i].a[6], src[i].a[7]);
        ca[2] = make_float4(src[i].a[8], src[i].a[9], 0.0f, 0.0f);
      }
    } else {
      n = 0;
    }
    fa.cnt_poly[q] = (uint32_t)n;
    if (!fa.fused)
      fa.tri_counts[t] = cnt;
  }
}

void launch_clip(const FrameArgs &fa, hipStream_t s) {
  if (!fa.n_tris || !fa.clip_cap)
    return;
  /* the queue length is known on the device only: a grid-stride loop */
  uint32_t grid = (fa.clip_cap + CLIP_WG - 1) / CLIP_WG;
  if (grid > 2048)
    grid = 2048;
  launch(k_clip, dim3(grid), dim3(CLIP_WG), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* exclusive scan of per-triangle screen-triangle counts                  */
/* --------------------------------------------------------------------- */

#define SCAN_ITEMS 16
#define SCAN_TILE (WG * SCAN_ITEMS)

DEV uint32_t block_excl_scan(uint32_t v, uint32_t *lds, uint32_t &total) {
  /* wave64 inclusive scan, then across the 4 waves */
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o)
      x += y;
  }
  if (lane == 63)
    lds[wid] = x;
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
  for (int w = 0; w < WG / 64; w++) {
    if (w < wid)
      wpre += lds[w];
    tot += lds[w];
  }
  __syncthreads();
  total = tot;
  return wpre + x - v;
}

__global__ __launch_bounds__(WG) void k_scan_reduce(const uint32_t *in, uint32_t n,
                                                    uint32_t *sums) {
  __shared__ uint32_t lds[WG / 64];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n)
      acc += in[base + k];
  uint32_t tot;
  block_excl_scan(acc, lds, tot);
  if (threadIdx.x == 0)
    sums[blockIdx.x] = tot;
}

/* single workgroup: exclusive scan of the block sums, in place */
__global__ __launch_bounds__(WG) void k_scan_sums(uint32_t *sums, uint32_t nb) {
  __shared__ uint32_t lds[WG / 64];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += WG) {
    uint32_t i = b0 + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0;
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, lds, tot);
    if (i < nb)
      sums[i] = carry + ex;
    carry += tot;
  }
}

__global__ __launch_bounds__(WG) void k_scan_apply(uint32_t *io, uint32_t n,
                                                   const uint32_t *sums,
                                                   uint32_t *counters,
                                                   uint32_t cap) {
  __shared__ uint32_t lds[WG / 64];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = (base + k < n) ? io[base + k] : 0;
    acc += v[k];
  }
  uint32_t tot;
  uint32_t run = sums[blockIdx.x] + block_excl_scan(acc, lds, tot);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n)
      io[base + k] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == WG - 1) {
    io[n] = run; /* total */
    if (counters) {
      counters[CNT_SUBTRI] = run;
      if (run > cap)
        atomicOr(&counters[CNT_OVERFLOW], 2u);
    }
  }
}

void launch_scan_counts(const FrameArgs &fa, uint32_t *block_sums, hipStream_t s) {
  uint32_t n = fa.n_tris;
  uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 0)
    nb = 1;
  launch(k_scan_reduce, dim3(nb), dim3(WG), 0, s, fa.tri_counts, n, block_sums);
  launch(k_scan_sums, dim3(1), dim3(WG), 0, s, block_sums, nb);
  launch(k_scan_apply, dim3(nb), dim3(WG), 0, s, fa.tri_counts, n, block_sums,
                     fa.counters, fa.subtri_cap);
}

/* --------------------------------------------------------------------- */
/* k_emit: write screen triangles at their scanned index                  */
/* --------------------------------------------------------------------- */

DEV void write_subtri(const FrameArgs &fa, uint32_t idx, uint32_t draw, float4 p0,
                      float4 p1, float4 p2, uint32_t a0, uint32_t a1, uint32_t a2) {
  if (idx >= fa.subtri_cap)
    return;
  const ViewGPU &vw = fa.views[fa.draws[draw].view];
  float half_w = (float)vw.w * 0.5f, half_h = (float)vw.h * 0.5f;
  SubTri st;
  screen_xform(p0, half_w, half_h, st.sx[0], st.sy[0], st.sz[0], st.iw[0]);
  screen_xform(p1, half_w, half_h, st.sx[1], st.sy[1], st.sz[1], st.iw[1]);
  screen_xform(p2, half_w, half_h, st.sx[2], st.sy[2], st.sz[2], st.iw[2]);
  st.draw = draw;
  st.a[0] = a0;
  st.a[1] = a1;
  st.a[2] = a2;
  fa.subtris[idx] = st;
}

__global__ __launch_bounds__(WG) void k_emit(FrameArgs fa) {
  uint32_t t = blockIdx.x * WG + threadIdx.x;
  if (t >= fa.n_tris)
    return;
  uint32_t off = fa.tri_counts[t];
  uint32_t cnt = fa.tri_counts[t + 1] - off;
  if (!cnt)
    return;
  uint32_t di = find_draw_by_tri(fa.draws, fa.ndraws, t);
  const DrawGPU &d = fa.draws[di];
  uint32_t slot = fa.clip_slot[t];
  if (slot == 0xffffffffu) {
    uint32_t k = t - d.tri_base;
    uint32_t v0 = d.vtx_base + d.ibuf[3 * k], v1 = d.vtx_base + d.ibuf[3 * k + 1],
             v2 = d.vtx_base + d.ibuf[3 * k + 2];
    write_subtri(fa, off, di, fa.vpos[v0], fa.vpos[v1], fa.vpos[v2], v0, v1, v2);
    return;
  }
  int n = (int)fa.cnt_poly[slot];
  bool cull = (d.flags & MOPH_DRAW_CULL) != 0;
  const ViewGPU &vw = fa.views[d.view];
  size_t base = (size_t)slot * MOPH_MAX_POLY;
  float4 p0 = fa.cpos[base];
  uint32_t o = off;
  for (int i = 1; i < n - 1; i++) {
    float4 p1 = fa.cpos[base + i], p2 = fa.cpos[base + i + 1];
    if (fan_ok(p0, p1, p2, cull, (d.flags & MOPH_DRAW_WIRE) != 0, vw.w, vw.h)) {
      write_subtri(fa, o, di, p0, p1, p2, fa.n_verts + (uint32_t)base,
                   fa.n_verts + (uint32_t)(base + i), fa.n_verts + (uint32_t)(base + i + 1));
      o++;
    }
  }
}

void launch_emit(const FrameArgs &fa, hipStream_t s) {
  if (!fa.n_tris)
    return;
  launch(k_emit, dim3((fa.n_tris + WG - 1) / WG), dim3(WG), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* raster: smooth multi-light coverage + depth (rasterizer.c:1926-1957)   */
/* --------------------------------------------------------------------- */

/* Coverage + min_d test + barycentric depth of one pixel; false if the
 * pixel is not drawn (edge test or AA distance test fails or z is NaN). */
DEV bool frag_depth(const RasterTri &r, float w0, float w1, float w2, float &z) {
  if (!(w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f))
    return false;
  float d0 = w0 * r.inv_e[0];
  float d1 = w1 * r.inv_e[1];
  float d2 = w2 * r.inv_e[2];
  float min_d = d0;
  if (d1 < min_d)
    min_d = d1;
  if (d2 < min_d)
    min_d = d2;
  if (!(min_d >= 0.0f))
    return false;
  float b0 = w0 * r.es.inv_area;
  float b1 = w1 * r.es.inv_area;
  float b2 = w2 * r.es.inv_area;
  z = b0 * r.sz[0] + b1 * r.sz[1] + b2 * r.sz[2];
  return z == z;
}

/* Visibility update: a 64-bit atomic min with no return value, so a lane
 * walking a triangle issues its pixels back to back without waiting on the
 * memory system.  It is issued as a global (not flat) atomic: a flat one
 * also counts against the LDS/scalar-load counter, so every later wait for
 * an LDS or scalar load would wait for the atomics in flight too.  Callers
 * hold the buffer pointer and pitch in registers: read through the view
 * record inside a walk, they would be reloaded after every atomic (which
 * might alias the record) and the load's wait would drain the atomics
 * issued before it, one memory round trip per pixel. */
typedef __attribute__((address_space(1))) unsigned long long gvis_t;
DEV void vis_update(unsigned long long *vis, size_t idx, unsigned long long key) {
  __hip_atomic_fetch_min((gvis_t *)(vis + idx), key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* The whole bbox walk of one small screen triangle by one lane, updating the
 * visibility buffer: the smooth path's coverage and barycentric depth, or
 * the flat path's edge test and incremental z (rasterizer.c:1000-1026). */
DEV void walk_small_vis(const ViewGPU &vw, const RasterTri &r, bool flat,
                        unsigned long long seq) {
  unsigned long long *const vis = vw.vis; /* registers, not the record (vis_update) */
  const size_t pitch = (size_t)vw.w;
  float w0r = r.es.w_row[0], w1r = r.es.w_row[1], w2r = r.es.w_row[2];
  if (flat) {
    FlatZ fz = flat_z_setup(r);
    float zr = fz.row;
    for (int y = r.es.min_y; y <= r.es.max_y; y++) {
      float w0 = w0r, w1 = w1r, w2 = w2r, z = zr;
      for (int x = r.es.min_x; x <= r.es.max_x; x++) {
        if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f && z == z) {
          size_t idx = (size_t)y * pitch + (size_t)x;
          vis_update(vis, idx, ((unsigned long long)depth_key(z) << 32) | seq);
        }
        w0 += r.es.e_dx[0];
        w1 += r.es.e_dx[1];
        w2 += r.es.e_dx[2];
        z += fz.dx;
      }
      w0r += r.es.e_dy[0];
      w1r += r.es.e_dy[1];
      w2r += r.es.e_dy[2];
      zr += fz.dy;
    }
    return;
  }
  for (int y = r.es.min_y; y <= r.es.max_y; y++) {
    float w0 = w0r, w1 = w1r, w2 = w2r;
    for (int x = r.es.min_x; x <= r.es.max_x; x++) {
      float z;
      if (frag_depth(r, w0, w1, w2, z)) {
        size_t idx = (size_t)y * pitch + (size_t)x;
        vis_update(vis, idx, ((unsigned long long)depth_key(z) << 32) | seq);
      }
      w0 += r.es.e_dx[0];
      w1 += r.es.e_dx[1];
      w2 += r.es.e_dx[2];
    }
    w0r += r.es.e_dy[0];
    w1r += r.es.e_dy[1];
    w2r += r.es.e_dy[2];
  }
}

__global__ __launch_bounds__(WG) void k_raster_small(FrameArgs fa) {
  uint32_t total = fa.counters[CNT_SUBTRI];
  if (total > fa.subtri_cap)
    total = fa.subtri_cap;
  for (uint32_t s = blockIdx.x * WG + threadIdx.x; s < total; s += gridDim.x * WG) {
    SubTri st = fa.subtris[s];
    const ViewGPU &vw = fa.views[fa.draws[st.draw].view];
    RasterTri r;
    if (!load_raster_tri(vw.w, vw.h, st, r))
      continue;
    if (bbox_area(r.es) > (long long)fa.small_area) {
      uint32_t q = atomicAdd(&fa.counters[CNT_BIG], 1u);
      fa.big_queue[q] = s;
      continue;
    }
    walk_small_vis(vw, r, sm_flat(draw_sm(fa.draws[st.draw])), (unsigned long long)(s + 1));
  }
}

/* Big screen triangles: one workgroup each, one lane per bbox row.  The row
 * starts are the serial chain w_row += e_dy (rasterizer.c:2148-2150); lane 0
 * extends it 256 rows at a time into LDS.  mode 0 updates the visibility
 * buffer; mode 1 (after all updates) stores the winner's edge values. */
/* One wave per large triangle (rows in chunks of 64 lanes). */
#define BIG_WG 64

/* Columns [xa, xb] of a bbox row outside which no pixel can pass the edge
 * test.  Along the row edge i takes the serial values c(n) = fl(c(n-1) +
 * e_dx[i]) from c(0) = w_i; each addition rounds by at most 2^-24 of its
 * result, so |c(n) - (w_i + n e_dx[i])| <= n 2^-23 M with M = |w_i| + N |e_dx[i]|
 * over a row of N columns.  A pixel whose exact value lies below minus that
 * margin cannot reach 0; the margin is doubled and two columns are added for
 * the float evaluation of the bound.  The walk still starts at min_x (the
 * values are serial) but skips the tests before xa and stops after xb. */
DEV void row_span(const RasterTri &r, const float w[3], int &xa, int &xb) {
  int lo = 0, n = r.es.max_x - r.es.min_x, hi = n;
  float N = (float)(n + 1);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float e = r.es.e_dx[i];
    float margin = N * (fabsf(w[i]) + N * fabsf(e)) * 2.384185791e-7f; /* 2^-22 */
    if (e < 0.0f) {
      float t = (w[i] + margin) / -e;
      int m = t < -4.0f ? -4 : (t > (float)n ? n : (int)t + 2);
      hi = m < hi ? m : hi;
    } else if (e > 0.0f) {
      float t = (-margin - w[i]) / e;
      int m = t < 0.0f ? 0 : (t > (float)n ? n + 4 : (int)t - 2);
      lo = m > lo ? m : lo;
    } else if (w[i] < -margin) {
      hi = -1;
    }
  }
  xa = r.es.min_x + (lo > 0 ? lo : 0);
  xb = r.es.min_x + hi;
}

template <int MODE>
__global__ __launch_bounds__(BIG_WG) void k_raster_big(FrameArgs fa) {
  __shared__ float rows[4][BIG_WG];
  __shared__ float carry[4];
  uint32_t nbig = fa.counters[CNT_BIG];
  for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
    uint32_t s = fa.big_queue[b];
    SubTri st;
    load_subtri(fa, s, st);
    const DrawGPU &dr = fa.draws[st.draw];
    const ViewGPU &vw = fa.views[dr.view];
    RasterTri r;
    if (!load_raster_tri_fast(vw.w, vw.h, st, r)) /* succeeded when queued; uniform */
      continue;
    bool flat = sm_flat(draw_sm(dr)); /* uniform */
    FlatZ fz = {0.0f, 0.0f, 0.0f};
    if (flat)
      fz = flat_z_setup(r);
    /* registers, not the record (vis_update) */
    unsigned long long *const vis = vw.vis;
    float4 *const wbuf = vw.wbuf;
    const size_t pitch = (size_t)vw.w;
    unsigned long long seq = (unsigned long long)(s + 1);
    int nrows = r.es.max_y - r.es.min_y + 1;
    /* order-free path: checkpoint store of this triangle (or none) */
    uint32_t nchk = chk_per_row(r.es);
    float4 *chk = (MODE == 0 && fa.fused && fa.big_chk[b] != ~0u) ? fa.chk + fa.big_chk[b]
                                                                    : nullptr;
    if (threadIdx.x == 0) {
      carry[0] = r.es.w_row[0];
      carry[1] = r.es.w_row[1];
      carry[2] = r.es.w_row[2];
      carry[3] = fz.row;
    }
    for (int c0 = 0; c0 < nrows; c0 += BIG_WG) {
      __syncthreads();
      if (threadIdx.x == 0) {
        float a0 = carry[0], a1 = carry[1], a2 = carry[2], a3 = carry[3];
        int lim = nrows - c0 < BIG_WG ? nrows - c0 : BIG_WG;
        for (int k = 0; k < lim; k++) {
          rows[0][k] = a0;
          rows[1][k] = a1;
          rows[2][k] = a2;
          rows[3][k] = a3;
          a0 += r.es.e_dy[0];
          a1 += r.es.e_dy[1];
          a2 += r.es.e_dy[2];
          if (flat)
            a3 += fz.dy;
        }
        carry[0] = a0;
        carry[1] = a1;
        carry[2] = a2;
        carry[3] = a3;
      }
      __syncthreads();
      int row = c0 + (int)threadIdx.x;
      if (row < nrows) {
        int y = r.es.min_y + row;
        float w0 = rows[0][threadIdx.x], w1 = rows[1][threadIdx.x], w2 = rows[2][threadIdx.x];
        float zf = rows[3][threadIdx.x];
        size_t rowbase = (size_t)y * pitch;
        float4 *crow = chk ? chk + (size_t)row * nchk : nullptr;
        int xa, xb;
        {
          float wr[3] = {w0, w1, w2};
          row_span(r, wr, xa, xb);
        }
        if (xa > xb)
          xb = r.es.min_x - 1; /* no pixel of this row passes */
        int x = r.es.min_x;
        for (; x < xa && x <= xb; x++) { /* no pixel of these passes the edge test */
          if (crow && ((x - r.es.min_x) & (CHK_SPAN - 1)) == 0)
            crow[(x - r.es.min_x) >> CHK_SHIFT] = make_float4(w0, w1, w2, zf);
          w0 += r.es.e_dx[0];
          w1 += r.es.e_dx[1];
          w2 += r.es.e_dx[2];
          zf += fz.dx;
        }
        for (; x <= xb; x++) {
          if (crow && ((x - r.es.min_x) & (CHK_SPAN - 1)) == 0)
            crow[(x - r.es.min_x) >> CHK_SHIFT] = make_float4(w0, w1, w2, zf);
          float z;
          bool hit;
          if (flat) {
            z = zf;
            hit = w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f && z == z;
          } else {
            hit = frag_depth(r, w0, w1, w2, z);
          }
          if (hit) {
            unsigned long long key = ((unsigned long long)depth_key(z) << 32) | seq;
            if (MODE == 0) {
              vis_update(vis, rowbase + x, key);
            } else if (vis[rowbase + x] == key) {
              wbuf[rowbase + x] = make_float4(w0, w1, w2, z);
            }
          }
          w0 += r.es.e_dx[0];
          w1 += r.es.e_dx[1];
          w2 += r.es.e_dx[2];
          zf += fz.dx;
        }
      }
    }
    __syncthreads();
  }
}

void launch_raster(const FrameArgs &fa, hipStream_t s) {
  uint32_t cap = fa.subtri_cap;
  uint32_t grid = (cap + WG - 1) / WG;
  uint32_t lim = fa.raster_grid ? fa.raster_grid : 8192;
  if (grid > lim)
    grid = lim;
  if (grid == 0)
    grid = 1;
  launch(k_raster_small, dim3(grid), dim3(WG), 0, s, fa);
  launch(k_raster_big<0>, dim3(8192), dim3(BIG_WG), 0, s, fa);
  launch(k_raster_big<1>, dim3(8192), dim3(BIG_WG), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* order-free path: classification, setup and small-triangle raster fused */
/* --------------------------------------------------------------------- */

/* Queue a big screen triangle or walk a small one; true if it reached the
 * raster (a non-empty clamped bbox and a non-zero area). */
DEV bool fused_setup(const FrameArgs &fa, const ViewGPU &vw, const SubTri &st, uint32_t e,
                     uint32_t okey, RasterTri &r, bool &walk) {
  walk = false;
  if (!edge_setup(st.sx, st.sy, vw.w, vw.h, r.es))
    return false;
  if (bbox_area(r.es) > (long long)fa.small_area) {
    uint32_t q = atomicAdd(&fa.counters[CNT_BIG], 1u);
    uint32_t n = (uint32_t)(r.es.max_y - r.es.min_y + 1) * chk_per_row(r.es);
    uint32_t off = atomicAdd(&fa.counters[CNT_CHK], n);
    if (off > fa.chk_cap || n > fa.chk_cap - off) {
      atomicOr(&fa.counters[CNT_OVERFLOW], 32u);
      off = ~0u;
    }
    fa.big_queue[q] = e;
    fa.big_chk[q] = off;
    fa.big_off[okey] = off;
    if (off != ~0u) {
      /* its chunks of 64 bbox rows for k_big_rows */
      uint32_t nch = (uint32_t)(r.es.max_y - r.es.min_y + 64) >> 6;
      uint32_t cb = atomicAdd(&fa.counters[CNT_BIGCHUNK], nch);
      if (cb > fa.chk_cap || nch > fa.chk_cap - cb)
        atomicOr(&fa.counters[CNT_OVERFLOW], 32u);
      else
        for (uint32_t j = 0; j < nch; j++)
          fa.big_map[cb + j] = make_uint2(q, j);
    }
    return true;
  }
  if (edges_nondegenerate(r.es)) {
    r.inv_e[0] = r.inv_e[1] = r.inv_e[2] = 1.0f; /* min_d implied, see above */
  } else {
    r.inv_e[0] = inv_edge_len(r.es.e_dx[0], r.es.e_dy[0]);
    r.inv_e[1] = inv_edge_len(r.es.e_dx[1], r.es.e_dy[1]);
    r.inv_e[2] = inv_edge_len(r.es.e_dx[2], r.es.e_dy[2]);
  }
  r.sz[0] = st.sz[0];
  r.sz[1] = st.sz[1];
  r.sz[2] = st.sz[2];
  walk = true;
  return true;
}

DEV bool fused_raster(const FrameArgs &fa, const ViewGPU &vw, const SubTri &st, bool flat,
                      uint32_t e, uint32_t okey) {
  RasterTri r;
  bool walk;
  bool emitted = fused_setup(fa, vw, st, e, okey, r, walk);
  if (walk)
    walk_small_vis(vw, r, flat, (unsigned long long)e + 1ull);
  return emitted;
}

/* Screen triangles that reached the raster, counted per workgroup into
 * blk_counts (summed by k_finish): one same-address atomic per wave from
 * every workgroup of the grid would serialize at the memory side. */
DEV void count_subtris(uint32_t *blk_counts, bool emitted) {
  __shared__ uint32_t wave_sum[WG / 64];
  unsigned long long m = __ballot(emitted);
  if ((threadIdx.x & 63u) == 0)
    wave_sum[threadIdx.x >> 6] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    for (int w = 0; w < WG / 64; w++)
      c += wave_sum[w];
    blk_counts[blockIdx.x] = c;
  }
}

/* Per triangle of the batch: index checks, trivial reject and clip-space
 * cull (rasterizer.c:2469-2485), trivial accept or a clip-queue entry
 * (2492-2504), then for an accepted triangle the fan checks (2515-2537) and
 * the visibility walk or a big-triangle queue entry. */
DEV bool tri_work(const FrameArgs &fa, const DrawGPU *draws, const ViewGPU *views, uint32_t b,
                  uint32_t di) {
  /* draw-major grid: the draw (and so the view) is uniform over the
   * workgroup (scalar loads) */
  const DrawGPU &d = draws[di];
  const ViewGPU &vw = views[d.view];
  uint32_t k = (b - d.tri_wg) * WG + threadIdx.x;
  uint32_t t = d.tri_base + k;
  bool emitted = false, walk = false;
  RasterTri r;
  if (k < d.tri_count) {
    uint32_t i0 = d.ibuf[3 * k], i1 = d.ibuf[3 * k + 1], i2 = d.ibuf[3 * k + 2];
    if (i0 < d.vertex_count && i1 < d.vertex_count && i2 < d.vertex_count) {
      uint32_t v0 = d.vtx_base + i0, v1 = d.vtx_base + i1, v2 = d.vtx_base + i2;
      float4 a = fa.vpos[v0], b = fa.vpos[v1], c = fa.vpos[v2];
      /* the screen positions are read with the clip positions (one memory
       * round trip, not two); only accepted triangles use them */
      float4 sc[3] = {fa.vscr[v0], fa.vscr[v1], fa.vscr[v2]};
      bool cull = (d.flags & MOPH_DRAW_CULL) != 0;
      bool alive = !trivially_rejected(a, b, c);
      if (alive && cull && a.w > 0.0f && b.w > 0.0f && c.w > 0.0f) {
        float ex = b.x * a.w - a.x * b.w;
        float ey = b.y * a.w - a.y * b.w;
        float fx = c.x * a.w - a.x * c.w;
        float fy = c.y * a.w - a.y * c.w;
        if (ex * fy - ey * fx <= 0.0f)
          alive = false;
      }
      if (alive) {
        if (a.w > 0.0f && b.w > 0.0f && c.w > 0.0f && inside_all(a) && inside_all(b) &&
            inside_all(c)) {
          if (!(fabsf(a.w) < 1e-7f || fabsf(b.w) < 1e-7f || fabsf(c.w) < 1e-7f)) {
            SubTri st;
#pragma unroll
            for (int i = 0; i < 3; i++) {
              st.sx[i] = sc[i].x;
              st.sy[i] = sc[i].y;
              st.sz[i] = sc[i].z;
              st.iw[i] = sc[i].w;
            }
            float signed_area = (st.sx[1] - st.sx[0]) * (st.sy[2] - st.sy[0]) -
                                (st.sx[2] - st.sx[0]) * (st.sy[1] - st.sy[0]);
            if (!(cull && signed_area >= 0.0f)) {
              fa.tri_rec[t] = make_uint4(v0, v1, v2, di);
              emitted = fused_setup(fa, vw, st, (t << 4) | FUSED_WHOLE, t, r, walk);
            }
          }
        } else {
          uint32_t q = atomicAdd(&fa.counters[CNT_CLIP], 1u);
          if (q < fa.clip_cap) {
            fa.clip_queue[q] = t;
            fa.tri_rec[t] = make_uint4(q, 0u, 0u, di);
          } else {
            atomicOr(&fa.counters[CNT_OVERFLOW], 1u);
          }
        }
      }
    }
  }
  if (walk)
    walk_small_vis(vw, r, sm_flat(draw_sm(d)), ((unsigned long long)t << 4 | FUSED_WHOLE) + 1ull);
  return emitted;
}

__global__ __launch_bounds__(WG) void k_tri(FrameArgs fa) {
  count_subtris(fa.blk_counts, tri_work(fa, fa.draws, fa.views, blockIdx.x, fa.wg_tri[blockIdx.x]));
}

/* Chunk-culled batches: the leading tri_direct per-triangle workgroups
 * (draws not culled by chunks) one per workgroup as in k_tri, then the ones
 * k_cull_chunks kept, walked by the remaining workgroups as in
 * k_vertex_list; workgroup g's count goes to blk_counts[g]. */
__global__ __launch_bounds__(WG) void k_tri_list(FrameArgs fa, const DrawGPU *__restrict__ draws,
                                                 const ViewGPU *__restrict__ views,
                                                 const uint32_t *__restrict__ wg_tri,
                                                 const uint2 *__restrict__ list) {
  if (blockIdx.x < fa.tri_direct) {
    count_subtris(fa.blk_counts, tri_work(fa, draws, views, blockIdx.x, wg_tri[blockIdx.x]));
    return;
  }
  __shared__ uint32_t acc;
  if (threadIdx.x == 0)
    acc = 0;
  __syncthreads();
  uint32_t n = fa.counters[CNT_TLIST], c = 0, e = blockIdx.x - fa.tri_direct;
  uint32_t stride = gridDim.x - fa.tri_direct;
  uint2 cur = e < n ? list[e] : make_uint2(0u, 0u);
  for (; e < n; e += stride) {
    uint2 nxt = e + stride < n ? list[e + stride] : make_uint2(0u, 0u);
    c += tri_work(fa, draws, views, cur.x, cur.y) ? 1u : 0u;
    cur = nxt;
  }
  if (c)
    atomicAdd(&acc, c);
  __syncthreads();
  if (threadIdx.x == 0)
    fa.blk_counts[blockIdx.x] = acc;
}

/* Fan triangles of the clipped polygons (after k_clip): the same checks and
 * raster as k_tri, one lane per polygon. */
#define FAN_MAX (MOPH_MAX_POLY - 2)

__global__ __launch_bounds__(WG) void k_clip_raster(FrameArgs fa, uint32_t *blk_counts) {
  __shared__ uint32_t acc;
  if (threadIdx.x == 0)
    acc = 0;
  __syncthreads();
  uint32_t nq = fa.counters[CNT_CLIP];
  if (nq > fa.clip_cap)
    nq = fa.clip_cap;
  uint32_t items = nq * FAN_MAX, emitted = 0;
  for (uint32_t it = blockIdx.x * WG + threadIdx.x; it < items; it += gridDim.x * WG) {
    uint32_t q = it / FAN_MAX;
    int i = (int)(it % FAN_MAX) + 1; /* fan triangle poly[0], poly[i], poly[i+1] */
    int n = (int)fa.cnt_poly[q];
    if (i >= n - 1)
      continue;
    uint32_t t = fa.clip_queue[q];
    const DrawGPU &d = fa.draws[fa.tri_rec[t].w];
    const ViewGPU &vw = fa.views[d.view];
    size_t base = (size_t)q * MOPH_MAX_POLY;
    if (!fan_ok(fa.cpos[base], fa.cpos[base + i], fa.cpos[base + i + 1],
                (d.flags & MOPH_DRAW_CULL) != 0, false, vw.w, vw.h))
      continue;
    uint32_t e = (t << 4) | (uint32_t)(i - 1);
    SubTri st;
    fused_subtri(fa, e, st);
    if (fused_raster(fa, vw, st, sm_flat(draw_sm(d)), e,
                     fa.n_tris + q * 16u + (uint32_t)(i - 1)))
      emitted++;
  }
  if (emitted)
    atomicAdd(&acc, emitted);
  __syncthreads();
  if (threadIdx.x == 0)
    blk_counts[blockIdx.x] = acc;
}

void launch_fused_setup(const FrameArgs &fa, hipStream_t s) {
  if (!fa.n_tris)
    return;
  if (fa.tri_list)
    launch(k_tri_list, dim3(fa.tri_grid), dim3(WG), 0, s, fa, fa.draws, fa.views, fa.wg_tri,
           (const uint2 *)fa.tri_list);
  else
    launch(k_tri, dim3(fa.n_wg_tri), dim3(WG), 0, s, fa);
  launch_clip(fa, s);
  /* blk_counts holds tri_grid + CLIP_RASTER_GRID entries */
  launch(k_clip_raster, dim3(CLIP_RASTER_GRID), dim3(WG), 0, s, fa,
                     fa.blk_counts + fa.tri_grid);
  /* the per-workgroup counts are summed by k_finish */
}

/* workgroups of k_tri_list and k_vertex_list that are resident at once on
 * the device: the fixed grids of the list walks (a larger grid would start
 * a second round late, holding its share of the list) */
void cull_list_grids(uint32_t *tri_grid, uint32_t *vtx_grid) {
  hipDeviceProp_t prop;
  int dev = 0, nt = 0, nv = 0;
  *tri_grid = *vtx_grid = 1024;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nt, k_tri_list, WG, 0) == hipSuccess && nt > 0)
    *tri_grid = (uint32_t)(nt * prop.multiProcessorCount);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nv, k_vertex_list, WG, 0) == hipSuccess && nv > 0)
    *vtx_grid = (uint32_t)(nv * prop.multiProcessorCount);
}

/* Order-free path, big screen triangles: one wave per chunk of 64 bbox rows,
 * so the chunks of a tall triangle are walked side by side rather than one
 * after another by a single wave.  Lane l takes row l of the chunk: its row
 * start is the serial chain w_row += e_dy (rasterizer.c:2148-2150) replayed
 * from the bbox top, then the row is walked as in k_raster_big: checkpoints
 * every 16 columns, tests only over the columns whose edge values can reach
 * 0 (row_span), a visibility atomic per covered pixel.  The shading pass
 * finds a big triangle's edge values from the checkpoints. */
__global__ __launch_bounds__(64) void k_big_rows(FrameArgs fa) {
  /* an overflowed checkpoint store or chunk map can leave entries of earlier
   * batches in the map: the frame is run again, and nothing is read here */
  if (fa.counters[CNT_OVERFLOW] & 32u)
    return;
  uint32_t nch = fa.counters[CNT_BIGCHUNK];
  if (nch > fa.chk_cap)
    nch = fa.chk_cap;
  uint32_t nbig = fa.counters[CNT_BIG];
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    uint2 m = fa.big_map[c];
    if (m.x >= nbig) /* uniform */
      continue;
    uint32_t s = fa.big_queue[m.x], off = fa.big_chk[m.x];
    SubTri st;
    load_subtri(fa, s, st);
    const DrawGPU &dr = fa.draws[st.draw];
    const ViewGPU &vw = fa.views[dr.view];
    RasterTri r;
    if (off == ~0u || !load_raster_tri_fast(vw.w, vw.h, st, r)) /* uniform */
      continue;
    bool flat = sm_flat(draw_sm(dr)); /* uniform */
    FlatZ fz = {0.0f, 0.0f, 0.0f};
    if (flat)
      fz = flat_z_setup(r);
    int row = (int)m.y * 64 + (int)threadIdx.x;
    if (row > r.es.max_y - r.es.min_y)
      continue;
    /* registers, not the record (vis_update) */
    unsigned long long *const vis = vw.vis;
    const size_t pitch = (size_t)vw.w;
    const unsigned long long seq = (unsigned long long)s + 1ull;
    float w0 = r.es.w_row[0], w1 = r.es.w_row[1], w2 = r.es.w_row[2], zf = fz.row;
    for (int j = 0; j < row; j++) {
      w0 += r.es.e_dy[0];
      w1 += r.es.e_dy[1];
      w2 += r.es.e_dy[2];
      if (flat)
        zf += fz.dy;
    }
    size_t rowbase = (size_t)(r.es.min_y + row) * pitch;
    float4 *crow = fa.chk + off + (size_t)row * chk_per_row(r.es);
    int xa, xb;
    {
      float wr[3] = {w0, w1, w2};
      row_span(r, wr, xa, xb);
    }
    if (xa > xb)
      xb = r.es.min_x - 1; /* no pixel of this row passes */
    int x = r.es.min_x;
    for (; x < xa && x <= xb; x++) { /* no pixel of these passes the edge test */
      if (((x - r.es.min_x) & (CHK_SPAN - 1)) == 0)
        crow[(x - r.es.min_x) >> CHK_SHIFT] = make_float4(w0, w1, w2, zf);
      w0 += r.es.e_dx[0];
      w1 += r.es.e_dx[1];
      w2 += r.es.e_dx[2];
      zf += fz.dx;
    }
    for (; x <= xb; x++) {
      if (((x - r.es.min_x) & (CHK_SPAN - 1)) == 0)
        crow[(x - r.es.min_x) >> CHK_SHIFT] = make_float4(w0, w1, w2, zf);
      float z;
      bool hit;
      if (flat) {
        z = zf;
        hit = w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f && z == z;
      } else {
        hit = frag_depth(r, w0, w1, w2, z);
      }
      if (hit)
        vis_update(vis, rowbase + (size_t)x, ((unsigned long long)depth_key(z) << 32) | seq);
      w0 += r.es.e_dx[0];
      w1 += r.es.e_dx[1];
      w2 += r.es.e_dx[2];
      zf += fz.dx;
    }
  }
}

/* raster_grid sets the number of waves working through the chunks (0 =
 * 8192). */
void launch_raster_big(const FrameArgs &fa, hipStream_t s) {
  uint32_t g = fa.raster_grid ? fa.raster_grid : 8192;
  launch(k_big_rows, dim3(g), dim3(64), 0, s, fa);
}

/* --------------------------------------------------------------------- */
/* shading (rasterizer.c:194-348, 1596-1814, 1950-2108)                   */
/* --------------------------------------------------------------------- */

/* shadow_test_pcf, rasterizer.c:305-348 */
DEV float shadow_pcf(const ViewGPU &fa, V3 wp) {
  float4 lp = mat4_mul_vec4(fa.shadow_vp, wp.x, wp.y, wp.z, 1.0f);
  if (lp.w <= 0.0f)
    return 1.0f;
  float inv_w = 1.0f / lp.w;
  float ndcx = lp.x * inv_w;
  float ndcy = lp.y * inv_w;
  float ndcz = (lp.z * inv_w + 1.0f) * 0.5f;
  float u = (ndcx + 1.0f) * 0.5f * (float)fa.shadow_w;
  float v = (1.0f - ndcy) * 0.5f * (float)fa.shadow_h;
  int cx = cvt_i32(u);
  int cy = cvt_i32(v);
  const float bias = 0.005f;
  if (cx >= 1 && cx < fa.shadow_w - 1 && cy >= 1 && cy < fa.shadow_h - 1) {
    /* all nine taps inside the map: three rows of three neighbours */
    const uint32_t *row = fa.shadow_map + (size_t)(cy - 1) * (size_t)fa.shadow_w + (cx - 1);
    int lit = 0;
#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
#pragma unroll
      for (int dx = 0; dx < 3; dx++)
        lit += (ndcz <= key_depth(row[dx]) + bias) ? 1 : 0;
      row += fa.shadow_w;
    }
    return mop_div_recip((float)lit, 9.0f, 1.0f / 9.0f);
  }
  int lit = 0, total = 0;
  for (int dy = -1; dy <= 1; dy++) {
    for (int dx = -1; dx <= 1; dx++) {
      int sx = (int)((uint32_t)cx + (uint32_t)dx);
      int sy = (int)((uint32_t)cy + (uint32_t)dy);
      if (sx < 0 || sx >= fa.shadow_w || sy < 0 || sy >= fa.shadow_h) {
        lit++;
        total++;
        continue;
      }
      /* the map holds the pass's ordered keys; key_depth returns the stored
       * depth (a -0.0 comes back as +0.0, which compares the same here) */
      float shadow_z = key_depth(fa.shadow_map[(size_t)sy * (size_t)fa.shadow_w + (size_t)sx]);
      if (ndcz <= shadow_z + bias)
        lit++;
      total++;
    }
  }
  return (float)lit / (float)total;
}

/* ibl_sample, rasterizer.c:187-207 (indices clamped: a NaN direction would
 * read outside the table in the reference) */
DEV void ibl_sample(const float *data, int w, int h, float u, float v, float out[3]) {
  u = u - floorf(u);
  v = v - floorf(v);
  float fx = u * (float)(w - 1);
  float fy = v * (float)(h - 1);
  int x0 = cvt_i32(fx), y0 = cvt_i32(fy);
  if (x0 < 0 || x0 >= w)
    x0 = 0;
  if (y0 < 0 || y0 >= h)
    y0 = 0;
  int x1 = (x0 + 1) % w;
  int y1 = y0 + 1 < h ? y0 + 1 : h - 1;
  float sx = fx - (float)x0, sy = fy - (float)y0;
  const float *p00 = &data[((size_t)y0 * w + x0) * 4];
  const float *p10 = &data[((size_t)y0 * w + x1) * 4];
  const float *p01 = &data[((size_t)y1 * w + x0) * 4];
  const float *p11 = &data[((size_t)y1 * w + x1) * 4];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float top = p00[c] * (1.0f - sx) + p10[c] * sx;
    float bot = p01[c] * (1.0f - sx) + p11[c] * sx;
    out[c] = top * (1.0f - sy) + bot * sy;
  }
}

/* dir_to_uv, rasterizer.c:210-216 */
DEV void dir_to_uv(float dx, float dy, float dz, float rot, float &u, float &v) {
  const float two_pi = 2.0f * 3.14159265f, pi = 3.14159265f;
  float phi = mop_atan2f(dz, dx) + rot;
  float cdy = dy < -1.0f ? -1.0f : (dy > 1.0f ? 1.0f : dy);
  u = mop_div_recip(phi, two_pi, 1.0f / two_pi) + 0.5f;
  v = mop_div_recip(mop_asinf(cdy), pi, 1.0f / pi) + 0.5f;
}

DEV void ibl_irradiance(const ViewGPU &fa, V3 n, float out[3]) {
  out[0] = out[1] = out[2] = 0.0f;
  float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
  if (len < 1e-6f)
    return;
  float u, v;
  float il = 1.0f / len; /* the three quotients n / len share it */
  dir_to_uv(mop_div_recip(n.x, len, il), mop_div_recip(n.y, len, il), mop_div_recip(n.z, len, il),
            fa.ibl_rot, u, v);
  ibl_sample(fa.irr, fa.irr_w, fa.irr_h, u, v, out);
  out[0] *= fa.ibl_int;
  out[1] *= fa.ibl_int;
  out[2] *= fa.ibl_int;
}

DEV void ibl_prefiltered(const ViewGPU &fa, V3 refl, float roughness, float out[3]) {
  out[0] = out[1] = out[2] = 0.0f;
  float len = sqrtf(refl.x * refl.x + refl.y * refl.y + refl.z * refl.z);
  if (len < 1e-6f)
    return;
  int level = cvt_i32(roughness * (float)(fa.pf_levels - 1));
  if (level >= fa.pf_levels)
    level = fa.pf_levels - 1;
  if (level < 0)
    level = 0;
  size_t offset = 0;
  for (int l = 0; l < level; l++) {
    int lw = fa.pf_w >> l;
    if (lw < 1)
      lw = 1;
    int lh = fa.pf_h >> l;
    if (lh < 1)
      lh = 1;
    offset += (size_t)lw * lh;
  }
  int lw = fa.pf_w >> level;
  if (lw < 1)
    lw = 1;
  int lh = fa.pf_h >> level;
  if (lh < 1)
    lh = 1;
  float u, v;
  float il = 1.0f / len;
  dir_to_uv(mop_div_recip(refl.x, len, il), mop_div_recip(refl.y, len, il),
            mop_div_recip(refl.z, len, il), fa.ibl_rot, u, v);
  ibl_sample(fa.pf + offset * 4, lw, lh, u, v, out);
  out[0] *= fa.ibl_int;
  out[1] *= fa.ibl_int;
  out[2] *= fa.ibl_int;
}

DEV void ibl_brdf(const ViewGPU &fa, float ndotv, float roughness, float &scale,
                  float &bias) {
  if (ndotv < 0.0f)
    ndotv = 0.0f;
  if (ndotv > 1.0f)
    ndotv = 1.0f;
  if (roughness < 0.0f)
    roughness = 0.0f;
  if (roughness > 1.0f)
    roughness = 1.0f;
  int n = fa.lut_size;
  int x = cvt_i32(ndotv * (float)(n - 1) + 0.5f);
  int y = cvt_i32(roughness * (float)(n - 1) + 0.5f);
  if (x >= n)
    x = n - 1;
  if (y >= n)
    y = n - 1;
  if (x < 0)
    x = 0;
  if (y < 0)
    y = 0;
  size_t idx = ((size_t)y * n + x) * 4;
  scale = fa.lut[idx + 0];
  bias = fa.lut[idx + 1];
}

DEV V3 light_pos(const MopHipLight &l) { return V3{l.position.x, l.position.y, l.position.z}; }
DEV V3 light_dir(const MopHipLight &l) { return V3{l.direction.x, l.direction.y, l.direction.z}; }

/* distance attenuation shared by point/spot (rasterizer.c:1625-1635) */
DEV float atten(const MopHipLight &l, float dist) {
  float a;
  if (l.range > 0.0f) {
    float r = dist / l.range;
    a = 1.0f - r;
    if (a < 0.0f)
      a = 0.0f;
    a *= a;
  } else {
    float d2 = dist * dist;
    if (d2 < 0.01f)
      d2 = 0.01f;
    a = 1.0f / d2;
  }
  return a;
}

struct ShadeOut {
  float r, g, b;
};

/* compute_multi_light_rgb (rasterizer.c:1596-1688) fused with
 * compute_multi_specular_ggx (rasterizer.c:1701-1814): the two loops touch the
 * same lights in the same order and accumulate into separate sums, so one
 * loop with the PCF lookup shared gives the same values.  Every operation is
 * the reference's, in its order, IEEE-rounded and unfused, so the sums come
 * out bit-identical. */
DEV void light_loop(const MopHipLight *L, uint32_t count, const ViewGPU &fa, V3 normal, V3 wp,
                    V3 view_dir, float ndotv, float alpha2, float k, float g1v, float f0_r,
                    float f0_g, float f0_b, float lit[3], float spec[3]) {
  for (uint32_t i = 0; i < count; i++) {
    const MopHipLight &l = L[i];
    if (!l.active)
      continue;
    /* diffuse part */
    float ndotl_d = 0.0f, att_d = 1.0f, spot = 1.0f;
    /* specular part */
    V3 l_dir = V3{0.0f, 0.0f, 0.0f};
    float att_s = 1.0f;
    if (l.type == MOP_HIP_LIGHT_DIRECTIONAL) {
      V3 dir = v3norm(light_dir(l));
      ndotl_d = -v3dot(normal, dir);
      l_dir = V3{-dir.x, -dir.y, -dir.z};
    } else if (l.type == MOP_HIP_LIGHT_POINT || l.type == MOP_HIP_LIGHT_SPOT) {
      V3 to_light = v3sub(light_pos(l), wp);
      float dist = v3len(to_light);
      if (dist < 1e-6f)
        dist = 1e-6f;
      V3 dir = v3scale(to_light, 1.0f / dist);
      ndotl_d = v3dot(normal, dir);
      if (l.type == MOP_HIP_LIGHT_SPOT) {
        V3 spot_dir = v3norm(light_dir(l));
        float cos_angle = -v3dot(dir, spot_dir);
        if (cos_angle < l.spot_outer_cos) {
          spot = 0.0f;
        } else if (cos_angle < l.spot_inner_cos) {
          float range_s = l.spot_inner_cos - l.spot_outer_cos;
          if (range_s > 1e-6f) {
            float t = (cos_angle - l.spot_outer_cos) / range_s;
            spot = t * t * (3.0f - 2.0f * t);
          }
        }
      }
      att_d = atten(l, dist);
      att_s = att_d;
      l_dir = dir;
    }
    if (ndotl_d < 0.0f)
      ndotl_d = 0.0f;
    bool dir_light = (l.type == MOP_HIP_LIGHT_DIRECTIONAL);
    float shadow = 1.0f;
    if (dir_light && fa.shadow_on)
      shadow = shadow_pcf(fa, wp);
    float contrib = ndotl_d * l.intensity * att_d * spot * shadow;
    lit[0] += contrib * l.color.r;
    lit[1] += contrib * l.color.g;
    lit[2] += contrib * l.color.b;

    float ndotl = v3dot(normal, l_dir);
    if (ndotl <= 0.0f)
      continue;
    V3 h = v3norm(v3add(l_dir, view_dir));
    float ndoth = v3dot(normal, h);
    if (ndoth < 0.0f)
      ndoth = 0.0f;
    float vdoth = v3dot(view_dir, h);
    if (vdoth < 0.0f)
      vdoth = 0.0f;
    float ndoth2 = ndoth * ndoth;
    float denom_d = ndoth2 * (alpha2 - 1.0f) + 1.0f;
    float D = alpha2 / (3.14159265358979323846f * denom_d * denom_d);
    float g1l = ndotl / (ndotl * (1.0f - k) + k);
    float G = g1v * g1l;
    float omv = 1.0f - vdoth;
    float omv2 = omv * omv;
    float omv5 = omv2 * omv2 * omv;
    float fr_r = f0_r + (1.0f - f0_r) * omv5;
    float fr_g = f0_g + (1.0f - f0_g) * omv5;
    float fr_b = f0_b + (1.0f - f0_b) * omv5;
    float spec_denom = 4.0f * ndotl * ndotv;
    if (spec_denom < 1e-6f)
      spec_denom = 1e-6f;
    float spec_term = D * G / spec_denom;
    float weight = spec_term * l.intensity * att_s * ndotl * shadow *
                   3.14159265358979323846f;
    spec[0] += weight * fr_r * l.color.r;
    spec[1] += weight * fr_g * l.color.g;
    spec[2] += weight * fr_b * l.color.b;
  }
}

DEV void lights_eval(const FrameArgs &frame, const ViewGPU &fa, const DrawGPU &d, V3 normal,
                     V3 wp, V3 view_dir,
                     float base_r, float base_g, float base_b, float lit[3], float spec[3]) {
  lit[0] = lit[1] = lit[2] = d.ambient;
  spec[0] = spec[1] = spec[2] = 0.0f;
  float roughness = d.roughness, metallic = d.metallic;
  float alpha = roughness * roughness;
  float alpha2 = alpha * alpha;
  if (alpha2 < 1e-7f)
    alpha2 = 1e-7f;
  float ndotv = v3dot(normal, view_dir);
  if (ndotv < 1e-4f)
    ndotv = 1e-4f;
  float f0_r = 0.04f * (1.0f - metallic) + base_r * metallic;
  float f0_g = 0.04f * (1.0f - metallic) + base_g * metallic;
  float f0_b = 0.04f * (1.0f - metallic) + base_b * metallic;
  float k = alpha * 0.5f;
  float g1v = ndotv / (ndotv * (1.0f - k) + k);
  /* the draws of one view share one light list: when every lane of the wave
   * has the same list, its address is wave-uniform and the light records are
   * read with scalar loads */
  uint32_t off = d.light_off, cnt = d.light_count;
  uint32_t off_u = __builtin_amdgcn_readfirstlane(off);
  uint32_t cnt_u = __builtin_amdgcn_readfirstlane(cnt);
  if (__all(off == off_u && cnt == cnt_u))
    light_loop(frame.lights + off_u, cnt_u, fa, normal, wp, view_dir, ndotv, alpha2, k, g1v,
               f0_r, f0_g, f0_b, lit, spec);
  else
    light_loop(frame.lights + off, cnt, fa, normal, wp, view_dir, ndotv, alpha2, k, g1v, f0_r,
               f0_g, f0_b, lit, spec);
}

/* Colour of one fragment of screen triangle st with edge values w0..w2:
 * perspective-correct interpolation, multi-light Lambert + GGX + PCF + IBL,
 * composition and clamp (rasterizer.c:1950-2092).  Also returns z. */
DEV void shade_fragment(const FrameArgs &frame, const ViewGPU &fa, const SubTri &st,
                        const RasterTri &r, const float4 *A0, const float4 *A1,
                        const float4 *A2, float w0, float w1, float w2, float &z, float &pr,
                        float &pg, float &pb) {
  float b0 = w0 * r.es.inv_area;
  float b1 = w1 * r.es.inv_area;
  float b2 = w2 * r.es.inv_area;
  z = b0 * st.sz[0] + b1 * st.sz[1] + b2 * st.sz[2];

  const DrawGPU &d = frame.draws[st.draw];
  float pc0 = b0 * st.iw[0];
  float pc1 = b1 * st.iw[1];
  float pc2 = b2 * st.iw[2];
  float inv_pc = 1.0f / (pc0 + pc1 + pc2);
  pc0 *= inv_pc;
  pc1 *= inv_pc;
  pc2 *= inv_pc;
  /* attribute k of vertex v: n=(A0.x,A0.y,A0.z) wp=(A0.w,A1.x,A1.y) c=(A1.z,A1.w,A2.x,A2.y) */
  V3 n = V3{pc0 * A0[0].x + pc1 * A0[1].x + pc2 * A0[2].x,
            pc0 * A0[0].y + pc1 * A0[1].y + pc2 * A0[2].y,
            pc0 * A0[0].z + pc1 * A0[1].z + pc2 * A0[2].z};
  n = v3norm(n);
  float fr = pc0 * A1[0].z + pc1 * A1[1].z + pc2 * A1[2].z;
  float fg = pc0 * A1[0].w + pc1 * A1[1].w + pc2 * A1[2].w;
  float fb_ = pc0 * A2[0].x + pc1 * A2[1].x + pc2 * A2[2].x;
  V3 wp = V3{pc0 * A0[0].w + pc1 * A0[1].w + pc2 * A0[2].w,
             pc0 * A1[0].x + pc1 * A1[1].x + pc2 * A1[2].x,
             pc0 * A1[0].y + pc1 * A1[1].y + pc2 * A1[2].y};
  V3 eye = V3{d.cam_eye[0], d.cam_eye[1], d.cam_eye[2]};
  V3 view_dir = v3norm(v3sub(eye, wp));

  float lit[3], spec[3];
  lights_eval(frame, fa, d, n, wp, view_dir, fr, fg, fb_, lit, spec);
  float metallic = d.metallic, roughness = d.roughness;
  float diffuse_scale = 1.0f - metallic;
  float env_r = 0.0f, env_g = 0.0f, env_b = 0.0f;
  float irr[3] = {0.0f, 0.0f, 0.0f};
  bool has_irr = fa.ibl_on && fa.irr;
  if (has_irr)
    ibl_irradiance(fa, n, irr);
  {
    float vdn = v3dot(view_dir, n);
    V3 refl = V3{2.0f * vdn * n.x - view_dir.x, 2.0f * vdn * n.y - view_dir.y,
                 2.0f * vdn * n.z - view_dir.z};
    float ndv = v3dot(n, view_dir);
    if (ndv < 0.0f)
      ndv = 0.0f;
    float f0_r = 0.04f * (1.0f - metallic) + fr * metallic;
    float f0_g = 0.04f * (1.0f - metallic) + fg * metallic;
    float f0_b = 0.04f * (1.0f - metallic) + fb_ * metallic;
    if (fa.ibl_on && fa.pf && fa.lut) {
      float pf[3];
      ibl_prefiltered(fa, refl, roughness, pf);
      float bs = 1.0f, bb = 0.0f;
      ibl_brdf(fa, ndv, roughness, bs, bb);
      env_r = pf[0] * (f0_r * bs + bb);
      env_g = pf[1] * (f0_g * bs + bb);
      env_b = pf[2] * (f0_b * bs + bb);
    } else if (metallic > 0.01f) {
      float sky_t = refl.y * 0.5f + 0.5f;
      if (sky_t < 0.0f)
        sky_t = 0.0f;
      if (sky_t > 1.0f)
        sky_t = 1.0f;
      float sky_rv = 0.35f, sky_gv = 0.38f, sky_bv = 0.45f;
      float gnd_rv = 0.08f, gnd_gv = 0.07f, gnd_bv = 0.06f;
      float he_r = gnd_rv + sky_t * (sky_rv - gnd_rv);
      float he_g = gnd_gv + sky_t * (sky_gv - gnd_gv);
      float he_b = gnd_bv + sky_t * (sky_bv - gnd_bv);
      float rough_mix = roughness * roughness;
      float avg_env = (he_r + he_g + he_b) * (1.0f / 3.0f);
      he_r += rough_mix * (avg_env - he_r);
      he_g += rough_mix * (avg_env - he_g);
      he_b += rough_mix * (avg_env - he_b);
      float om = 1.0f - ndv;
      float om2 = om * om;
      float om5 = om2 * om2 * om;
      float fe_r = f0_r + (1.0f - f0_r) * om5;
      float fe_g = f0_g + (1.0f - f0_g) * om5;
      float fe_b = f0_b + (1.0f - f0_b) * om5;
      env_r = he_r * fe_r * d.ambient * 3.0f;
      env_g = he_g * fe_g * d.ambient * 3.0f;
      env_b = he_b * fe_b * d.ambient * 3.0f;
    }
  }
  if (has_irr) {
    pr = fr * (lit[0] + irr[0]) * diffuse_scale + env_r + spec[0];
    pg = fg * (lit[1] + irr[1]) * diffuse_scale + env_g + spec[1];
    pb = fb_ * (lit[2] + irr[2]) * diffuse_scale + env_b + spec[2];
  } else {
    pr = fr * lit[0] * diffuse_scale + env_r + spec[0];
    pg = fg * lit[1] * diffuse_scale + env_g + spec[1];
    pb = fb_ * lit[2] * diffuse_scale + env_b + spec[2];
  }
  if (pr < 0.0f)
    pr = 0.0f;
  if (pg < 0.0f)
    pg = 0.0f;
  if (pb < 0.0f)
    pb = 0.0f;
}

DEV float clamp01f(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

/* compute_multi_light (rasterizer.c:1498-1589): scalar intensity of the flat
 * path at one point; spot lights fall off as 1/(1+d^2). */
DEV float multi_light_scalar(const FrameArgs &frame, const ViewGPU &fa, const DrawGPU &d,
                             V3 normal, V3 wp) {
  float total = d.ambient;
  const MopHipLight *L = frame.lights + d.light_off;
  for (uint32_t i = 0; i < d.light_count; i++) {
    const MopHipLight &l = L[i];
    if (!l.active)
      continue;
    float ndotl = 0.0f, att = 1.0f, spot = 1.0f;
    if (l.type == MOP_HIP_LIGHT_DIRECTIONAL) {
      V3 dir = v3norm(light_dir(l));
      ndotl = -v3dot(normal, dir);
    } else if (l.type == MOP_HIP_LIGHT_POINT || l.type == MOP_HIP_LIGHT_SPOT) {
      V3 to_light = v3sub(light_pos(l), wp);
      float dist = v3len(to_light);
      if (dist < 1e-6f)
        dist = 1e-6f;
      V3 dir = v3scale(to_light, 1.0f / dist);
      ndotl = v3dot(normal, dir);
      if (l.type == MOP_HIP_LIGHT_SPOT) {
        V3 spot_dir = v3norm(light_dir(l));
        float cos_angle = -v3dot(dir, spot_dir);
        if (cos_angle < l.spot_outer_cos) {
          spot = 0.0f;
        } else if (cos_angle < l.spot_inner_cos) {
          float range_s = l.spot_inner_cos - l.spot_outer_cos;
          if (range_s > 1e-6f) {
            float t = (cos_angle - l.spot_outer_cos) / range_s;
            spot = t * t * (3.0f - 2.0f * t);
          }
        }
        if (l.range > 0.0f) {
          float r = dist / l.range;
          att = 1.0f - r;
          if (att < 0.0f)
            att = 0.0f;
          att *= att;
        } else {
          att = 1.0f / (1.0f + dist * dist);
        }
      } else {
        att = atten(l, dist);
      }
    }
    if (ndotl < 0.0f)
      ndotl = 0.0f;
    float shadow = 1.0f;
    if (l.type == MOP_HIP_LIGHT_DIRECTIONAL && fa.shadow_on)
      shadow = shadow_pcf(fa, wp);
    total += ndotl * l.intensity * att * spot * shadow;
  }
  return clamp01f(total);
}

DEV V3 draw_light_dir(const DrawGPU &d) { return V3{d.light_dir[0], d.light_dir[1], d.light_dir[2]}; }

/* Constant colour of a flat-shaded screen triangle: face normal from the sum
 * of the fan vertices' normals, colour averaged; lit draws evaluate
 * compute_multi_light at the centroid and add the metallic term
 * (rasterizer.c:2556-2596), unlit draws use light_dir Lambert
 * (rasterizer.c:1199-1220). */
DEV void flat_color(const FrameArgs &frame, const ViewGPU &fa, const SubTri &st,
                    const DrawGPU &d, int sm, const float4 *A0, const float4 *A1,
                    const float4 *A2, float &cr, float &cg, float &cb) {
  V3 fn = v3norm(V3{(A0[0].x + A0[1].x + A0[2].x), (A0[0].y + A0[1].y + A0[2].y),
                    (A0[0].z + A0[1].z + A0[2].z)});
  float avg_r = (A1[0].z + A1[1].z + A1[2].z) * (1.0f / 3.0f);
  float avg_g = (A1[0].w + A1[1].w + A1[2].w) * (1.0f / 3.0f);
  float avg_b = (A2[0].x + A2[1].x + A2[2].x) * (1.0f / 3.0f);
  if (sm == MOPH_SM_LIT_FLAT || sm == MOPH_SM_LIT_SMOOTH) { /* lit (wireframe: any mode) */
    V3 wp = V3{(A0[0].w + A0[1].w + A0[2].w) / 3.0f, (A1[0].x + A1[1].x + A1[2].x) / 3.0f,
               (A1[0].y + A1[1].y + A1[2].y) / 3.0f};
    float lighting = multi_light_scalar(frame, fa, d, fn, wp);
    float diffuse_scale = 1.0f - d.metallic;
    float env_flat = d.ambient * d.metallic * 0.3f;
    cr = avg_r * (lighting * diffuse_scale + env_flat);
    cg = avg_g * (lighting * diffuse_scale + env_flat);
    cb = avg_b * (lighting * diffuse_scale + env_flat);
  } else {
    float ndotl = v3dot(fn, v3norm(draw_light_dir(d)));
    if (ndotl < 0.0f)
      ndotl = 0.0f;
    float lighting = clamp01f(d.ambient + (1.0f - d.ambient) * ndotl);
    cr = avg_r * lighting;
    cg = avg_g * lighting;
    cb = avg_b * lighting;
  }
  if (cr < 0.0f)
    cr = 0.0f;
  if (cg < 0.0f)
    cg = 0.0f;
  if (cb < 0.0f)
    cb = 0.0f;
}

/* Colour of one fragment of the unlit smooth path: perspective-correct
 * normal and colour, light_dir Lambert plus the fixed Blinn-Phong term
 * (rasterizer.c:1368-1419). */
DEV void unlit_smooth_color(const FrameArgs &frame, const SubTri &st, const RasterTri &r,
                            const DrawGPU &d, float w0, float w1, float w2, float &pr, float &pg,
                            float &pb) {
  float b0 = w0 * r.es.inv_area;
  float b1 = w1 * r.es.inv_area;
  float b2 = w2 * r.es.inv_area;
  float pc0 = b0 * st.iw[0];
  float pc1 = b1 * st.iw[1];
  float pc2 = b2 * st.iw[2];
  float inv_pc = 1.0f / (pc0 + pc1 + pc2);
  pc0 *= inv_pc;
  pc1 *= inv_pc;
  pc2 *= inv_pc;
  float4 A0[3], A1[3], A2[3];
  load_tri_attr(frame, st, A0, A1, A2);
  V3 n = V3{pc0 * A0[0].x + pc1 * A0[1].x + pc2 * A0[2].x,
            pc0 * A0[0].y + pc1 * A0[1].y + pc2 * A0[2].y,
            pc0 * A0[0].z + pc1 * A0[1].z + pc2 * A0[2].z};
  n = v3norm(n);
  float fr = pc0 * A1[0].z + pc1 * A1[1].z + pc2 * A1[2].z;
  float fg = pc0 * A1[0].w + pc1 * A1[1].w + pc2 * A1[2].w;
  float fb_ = pc0 * A2[0].x + pc1 * A2[1].x + pc2 * A2[2].x;
  V3 nl = v3norm(draw_light_dir(d));
  float ndotl = v3dot(n, nl);
  if (ndotl < 0.0f)
    ndotl = 0.0f;
  float lit = d.ambient + (1.0f - d.ambient) * ndotl;
  V3 half_v = v3norm(V3{nl.x, nl.y + 1.0f, nl.z});
  float ndoth = v3dot(n, half_v);
  if (ndoth < 0.0f)
    ndoth = 0.0f;
  float spec = ndoth * ndoth;
  spec *= spec;
  spec *= spec;
  spec *= spec;
  spec *= spec;
  spec *= 0.25f;
  pr = fr * lit + spec;
  pg = fg * lit + spec;
  pb = fb_ * lit + spec;
  if (pr < 0.0f)
    pr = 0.0f;
  if (pg < 0.0f)
    pg = 0.0f;
  if (pb < 0.0f)
    pb = 0.0f;
}

DEV uint32_t find_view_by_tile(const ViewGPU *v, uint32_t n, uint32_t t) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (v[mid].tile_base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

DEV uint32_t find_view_by_stile(const ViewGPU *v, uint32_t n, uint32_t t) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (v[mid].stile_base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
 * XCD b % 8), each with its own L2.  xcd_tile(b, n) permutes n tile
 * numbers so that XCD x takes one contiguous run of tiles, in raster order:
 * neighbouring tiles share the triangles, vertices and halo pixels they read
 * and the partial cache lines they write in one L2. */
DEV uint32_t xcd_tile(uint32_t b, uint32_t n) {
  uint32_t x = b & 7u, j = b >> 3, q = n >> 3, r = n & 7u;
  return x * q + (x < r ? x : r) + j;
}

/* Pixel idx of a cleared frame that no triangle of its first segment
 * covers: the clear values of mop_sw_framebuffer_clear (rasterizer.c:114-134). */
DEV void shade_clear_px(const ViewGPU &fa, size_t idx) {
  if (!fa.shade_clear)
    return;
  if (fa.shade_clear == 1)
    reinterpret_cast<float4 *>(fa.hdr)[idx] =
        make_float4(fa.clear[0], fa.clear[1], fa.clear[2], fa.clear[3]);
  fa.depth[idx] = 1.0f;
  fa.ids[idx] = 0u;
}

/* Colour, depth and object id of the pixel (px, py) whose visibility key
 * names screen triangle s; false when the key does not name a triangle of
 * this batch that covers the pixel (shade_key_fetch). */
DEV bool shade_px(const FrameArgs &frame, const ViewGPU &fa, uint32_t s, int px, int py,
                  size_t idx, float4 &out, float &z, uint32_t &id) {
  KeyFetch k;
  if (!shade_key_fetch(frame, fa, s, px, py, idx, k))
    return false;
  const SubTri &st = k.st;
  const DrawGPU &d = frame.draws[st.draw];
  int sm = k.sm;
  bool flat = k.flat;
  const RasterTri &r = k.r;
  float w0 = k.w0, w1 = k.w1, w2 = k.w2, zf = k.zf;
  float pr, pg, pb;
  if (flat) {
    z = zf;
    flat_color(frame, fa, st, d, sm, k.A0, k.A1, k.A2, pr, pg, pb);
  } else {
    shade_fragment(frame, fa, st, r, k.A0, k.A1, k.A2, w0, w1, w2, z, pr, pg, pb);
  }
  out = make_float4(pr, pg, pb, 1.0f);
  id = d.object_id;
  return true;
}

/* One workgroup per shading tile of one view; tiles of all views of the
 * batch are numbered consecutively.  Every pixel's visibility entry is read
 * and, after the pixel is shaded, set back to the cleared key, so the buffer
 * is clear for the next segment or frame without a clearing pass.  All
 * stores come after all loads.  The kernel is bound by VALU issue and the
 * latency of its chain of dependent gathers.  It is compiled for 4, 5, 6
 * and 8 waves per SIMD (at most 128, 96, 80 and 64 VGPRs, the tighter ones
 * spilling); the autotuner picks one per batch (mop_hip_set_shade_waves),
 * five by default. */
template <int WAVES>
__global__ __launch_bounds__(1024, WAVES) void k_shade(FrameArgs frame) {
  uint32_t vi = find_view_by_stile(frame.views, frame.nviews, blockIdx.x);
  const ViewGPU &fa = frame.views[vi];
  uint32_t T = frame.shade_tile;
  uint32_t t = blockIdx.x - fa.stile_base;
  uint32_t tiles_x = ((uint32_t)fa.w + T - 1) / T;
  int px = (int)((t % tiles_x) * T + threadIdx.x % T);
  int py = (int)((t / tiles_x) * T + threadIdx.x / T);
  if (px >= fa.w || py >= fa.h)
    return;
  size_t idx = (size_t)py * (size_t)fa.w + (size_t)px;
  uint32_t s1 = (uint32_t)fa.vis[idx];
  float4 c;
  float z;
  uint32_t id;
  bool hit = s1 != 0 && shade_px(frame, fa, s1 - 1, px, py, idx, c, z, id);
  fa.vis[idx] = (unsigned long long)depth_key(1.0f) << 32;
  if (hit) {
    reinterpret_cast<float4 *>(fa.hdr)[idx] = c;
    fa.depth[idx] = z;
    fa.ids[idx] = id;
  } else {
    shade_clear_px(fa, idx); /* uncovered */
  }
}

/* End of a batch: add the per-workgroup screen-triangle counts (order-free
 * path) to the batch counters -- each workgroup sums a chunk and adds it
 * once -- and the last workgroup to finish copies the counters to every
 * view's host-visible counter slot. */
__global__ __launch_bounds__(WG) void k_finish(FrameArgs fa, uint32_t nblk, uint32_t chunk) {
  __shared__ uint32_t part[WG];
  __shared__ bool last;
  uint32_t lo = blockIdx.x * chunk, hi = lo + chunk < nblk ? lo + chunk : nblk;
  uint32_t c = 0;
  for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += 8 * WG) {
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) { /* eight loads in flight per thread */
      uint32_t i = i0 + (uint32_t)j * WG;
      v[j] = i < hi ? fa.blk_counts[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
      c += v[j];
  }
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = WG / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (part[0])
      atomicAdd(&fa.counters[CNT_SUBTRI], part[0]);
    __threadfence();
    last = atomicAdd(&fa.counters[CNT_TICKET], 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last)
    return;
  __threadfence();
  for (uint32_t v = threadIdx.x; v < fa.nviews; v += WG) {
    uint32_t *o = fa.views[v].cnt_out;
    for (int k = 0; k < CNT_N; k++)
      o[k] = __hip_atomic_load(&fa.counters[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    /* chunks of the culled draws that k_cull_chunks did not keep */
    o[CNT_CULLED] = fa.tri_list ? fa.n_wg_tri - fa.tri_direct -
                                      __hip_atomic_load(&fa.counters[CNT_TLIST], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                : 0u;
  }
  if (threadIdx.x == 0 && fa.sticky) {
    /* word 0: overflow bits; words 1-3: the largest clip-queue,
     * checkpoint-store and tile-bin needs seen (the host sizes the next
     * batches from them) */
    const int need[3] = {CNT_CLIP, CNT_CHK, CNT_BINS};
    uint32_t ovf =
        __hip_atomic_load(&fa.counters[CNT_OVERFLOW], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t acc = atomicOr(fa.sticky, ovf) | ovf;
    uint32_t mx[3];
    for (int k = 0; k < 3; k++) {
      uint32_t v =
          __hip_atomic_load(&fa.counters[need[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t old = atomicMax(fa.sticky + 1 + k, v);
      mx[k] = old > v ? old : v;
    }
    fa.sticky_out[1] = mx[0];
    fa.sticky_out[2] = mx[1];
    fa.sticky_out[3] = mx[2];
    fa.sticky_out[0] = acc;
  }
}

void launch_export_counters(const FrameArgs &fa, hipStream_t s) {
  if (!fa.nviews)
    return;
  uint32_t nblk = fa.fused ? fa.tri_grid + CLIP_RASTER_GRID : 0u;
  uint32_t grid = (nblk + 4095) / 4096;
  if (grid < 1)
    grid = 1;
  if (grid > 256)
    grid = 256;
  uint32_t chunk = (nblk + grid - 1) / grid;
  launch(k_finish, dim3(grid), dim3(WG), 0, s, fa, nblk, chunk);
}

void launch_shade(const FrameArgs &fa, hipStream_t s) {
  if (!fa.n_stiles)
    return;
  dim3 g(fa.n_stiles), b(fa.shade_tile * fa.shade_tile);
  switch (fa.shade_waves) {
  case 4: launch(k_shade<4>, g, b, 0, s, fa); break;
  case 6: launch(k_shade<6>, g, b, 0, s, fa); break;
  case 8: launch(k_shade<8>, g, b, 0, s, fa); break;
  default: launch(k_shade<5>, g, b, 0, s, fa); break;
  }
}

/* --------------------------------------------------------------------- */
/* clear (rasterizer.c:114-134), resolve (2770-2803), post (postprocess.c) */
/* --------------------------------------------------------------------- */

__global__ __launch_bounds__(WG) void k_clear(int npx, float r, float g, float b, float a,
                                              uint8_t *color, float *hdr, float *depth,
                                              uint32_t *ids, unsigned long long *vis) {
  int i = blockIdx.x * WG + threadIdx.x;
  if (i >= npx)
    return;
  uchar4 c;
  c.x = cvt_u8(r * 255.0f);
  c.y = cvt_u8(g * 255.0f);
  c.z = cvt_u8(b * 255.0f);
  c.w = cvt_u8(a * 255.0f);
  reinterpret_cast<uchar4 *>(color)[i] = c;
  reinterpret_cast<float4 *>(hdr)[i] = make_float4(r, g, b, a);
  depth[i] = 1.0f;
  ids[i] = 0u;
  vis[i] = (unsigned long long)depth_key(1.0f) << 32;
}

DEV uint32_t find_clear_view(const ClearViewGPU *v, uint32_t n, uint32_t px) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (v[mid].px_base <= px)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* mop_sw_framebuffer_clear (rasterizer.c:114-134) for every view of a
 * frame execution in one launch; the RGBA8 image is left alone (the tail
 * writes every pixel of it). */
__global__ __launch_bounds__(WG) void k_clear_views(const ClearViewGPU *views, uint32_t n,
                                                    uint32_t total) {
  uint32_t p = blockIdx.x * WG + threadIdx.x;
  if (p >= total)
    return;
  const ClearViewGPU &v = views[find_clear_view(views, n, p)];
  uint32_t i = p - v.px_base;
  reinterpret_cast<float4 *>(v.hdr)[i] = make_float4(v.clear[0], v.clear[1], v.clear[2],
                                                     v.clear[3]);
  v.depth[i] = 1.0f;
  v.ids[i] = 0u;
  v.vis[i] = (unsigned long long)depth_key(1.0f) << 32;
}

void launch_clear_views(const ClearViewGPU *views, uint32_t n, uint32_t total, hipStream_t s) {
  if (!n || !total)
    return;
  launch(k_clear_views, dim3((total + WG - 1) / WG), dim3(WG), 0, s, views, n, total);
}

void launch_clear(int npx, const float clear[4], uint8_t *color, float *hdr, float *depth,
                  uint32_t *ids, unsigned long long *vis, hipStream_t s) {
  launch(k_clear, dim3((npx + WG - 1) / WG), dim3(WG), 0, s, npx, clear[0],
                     clear[1], clear[2], clear[3], color, hdr, depth, ids, vis);
}

__global__ __launch_bounds__(WG) void k_vis_from_depth(int npx, const float *depth,
                                                       unsigned long long *vis) {
  int i = blockIdx.x * WG + threadIdx.x;
  if (i >= npx)
    return;
  float dz = depth[i];
  /* a NaN depth can never be beaten by `z < depth` */
  uint32_t k = (dz == dz) ? depth_key(dz) : 0u;
  vis[i] = (unsigned long long)k << 32;
}

void launch_vis_from_depth(int npx, const float *depth, unsigned long long *vis,
                           hipStream_t s) {
  launch(k_vis_from_depth, dim3((npx + WG - 1) / WG), dim3(WG), 0, s, npx, depth,
                     vis);
}

DEV float aces(float x) {
  float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
  float m = (x * (a * x + b)) / (x * (c * x + d) + e);
  return m < 0.0f ? 0.0f : (m > 1.0f ? 1.0f : m);
}

DEV uint32_t find_post_view(const PostViewGPU *v, uint32_t n, uint32_t t) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (v[mid].tile_base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* mop_sw_hdr_resolve for one pixel (rasterizer.c:2776-2803): background
 * (id 0) clamped, geometry ACES-tonemapped with the exposure; packed RGBA8. */
DEV uint32_t resolve_px(const PostViewGPU &v, size_t i) {
  float4 h = reinterpret_cast<const float4 *>(v.hdr)[i];
  uint32_t r, g, b, a;
  if (v.ids[i] == 0) {
    r = cvt_u8(fminf(h.x, 1.0f) * 255.0f + 0.5f);
    g = cvt_u8(fminf(h.y, 1.0f) * 255.0f + 0.5f);
    b = cvt_u8(fminf(h.z, 1.0f) * 255.0f + 0.5f);
    a = 255;
  } else {
    r = cvt_u8(aces(h.x * v.exposure) * 255.0f + 0.5f);
    g = cvt_u8(aces(h.y * v.exposure) * 255.0f + 0.5f);
    b = cvt_u8(aces(h.z * v.exposure) * 255.0f + 0.5f);
    a = cvt_u8(h.w * 255.0f + 0.5f);
  }
  return r | (g << 8) | (b << 16) | (a << 24);
}

DEV uint32_t ch(uint32_t c, int k) { return (c >> (8 * k)) & 255u; }

/* luma of a resolved RGBA8 pixel (rasterizer.c:2633-2636) */
DEV float luma_px(uint32_t c) {
  return 0.299f * (float)ch(c, 0) + 0.587f * (float)ch(c, 1) + 0.114f * (float)ch(c, 2);
}

DEV float srgb_enc(float c) {
  return c <= 0.0031308f ? c * 12.92f : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}

#define OUTLINE_CHROME 0xFFFD0000u

DEV bool sel_has(const uint32_t *sel, uint32_t n, uint32_t id) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (sel[mid] < id)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && sel[lo] == id;
}

/* one paint of the outline overlay on a channel value (overlay_builtin.c:
 * 1383-1395): (u8)(c (1 - alpha) + accent alpha) */
DEV uint32_t outline_paint(uint32_t c, uint32_t acc, float alpha) {
  return cvt_u8((float)c * (1.0f - alpha) + (float)acc * alpha);
}

/* The object outline of mop_overlay_builtin_outline (overlay_builtin.c:
 * 1309-1400) at one destination pixel.  The overlay scans the image in row
 * order; an edge pixel of a selected object paints the 5x5 block around it,
 * an edge pixel of another object paints itself, and every paint blends
 * into the value the earlier paints left.  Paints with one alpha commute, so
 * a destination's result depends only on how many selected edges of its
 * 5x5 block come before it in scan order, whether it is an edge itself, and
 * how many selected edges come after it.  flag: bit 0 edge, bit 1 selected;
 * (fx, fy) is the destination in the flag array, whose rows are `pitch`
 * bytes. */
DEV uint32_t outline_px(uint32_t c, const uint8_t *flag, int pitch, int fx, int fy,
                        const PostViewGPU &v) {
  int before = 0, after = 0;
  for (int dy = -2; dy <= 2; dy++)
    for (int dx = -2; dx <= 2; dx++) {
      if (flag[(fy + dy) * pitch + fx + dx] != 3)
        continue;
      if (dy < 0 || (dy == 0 && dx < 0))
        before++;
      else if (dy > 0 || dx > 0)
        after++;
    }
  uint32_t self = flag[fy * pitch + fx];
  if (!before && !after && self != 1 && self != 3)
    return c;
  uint32_t out = c & 0xff000000u; /* alpha is left alone */
  for (int k = 0; k < 3; k++) {
    uint32_t x = ch(c, k), acc = (v.accent >> (8 * k)) & 255u;
    for (int j = 0; j < before; j++)
      x = outline_paint(x, acc, v.alpha_sel);
    if (self == 3)
      x = outline_paint(x, acc, v.alpha_sel);
    else if (self == 1)
      x = outline_paint(x, acc, v.alpha_unsel);
    for (int j = 0; j < after; j++)
      x = outline_paint(x, acc, v.alpha_sel);
    out |= x << (8 * k);
  }
  return out;
}

/* The frame tail of one view tile: mop_sw_hdr_resolve, the object outline
 * when requested, then (when post effects are requested) mop_sw_fxaa
 * (rasterizer.c:2625-2761) and the per-pixel effects of
 * mop_postprocess_apply (postprocess.c:36-116).  FXAA reads only resolved
 * colours of the pixel's 3x3 neighbourhood, so each workgroup resolves its
 * 16x16 tile plus a one-pixel halo into LDS and never writes the
 * intermediate image; the outline of that 18x18 block needs the edge flags
 * of a 22x22 block, which need the ids of a 24x24 block.  Without fog or
 * vignette the colour effects are a per-channel table (v.lut). */
/* LDS of the outline: the ids of a 24x24 block and the edge flags of a
 * 22x22 block; a stand-in of one row when no view asks for the outline
 * (the code that indexes it is then never compiled in) */
template <bool OUTLINE> struct OutlineLds {
  uint32_t oid[24][24];
  uint8_t oflag[22 * 22];
};
template <> struct OutlineLds<false> {
  uint32_t oid[1][24];
  uint8_t oflag[4];
};

/* OUTLINE: some view of the launch asks for the outline (the LDS for its
 * ids and edge flags is only reserved then) */
template <bool OUTLINE>
__global__ __launch_bounds__(WG) void k_tail(const PostViewGPU *views, uint32_t nviews) {
  __shared__ uint32_t col[18][18];
  __shared__ float lum[18][18];
  __shared__ OutlineLds<OUTLINE> ol_lds;
  auto &oid = ol_lds.oid;
  uint8_t *oflag = ol_lds.oflag;
  uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const PostViewGPU &v = views[find_post_view(views, nviews, tile)];
  uint32_t t = tile - v.tile_base;
  uint32_t tiles_x = ((uint32_t)v.w + 15) / 16;
  int x0 = (int)(t % tiles_x) * 16, y0 = (int)(t / tiles_x) * 16;
  int x = x0 + (int)(threadIdx.x & 15);
  int y = y0 + (int)(threadIdx.x >> 4);
  const int w = v.w, h = v.h;
  const uint32_t effects = v.effects;
  const bool outline = OUTLINE && v.outline != 0;
  if (effects == 0 && !outline) {
    if (x < w && y < h) {
      size_t i = (size_t)y * (size_t)w + (size_t)x;
      reinterpret_cast<uint32_t *>(v.dst)[i] = resolve_px(v, i);
    }
    return; /* uniform: effects and outline are per view */
  }
  if (outline) {
    for (int i = (int)threadIdx.x; i < 24 * 24; i += WG) {
      int gx = x0 + i % 24 - 4, gy = y0 + i / 24 - 4;
      uint32_t id = 0;
      if (gx >= 0 && gx < w && gy >= 0 && gy < h)
        id = v.ids[(size_t)gy * (size_t)w + (size_t)gx];
      oid[i / 24][i % 24] = id;
    }
    __syncthreads();
    /* edge pixels: a non-background, non-chrome id with a 4-neighbour in the
     * image whose id differs and is not chrome (overlay_builtin.c:1342-1370) */
    for (int i = (int)threadIdx.x; i < 22 * 22; i += WG) {
      int fx = i % 22, fy = i / 22;
      int gx = x0 + fx - 3, gy = y0 + fy - 3;
      uint8_t f = 0;
      if (gx >= 0 && gx < w && gy >= 0 && gy < h) {
        int ox = fx + 1, oy = fy + 1;
        uint32_t id = oid[oy][ox];
        if (id != 0 && id < OUTLINE_CHROME) {
          bool edge = false;
          uint32_t nid;
          if (gx > 0) {
            nid = oid[oy][ox - 1];
            edge |= nid != id && nid < OUTLINE_CHROME;
          }
          if (gx < w - 1) {
            nid = oid[oy][ox + 1];
            edge |= nid != id && nid < OUTLINE_CHROME;
          }
          if (gy > 0) {
            nid = oid[oy - 1][ox];
            edge |= nid != id && nid < OUTLINE_CHROME;
          }
          if (gy < h - 1) {
            nid = oid[oy + 1][ox];
            edge |= nid != id && nid < OUTLINE_CHROME;
          }
          if (edge)
            f = sel_has(v.sel, v.n_sel, id) ? 3 : 1;
        }
      }
      oflag[i] = f;
    }
    __syncthreads();
  }
  for (int i = (int)threadIdx.x; i < 18 * 18; i += WG) {
    int lx = x0 + i % 18 - 1, ly = y0 + i / 18 - 1;
    uint32_t c = 0;
    if (lx >= 0 && lx < w && ly >= 0 && ly < h) {
      c = resolve_px(v, (size_t)ly * (size_t)w + (size_t)lx);
      /* chrome destinations are never painted (overlay_builtin.c:1380-1381) */
      if (outline && oid[i / 18 + 3][i % 18 + 3] < OUTLINE_CHROME)
        c = outline_px(c, oflag, 22, i % 18 + 2, i / 18 + 2, v);
    }
    col[i / 18][i % 18] = c;
    lum[i / 18][i % 18] = luma_px(c);
  }
  __syncthreads();
  if (x >= w || y >= h)
    return;
  const int tx = (int)(threadIdx.x & 15) + 1, ty = (int)(threadIdx.x >> 4) + 1;
  uint32_t cM = col[ty][tx];
  if (effects == 0) {
    reinterpret_cast<uint32_t *>(v.dst)[(size_t)y * (size_t)w + (size_t)x] = cM;
    return;
  }
  uint32_t o0 = ch(cM, 0), o1 = ch(cM, 1), o2 = ch(cM, 2), o3 = ch(cM, 3);
  if ((effects & MOP_HIP_POST_FXAA) && w >= 3 && h >= 3 && x >= 1 && x < w - 1 && y >= 1 &&
      y < h - 1) {
    const float EDGE_THRESHOLD = 0.0625f, EDGE_THRESHOLD_MIN = 0.004f, SUBPIX_QUALITY = 0.75f;
    float lM = lum[ty][tx];
    float lN = lum[ty - 1][tx];
    float lS = lum[ty + 1][tx];
    float lW = lum[ty][tx - 1];
    float lE = lum[ty][tx + 1];
    float range_min = lM;
    if (lN < range_min) range_min = lN;
    if (lS < range_min) range_min = lS;
    if (lW < range_min) range_min = lW;
    if (lE < range_min) range_min = lE;
    float range_max = lM;
    if (lN > range_max) range_max = lN;
    if (lS > range_max) range_max = lS;
    if (lW > range_max) range_max = lW;
    if (lE > range_max) range_max = lE;
    float range = range_max - range_min;
    if (!(range < EDGE_THRESHOLD * range_max + EDGE_THRESHOLD_MIN)) {
      float lNW = lum[ty - 1][tx - 1];
      float lNE = lum[ty - 1][tx + 1];
      float lSW = lum[ty + 1][tx - 1];
      float lSE = lum[ty + 1][tx + 1];
      float luma_avg = (lN + lS + lW + lE) * (2.0f / 12.0f) + (lNW + lNE + lSW + lSE) * (1.0f / 12.0f);
      float subpix = fabsf(luma_avg - lM);
      subpix = subpix / range;
      if (subpix > 1.0f)
        subpix = 1.0f;
      subpix = subpix * subpix * (3.0f - 2.0f * subpix);
      subpix *= SUBPIX_QUALITY;
      float edge_h = fabsf(-2.0f * lW + lNW + lSW) + fabsf(-2.0f * lM + lN + lS) * 2.0f +
                     fabsf(-2.0f * lE + lNE + lSE);
      float edge_v = fabsf(-2.0f * lN + lNW + lNE) + fabsf(-2.0f * lM + lW + lE) * 2.0f +
                     fabsf(-2.0f * lS + lSW + lSE);
      bool is_horiz = (edge_h >= edge_v);
      int step_x = is_horiz ? 0 : 1;
      int step_y = is_horiz ? 1 : 0;
      float luma_pos = is_horiz ? lS : lE;
      float luma_neg = is_horiz ? lN : lW;
      float grad_pos = fabsf(luma_pos - lM);
      float grad_neg = fabsf(luma_neg - lM);
      int off_x, off_y;
      if (grad_pos >= grad_neg) {
        off_x = step_x;
        off_y = step_y;
      } else {
        off_x = -step_x;
        off_y = -step_y;
      }
      float blend = subpix;
      /* x, y lie in 1..w-2, 1..h-2 here, so the neighbour is in the image and in
       * the tile's halo */
      uint32_t cN = col[ty + off_y][tx + off_x];
      float inv_blend = 1.0f - blend;
      o0 = cvt_u8((float)ch(cM, 0) * inv_blend + (float)ch(cN, 0) * blend);
      o1 = cvt_u8((float)ch(cM, 1) * inv_blend + (float)ch(cN, 1) * blend);
      o2 = cvt_u8((float)ch(cM, 2) * inv_blend + (float)ch(cN, 2) * blend);
    }
  }
  if (v.lut) {
    o0 = v.lut[o0];
    o1 = v.lut[o1];
    o2 = v.lut[o2];
  } else {
    /* postprocess.c runs its per-pixel loop whenever effects != 0 */
    uint32_t rest = effects & (MOP_HIP_POST_GAMMA | MOP_HIP_POST_TONEMAP |
                               MOP_HIP_POST_VIGNETTE | MOP_HIP_POST_FOG);
    float r = (float)o0 / 255.0f;
    float g = (float)o1 / 255.0f;
    float b = (float)o2 / 255.0f;
    if ((rest & MOP_HIP_POST_FOG) && v.has_fog) {
      float dz = v.depth[(size_t)y * (size_t)w + (size_t)x];
      float range = v.fog_far - v.fog_near;
      float factor = 0.0f;
      if (range > 0.0f) {
        factor = (dz - v.fog_near) / range;
        if (factor < 0.0f) factor = 0.0f;
        if (factor > 1.0f) factor = 1.0f;
      }
      r = r + (v.fog_r - r) * factor;
      g = g + (v.fog_g - g) * factor;
      b = b + (v.fog_b - b) * factor;
    }
    if (rest & MOP_HIP_POST_TONEMAP) {
      r = r / (r + 1.0f);
      g = g / (g + 1.0f);
      b = b / (b + 1.0f);
    }
    if (rest & MOP_HIP_POST_GAMMA) {
      r = srgb_enc(r > 0.0f ? r : 0.0f);
      g = srgb_enc(g > 0.0f ? g : 0.0f);
      b = srgb_enc(b > 0.0f ? b : 0.0f);
    }
    if (rest & MOP_HIP_POST_VIGNETTE) {
      float cx = (float)w * 0.5f, cy = (float)h * 0.5f;
      float max_dist = sqrtf(cx * cx + cy * cy);
      float dx = (float)x - cx, dy = (float)y - cy;
      float dist = sqrtf(dx * dx + dy * dy) / max_dist;
      float vig = 1.0f - dist * dist * 0.5f;
      if (vig < 0.0f) vig = 0.0f;
      r *= vig;
      g *= vig;
      b *= vig;
    }
    if (r > 1.0f) r = 1.0f;
    if (g > 1.0f) g = 1.0f;
    if (b > 1.0f) b = 1.0f;
    if (r < 0.0f) r = 0.0f;
    if (g < 0.0f) g = 0.0f;
    if (b < 0.0f) b = 0.0f;
    o0 = cvt_u8(r * 255.0f + 0.5f);
    o1 = cvt_u8(g * 255.0f + 0.5f);
    o2 = cvt_u8(b * 255.0f + 0.5f);
  }
  reinterpret_cast<uint32_t *>(v.dst)[(size_t)y * (size_t)w + (size_t)x] =
      o0 | (o1 << 8) | (o2 << 16) | (o3 << 24);
}

void launch_tail_views(const PostViewGPU *views, uint32_t nviews, uint32_t n_tiles,
                       bool outline, hipStream_t s) {
  if (!n_tiles)
    return;
  if (outline)
    launch(k_tail<true>, dim3(n_tiles), dim3(WG), 0, s, views, nviews);
  else
    launch(k_tail<false>, dim3(n_tiles), dim3(WG), 0, s, views, nviews);
}

__global__ void k_fill_u32(size_t n, uint32_t *p, uint32_t v) {
  size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
  if (i < n)
    p[i] = v;
}

void launch_fill_u32(size_t n, uint32_t *p, uint32_t v, hipStream_t s) {
  if (!n)
    return;
  launch(k_fill_u32, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, s, n, p, v);
}

/* --------------------------------------------------------------------- */
/* shadow depth pass (mop_sw_shadow_render_mesh, rasterizer.c:357-493)    */
/* --------------------------------------------------------------------- */

__global__ __launch_bounds__(WG) void k_shadow_clear(int n, uint32_t *keys) {
  int i = blockIdx.x * WG + threadIdx.x;
  if (i < n)
    keys[i] = depth_key(1.0f);
}

void launch_shadow_clear(int n, uint32_t *keys, hipStream_t s) {
  launch(k_shadow_clear, dim3((n + WG - 1) / WG), dim3(WG), 0, s, n, keys);
}

DEV uint32_t find_shadow_draw(const ShadowDrawGPU *d, uint32_t n, uint32_t t) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (d[mid].tri_base <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

struct ShadowTri {
  EdgeSetup es;
  float sz[3];
};

DEV uint32_t find_shadow_draw_vtx(const ShadowDrawGPU *d, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (d[mid].vtx_base <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* wave-uniform first search, as find_draw_by_tri_wave */
DEV uint32_t find_shadow_draw_wave(const ShadowDrawGPU *d, uint32_t n, uint32_t t) {
  uint32_t t0 = __builtin_amdgcn_readfirstlane(t);
  uint32_t i = find_shadow_draw(d, n, t0);
  uint32_t hi = (i + 1 < n) ? d[i + 1].tri_base : 0xffffffffu;
  if (t >= d[i].tri_base && t < hi)
    return i;
  return find_shadow_draw(d, n, t);
}

DEV uint32_t find_shadow_draw_vtx_wave(const ShadowDrawGPU *d, uint32_t n, uint32_t v) {
  uint32_t v0 = __builtin_amdgcn_readfirstlane(v);
  uint32_t i = find_shadow_draw_vtx(d, n, v0);
  uint32_t hi = (i + 1 < n) ? d[i + 1].vtx_base : 0xffffffffu;
  if (v >= d[i].vtx_base && v < hi)
    return i;
  return find_shadow_draw_vtx(d, n, v);
}

/* Per caster vertex, once: light clip position and shadow-map position
 * (the per-vertex part of mop_sw_shadow_render_mesh, rasterizer.c:383-420;
 * a vertex shared by several triangles gets the same values each time). */
__global__ __launch_bounds__(WG) void k_shadow_vertex(ShadowArgsGPU sa) {
  uint32_t v = blockIdx.x * WG + threadIdx.x;
  if (v >= sa.n_verts)
    return;
  const ShadowDrawGPU &d = sa.draws[find_shadow_draw_vtx_wave(sa.draws, sa.ndraws, v)];
  const float *p = (const float *)(d.vbuf + (size_t)(v - d.vtx_base) * sizeof(MopHipVertex));
  float4 c = mat4_mul_vec4(d.mvp, p[0], p[1], p[2], 1.0f);
  float half = (float)sa.size * 0.5f;
  float4 sc;
  screen_xform(c, half, half, sc.x, sc.y, sc.z, sc.w);
  sa.lclip[v] = c;
  sa.lscr[v] = sc;
}

DEV bool shadow_setup(const ShadowArgsGPU &sa, uint32_t t, ShadowTri &st, bool wave = false) {
  uint32_t di = wave ? find_shadow_draw_wave(sa.draws, sa.ndraws, t)
                     : find_shadow_draw(sa.draws, sa.ndraws, t);
  const ShadowDrawGPU &d = sa.draws[di];
  uint32_t k = t - d.tri_base;
  uint32_t i0 = d.ibuf[3 * k], i1 = d.ibuf[3 * k + 1], i2 = d.ibuf[3 * k + 2];
  if (i0 >= d.vertex_count || i1 >= d.vertex_count || i2 >= d.vertex_count)
    return false;
  uint32_t v0 = d.vtx_base + i0, v1 = d.vtx_base + i1, v2 = d.vtx_base + i2;
  float4 c0 = sa.lclip[v0], c1 = sa.lclip[v1], c2 = sa.lclip[v2];
  if (trivially_rejected(c0, c1, c2))
    return false;
  if (c0.w <= 1e-6f || c1.w <= 1e-6f || c2.w <= 1e-6f)
    return false;
  float4 s0 = sa.lscr[v0], s1 = sa.lscr[v1], s2 = sa.lscr[v2];
  float sx[3] = {s0.x, s1.x, s2.x}, sy[3] = {s0.y, s1.y, s2.y};
  st.sz[0] = s0.z;
  st.sz[1] = s1.z;
  st.sz[2] = s2.z;
  return edge_setup(sx, sy, sa.size, sa.size, st.es);
}

DEV void shadow_frag(const ShadowArgsGPU &sa, const ShadowTri &st, int x, int y, float w0,
                     float w1, float w2) {
  float b0 = w0 * st.es.inv_area, b1 = w1 * st.es.inv_area, b2 = w2 * st.es.inv_area;
  float z = b0 * st.sz[0] + b1 * st.sz[1] + b2 * st.sz[2];
  if (z != z)
    return;
  uint32_t k = depth_key(z);
  atomicMin(sa.keys + (size_t)y * (size_t)sa.size + (size_t)x, k);
}

__global__ __launch_bounds__(WG) void k_shadow_small(ShadowArgsGPU sa) {
  uint32_t t = blockIdx.x * WG + threadIdx.x;
  if (t >= sa.n_tris)
    return;
  ShadowTri st;
  if (!shadow_setup(sa, t, st, true))
    return;
  if (bbox_area(st.es) > MOPH_SMALL_AREA) {
    uint32_t q = atomicAdd(&sa.counters[CNT_SHADOW_BIG], 1u);
    if (q < sa.big_cap)
      sa.big_queue[q] = t;
    else
      atomicOr(&sa.counters[CNT_OVERFLOW], 4u);
    return;
  }
  float w0r = st.es.w_row[0], w1r = st.es.w_row[1], w2r = st.es.w_row[2];
  for (int y = st.es.min_y; y <= st.es.max_y; y++) {
    float w0 = w0r, w1 = w1r, w2 = w2r;
    for (int x = st.es.min_x; x <= st.es.max_x; x++) {
      if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f)
        shadow_frag(sa, st, x, y, w0, w1, w2);
      w0 += st.es.e_dx[0];
      w1 += st.es.e_dx[1];
      w2 += st.es.e_dx[2];
    }
    w0r += st.es.e_dy[0];
    w1r += st.es.e_dy[1];
    w2r += st.es.e_dy[2];
  }
}

__global__ __launch_bounds__(WG) void k_shadow_big(ShadowArgsGPU sa) {
  __shared__ float rows[3][WG];
  __shared__ float carry[3];
  uint32_t nbig = sa.counters[CNT_SHADOW_BIG];
  if (nbig > sa.big_cap)
    nbig = sa.big_cap;
  for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
    ShadowTri st;
    if (!shadow_setup(sa, sa.big_queue[b], st)) /* succeeded when queued; uniform */
      continue;
    int nrows = st.es.max_y - st.es.min_y + 1;
    if (threadIdx.x == 0) {
      carry[0] = st.es.w_row[0];
      carry[1] = st.es.w_row[1];
      carry[2] = st.es.w_row[2];
    }
    for (int c0 = 0; c0 < nrows; c0 += WG) {
      __syncthreads();
      if (threadIdx.x == 0) {
        float a0 = carry[0], a1 = carry[1], a2 = carry[2];
        int lim = nrows - c0 < WG ? nrows - c0 : WG;
        for (int k = 0; k < lim; k++) {
          rows[0][k] = a0;
          rows[1][k] = a1;
          rows[2][k] = a2;
          a0 += st.es.e_dy[0];
          a1 += st.es.e_dy[1];
          a2 += st.es.e_dy[2];
        }
        carry[0] = a0;
        carry[1] = a1;
        carry[2] = a2;
      }
      __syncthreads();
      int row = c0 + (int)threadIdx.x;
      if (row < nrows) {
        int y = st.es.min_y + row;
        float w0 = rows[0][threadIdx.x], w1 = rows[1][threadIdx.x], w2 = rows[2][threadIdx.x];
        for (int x = st.es.min_x; x <= st.es.max_x; x++) {
          if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f)
            shadow_frag(sa, st, x, y, w0, w1, w2);
          w0 += st.es.e_dx[0];
          w1 += st.es.e_dx[1];
          w2 += st.es.e_dx[2];
        }
      }
    }
    __syncthreads();
  }
}

/* SSAA box downsample of mop_viewport_read_color (viewport.c:3385-3411): each
 * output pixel is the integer mean (truncating division) of its sf x sf
 * block of the internal RGBA8 image, channel by channel. */
__global__ __launch_bounds__(WG) void k_downsample(const uint8_t *src, int src_w, uint8_t *dst,
                                                   int pw, int ph, int sf) {
  int i = (int)(blockIdx.x * WG + threadIdx.x);
  if (i >= pw * ph)
    return;
  int x = i % pw, y = i / pw;
  int r = 0, g = 0, b = 0, a = 0;
  for (int dy = 0; dy < sf; dy++) {
    const uchar4 *row = reinterpret_cast<const uchar4 *>(src) + (size_t)(y * sf + dy) * src_w;
    for (int dx = 0; dx < sf; dx++) {
      uchar4 p = row[x * sf + dx];
      r += p.x;
      g += p.y;
      b += p.z;
      a += p.w;
    }
  }
  int n = sf * sf;
  reinterpret_cast<uchar4 *>(dst)[i] =
      make_uchar4((uint8_t)(r / n), (uint8_t)(g / n), (uint8_t)(b / n), (uint8_t)(a / n));
}

void launch_downsample(const uint8_t *src, int src_w, uint8_t *dst, int pw, int ph, int sf,
                       hipStream_t s) {
  int n = pw * ph;
  if (n > 0)
    launch(k_downsample, dim3((n + WG - 1) / WG), dim3(WG), 0, s, src, src_w, dst, pw, ph, sf);
}

/* The environment background (pass_background_hdri_cpu, viewport.c:2128-2180):
 * per pixel the camera ray through its centre, normalized, mapped to the
 * equirect map with the host library's atan2f/asinf, sampled bilinearly and
 * scaled by the intensity; alpha 1.  Mode 1 forms the ray from the camera
 * basis and tan(fov/2) the host computed in the CPU path's order (bit-exact
 * with it); mode 2 from the inverse view-projection and the camera position
 * (the RHI draw_skybox arguments). */
__global__ __launch_bounds__(WG) void k_background(BackgroundGPU bg) {
  int i = (int)(blockIdx.x * WG + threadIdx.x);
  if (i >= bg.w * bg.h)
    return;
  int x = i % bg.w, y = i / bg.w;
  float dx, dy, dz;
  if (bg.mode == 1) {
    float ndc_y = 1.0f - 2.0f * ((float)y + 0.5f) / (float)bg.h;
    float ndc_x = 2.0f * ((float)x + 0.5f) / (float)bg.w - 1.0f;
    dx = bg.fwd[0] + bg.right[0] * ndc_x * bg.half_w + bg.up[0] * ndc_y * bg.half_h;
    dy = bg.fwd[1] + bg.right[1] * ndc_x * bg.half_w + bg.up[1] * ndc_y * bg.half_h;
    dz = bg.fwd[2] + bg.right[2] * ndc_x * bg.half_w + bg.up[2] * ndc_y * bg.half_h;
  } else {
    float ndc_x = 2.0f * ((float)x + 0.5f) / (float)bg.w - 1.0f;
    float ndc_y = 1.0f - 2.0f * ((float)y + 0.5f) / (float)bg.h;
    float4 p = mat4_mul_vec4(bg.inv_vp, ndc_x, ndc_y, 1.0f, 1.0f);
    float iw = 1.0f / p.w;
    dx = p.x * iw - bg.cam[0];
    dy = p.y * iw - bg.cam[1];
    dz = p.z * iw - bg.cam[2];
  }
  float len = sqrtf(dx * dx + dy * dy + dz * dz);
  float il = 1.0f / len;
  dx = mop_div_recip(dx, len, il);
  dy = mop_div_recip(dy, len, il);
  dz = mop_div_recip(dz, len, il);
  const float pi = 3.14159265358979f;
  float phi_val = mop_atan2f(dz, dx) - bg.rotation;
  float clamped_dy = dy < -1.0f ? -1.0f : (dy > 1.0f ? 1.0f : dy);
  float theta_val = mop_asinf(clamped_dy);
  float eu = mop_div_recip(phi_val, 2.0f * pi, 1.0f / (2.0f * pi)) + 0.5f;
  float ev = mop_div_recip(theta_val, pi, 1.0f / pi) + 0.5f;
  float smp[3];
  ibl_sample(bg.env, bg.env_w, bg.env_h, eu, ev, smp);
  reinterpret_cast<float4 *>(bg.hdr)[i] =
      make_float4(smp[0] * bg.intensity, smp[1] * bg.intensity, smp[2] * bg.intensity, 1.0f);
}

void launch_background(const BackgroundGPU &bg, hipStream_t s) {
  int n = bg.w * bg.h;
  if (n > 0 && bg.env)
    launch(k_background, dim3((n + WG - 1) / WG), dim3(WG), 0, s, bg);
}

/* One caster record written from kernel arguments (no host staging). */
__global__ void k_put_shadow_draw(ShadowDrawGPU g, ShadowDrawGPU *dst) {
  if (threadIdx.x == 0)
    *dst = g;
}

void launch_put_shadow_draw(const ShadowDrawGPU &g, ShadowDrawGPU *dst, hipStream_t s) {
  launch(k_put_shadow_draw, dim3(1), dim3(64), 0, s, g, dst);
}

void launch_shadow_raster(const ShadowArgsGPU &sa, hipStream_t s) {
  if (!sa.n_tris)
    return;
  launch(k_shadow_vertex, dim3((sa.n_verts + WG - 1) / WG), dim3(WG), 0, s, sa);
  launch(k_shadow_small, dim3((sa.n_tris + WG - 1) / WG), dim3(WG), 0, s, sa);
  launch(k_shadow_big, dim3(2048), dim3(WG), 0, s, sa);
}

__global__ __launch_bounds__(WG) void k_shadow_finalize(int n, const uint32_t *keys,
                                                        float *depth) {
  int i = blockIdx.x * WG + threadIdx.x;
  if (i < n)
    depth[i] = key_depth(keys[i]);
}

void launch_shadow_finalize(int n, const uint32_t *keys, float *depth, hipStream_t s) {
  launch(k_shadow_finalize, dim3((n + WG - 1) / WG), dim3(WG), 0, s, n, keys,
                     depth);
}

__global__ void k_copy_u8(size_t n, const uint8_t *src, uint8_t *dst) {
  size_t i = ((size_t)blockIdx.x * WG + threadIdx.x) * 16;
  if (i + 16 <= n) {
    *reinterpret_cast<uint4 *>(dst + i) = *reinterpret_cast<const uint4 *>(src + i);
  } else {
    for (; i < n; i++)
      dst[i] = src[i];
  }
}

/* Host-to-device upload of a frame's small tables (draws, lights, views,
 * workgroup maps) from pinned host memory: many workgroups reading 16 bytes
 * per lane over PCIe.  Launched directly (also while a plan is issued or
 * captured), never recorded; n is rounded up to 16 bytes, which both the
 * pinned sources and the device tables have room for. */
__global__ __launch_bounds__(WG) void k_upload(const uint4 *src, uint4 *dst, uint32_t n16) {
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n16; i += gridDim.x * WG)
    dst[i] = src[i];
}

void launch_upload(const void *src, void *dst, size_t n, hipStream_t s) {
  uint32_t n16 = (uint32_t)((n + 15) / 16);
  if (!n16)
    return;
  uint32_t grid = (n16 + WG - 1) / WG;
  if (grid > 1024)
    grid = 1024;
  hipLaunchKernelGGL(k_upload, dim3(grid), dim3(WG), 0, s, (const uint4 *)src, (uint4 *)dst,
                     n16);
}

void launch_copy_u8(size_t n, const uint8_t *src, uint8_t *dst, hipStream_t s) {
  size_t blocks = (n + WG * 16 - 1) / (WG * 16);
  launch(k_copy_u8, dim3((unsigned)blocks), dim3(WG), 0, s, n, src, dst);
}

/* --------------------------------------------------------------------- */
/* ordered path: draws whose result depends on submission order          */
/* (blending, depth write off -- the core's instanced draws --, opacity  */
/* below 1).  Screen triangles are binned to 16x16 tiles (count, scan,   */
/* fill), each tile's list is sorted by submission index in LDS, and one */
/* workgroup per tile replays the CPU's per-triangle loop with the tile's */
/* colour, depth and object-ID held in LDS (rasterizer.c:1926-2140,      */
/* cpu_backend.c:539-545,622-628 for depth_write = false).               */
/* --------------------------------------------------------------------- */

#define OT 16
#define OT_CAP 2048

/* Pixels a wireframe triangle's three lines may plot: Wu's algorithm plots
 * at round(x) and floor(y), floor(y)+1 of points on the edges (the span's
 * intercept is a serial float sum whose drift stays well under a pixel), so
 * the vertex bbox widened by 4 px covers them.  Clamped to the view; false when
 * empty (aa_plot drops off-screen pixels, rasterizer.c:691-692). */
DEV bool wire_bbox(const SubTri &st, int W, int H, int &x0, int &y0, int &x1, int &y1) {
  float mnx = fminf(fminf(st.sx[0], st.sx[1]), st.sx[2]);
  float mny = fminf(fminf(st.sy[0], st.sy[1]), st.sy[2]);
  float mxx = fmaxf(fmaxf(st.sx[0], st.sx[1]), st.sx[2]);
  float mxy = fmaxf(fmaxf(st.sy[0], st.sy[1]), st.sy[2]);
  if (!(mnx == mnx && mny == mny && mxx == mxx && mxy == mxy))
    return false;
  float lo_x = fmaxf(floorf(mnx) - 4.0f, 0.0f), lo_y = fmaxf(floorf(mny) - 4.0f, 0.0f);
  float hi_x = fminf(ceilf(mxx) + 4.0f, (float)(W - 1));
  float hi_y = fminf(ceilf(mxy) + 4.0f, (float)(H - 1));
  if (lo_x > hi_x || lo_y > hi_y)
    return false;
  x0 = (int)lo_x;
  y0 = (int)lo_y;
  x1 = (int)hi_x;
  y1 = (int)hi_y;
  return true;
}

/* Screen bbox of an ordered primitive: the clamped triangle bbox, or the
 * wireframe bbox above. */
DEV bool prim_bbox(const FrameArgs &fa, const SubTri &st, int W, int H, int &x0, int &y0,
                   int &x1, int &y1) {
  if (fa.draws[st.draw].flags & MOPH_DRAW_WIRE)
    return wire_bbox(st, W, H, x0, y0, x1, y1);
  EdgeSetup es;
  if (!edge_setup(st.sx, st.sy, W, H, es))
    return false;
  x0 = es.min_x;
  y0 = es.min_y;
  x1 = es.max_x;
  y1 = es.max_y;
  return true;
}

/* Adds one to counts[tile] for every 16x16 tile of the bbox [tx0..tx0+ntx)
 * x [ty0..) (nt tiles in all) of each active lane, with one atomic per tile
 * and wave: the lanes on the same tile take consecutive slots after the
 * leader's atomic (neighbouring screen triangles share tiles; same-address
 * atomics from every lane would queue at the memory side).  With `at`, the
 * slot of each (lane, tile) is passed to at(tile, slot) in the order the
 * tiles are visited. */
template <class At>
DEV void wave_tile_add(uint32_t *counts, uint32_t tile_base, int tiles_x, int tx0, int ty0, int ntx,
                       int nt, At at) {
  uint32_t lane = __lane_id();
  unsigned long long below = (1ull << lane) - 1ull;
  for (int k = 0;;) {
    bool pend = k < nt;
    unsigned long long pm = __ballot(pend);
    if (!pm)
      break;
    int leader = __ffsll((long long)pm) - 1;
    uint32_t mine = pend ? tile_base + (uint32_t)((ty0 + k / ntx) * tiles_x + tx0 + k % ntx) : ~0u;
    uint32_t tile = __shfl(mine, leader);
    bool match = pend && mine == tile;
    unsigned long long mm = __ballot(match);
    uint32_t base = 0;
    if ((int)lane == leader)
      base = atomicAdd(&counts[tile], (uint32_t)__popcll(mm));
    base = __shfl(base, leader);
    if (match) {
      at(tile, base + (uint32_t)__popcll(mm & below));
      k++;
    }
  }
}

__global__ __launch_bounds__(WG) void k_bin_count(FrameArgs fa, uint32_t *tile_counts) {
  uint32_t total = fa.counters[CNT_SUBTRI];
  if (total > fa.subtri_cap)
    total = fa.subtri_cap;
  for (uint32_t s = blockIdx.x * WG + threadIdx.x; s < total; s += gridDim.x * WG) {
    SubTri st = fa.subtris[s];
    const ViewGPU &vw = fa.views[fa.draws[st.draw].view];
    int bx0, by0, bx1, by1;
    if (!prim_bbox(fa, st, vw.w, vw.h, bx0, by0, bx1, by1))
      continue;
    int tx0 = bx0 / OT, ty0 = by0 / OT, ntx = bx1 / OT - tx0 + 1;
    wave_tile_add(tile_counts, vw.tile_base, (vw.w + OT - 1) / OT, tx0, ty0, ntx,
                  ntx * (by1 / OT - ty0 + 1), [](uint32_t, uint32_t) {});
  }
}

__global__ __launch_bounds__(WG) void k_bin_fill(FrameArgs fa, const uint32_t *tile_start,
                                                 uint32_t *tile_cursor, uint32_t *bins,
                                                 uint32_t bin_cap) {
  uint32_t total = fa.counters[CNT_SUBTRI];
  if (total > fa.subtri_cap)
    total = fa.subtri_cap;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    fa.counters[CNT_BINS] = tile_start[fa.n_tiles]; /* the capacity this batch needs */
  for (uint32_t s = blockIdx.x * WG + threadIdx.x; s < total; s += gridDim.x * WG) {
    SubTri st = fa.subtris[s];
    const ViewGPU &vw = fa.views[fa.draws[st.draw].view];
    int bx0, by0, bx1, by1;
    if (!prim_bbox(fa, st, vw.w, vw.h, bx0, by0, bx1, by1))
      continue;
    int tx0 = bx0 / OT, ty0 = by0 / OT, ntx = bx1 / OT - tx0 + 1;
    wave_tile_add(tile_cursor, vw.tile_base, (vw.w + OT - 1) / OT, tx0, ty0, ntx,
                  ntx * (by1 / OT - ty0 + 1), [&](uint32_t t, uint32_t slot) {
                    uint32_t pos = tile_start[t] + slot;
                    if (pos < bin_cap)
                      bins[pos] = s;
                    else
                      atomicOr(&fa.counters[CNT_OVERFLOW], 16u);
                  });
  }
}

/* in-LDS bitonic sort of n (<= OT_CAP) keys, ascending; pads with ~0 */
DEV void lds_sort(uint32_t *k, int n) {
  int m = 1;
  while (m < n)
    m <<= 1;
  for (int i = n + (int)threadIdx.x; i < m; i += WG)
    k[i] = 0xffffffffu;
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < m; i += WG) {
        int j = i ^ stride;
        if (j > i) {
          bool up = ((i & size) == 0);
          uint32_t a = k[i], b = k[j];
          if ((a > b) == up) {
            k[i] = b;
            k[j] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

/* Fragments whose colour is still to be computed, per wave of the tile (one
 * entry per fragment, in the order the fragments were met). */
#define OQ 64
#define OQ_WAVES (OT * OT / 64)

/* A list entry set up once for the wave: the screen triangle's raster
 * setup, its index and draw.  kind 0: nothing to raster (empty bbox or zero
 * area), 1: a triangle, 2: a wireframe primitive. */
struct OrdRec {
  RasterTri r;
  uint32_t s, draw, kind;
};

struct OrderedLds {
  OrdRec rec[OQ_WAVES][64]; /* the entries of the wave's current list chunk */
  float4 hdr[OT * OT];
  float depth[OT * OT];
  float saved[OT * OT];
  uint32_t ids[OT * OT];
  uint32_t list[OT_CAP];
  uint32_t q_s[OQ_WAVES][OQ];    /* screen triangle */
  uint32_t q_meta[OQ_WAVES][OQ]; /* lane | kind << 8 | blend << 12 */
  float q_fa[OQ_WAVES][OQ];      /* the blend's alpha */
  float4 q_v[OQ_WAVES][OQ];      /* edge values in, colour out */
  uint32_t q_n[OQ_WAVES];
};

/* What the in-order pass decides for one fragment of this lane's pixel:
 * whether a colour is to be blended, how, and how to get it.  kind 0: the
 * lit smooth model's colour from the edge values in v; kind 1: the flat
 * model's colour; kind 2: the colour is v already (unlit smooth). */
struct OrdFrag {
  bool q;
  uint32_t kind, blend;
  float fa;
  float4 v;
};

/* One screen triangle against this lane's pixel, the way one iteration of
 * the smooth multi-light loop treats it (rasterizer.c:1929-2140) or the
 * flat loop (rasterizer.c:1000-1085): coverage, the depth test, and the
 * depth and id writes, all against the tile as the earlier fragments left
 * it.  None of these depends on a fragment's colour, so the colour and its
 * blend are deferred (f): ordered_flush computes the colours of a wave's
 * pending fragments with one lane each and applies the blends in order. */
DEV void ordered_fragment(const FrameArgs &frame, const ViewGPU &vw, uint32_t s, uint32_t draw,
                          const RasterTri &r, int px, int py, int li, OrderedLds &L,
                          OrdFrag &f) {
  const DrawGPU &d = frame.draws[draw];
  int sm = draw_sm(d);
  bool flat = sm_flat(sm);
  FlatZ fz = {0.0f, 0.0f, 0.0f};
  if (flat)
    fz = flat_z_setup(r);
  float w0, w1, w2, zf;
  replay_walk(r, flat, fz, px, py, w0, w1, w2, zf);
  if (!(w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f))
    return;
  bool depth_test = (d.flags & MOPH_DRAW_DEPTH_TEST) != 0;
  uint32_t blend = (d.flags >> MOPH_DRAW_BLEND_SHIFT) & 3u;
  f.blend = blend;
  if (flat) {
    /* rasterize_filled_triangle: the blended branch writes colour only,
     * never depth or id */
    if (depth_test && !(zf < L.depth[li]))
      return;
    float ca = clamp01f(d.opacity);
    if (sm == MOPH_SM_UNLIT_FLAT)
      ca = (float)cvt_u8(ca * 255.0f) / 255.0f;
    if (blend == MOP_HIP_BLEND_OPAQUE && ca >= 1.0f) {
      L.depth[li] = zf;
      L.ids[li] = d.object_id;
    }
    f.q = true;
    f.kind = 1;
    f.fa = ca;
    return;
  }
  float d0 = w0 * r.inv_e[0], d1 = w1 * r.inv_e[1], d2 = w2 * r.inv_e[2];
  float min_d = d0;
  if (d1 < min_d)
    min_d = d1;
  if (d2 < min_d)
    min_d = d2;
  if (!(min_d >= 0.0f))
    return;
  float b0 = w0 * r.es.inv_area, b1 = w1 * r.es.inv_area, b2 = w2 * r.es.inv_area;
  float zt = b0 * r.sz[0] + b1 * r.sz[1] + b2 * r.sz[2];
  if (depth_test && !(zt < L.depth[li]))
    return;
  float cov;
  float op = d.opacity < 0.0f ? 0.0f : (d.opacity > 1.0f ? 1.0f : d.opacity);
  if (sm == MOPH_SM_UNLIT_SMOOTH) {
    /* silhouette AA against the depth already in the buffer
     * (rasterizer.c:1421-1426) */
    float pr, pg, pb;
    unlit_smooth_color(frame, frame.subtris[s], r, d, w0, w1, w2, pr, pg, pb);
    bool silhouette = (L.depth[li] - zt) > 0.005f;
    cov = (min_d < 1.0f && silhouette) ? min_d : 1.0f;
    f.kind = 2;
    f.v = make_float4(pr, pg, pb, 0.0f);
  } else {
    cov = (min_d < 1.0f && blend != MOP_HIP_BLEND_OPAQUE) ? min_d : 1.0f;
    f.kind = 0;
    f.v = make_float4(w0, w1, w2, 0.0f);
  }
  float fa_ = op * cov;
  /* the shaded z is the barycentric zt (shade_fragment's same expression) */
  if ((blend == MOP_HIP_BLEND_OPAQUE && fa_ >= 1.0f) || fa_ > 0.5f) {
    L.depth[li] = zt;
    L.ids[li] = d.object_id;
  }
  f.q = true;
  f.fa = fa_;
}

/* The colours of this wave's pending fragments, one lane per fragment, then
 * their blends into the tile in the order the fragments were met (each lane
 * applies those of its own pixel): an opaque one with alpha 1 replaces the
 * colour, the others blend (rasterizer.c:2101-2131, 1054-1083).  Called by
 * all lanes of the wave. */
DEV void ordered_flush(const FrameArgs &frame, const ViewGPU &vw, int li, OrderedLds &L) {
  uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t n = L.q_n[w];
  if (n == 0)
    return;
  if (lane < n) {
    uint32_t kind = (L.q_meta[w][lane] >> 8) & 15u;
    if (kind != 2u) {
      SubTri st = frame.subtris[L.q_s[w][lane]];
      const DrawGPU &d = frame.draws[st.draw];
      float4 A0[3], A1[3], A2[3];
      load_tri_attr(frame, st, A0, A1, A2);
      float pr, pg, pb;
      if (kind == 0u) {
        RasterTri r;
        load_raster_tri(vw.w, vw.h, st, r);
        float4 v = L.q_v[w][lane];
        float z;
        shade_fragment(frame, vw, st, r, A0, A1, A2, v.x, v.y, v.z, z, pr, pg, pb);
      } else {
        flat_color(frame, vw, st, d, draw_sm(d), A0, A1, A2, pr, pg, pb);
      }
      L.q_v[w][lane] = make_float4(pr, pg, pb, 0.0f);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (uint32_t k = 0; k < n; k++) {
    uint32_t meta = L.q_meta[w][k];
    if ((meta & 63u) != lane)
      continue;
    uint32_t blend = meta >> 12;
    float fa_ = L.q_fa[w][k];
    float4 c = L.q_v[w][k];
    float pr = c.x, pg = c.y, pb = c.z;
    if (blend == MOP_HIP_BLEND_OPAQUE && fa_ >= 1.0f) {
      L.hdr[li] = make_float4(pr, pg, pb, 1.0f);
      continue;
    }
    float inv_fa = 1.0f - fa_;
    float4 dc = L.hdr[li];
    if (blend == MOP_HIP_BLEND_ADDITIVE) {
      pr = dc.x + pr * fa_;
      pg = dc.y + pg * fa_;
      pb = dc.z + pb * fa_;
    } else if (blend == MOP_HIP_BLEND_MULTIPLY) {
      pr = dc.x * pr;
      pg = dc.y * pg;
      pb = dc.z * pb;
    } else {
      pr = pr * fa_ + dc.x * inv_fa;
      pg = pg * fa_ + dc.y * inv_fa;
      pb = pb * fa_ + dc.z * inv_fa;
    }
    L.hdr[li] = make_float4(pr, pg, pb, 1.0f);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  L.q_n[w] = 0u;
}

/* Queue the fragment f of screen triangle s met by this lane (all lanes of
 * the wave call this; a full queue is flushed first). */
DEV void ordered_enqueue(const FrameArgs &frame, const ViewGPU &vw, uint32_t s, int li,
                         OrderedLds &L, const OrdFrag &f) {
  unsigned long long m = __ballot(f.q);
  if (m == 0ull)
    return;
  uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t cnt = (uint32_t)__popcll(m);
  if (L.q_n[w] + cnt > OQ)
    ordered_flush(frame, vw, li, L);
  uint32_t n = L.q_n[w];
  if (f.q) {
    uint32_t pos = n + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    L.q_s[w][pos] = s;
    L.q_meta[w][pos] = lane | (f.kind << 8) | (f.blend << 12);
    L.q_fa[w][pos] = f.fa;
    L.q_v[w][pos] = f.v;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  L.q_n[w] = n + cnt;
}

/* aa_plot (rasterizer.c:688-730): coverage blend of the line colour, depth
 * and id written above coverage 0.5. */
DEV void wire_plot(const DrawGPU &d, float z, float cov, float fr, float fg, float fb_, int li,
                   OrderedLds &L) {
  if (cov < 0.004f)
    return;
  if ((d.flags & MOPH_DRAW_DEPTH_TEST) && z >= L.depth[li])
    return;
  float4 dc = L.hdr[li];
  float inv = 1.0f - cov;
  L.hdr[li] = make_float4(fr * cov + dc.x * inv, fg * cov + dc.y * inv, fb_ * cov + dc.z * inv,
                          1.0f);
  if (cov > 0.5f) {
    L.depth[li] = z;
    L.ids[li] = d.object_id;
  }
}

DEV float fpartf(float x) { return x - floorf(x); }
DEV float rfpartf(float x) { return 1.0f - fpartf(x); }

/* The plots one Xiaolin Wu line (mop_sw_draw_line_aa, thickness 1,
 * rasterizer.c:779-871) makes at pixel (px, py), in the line's order:
 * first endpoint, second endpoint, then the span between them, whose
 * intercept is the serial sum intery += gradient. */
DEV void wire_line(const DrawGPU &d, float x0, float y0, float z0, float x1, float y1, float z1,
                   int px, int py, float fr, float fg, float fb_, int li, OrderedLds &L) {
  bool steep = fabsf(y1 - y0) > fabsf(x1 - x0);
  if (steep) {
    float t = x0;
    x0 = y0;
    y0 = t;
    t = x1;
    x1 = y1;
    y1 = t;
  }
  if (x0 > x1) {
    float t = x0;
    x0 = x1;
    x1 = t;
    t = y0;
    y0 = y1;
    y1 = t;
    t = z0;
    z0 = z1;
    z1 = t;
  }
  int X = steep ? py : px, Y = steep ? px : py;
  float dx = x1 - x0, dy = y1 - y0;
  float gradient = (dx < 0.001f) ? 1.0f : dy / dx;
  float xend = roundf(x0);
  float yend1 = y0 + gradient * (xend - x0);
  float xgap = rfpartf(x0 + 0.5f);
  int xp1 = cvt_i32(xend), yp1 = cvt_i32(floorf(yend1));
  if (X == xp1) {
    if (Y == yp1)
      wire_plot(d, z0, rfpartf(yend1) * xgap, fr, fg, fb_, li, L);
    else if (Y == yp1 + 1)
      wire_plot(d, z0, fpartf(yend1) * xgap, fr, fg, fb_, li, L);
  }
  xend = roundf(x1);
  float yend2 = y1 + gradient * (xend - x1);
  xgap = fpartf(x1 + 0.5f);
  int xp2 = cvt_i32(xend), yp2 = cvt_i32(floorf(yend2));
  if (X == xp2) {
    if (Y == yp2)
      wire_plot(d, z1, rfpartf(yend2) * xgap, fr, fg, fb_, li, L);
    else if (Y == yp2 + 1)
      wire_plot(d, z1, fpartf(yend2) * xgap, fr, fg, fb_, li, L);
  }
  if (X > xp1 && X < xp2) {
    float total_steps = (float)(xp2 - xp1);
    if (total_steps < 1.0f)
      total_steps = 1.0f;
    float intery = yend1 + gradient;
    for (int x = xp1 + 1; x < X; x++)
      intery += gradient;
    int iy = cvt_i32(floorf(intery));
    if (Y == iy || Y == iy + 1) {
      float t = (float)(X - xp1) / total_steps;
      float z = z0 + t * (z1 - z0);
      wire_plot(d, z, Y == iy ? rfpartf(intery) : fpartf(intery), fr, fg, fb_, li, L);
    }
  }
}

/* A wireframe triangle: its flat colour truncated to 8 bits, then edges
 * v0-v1, v1-v2, v2-v0 (rasterizer.c:2590-2603, 1222-1233). */
DEV void wire_fragment(const FrameArgs &frame, const ViewGPU &vw, const SubTri &st, int px,
                       int py, int li, OrderedLds &L) {
  const DrawGPU &d = frame.draws[st.draw];
  float cr, cg, cb;
  float4 A0[3], A1[3], A2[3];
  load_tri_attr(frame, st, A0, A1, A2);
  flat_color(frame, vw, st, d, draw_sm(d), A0, A1, A2, cr, cg, cb);
  float fr = (float)cvt_u8((cr > 1.0f ? 1.0f : cr) * 255.0f) / 255.0f;
  float fg = (float)cvt_u8((cg > 1.0f ? 1.0f : cg) * 255.0f) / 255.0f;
  float fb_ = (float)cvt_u8((cb > 1.0f ? 1.0f : cb) * 255.0f) / 255.0f;
  wire_line(d, st.sx[0], st.sy[0], st.sz[0], st.sx[1], st.sy[1], st.sz[1], px, py, fr, fg, fb_,
            li, L);
  wire_line(d, st.sx[1], st.sy[1], st.sz[1], st.sx[2], st.sy[2], st.sz[2], px, py, fr, fg, fb_,
            li, L);
  wire_line(d, st.sx[2], st.sy[2], st.sz[2], st.sx[0], st.sy[0], st.sz[0], px, py, fr, fg, fb_,
            li, L);
}

/* One screen triangle of the tile's submission-ordered list against this
 * lane's pixel: draw changes (depth_write = false saves and restores the
 * tile's depth), the wave-level bbox skip, then the fragment. */
struct OrderedState {
  uint32_t cur_draw;
  bool saved_live;
};

/* The CPU loop's per-draw state at a change of draw: the end of a draw with
 * depth_write = false restores the depth it saved at its start
 * (cpu_backend.c:539-545,622-628). */
DEV void ordered_draw(const FrameArgs &frame, uint32_t draw, bool inside, int li, OrderedLds &L,
                      OrderedState &os) {
  if (draw == os.cur_draw)
    return;
  if (os.saved_live && inside)
    L.depth[li] = L.saved[li];
  os.saved_live = false;
  os.cur_draw = draw;
  if (!(frame.draws[draw].flags & MOPH_DRAW_DEPTH_WRITE)) {
    if (inside)
      L.saved[li] = L.depth[li];
    os.saved_live = true;
  }
}

/* One list entry (set up in R) against the wave's pixels, in order. */
DEV void ordered_rec(const FrameArgs &frame, const ViewGPU &vw, const OrdRec &R, int px, int py,
                     bool inside, int li, OrderedLds &L, OrderedState &os) {
  ordered_draw(frame, R.draw, inside, li, L, os);
  if (R.kind == 2u) {
    ordered_flush(frame, vw, li, L); /* the line blends after the pending colours */
    SubTri st = frame.subtris[R.s];
    int bx0, by0, bx1, by1;
    bool ok = wire_bbox(st, vw.w, vw.h, bx0, by0, bx1, by1);
    bool in_bbox = ok && inside && px >= bx0 && px <= bx1 && py >= by0 && py <= by1;
    if (__ballot(in_bbox) == 0ull)
      return;
    if (in_bbox)
      wire_fragment(frame, vw, st, px, py, li, L);
    return;
  }
  bool in_bbox = R.kind == 1u && inside && px >= R.r.es.min_x && px <= R.r.es.max_x &&
                 py >= R.r.es.min_y && py <= R.r.es.max_y;
  /* coverage mask of the wave: skip the triangle when no lane is in its bbox */
  if (__ballot(in_bbox) == 0ull)
    return;
  OrdFrag f;
  f.q = false;
  f.kind = f.blend = 0u;
  f.fa = 0.0f;
  f.v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (in_bbox)
    ordered_fragment(frame, vw, R.s, R.draw, R.r, px, py, li, L, f);
  ordered_enqueue(frame, vw, R.s, li, L, f);
}

/* Entry s set up by this lane. */
DEV void ordered_setup(const FrameArgs &frame, const ViewGPU &vw, uint32_t s, OrdRec &R) {
  SubTri st = frame.subtris[s];
  R.s = s;
  R.draw = st.draw;
  if (frame.draws[st.draw].flags & MOPH_DRAW_WIRE)
    R.kind = 2u;
  else
    R.kind = load_raster_tri(vw.w, vw.h, st, R.r) ? 1u : 0u;
}

/* Entry s of an untagged list: set up by every lane of the wave. */
DEV void ordered_step(const FrameArgs &frame, const ViewGPU &vw, uint32_t s, int px, int py,
                      bool inside, int li, OrderedLds &L, OrderedState &os) {
  OrdRec R;
  ordered_setup(frame, vw, s, R);
  ordered_rec(frame, vw, R, px, py, inside, li, L, os);
}

/* Tags each sorted list entry with the waves of the tile it concerns: wave w
 * holds rows 4w..4w+3 of the 16x16 tile, and an entry whose primitive's bbox
 * (the bbox ordered_step tests, prim_bbox) misses those rows cannot touch the
 * wave's pixels, so the wave passes it over without loading it.  The bboxes
 * of the list are computed by all threads at once.  Bits 28-31 of an entry
 * hold the mask (indices stay below 2^28: the caller checks); an entry
 * outside the store gets an empty mask. */
DEV void tag_waves(const FrameArgs &frame, const ViewGPU &vw, uint32_t *list, int m,
                   uint32_t total, int tile_y0) {
  for (int i = (int)threadIdx.x; i < m; i += WG) {
    uint32_t s = list[i];
    uint32_t mask = 0;
    if (s < total) {
      int bx0, by0, bx1, by1;
      if (prim_bbox(frame, frame.subtris[s], vw.w, vw.h, bx0, by0, bx1, by1)) {
        int r0 = by0 - tile_y0, r1 = by1 - tile_y0;
#pragma unroll
        for (int w = 0; w < OT * OT / 64; w++)
          if (r0 <= 4 * w + 3 && r1 >= 4 * w)
            mask |= 1u << w;
      }
    }
    list[i] = (s & 0x0fffffffu) | (mask << 28);
  }
  __syncthreads();
}

/* The entries of a tagged list this wave walks, in order. */
DEV void walk_tagged(const FrameArgs &frame, const ViewGPU &vw, const uint32_t *list, int m,
                     int px, int py, bool inside, int li, OrderedLds &L, OrderedState &os) {
  uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t wbit = 1u << (28 + w);
  /* 64 list positions at a time: the entries tagged for this wave are set
   * up once each, one per lane, into the wave's LDS records, then walked in
   * order from there */
  for (int base = 0; base < m; base += 64) {
    int e = base + (int)lane;
    uint32_t v = e < m ? list[e] : 0u;
    bool mine = (v & wbit) != 0u;
    unsigned long long mm = __ballot(mine);
    if (mm == 0ull)
      continue;
    if (mine)
      ordered_setup(frame, vw, v & 0x0fffffffu,
                    L.rec[w][__popcll(mm & ((1ull << lane) - 1ull))]);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int n = __popcll(mm);
    for (int k = 0; k < n; k++)
      ordered_rec(frame, vw, L.rec[w][k], px, py, inside, li, L, os);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

/* One workgroup per 16x16 tile.  The tile's bin holds the indices of the
 * screen triangles whose bbox overlaps it, in arbitrary order; they are
 * walked in submission order.  A list of at most OT_CAP entries is sorted
 * in LDS at once.  A longer list is walked in windows of the index space
 * [lo, hi) holding at most OT_CAP entries each: the window is halved until
 * its count fits, gathered into LDS, sorted and walked, then the next window
 * starts at hi (every pass reads the bin once; entries are distinct, so a
 * window of one index holds at most one).  A batch whose bins overflowed
 * (flag 16) is run again by the host; its tiles return at once, and a tile
 * never reads past the bins that were written or an index outside the
 * screen-triangle store. */
__global__ __launch_bounds__(WG) void k_tile_ordered(FrameArgs frame, const uint32_t *tile_start,
                                                     const uint32_t *bins, uint32_t bin_cap) {
  __shared__ OrderedLds L;
  __shared__ uint32_t count;
  if (frame.counters[CNT_OVERFLOW] & 16u)
    return; /* bins incomplete; uniform */
  uint32_t vi = find_view_by_tile(frame.views, frame.nviews, blockIdx.x);
  const ViewGPU &vw = frame.views[vi];
  uint32_t t = blockIdx.x - vw.tile_base;
  uint32_t tiles_x = ((uint32_t)vw.w + OT - 1) / OT;
  int px = (int)(t % tiles_x) * OT + (int)(threadIdx.x % OT);
  int py = (int)(t / tiles_x) * OT + (int)(threadIdx.x / OT);
  bool inside = px < vw.w && py < vw.h;
  int li = threadIdx.x;
  size_t idx = (size_t)py * (size_t)vw.w + (size_t)px;
  uint32_t b0 = tile_start[blockIdx.x], b1 = tile_start[blockIdx.x + 1];
  if (b1 > bin_cap)
    b1 = bin_cap;
  uint32_t n = b1 > b0 ? b1 - b0 : 0u;
  if (n == 0)
    return; /* uniform */
  uint32_t total = frame.counters[CNT_SUBTRI];
  if (total > frame.subtri_cap)
    total = frame.subtri_cap;
  bool tag = total <= 0x10000000u; /* indices fit the 28 bits of a tagged entry */
  int tile_y0 = (int)(t / tiles_x) * OT;
  if (inside) {
    L.hdr[li] = reinterpret_cast<const float4 *>(vw.hdr)[idx];
    L.depth[li] = vw.depth[idx];
    L.ids[li] = vw.ids[idx];
  }
  if ((threadIdx.x & 63u) == 0)
    L.q_n[threadIdx.x >> 6] = 0u;
  OrderedState os = {0xffffffffu, false};
  if (n <= OT_CAP) {
    for (uint32_t i = threadIdx.x; i < n; i += WG)
      L.list[i] = bins[b0 + i];
    __syncthreads();
    lds_sort(L.list, (int)n);
    if (tag) {
      tag_waves(frame, vw, L.list, (int)n, total, tile_y0);
      walk_tagged(frame, vw, L.list, (int)n, px, py, inside, li, L, os);
    } else {
      for (uint32_t e = 0; e < n; e++) {
        uint32_t s = L.list[e];
        if (s < total)
          ordered_step(frame, vw, s, px, py, inside, li, L, os);
      }
    }
  } else {
    uint32_t lo = 0, span = total;
    while (lo < total) {
      uint32_t hi, c;
      for (;;) {
        hi = (total - lo > span) ? lo + span : total;
        __syncthreads();
        if (threadIdx.x == 0)
          count = 0;
        __syncthreads();
        uint32_t mine = 0;
        for (uint32_t i = threadIdx.x; i < n; i += WG) {
          uint32_t v = bins[b0 + i];
          mine += (v >= lo && v < hi) ? 1u : 0u;
        }
        if (mine)
          atomicAdd(&count, mine);
        __syncthreads();
        c = count;
        if (c <= OT_CAP || hi - lo <= 1u)
          break;
        span = (hi - lo) >> 1;
      }
      if (c) {
        __syncthreads();
        if (threadIdx.x == 0)
          count = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += WG) {
          uint32_t v = bins[b0 + i];
          if (v >= lo && v < hi) {
            uint32_t at = atomicAdd(&count, 1u);
            if (at < OT_CAP)
              L.list[at] = v;
          }
        }
        __syncthreads();
        int m = (int)(c < OT_CAP ? c : OT_CAP);
        lds_sort(L.list, m);
        if (tag) {
          tag_waves(frame, vw, L.list, m, total, tile_y0);
          walk_tagged(frame, vw, L.list, m, px, py, inside, li, L, os);
        } else {
          for (int e = 0; e < m; e++) {
            uint32_t s = L.list[e];
            if (s < total)
              ordered_step(frame, vw, s, px, py, inside, li, L, os);
          }
        }
      }
      /* a sparse window lets the next one grow */
      if (c < OT_CAP / 2 && span < 0x40000000u)
        span <<= 1;
      lo = hi;
    }
  }
  ordered_flush(frame, vw, li, L);
  if (os.saved_live && inside)
    L.depth[li] = L.saved[li];
  if (inside) {
    reinterpret_cast<float4 *>(vw.hdr)[idx] = L.hdr[li];
    vw.depth[idx] = L.depth[li];
    vw.ids[idx] = L.ids[li];
  }
}

void launch_ordered(const FrameArgs &fa, uint32_t *tile_counts, uint32_t *tile_cursor,
                    uint32_t *bins, uint32_t bin_cap, uint32_t *block_sums, hipStream_t s) {
  uint32_t nt = fa.n_tiles;
  if (!nt)
    return;
  launch_fill_u32((size_t)nt + 1, tile_counts, 0u, s);
  launch_fill_u32(nt, tile_cursor, 0u, s);
  uint32_t grid = (fa.subtri_cap + WG - 1) / WG;
  if (grid > 4096)
    grid = 4096;
  if (grid == 0)
    grid = 1;
  launch(k_bin_count, dim3(grid), dim3(WG), 0, s, fa, tile_counts);
  /* exclusive scan of tile counts in place (total lands in tile_counts[nt]) */
  uint32_t nb = (nt + SCAN_TILE - 1) / SCAN_TILE;
  launch(k_scan_reduce, dim3(nb), dim3(WG), 0, s, tile_counts, nt, block_sums);
  launch(k_scan_sums, dim3(1), dim3(WG), 0, s, block_sums, nb);
  launch(k_scan_apply, dim3(nb), dim3(WG), 0, s, tile_counts, nt, block_sums,
                     (uint32_t *)nullptr, 0u);
  launch(k_bin_fill, dim3(grid), dim3(WG), 0, s, fa, tile_counts, tile_cursor, bins,
                     bin_cap);
  launch(k_tile_ordered, dim3(nt), dim3(WG), 0, s, fa, tile_counts, bins, bin_cap);
}
