/*
 * shade_key.h -- the part of the order-free shading pass that turns a
 * visibility key into the triangle and edge values to shade with: key
 * validation, the screen triangle's load, its edge setup, and the edge
 * values at the pixel (replayed from the bbox corner, or from the nearest
 * big-triangle checkpoint).  Every read that depends on the key happens
 * here, each one after the checks that bound it.
 *
 * Compiled for the device by kernels.hip and for the host by
 * tests/key_check.cpp, which feeds it random and stale keys against tables
 * of garbage under AddressSanitizer.
 */
#ifndef MOP_HIP_SHADE_KEY_H
#define MOP_HIP_SHADE_KEY_H

#include "devmath.h"
#include "mop_hip_internal.h"

struct RasterTri {
  EdgeSetup es;
  float inv_e[3];
  float sz[3];
};

HD bool load_raster_tri(int W, int H, const SubTri &st, RasterTri &r) {
  if (!edge_setup(st.sx, st.sy, W, H, r.es))
    return false;
  r.inv_e[0] = inv_edge_len(r.es.e_dx[0], r.es.e_dy[0]);
  r.inv_e[1] = inv_edge_len(r.es.e_dx[1], r.es.e_dy[1]);
  r.inv_e[2] = inv_edge_len(r.es.e_dx[2], r.es.e_dy[2]);
  r.sz[0] = st.sz[0];
  r.sz[1] = st.sz[1];
  r.sz[2] = st.sz[2];
  return true;
}

/* When is the smooth walk's AA distance test (min_i w_i * inv_e_i >= 0,
 * rasterizer.c:1929-1936) implied by its edge test (all w_i >= 0)?  With
 * inv_e_i finite and positive, w_i * inv_e_i is >= 0 (or -0.0, or +inf)
 * for every w_i >= 0, so the min passes too.  inv_e_i is 1/len (finite) when
 * len > 1e-6 and 0 otherwise; a squared length above 1e-10 guarantees the
 * former (and one below 1e30 keeps 1/len positive).  Only a triangle with a
 * near-degenerate or enormous edge needs the full test
 * (where 0 * inf could make it fail); the others skip the three square roots
 * and divisions.  The order-free kernels use this; the ordered path also
 * needs min_d as coverage and always computes it. */
HD bool edges_nondegenerate(const EdgeSetup &es) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float l2 = es.e_dx[k] * es.e_dx[k] + es.e_dy[k] * es.e_dy[k];
    if (!(l2 > 1e-10f && l2 < 1e30f)) /* and 1/len stays a normal positive float */
      return false;
  }
  return true;
}

/* load_raster_tri for the order-free kernels: inv_e only where the AA
 * distance test is not implied (then inv_e holds the real values, else
 * ones, which make frag_depth's min_d test the edge test's consequence). */
HD bool load_raster_tri_fast(int W, int H, const SubTri &st, RasterTri &r) {
  if (!edge_setup(st.sx, st.sy, W, H, r.es))
    return false;
  if (edges_nondegenerate(r.es)) {
    r.inv_e[0] = r.inv_e[1] = r.inv_e[2] = 1.0f;
  } else {
    r.inv_e[0] = inv_edge_len(r.es.e_dx[0], r.es.e_dy[0]);
    r.inv_e[1] = inv_edge_len(r.es.e_dx[1], r.es.e_dy[1]);
    r.inv_e[2] = inv_edge_len(r.es.e_dx[2], r.es.e_dy[2]);
  }
  r.sz[0] = st.sz[0];
  r.sz[1] = st.sz[1];
  r.sz[2] = st.sz[2];
  return true;
}

HD int draw_sm(const DrawGPU &d) { return (int)((d.flags >> MOPH_DRAW_SM_SHIFT) & 3u); }
HD bool sm_flat(int sm) { return sm == MOPH_SM_LIT_FLAT || sm == MOPH_SM_UNLIT_FLAT; }

/* Incremental depth of the flat path, z_row += z_dy per row and z += z_dx
 * per pixel (rasterizer.c:973-976, 1019-1022). */
struct FlatZ {
  float row, dx, dy;
};

HD FlatZ flat_z_setup(const RasterTri &r) {
  float ia = r.es.inv_area;
  FlatZ f;
  f.dx = (r.es.e_dx[0] * r.sz[0] + r.es.e_dx[1] * r.sz[1] + r.es.e_dx[2] * r.sz[2]) * ia;
  f.dy = (r.es.e_dy[0] * r.sz[0] + r.es.e_dy[1] * r.sz[1] + r.es.e_dy[2] * r.sz[2]) * ia;
  f.row = (r.es.w_row[0] * r.sz[0] + r.es.w_row[1] * r.sz[1] + r.es.w_row[2] * r.sz[2]) * ia;
  return f;
}

/* Edge values (and the flat path's z) at pixel (px, py), replaying the
 * serial increments from the bbox corner in the reference's order. */
HD void replay_walk(const RasterTri &r, bool flat, const FlatZ &fz, int px, int py, float &w0,
                     float &w1, float &w2, float &z) {
  float w0r = r.es.w_row[0], w1r = r.es.w_row[1], w2r = r.es.w_row[2];
  float zr = fz.row;
  for (int y = r.es.min_y; y < py; y++) {
    w0r += r.es.e_dy[0];
    w1r += r.es.e_dy[1];
    w2r += r.es.e_dy[2];
    if (flat)
      zr += fz.dy;
  }
  w0 = w0r;
  w1 = w1r;
  w2 = w2r;
  z = zr;
  for (int x = r.es.min_x; x < px; x++) {
    w0 += r.es.e_dx[0];
    w1 += r.es.e_dx[1];
    w2 += r.es.e_dx[2];
    if (flat)
      z += fz.dx;
  }
}


/* Screen triangle e of the order-free path, e = (t << 4) | k: k = 15 is
 * triangle t drawn whole, k < 15 the fan triangle k+1 of its clipped polygon
 * (poly[0], poly[k+1], poly[k+2]; rasterizer.c:2510-2513).  e + 1 is its
 * submission sequence number: triangles in submission order, fan triangles
 * in fan order. */
#define FUSED_WHOLE 15u

HD void fused_subtri(const FrameArgs &fa, uint32_t e, SubTri &st) {
  uint32_t t = e >> 4, k = e & 15u;
  uint4 rec = fa.tri_rec[t];
  st.draw = rec.w;
  if (k == FUSED_WHOLE) {
    uint32_t v[3] = {rec.x, rec.y, rec.z};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      float4 sc = fa.vscr[v[i]];
      st.sx[i] = sc.x;
      st.sy[i] = sc.y;
      st.sz[i] = sc.z;
      st.iw[i] = sc.w;
      st.a[i] = v[i];
    }
  } else {
    const ViewGPU &vw = fa.views[fa.draws[rec.w].view];
    float hw = (float)vw.w * 0.5f, hh = (float)vw.h * 0.5f;
    uint32_t base = rec.x * MOPH_MAX_POLY;
    uint32_t v[3] = {base, base + k + 1, base + k + 2};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      screen_xform(fa.cpos[v[i]], hw, hh, st.sx[i], st.sy[i], st.sz[i], st.iw[i]);
      st.a[i] = fa.n_verts + v[i];
    }
  }
}

/* Screen triangle s of either path: an index into the screen-triangle store,
 * or an order-free entry (above). */
HD void load_subtri(const FrameArgs &fa, uint32_t s, SubTri &st) {
  if (fa.fused)
    fused_subtri(fa, s, st);
  else
    st = fa.subtris[s];
}


/* The three attribute float4 of attribute slot `slot` (vertex cache or
 * clip store). */
HD void load_attr(const FrameArgs &fa, uint32_t slot, float4 &a0, float4 &a1, float4 &a2) {
  const float4 *p = (slot < fa.n_verts) ? fa.vattr + (size_t)slot * 3
                                        : fa.cattr + (size_t)(slot - fa.n_verts) * 3;
  a0 = p[0];
  a1 = p[1];
  a2 = p[2];
}

/* The attributes of a sub-triangle's three vertices. */
HD void load_tri_attr(const FrameArgs &fa, const SubTri &st, float4 *A0, float4 *A1, float4 *A2) {
  for (int i = 0; i < 3; i++)
    load_attr(fa, st.a[i], A0[i], A1[i], A2[i]);
}

/* What the shading of one pixel needs from its key. */
struct KeyFetch {
  SubTri st;
  RasterTri r; /* inv_e is not set: shading does not use it */
  float4 A0[3], A1[3], A2[3]; /* the vertices' attributes */
  float w0, w1, w2, zf;
  int sm;
  bool flat;
};

/* The screen triangle and edge values of pixel (px, py) (buffer index idx)
 * whose visibility key names screen triangle s; false when the key does not
 * name a triangle of this batch that covers the pixel.  The visibility
 * buffer holds only keys written by this batch's raster (it is clear at the
 * start of every frame); the checks keep every read in bounds even for a
 * key that breaks that: a triangle index outside the batch, record fields
 * outside their stores (records of triangles the raster culled keep earlier
 * frames' contents), a pixel outside the triangle's bbox, a checkpoint
 * outside the store. */
HD bool shade_key_fetch(const FrameArgs &frame, const ViewGPU &fa, uint32_t s, int px, int py,
                        size_t idx, KeyFetch &k) {
  if (frame.fused ? (s >> 4) >= frame.n_tris : s >= frame.subtri_cap)
    return false; /* not a key of this batch */
  if (frame.fused) {
    uint4 rec = frame.tri_rec[s >> 4];
    bool ok = rec.w < frame.ndraws &&
              ((s & 15u) == FUSED_WHOLE
                   ? (rec.x < frame.n_verts && rec.y < frame.n_verts && rec.z < frame.n_verts)
                   : (rec.x < frame.clip_cap && (s & 15u) < MOPH_MAX_POLY - 2));
    if (!ok)
      return false;
  }
  load_subtri(frame, s, k.st);
  uint32_t na = frame.n_verts + frame.clip_cap * MOPH_MAX_POLY;
  if (k.st.draw >= frame.ndraws || k.st.a[0] >= na || k.st.a[1] >= na || k.st.a[2] >= na)
    return false;
  /* the attributes depend on the record alone: issued now, they are in
   * flight while the edge setup and the walk replay run */
  load_tri_attr(frame, k.st, k.A0, k.A1, k.A2);
  RasterTri &r = k.r;
  if (!load_raster_tri_fast(fa.w, fa.h, k.st, r) || px < r.es.min_x || px > r.es.max_x ||
      py < r.es.min_y || py > r.es.max_y)
    return false;
  k.sm = draw_sm(frame.draws[k.st.draw]);
  k.flat = sm_flat(k.sm);
  if (bbox_area(r.es) > (long long)frame.small_area && frame.fused) {
    /* big screen triangle: walk state at the last checkpoint column of the
     * row, then the remaining serial steps */
    uint32_t t = s >> 4, f = s & 15u;
    uint32_t off = frame.big_off[f == FUSED_WHOLE ? t
                                                  : frame.n_tris + frame.tri_rec[t].x * 16u + f];
    if (off == ~0u) /* checkpoint store overflowed: the frame is run again */
      return false;
    uint32_t dx = (uint32_t)(px - r.es.min_x);
    size_t ci = off + (size_t)(py - r.es.min_y) * chk_per_row(r.es) + (dx >> CHK_SHIFT);
    if (ci >= frame.chk_cap)
      return false;
    float4 c = frame.chk[ci];
    k.w0 = c.x;
    k.w1 = c.y;
    k.w2 = c.z;
    k.zf = c.w;
    FlatZ fz = {0.0f, 0.0f, 0.0f};
    if (k.flat)
      fz = flat_z_setup(r);
    for (uint32_t j = dx & ~(CHK_SPAN - 1u); j < dx; j++) {
      k.w0 += r.es.e_dx[0];
      k.w1 += r.es.e_dx[1];
      k.w2 += r.es.e_dx[2];
      k.zf += fz.dx;
    }
  } else if (bbox_area(r.es) > (long long)frame.small_area) {
    float4 w = fa.wbuf[idx];
    k.w0 = w.x;
    k.w1 = w.y;
    k.w2 = w.z;
    k.zf = w.w;
  } else {
    /* replay the reference's incremental walk up to this pixel */
    FlatZ fz = {0.0f, 0.0f, 0.0f};
    if (k.flat)
      fz = flat_z_setup(r);
    replay_walk(r, k.flat, fz, px, py, k.w0, k.w1, k.w2, k.zf);
  }
  return true;
}

#endif
