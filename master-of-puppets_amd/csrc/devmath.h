/*
 * devmath.h -- gfx950 device restatement of the scalar arithmetic the
 * reference rasterizer uses, written so each expression keeps the reference's
 * evaluation order.  Built with -ffp-contract=off (no FMA fusion), IEEE
 * division and square root, denormals preserved; the host reference is SSE
 * binary32 with no contraction, so each operation rounds identically.
 */
#ifndef MOP_HIP_DEVMATH_H
#define MOP_HIP_DEVMATH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEV __device__ __forceinline__
/* also compiled for the host (the shading pass's key checks, shade_key.h) */
#define HD __host__ __device__ __forceinline__

struct V3 {
  float x, y, z;
};

DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
/* math.c:18-50 */
DEV V3 v3add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEV V3 v3sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEV V3 v3scale(V3 v, float s) { return V3{v.x * s, v.y * s, v.z * s}; }
DEV float v3dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV float v3len(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
DEV V3 v3norm(V3 v) {
  float len = v3len(v);
  if (len < 1e-8f)
    return V3{0.0f, 0.0f, 0.0f};
  float inv = 1.0f / len;
  return V3{v.x * inv, v.y * inv, v.z * inv};
}

/* mop_mat4_mul_vec4, math.c:196-203: row r = ((m[r]x + m[4+r]y) + m[8+r]z) + m[12+r]w */
HD float4 mat4_mul_vec4(const float *m, float x, float y, float z, float w) {
  float4 r;
  r.x = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
  r.y = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
  r.z = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
  r.w = m[3] * x + m[7] * y + m[11] * z + m[15] * w;
  return r;
}

/* C's (int) cast of a float as x86-64 executes it (cvttss2si): truncation,
 * and 0x80000000 for NaN or anything outside int32 range. */
HD int cvt_i32(float f) {
  if (!(f >= -2147483648.0f && f < 2147483648.0f))
    return (int)0x80000000;
  return (int)f;
}
/* (uint8_t)(float) as gcc emits it on x86-64: cvttss2si then the low byte. */
DEV uint8_t cvt_u8(float f) { return (uint8_t)(uint32_t)cvt_i32(f); }

/* Order-preserving map of a float to uint32 (NaN excluded by callers);
 * -0.0 is folded onto +0.0 so that ties compare equal as `z < depth` does. */
DEV uint32_t depth_key(float z) {
  uint32_t u = __float_as_uint(z == 0.0f ? 0.0f : z);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
DEV float key_depth(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

/* Perspective divide + viewport, rasterizer.c:2520-2532 (also 396-407). */
HD void screen_xform(float4 p, float half_w, float half_h, float &sx,
                      float &sy, float &sz, float &iw) {
  iw = 1.0f / p.w;
  sx = (p.x * iw + 1.0f) * half_w;
  sy = (1.0f - p.y * iw) * half_h;
  sz = (p.z * iw + 1.0f) * 0.5f;
}

/* Setup shared by the shadow walk (rasterizer.c:409-462) and the smooth
 * multi-light walk (rasterizer.c:1842-1924): clamped bbox, area, flipped
 * edge steps, row-start edge values at the (min_x+.5, min_y+.5) centre. */
struct EdgeSetup {
  int min_x, min_y, max_x, max_y;
  float area, inv_area;
  float e_dx[3], e_dy[3];
  float w_row[3];
};

HD bool edge_setup(const float *sx, const float *sy, int W, int H,
                    EdgeSetup &s) {
  float fmin_x = sx[0];
  if (sx[1] < fmin_x)
    fmin_x = sx[1];
  if (sx[2] < fmin_x)
    fmin_x = sx[2];
  float fmin_y = sy[0];
  if (sy[1] < fmin_y)
    fmin_y = sy[1];
  if (sy[2] < fmin_y)
    fmin_y = sy[2];
  float fmax_x = sx[0];
  if (sx[1] > fmax_x)
    fmax_x = sx[1];
  if (sx[2] > fmax_x)
    fmax_x = sx[2];
  float fmax_y = sy[0];
  if (sy[1] > fmax_y)
    fmax_y = sy[1];
  if (sy[2] > fmax_y)
    fmax_y = sy[2];
  int min_x = cvt_i32(floorf(fmin_x));
  int min_y = cvt_i32(floorf(fmin_y));
  int max_x = cvt_i32(ceilf(fmax_x));
  int max_y = cvt_i32(ceilf(fmax_y));
  if (min_x < 0)
    min_x = 0;
  if (min_y < 0)
    min_y = 0;
  if (max_x >= W)
    max_x = W - 1;
  if (max_y >= H)
    max_y = H - 1;
  if (min_x > max_x || min_y > max_y)
    return false;
  float sx0 = sx[0], sy0 = sy[0], sx1 = sx[1], sy1 = sy[1], sx2 = sx[2],
        sy2 = sy[2];
  float area = (sx1 - sx0) * (sy2 - sy0) - (sx2 - sx0) * (sy1 - sy0);
  if (fabsf(area) < 1e-6f)
    return false;
  bool flip = (area < 0.0f);
  s.area = area;
  s.inv_area = 1.0f / fabsf(area);
  float e0_dx = sy1 - sy2, e0_dy = sx2 - sx1;
  float e1_dx = sy2 - sy0, e1_dy = sx0 - sx2;
  float e2_dx = sy0 - sy1, e2_dy = sx1 - sx0;
  if (flip) {
    e0_dx = -e0_dx;
    e0_dy = -e0_dy;
    e1_dx = -e1_dx;
    e1_dy = -e1_dy;
    e2_dx = -e2_dx;
    e2_dy = -e2_dy;
  }
  float px0 = (float)min_x + 0.5f;
  float py0 = (float)min_y + 0.5f;
  float w0_row = (sx2 - sx1) * (py0 - sy1) - (sy2 - sy1) * (px0 - sx1);
  float w1_row = (sx0 - sx2) * (py0 - sy2) - (sy0 - sy2) * (px0 - sx2);
  float w2_row = (sx1 - sx0) * (py0 - sy0) - (sy1 - sy0) * (px0 - sx0);
  if (flip) {
    w0_row = -w0_row;
    w1_row = -w1_row;
    w2_row = -w2_row;
  }
  s.min_x = min_x;
  s.min_y = min_y;
  s.max_x = max_x;
  s.max_y = max_y;
  s.e_dx[0] = e0_dx;
  s.e_dx[1] = e1_dx;
  s.e_dx[2] = e2_dx;
  s.e_dy[0] = e0_dy;
  s.e_dy[1] = e1_dy;
  s.e_dy[2] = e2_dy;
  s.w_row[0] = w0_row;
  s.w_row[1] = w1_row;
  s.w_row[2] = w2_row;
  return true;
}

/* AA edge scale, rasterizer.c:1898-1903 */
HD float inv_edge_len(float dx, float dy) {
  float len = sqrtf(dx * dx + dy * dy);
  return (len > 1e-6f) ? 1.0f / len : 0.0f;
}

HD long long bbox_area(const EdgeSetup &s) {
  return (long long)(s.max_x - s.min_x + 1) * (long long)(s.max_y - s.min_y + 1);
}

/* checkpoints per bbox row of a big screen triangle: one per CHK_SPAN columns */
#define CHK_SHIFT 4
#define CHK_SPAN (1 << CHK_SHIFT)
HD uint32_t chk_per_row(const EdgeSetup &es) {
  return (uint32_t)(es.max_x - es.min_x + CHK_SPAN) >> CHK_SHIFT;
}

#endif
