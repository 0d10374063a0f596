/*
 * cull.h -- exact culling of a triangle chunk against the trivial frustum
 * reject of the raster (rasterizer.c:2469-2475): a triangle whose three
 * clip-space vertices all lie outside one of the six planes x < -w, x > w,
 * y < -w, y > w, z < -w, z > w produces nothing.  box_rejected answers for
 * every triangle whose vertices lie in an object-space box at once, so a
 * chunk whose box passes it can be skipped with the same result.
 *
 * Compiled for the device by kernels.hip and for the host by
 * tests/cull_check.cpp, which checks the margin against the vertex
 * transform's own arithmetic on random boxes, matrices and points.
 */
#ifndef MOP_HIP_CULL_H
#define MOP_HIP_CULL_H

#include "devmath.h"

/* The plane (0..5 in the order above) that every point of the box fails,
 * or -1.  The clip-space plane functions are affine in the object-space point, so
 * over a box each takes its largest value at a corner.  The vertex
 * transform (mat4_mul_vec4, then z += bias * w) rounds: a row
 * ((m0 x + m4 y) + m8 z) + m12 is within 4u (u = 2^-24) of its exact value
 * times the row's magnitude bound A = |m0| X + |m4| Y + |m8| Z + |m12| over
 * the box (X, Y, Z the largest coordinate magnitudes).  A plane is taken as
 * failed by the whole box when every corner, computed the same way, fails
 * it by more than 2e-6 times the magnitude bounds of the rows involved --
 * over three times what the rounding of a corner and of a vertex can move
 * the two values together -- plus a tiny absolute term for denormals.
 * Boxes or matrices too large for the bounds, and NaN, never cull. */
HD int box_rejected_plane(const float *m, float bias, float4 lo, float4 hi) {
  if (!(lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z))
    return -1;
  float X = fmaxf(fabsf(lo.x), fabsf(hi.x)), Y = fmaxf(fabsf(lo.y), fabsf(hi.y)),
        Z = fmaxf(fabsf(lo.z), fabsf(hi.z));
  float A[4];
  for (int r = 0; r < 4; r++)
    A[r] = fabsf(m[r]) * X + fabsf(m[4 + r]) * Y + fabsf(m[8 + r]) * Z + fabsf(m[12 + r]);
  float ab = fabsf(bias);
  float mx = 2e-6f * (A[0] + A[3]) + 1e-30f;
  float my = 2e-6f * (A[1] + A[3]) + 1e-30f;
  float mz = 2e-6f * (A[2] + (1.0f + ab) * A[3]) + 1e-30f;
  if (!(mx < 1e30f && my < 1e30f && mz < 1e30f))
    return -1;
  bool o0 = true, o1 = true, o2 = true, o3 = true, o4 = true, o5 = true;
  for (int c = 0; c < 8; c++) {
    float4 v = mat4_mul_vec4(m, (c & 1) ? hi.x : lo.x, (c & 2) ? hi.y : lo.y,
                             (c & 4) ? hi.z : lo.z, 1.0f);
    float z = bias != 0.0f ? v.z + bias * v.w : v.z;
    o0 = o0 && v.x + v.w < -mx; /* x < -w */
    o1 = o1 && v.x - v.w > mx;  /* x > w */
    o2 = o2 && v.y + v.w < -my;
    o3 = o3 && v.y - v.w > my;
    o4 = o4 && z + v.w < -mz;
    o5 = o5 && z - v.w > mz;
  }
  return o0 ? 0 : o1 ? 1 : o2 ? 2 : o3 ? 3 : o4 ? 4 : o5 ? 5 : -1;
}

/* true when every triangle with its vertices in the box is trivially
 * rejected (the plane box_rejected_plane names fails at all of them) */
HD bool box_rejected(const float *m, float bias, float4 lo, float4 hi) {
  return box_rejected_plane(m, bias, lo, hi) >= 0;
}

#endif
