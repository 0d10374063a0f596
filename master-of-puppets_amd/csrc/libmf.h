/*
 * libmf.h -- single-precision atan2f and asinf that return the same bits as
 * the host C library the reference is built against, and a division by a
 * precomputed reciprocal that returns IEEE division's bits (mop_div_recip).
 *
 * atan2f, asinf: the host library is glibc 2.35 on x86-64,
 * whose float versions are not correctly rounded: a correctly rounded
 * atan2f differs from it on about 16% of arguments, asinf on about 7%).
 *
 * The reference samples its IBL tables at atan2f/asinf of the shading
 * normal and reflection vector (rasterizer.c:210-216) and builds the
 * tables from the same calls (environment.c:62-68); bit-identical shading
 * needs bit-identical angles.
 *
 * Methods (published, restated here, not copied):
 *  - atanf: fdlibm's argument reduction to |x| < 7/16 around the breakpoints
 *    atan(0.5), atan(1), atan(1.5), atan(inf) (hi/lo split constants) and an
 *    odd polynomial of degree 23 split into even/odd halves in z = x^2;
 *  - atan2f: fdlibm's quadrant logic on top of atanf(|y/x|) with pi and a
 *    pi lo-word correction;
 *  - asinf: x + x^3 p(x^2) with the degree-4 Cephes minimax polynomial for
 *    |x| < 0.5, and for |x| >= 0.5 the half-angle identity
 *    asin(x) = pi/2 - 2 asin(sqrt((1-|x|)/2)) with fdlibm's hi/lo split of
 *    the square root below |x| = 0.975.
 * Every operation is a plain IEEE single-precision add, multiply, divide or
 * square root with no fused multiply-add (the translation unit is built with
 * contraction off).  Agreement with the host library was checked bit for
 * bit: asinf on every float in [-1, 1], atanf on every finite float, atan2f
 * on 4e8 random argument pairs (tests/test_libmf.py repeats a sample).
 *
 * Usable from C (host checks) and from HIP device code.
 */
#ifndef MOP_HIP_LIBMF_H
#define MOP_HIP_LIBMF_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
#include <hip/hip_runtime.h>
#define LMF_FN __device__ __forceinline__
#define LMF_CONST __device__ static const
static __device__ __forceinline__ uint32_t lmf_bits(float f) { return __float_as_uint(f); }
static __device__ __forceinline__ float lmf_float(uint32_t u) { return __uint_as_float(u); }
#define LMF_SQRTF sqrtf
#define LMF_FABSF fabsf
#define LMF_FMAF __builtin_fmaf
#else
#include <math.h>
#include <string.h>
#define LMF_FN static inline
#define LMF_CONST static const
static inline uint32_t lmf_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static inline float lmf_float(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
#define LMF_SQRTF sqrtf
#define LMF_FABSF fabsf
#define LMF_FMAF fmaf
#endif

/* a / b from y = 1 / b (an IEEE division, computed once for a constant or
 * shared divisor): q = a y, then q + (a - q b) y with the residual exact by
 * fused multiply-add (Markstein's correction), which is the correctly
 * rounded quotient when y is the correctly rounded reciprocal and no
 * intermediate leaves the normal range.  Quotients of magnitude outside
 * [2^-99, 2^97] (zeros, tiny, huge, NaN), numerators outside [2^-100, 2^100]
 * and reciprocals outside [2^-120, 2^120] take the division (there the
 * residual or the reciprocal would lose bits to underflow).  Checked against
 * IEEE division on every float numerator for the constant divisors the
 * shading uses and on random pairs (tests/test_divide.py). */
LMF_FN float mop_div_recip(float a, float b, float y) {
  float q = a * y;
  float aq = LMF_FABSF(q), ay = LMF_FABSF(y), aa = LMF_FABSF(a);
  if (!(aq >= 0x1p-99f && aq <= 0x1p97f && ay >= 0x1p-120f && ay <= 0x1p120f &&
        aa >= 0x1p-100f && aa <= 0x1p100f))
    return a / b;
  float r = LMF_FMAF(-q, b, a);
  return LMF_FMAF(r, y, q);
}

/* atan at the reduction breakpoints, split hi + lo */
LMF_CONST float lmf_atan_hi[4] = {4.6364760399e-01f, 7.8539812565e-01f, 9.8279368877e-01f,
                                  1.5707962513e+00f};
LMF_CONST float lmf_atan_lo[4] = {5.0121582440e-09f, 3.7748947079e-08f, 3.4473217170e-08f,
                                  7.5497894159e-08f};
/* odd series coefficients of atan on |x| < 7/16 */
LMF_CONST float lmf_at[11] = {3.3333334327e-01f,  -2.0000000298e-01f, 1.4285714924e-01f,
                              -1.1111110449e-01f, 9.0908870101e-02f,  -7.6918758452e-02f,
                              6.6610731184e-02f,  -5.8335702866e-02f, 4.9768779427e-02f,
                              -3.6531571299e-02f, 1.6285819933e-02f};

/* The three functions below evaluate every path and select the result, so
 * that the lanes of a wave never diverge: one division and one polynomial
 * per atanf (the argument reduction of each interval is a quotient, |x|/1
 * for the smallest), one division, one square root and one polynomial per
 * asinf.  Each selected value is the same expression as the branch it
 * stands for, so the results are the branchy forms' bit for bit. */

LMF_FN float mop_atanf(float x) {
  int32_t hx = (int32_t)lmf_bits(x), ix = hx & 0x7fffffff;
  float ax = LMF_FABSF(x);
  /* reduction interval: -1 below 7/16 (no reduction), else 0..3 */
  int id = ix < 0x3ee00000 ? -1
           : ix < 0x3f300000 ? 0    /* 7/16 <= |x| < 11/16 */
           : ix < 0x3f980000 ? 1    /* 11/16 <= |x| < 19/16 */
           : ix < 0x401c0000 ? 2    /* 19/16 <= |x| < 2.4375 */
                             : 3;   /* 2.4375 <= |x| */
  float num = id < 0 ? x : id == 0 ? 2.0f * ax - 1.0f : id == 1 ? ax - 1.0f
            : id == 2 ? ax - 1.5f : -1.0f;
  float den = id < 0 ? 1.0f : id == 0 ? 2.0f + ax : id == 1 ? ax + 1.0f
            : id == 2 ? 1.0f + 1.5f * ax : ax;
  float t = num / den; /* x / 1 is x */
  float z = t * t;
  float w = z * z;
  float s1 = z * (lmf_at[0] +
                  w * (lmf_at[2] + w * (lmf_at[4] + w * (lmf_at[6] + w * (lmf_at[8] +
                                                                          w * lmf_at[10])))));
  float s2 = w * (lmf_at[1] + w * (lmf_at[3] + w * (lmf_at[5] + w * (lmf_at[7] + w * lmf_at[9]))));
  float hi = id <= 0 ? lmf_atan_hi[0] : id == 1 ? lmf_atan_hi[1] : id == 2 ? lmf_atan_hi[2]
                                                                          : lmf_atan_hi[3];
  float lo = id <= 0 ? lmf_atan_lo[0] : id == 1 ? lmf_atan_lo[1] : id == 2 ? lmf_atan_lo[2]
                                                                          : lmf_atan_lo[3];
  float big = hi - ((t * (s1 + s2) - lo) - t);
  float r = id < 0 ? t - t * (s1 + s2) : (hx < 0 ? -big : big);
  if (ix < 0x31000000) /* |x| < 2^-29 */
    r = x;
  if (ix >= 0x4c000000) /* |x| >= 2^25, or NaN */
    r = ix > 0x7f800000 ? x + x
                        : (hx > 0 ? lmf_atan_hi[3] + lmf_atan_lo[3] : -lmf_atan_hi[3] - lmf_atan_lo[3]);
  return r;
}

LMF_FN float mop_atan2f(float y, float x) {
  const float pi_o_4 = 7.8539818525e-01f, pi_o_2 = 1.5707963705e+00f,
              pi = 3.1415927410e+00f, pi_lo = -8.7422776573e-08f, tiny = 1.0e-30f;
  int32_t hx = (int32_t)lmf_bits(x), ix = hx & 0x7fffffff;
  int32_t hy = (int32_t)lmf_bits(y), iy = hy & 0x7fffffff;
  int m = ((hy >> 31) & 1) | ((hx >> 30) & 2); /* 2 sign(x) + sign(y) */
  int x_one = hx == 0x3f800000;
  /* atanf(y) when x = 1, else atanf(|y/x|) */
  float za = mop_atanf(x_one ? y : LMF_FABSF(y / x));
  int32_t k = (iy - ix) >> 23;
  float z = k > 60 ? pi_o_2 + 0.5f * pi_lo : ((hx < 0 && k < -60) ? 0.0f : za);
  float r = m == 0 ? z : m == 1 ? lmf_float(lmf_bits(z) ^ 0x80000000u)
          : m == 2 ? pi - (z - pi_lo) : (z - pi_lo) - pi;
  /* the special cases, lowest precedence first */
  if (iy == 0x7f800000)
    r = hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (ix == 0x7f800000) {
    if (iy == 0x7f800000)
      r = m == 0 ? pi_o_4 + tiny : m == 1 ? -pi_o_4 - tiny
        : m == 2 ? 3.0f * pi_o_4 + tiny : -3.0f * pi_o_4 - tiny;
    else
      r = m == 0 ? 0.0f : m == 1 ? -0.0f : m == 2 ? pi + tiny : -pi - tiny;
  }
  if (ix == 0)
    r = hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (iy == 0)
    r = m <= 1 ? y : (m == 2 ? pi + tiny : -pi - tiny);
  if (x_one)
    r = za;
  if (ix > 0x7f800000 || iy > 0x7f800000)
    r = x + y; /* NaN */
  return r;
}

LMF_FN float mop_asinf(float x) {
  const float pio2_hi = 1.57079637050628662109375f, pio2_lo = -4.37113900018624283e-8f,
              pio4_hi = 0.785398185253143310546875f;
  /* asin x = x + x^3 p(x^2) on |x| <= 0.5 */
  const float p0 = 1.666675248e-1f, p1 = 7.495297643e-2f, p2 = 4.547037598e-2f,
              p3 = 2.417951451e-2f, p4 = 4.216630880e-2f;
  int32_t hx = (int32_t)lmf_bits(x), ix = hx & 0x7fffffff;
  int small = ix < 0x3f000000; /* |x| < 0.5 */
  /* the polynomial's argument: x^2 below 0.5, (1 - |x|) / 2 above */
  float t = small ? x * x : (1.0f - LMF_FABSF(x)) * 0.5f;
  float p = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
  float s = LMF_SQRTF(t);
  /* |x| > 0.975 */
  float ra = pio2_hi - (2.0f * (s + s * p) - pio2_lo);
  /* 0.5 <= |x| <= 0.975: s with its low 12 bits cleared, and the correction */
  float wh = lmf_float(lmf_bits(s) & 0xfffff000u);
  float c = (t - wh * wh) / (s + wh);
  float pp = 2.0f * s * p - (pio2_lo - 2.0f * c);
  float q = pio4_hi - 2.0f * wh;
  float rb = pio4_hi - (pp - q);
  float large = ix >= 0x3F79999A ? ra : rb;
  float r = small ? (ix < 0x32000000 ? x : x + x * p) : (hx > 0 ? large : -large);
  if (ix == 0x3f800000)
    r = x * pio2_hi + x * pio2_lo; /* +-1 */
  if (ix > 0x3f800000)
    r = (x - x) / (x - x); /* |x| > 1: NaN */
  return r;
}

#endif
