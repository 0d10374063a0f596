/*
 * libmf.h -- single-precision atan2f and asinf that return the same bits as
 * the host C library the reference is built against (glibc 2.35 on x86-64,
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
#endif

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

LMF_FN float mop_atanf(float x) {
  int32_t hx = (int32_t)lmf_bits(x), ix = hx & 0x7fffffff, id;
  if (ix >= 0x4c000000) { /* |x| >= 2^25 */
    if (ix > 0x7f800000)
      return x + x; /* NaN */
    return hx > 0 ? lmf_atan_hi[3] + lmf_atan_lo[3] : -lmf_atan_hi[3] - lmf_atan_lo[3];
  }
  if (ix < 0x3ee00000) { /* |x| < 0.4375 */
    if (ix < 0x31000000) /* |x| < 2^-29 */
      return x;
    id = -1;
  } else {
    x = LMF_FABSF(x);
    if (ix < 0x3f980000) {   /* |x| < 1.1875 */
      if (ix < 0x3f300000) { /* 7/16 <= |x| < 11/16 */
        id = 0;
        x = (2.0f * x - 1.0f) / (2.0f + x);
      } else { /* 11/16 <= |x| < 19/16 */
        id = 1;
        x = (x - 1.0f) / (x + 1.0f);
      }
    } else if (ix < 0x401c0000) { /* |x| < 2.4375 */
      id = 2;
      x = (x - 1.5f) / (1.0f + 1.5f * x);
    } else { /* 2.4375 <= |x| < 2^25 */
      id = 3;
      x = -1.0f / x;
    }
  }
  float z = x * x;
  float w = z * z;
  float s1 = z * (lmf_at[0] +
                  w * (lmf_at[2] + w * (lmf_at[4] + w * (lmf_at[6] + w * (lmf_at[8] +
                                                                          w * lmf_at[10])))));
  float s2 = w * (lmf_at[1] + w * (lmf_at[3] + w * (lmf_at[5] + w * (lmf_at[7] + w * lmf_at[9]))));
  if (id < 0)
    return x - x * (s1 + s2);
  z = lmf_atan_hi[id] - ((x * (s1 + s2) - lmf_atan_lo[id]) - x);
  return hx < 0 ? -z : z;
}

LMF_FN float mop_atan2f(float y, float x) {
  const float pi_o_4 = 7.8539818525e-01f, pi_o_2 = 1.5707963705e+00f,
              pi = 3.1415927410e+00f, pi_lo = -8.7422776573e-08f, tiny = 1.0e-30f;
  int32_t hx = (int32_t)lmf_bits(x), ix = hx & 0x7fffffff;
  int32_t hy = (int32_t)lmf_bits(y), iy = hy & 0x7fffffff;
  if (ix > 0x7f800000 || iy > 0x7f800000)
    return x + y; /* NaN */
  if (hx == 0x3f800000)
    return mop_atanf(y); /* x = 1 */
  int m = ((hy >> 31) & 1) | ((hx >> 30) & 2); /* 2 sign(x) + sign(y) */
  if (iy == 0) {
    if (m <= 1)
      return y;
    return m == 2 ? pi + tiny : -pi - tiny;
  }
  if (ix == 0)
    return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  if (ix == 0x7f800000) {
    if (iy == 0x7f800000) {
      switch (m) {
      case 0: return pi_o_4 + tiny;
      case 1: return -pi_o_4 - tiny;
      case 2: return 3.0f * pi_o_4 + tiny;
      default: return -3.0f * pi_o_4 - tiny;
      }
    }
    switch (m) {
    case 0: return 0.0f;
    case 1: return -0.0f;
    case 2: return pi + tiny;
    default: return -pi - tiny;
    }
  }
  if (iy == 0x7f800000)
    return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
  int32_t k = (iy - ix) >> 23;
  float z;
  if (k > 60)
    z = pi_o_2 + 0.5f * pi_lo; /* |y/x| > 2^60 */
  else if (hx < 0 && k < -60)
    z = 0.0f;
  else
    z = mop_atanf(LMF_FABSF(y / x));
  switch (m) {
  case 0: return z;
  case 1: return lmf_float(lmf_bits(z) ^ 0x80000000u);
  case 2: return pi - (z - pi_lo);
  default: return (z - pi_lo) - pi;
  }
}

LMF_FN float mop_asinf(float x) {
  const float pio2_hi = 1.57079637050628662109375f, pio2_lo = -4.37113900018624283e-8f,
              pio4_hi = 0.785398185253143310546875f;
  /* asin x = x + x^3 p(x^2) on |x| <= 0.5 */
  const float p0 = 1.666675248e-1f, p1 = 7.495297643e-2f, p2 = 4.547037598e-2f,
              p3 = 2.417951451e-2f, p4 = 4.216630880e-2f;
  int32_t hx = (int32_t)lmf_bits(x), ix = hx & 0x7fffffff;
  if (ix == 0x3f800000)
    return x * pio2_hi + x * pio2_lo; /* +-1 */
  if (ix > 0x3f800000)
    return (x - x) / (x - x); /* |x| > 1: NaN */
  if (ix < 0x3f000000) {      /* |x| < 0.5 */
    if (ix < 0x32000000)      /* |x| < 2^-27 */
      return x;
    float t = x * x;
    float w = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
    return x + x * w;
  }
  float w = 1.0f - LMF_FABSF(x);
  float t = w * 0.5f;
  float p = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
  float s = LMF_SQRTF(t);
  if (ix >= 0x3F79999A) { /* |x| > 0.975 */
    t = pio2_hi - (2.0f * (s + s * p) - pio2_lo);
  } else {
    float wh = lmf_float(lmf_bits(s) & 0xfffff000u); /* s with its low 12 bits cleared */
    float c = (t - wh * wh) / (s + wh);
    float r = p;
    p = 2.0f * s * r - (pio2_lo - 2.0f * c);
    float q = pio4_hi - 2.0f * wh;
    t = pio4_hi - (p - q);
  }
  return hx > 0 ? t : -t;
}

#endif
