/*
 * ibl.hip -- device precompute of the image-based-lighting tables from an
 * equirectangular HDR environment, restating environment.c:115-484:
 *   irradiance   64x32,  128 cosine-weighted Hammersley samples per texel
 *   prefiltered  128x64 and 4 halved levels (roughness l/4), 64 GGX samples
 *   BRDF LUT     256x256 split-sum (scale, bias), 64 GGX samples
 * One lane per output texel.  The tables feed the shading kernel exactly as
 * mop_sw_ibl_set's arrays do, and come out bit-identical to the host
 * precompute: every sinf/cosf the precompute calls takes one of a fixed set
 * of arguments (texel-centre angles per row and column of each table, and
 * 2 pi xi2 of each Hammersley sample), which the host evaluates with its C
 * library into a small table (ibl_trig_fill); atan2f/asinf of the sample
 * directions use the host library's algorithms restated in libmf.h.
 */
#include <math.h>

#include "devmath.h"
#include "libmf.h"
#include "mop_hip_internal.h"

#define IRR_W 64
#define IRR_H 32
#define IRR_SAMPLES 128
#define PF_W 128
#define PF_H 64
#define PF_LEVELS 5
#define PF_SAMPLES 64
#define LUT_N 256
#define LUT_SAMPLES 64

static const float PI_F = 3.14159265358979323846f;

/* sample_hdr_bilinear, environment.c:32-59 (RGB only is used) */
DEV void env_bilinear(const float *data, int w, int h, float u, float v, float out[3]) {
  u = u - floorf(u);
  v = v - floorf(v);
  float fx = u * (float)(w - 1);
  float fy = v * (float)(h - 1);
  int x0 = cvt_i32(fx), y0 = cvt_i32(fy);
  if (x0 < 0 || x0 >= w) x0 = 0;
  if (y0 < 0 || y0 >= h) y0 = 0;
  int x1 = (x0 + 1) % w;
  int y1 = y0 + 1;
  if (y1 >= h)
    y1 = h - 1;
  float sx = fx - (float)x0, sy = fy - (float)y0;
  const float *p00 = &data[((size_t)y0 * w + x0) * 4];
  const float *p10 = &data[((size_t)y0 * w + x1) * 4];
  const float *p01 = &data[((size_t)y1 * w + x0) * 4];
  const float *p11 = &data[((size_t)y1 * w + x1) * 4];
  for (int c = 0; c < 3; c++) {
    float top = p00[c] * (1.0f - sx) + p10[c] * sx;
    float bot = p01[c] * (1.0f - sx) + p11[c] * sx;
    out[c] = top * (1.0f - sy) + bot * sy;
  }
}

/* dir_to_equirect, environment.c:62-68 */
DEV void dir_to_equirect(float dx, float dy, float dz, float &u, float &v) {
  float phi = mop_atan2f(dz, dx) + 0.0f;
  float theta = mop_asinf(dy < -1.0f ? -1.0f : (dy > 1.0f ? 1.0f : dy));
  u = phi / (2.0f * PI_F) + 0.5f;
  v = theta / PI_F + 0.5f;
}

/* Layout of the host-evaluated sine/cosine table (floats):
 *   [0, 128)  cosf(2 pi xi2(s)), s < 128   [128, 256) sinf of the same
 *   then per table (irradiance 64x32, prefiltered level 0..4): cosf and
 *   sinf of each column's phi, then cosf and sinf of each row's theta. */
#define TRIG_SAMPLES 128
#define TRIG_IRR (2 * TRIG_SAMPLES)
#define TRIG_PF0 (TRIG_IRR + 2 * IRR_W + 2 * IRR_H)

DEV size_t trig_level(int level) {
  size_t off = TRIG_PF0;
  for (int l = 0; l < level; l++) {
    int w = PF_W >> l, h = PF_H >> l;
    off += (size_t)(2 * (w < 1 ? 1 : w) + 2 * (h < 1 ? 1 : h));
  }
  return off;
}

/* texel -> unit direction + tangent frame (environment.c:127-161); tab: the
 * table's block of the sine/cosine table (cos phi[w], sin phi[w], cos
 * theta[h], sin theta[h]) */
DEV void texel_frame(const float *tab, int x, int y, int w, int h, float n[3], float t[3],
                     float b[3]) {
  float ct = tab[2 * w + y];
  float nx = ct * tab[x], ny = tab[2 * w + h + y], nz = ct * tab[w + x];
  float len = sqrtf(nx * nx + ny * ny + nz * nz);
  nx /= len;
  ny /= len;
  nz /= len;
  float upx = 0.0f, upy = 1.0f, upz = 0.0f;
  if (fabsf(ny) > 0.99f) {
    upx = 0.0f;
    upy = 0.0f;
    upz = 1.0f;
  }
  float tx = upy * nz - upz * ny;
  float ty = upz * nx - upx * nz;
  float tz = upx * ny - upy * nx;
  float tlen = sqrtf(tx * tx + ty * ty + tz * tz);
  if (tlen > 1e-6f) {
    tx /= tlen;
    ty /= tlen;
    tz /= tlen;
  }
  n[0] = nx; n[1] = ny; n[2] = nz;
  t[0] = tx; t[1] = ty; t[2] = tz;
  b[0] = ny * tz - nz * ty;
  b[1] = nz * tx - nx * tz;
  b[2] = nx * ty - ny * tx;
}

__global__ void k_irradiance(const float *env, int ew, int eh, float *irr, const float *trig) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= IRR_W * IRR_H)
    return;
  int x = i % IRR_W, y = i / IRR_W;
  float n[3], t[3], b[3];
  texel_frame(trig + TRIG_IRR, x, y, IRR_W, IRR_H, n, t, b);
  float acc[3] = {0, 0, 0};
  float total = 0.0f;
  for (int s = 0; s < IRR_SAMPLES; s++) {
    float xi1 = (float)s / (float)IRR_SAMPLES;
    float cos_t = sqrtf(xi1);
    float sin_t = sqrtf(1.0f - xi1);
    float lx = sin_t * trig[s], ly = cos_t, lz = sin_t * trig[TRIG_SAMPLES + s];
    float wx = t[0] * lx + n[0] * ly + b[0] * lz;
    float wy = t[1] * lx + n[1] * ly + b[1] * lz;
    float wz = t[2] * lx + n[2] * ly + b[2] * lz;
    float su, sv, smp[3];
    dir_to_equirect(wx, wy, wz, su, sv);
    env_bilinear(env, ew, eh, su, sv, smp);
    acc[0] += smp[0];
    acc[1] += smp[1];
    acc[2] += smp[2];
    total += 1.0f;
  }
  float *o = irr + (size_t)i * 4;
  o[0] = acc[0] / total;
  o[1] = acc[1] / total;
  o[2] = acc[2] / total;
  o[3] = 1.0f;
}

__global__ void k_prefiltered(const float *env, int ew, int eh, float *pf, const float *trig) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int level = 0, w = PF_W, h = PF_H;
  size_t offset = 0;
  int rem = i;
  while (level < PF_LEVELS) {
    w = PF_W >> level;
    h = PF_H >> level;
    if (w < 1) w = 1;
    if (h < 1) h = 1;
    if (rem < w * h)
      break;
    rem -= w * h;
    offset += (size_t)w * h;
    level++;
  }
  if (level >= PF_LEVELS)
    return;
  int x = rem % w, y = rem / w;
  float roughness = (float)level / (float)(PF_LEVELS - 1);
  float alpha = roughness * roughness;
  float alpha2 = alpha * alpha;
  if (alpha2 < 1e-7f)
    alpha2 = 1e-7f;
  float n[3], t[3], b[3];
  texel_frame(trig + trig_level(level), x, y, w, h, n, t, b);
  float vx = n[0], vy = n[1], vz = n[2];
  float acc[3] = {0, 0, 0};
  float total = 0.0f;
  for (int s = 0; s < PF_SAMPLES; s++) {
    float xi1 = (float)s / (float)PF_SAMPLES;
    float cos_h = sqrtf((1.0f - xi1) / (1.0f + (alpha2 - 1.0f) * xi1));
    float sin_h = sqrtf(1.0f - cos_h * cos_h);
    float hxl = sin_h * trig[s], hyl = cos_h, hzl = sin_h * trig[TRIG_SAMPLES + s];
    float hx = t[0] * hxl + n[0] * hyl + b[0] * hzl;
    float hy = t[1] * hxl + n[1] * hyl + b[1] * hzl;
    float hz = t[2] * hxl + n[2] * hyl + b[2] * hzl;
    float vdh = vx * hx + vy * hy + vz * hz;
    float lx = 2.0f * vdh * hx - vx;
    float ly = 2.0f * vdh * hy - vy;
    float lz = 2.0f * vdh * hz - vz;
    float ndl = n[0] * lx + n[1] * ly + n[2] * lz;
    if (ndl <= 0.0f)
      continue;
    float su, sv, smp[3];
    dir_to_equirect(lx, ly, lz, su, sv);
    env_bilinear(env, ew, eh, su, sv, smp);
    acc[0] += smp[0] * ndl;
    acc[1] += smp[1] * ndl;
    acc[2] += smp[2] * ndl;
    total += ndl;
  }
  float *o = pf + (offset + (size_t)y * w + x) * 4;
  if (total > 0.0f) {
    o[0] = acc[0] / total;
    o[1] = acc[1] / total;
    o[2] = acc[2] / total;
  } else {
    o[0] = o[1] = o[2] = 0.0f;
  }
  o[3] = 1.0f;
}

__global__ void k_brdf_lut(float *lut, const float *trig) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= LUT_N * LUT_N)
    return;
  int x = i % LUT_N, y = i / LUT_N;
  float roughness = ((float)y + 0.5f) / (float)LUT_N;
  float alpha = roughness * roughness;
  float alpha2 = alpha * alpha;
  if (alpha2 < 1e-7f)
    alpha2 = 1e-7f;
  float k = alpha * 0.5f;
  float ndotv = ((float)x + 0.5f) / (float)LUT_N;
  if (ndotv < 1e-4f)
    ndotv = 1e-4f;
  float sin_v = sqrtf(1.0f - ndotv * ndotv);
  float vx_t = sin_v, vy_t = 0.0f, vz_t = ndotv;
  float scale = 0.0f, bias = 0.0f;
  for (int s = 0; s < LUT_SAMPLES; s++) {
    float xi1 = (float)s / (float)LUT_SAMPLES;
    float cos_h = sqrtf((1.0f - xi1) / (1.0f + (alpha2 - 1.0f) * xi1));
    float sin_h = sqrtf(1.0f - cos_h * cos_h);
    float hx = sin_h * trig[s], hy = sin_h * trig[TRIG_SAMPLES + s], hz = cos_h;
    float vdh = vx_t * hx + vy_t * hy + vz_t * hz;
    float lz = 2.0f * vdh * hz - vz_t;
    float ndl = lz;
    if (ndl <= 0.0f)
      continue;
    float ndh = hz;
    float vdh_c = vdh > 0.0f ? vdh : 0.0f;
    float g1v_t = ndotv / (ndotv * (1.0f - k) + k);
    float g1l_t = ndl / (ndl * (1.0f - k) + k);
    float g = g1v_t * g1l_t;
    float vis = (g * vdh_c) / (ndh * ndotv + 1e-6f);
    float fc = 1.0f - vdh_c;
    float fc2 = fc * fc;
    float fc5 = fc2 * fc2 * fc;
    scale += vis * (1.0f - fc5);
    bias += vis * fc5;
  }
  float *o = lut + (size_t)i * 4;
  o[0] = scale / (float)LUT_SAMPLES;
  o[1] = bias / (float)LUT_SAMPLES;
  o[2] = 0.0f;
  o[3] = 1.0f;
}

size_t ibl_pf_texels(void) {
  size_t n = 0;
  for (int l = 0; l < PF_LEVELS; l++) {
    int w = PF_W >> l, h = PF_H >> l;
    if (w < 1) w = 1;
    if (h < 1) h = 1;
    n += (size_t)w * h;
  }
  return n;
}

/* The sine/cosine table the precompute kernels read, evaluated on the host
 * with the host C library (called through pointers the compiler cannot see
 * through, so nothing is folded at build time).  Each argument is formed
 * with the precompute's own single-precision expressions
 * (environment.c:126-130, 181, 329, 434). */
size_t ibl_trig_floats(void) {
  size_t n = TRIG_PF0;
  for (int l = 0; l < PF_LEVELS; l++) {
    int w = PF_W >> l, h = PF_H >> l;
    n += (size_t)(2 * (w < 1 ? 1 : w) + 2 * (h < 1 ? 1 : h));
  }
  return n;
}

static float (*volatile host_cosf)(float) = cosf;
static float (*volatile host_sinf)(float) = sinf;

static void trig_block(float *tab, int w, int h) {
  for (int x = 0; x < w; x++) {
    float u = ((float)x + 0.5f) / (float)w;
    float phi = (u - 0.5f) * 2.0f * PI_F;
    tab[x] = host_cosf(phi);
    tab[w + x] = host_sinf(phi);
  }
  for (int y = 0; y < h; y++) {
    float v = ((float)y + 0.5f) / (float)h;
    float theta = (v - 0.5f) * PI_F;
    tab[2 * w + y] = host_cosf(theta);
    tab[2 * w + h + y] = host_sinf(theta);
  }
}

void ibl_trig_fill(float *tab) {
  for (int s = 0; s < TRIG_SAMPLES; s++) {
    uint32_t bits = (uint32_t)s;
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    float xi2 = (float)bits * 2.3283064365386963e-10f;
    float phi = 2.0f * PI_F * xi2;
    tab[s] = host_cosf(phi);
    tab[TRIG_SAMPLES + s] = host_sinf(phi);
  }
  trig_block(tab + TRIG_IRR, IRR_W, IRR_H);
  size_t off = TRIG_PF0;
  for (int l = 0; l < PF_LEVELS; l++) {
    int w = PF_W >> l, h = PF_H >> l;
    if (w < 1) w = 1;
    if (h < 1) h = 1;
    trig_block(tab + off, w, h);
    off += (size_t)(2 * w + 2 * h);
  }
}

void launch_ibl_precompute(const float *env, int ew, int eh, float *irr, float *pf, float *lut,
                           const float *trig, hipStream_t s) {
  hipLaunchKernelGGL(k_irradiance, dim3((IRR_W * IRR_H + 63) / 64), dim3(64), 0, s, env, ew,
                     eh, irr, trig);
  size_t npf = ibl_pf_texels();
  hipLaunchKernelGGL(k_prefiltered, dim3((unsigned)((npf + 63) / 64)), dim3(64), 0, s, env, ew,
                     eh, pf, trig);
  hipLaunchKernelGGL(k_brdf_lut, dim3((LUT_N * LUT_N + 255) / 256), dim3(256), 0, s, lut, trig);
}
