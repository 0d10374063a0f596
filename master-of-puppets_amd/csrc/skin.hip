// skin.hip -- linear-blend skinning and morph-target kernels for gfx950.
//
// The device side of the skinning and morph dispatch hooks: one lane per
// vertex rewrites the position (and normal) of a copy of the bind-pose
// vertex in the draw's own vertex buffer, with the arithmetic of
// mop_skin_apply (viewport.c:1572-1680) and mop_morph_apply
// (viewport.c:1763-1805) kept in the same order.  Attributes are read
// through byte copies, so any stride and offset the vertex format declares
// is valid.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "mop_hip_internal.h"

#define DEV __device__ __forceinline__

namespace {

DEV float ld_f(const uint8_t *p) {
  float v;
  memcpy(&v, p, 4);
  return v;
}

DEV void st_f(uint8_t *p, float v) { memcpy(p, &v, 4); }

/* copy one vertex record; dword copies when the record is dword aligned */
DEV void copy_vertex(const uint8_t *src, uint8_t *dst, uint32_t stride, bool aligned) {
  if (aligned) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    for (uint32_t k = 0; k < stride / 4; k++)
      d[k] = s[k];
  } else {
    for (uint32_t k = 0; k < stride; k++)
      dst[k] = src[k];
  }
}

__global__ __launch_bounds__(256) void k_skin(SkinArgs a) {
  uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= a.vertex_count)
    return;
  const uint8_t *vert = a.bind + (size_t)v * a.stride;
  uint8_t *out = a.out + (size_t)v * a.stride;
  copy_vertex(vert, out, a.stride, a.aligned);
  const uint8_t *jd = vert + a.joints_off;
  uint32_t jb[4] = {jd[0], jd[1], jd[2], jd[3]};
  const uint8_t *wd = vert + a.weights_off;
  float wb[4] = {ld_f(wd), ld_f(wd + 4), ld_f(wd + 8), ld_f(wd + 12)};
  const uint8_t *pd = vert + a.pos_off;
  float px = ld_f(pd), py = ld_f(pd + 4), pz = ld_f(pd + 8);
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int b = 0; b < 4; b++) {
    if (wb[b] <= 0.0f)
      continue;
    uint32_t bi = jb[b];
    if (bi >= a.bone_count)
      continue;
    const float *m = a.bones + (size_t)bi * 16;
    float rx = m[0] * px + m[4] * py + m[8] * pz + m[12];
    float ry = m[1] * px + m[5] * py + m[9] * pz + m[13];
    float rz = m[2] * px + m[6] * py + m[10] * pz + m[14];
    sx += wb[b] * rx;
    sy += wb[b] * ry;
    sz += wb[b] * rz;
  }
  uint8_t *op = out + a.pos_off;
  st_f(op, sx);
  st_f(op + 4, sy);
  st_f(op + 8, sz);
  if (a.normal_off >= 0) {
    const uint8_t *nd = vert + a.normal_off;
    float nx = ld_f(nd), ny = ld_f(nd + 4), nz = ld_f(nd + 8);
    float snx = 0.0f, sny = 0.0f, snz = 0.0f;
    for (int b = 0; b < 4; b++) {
      if (wb[b] <= 0.0f)
        continue;
      uint32_t bi = jb[b];
      if (bi >= a.bone_count)
        continue;
      const float *m = a.bones + (size_t)bi * 16;
      float rx = m[0] * nx + m[4] * ny + m[8] * nz;
      float ry = m[1] * nx + m[5] * ny + m[9] * nz;
      float rz = m[2] * nx + m[6] * ny + m[10] * nz;
      snx += wb[b] * rx;
      sny += wb[b] * ry;
      snz += wb[b] * rz;
    }
    float len = sqrtf(snx * snx + sny * sny + snz * snz);
    if (len > 1e-8f) {
      snx /= len;
      sny /= len;
      snz /= len;
    }
    uint8_t *on = out + a.normal_off;
    st_f(on, snx);
    st_f(on + 4, sny);
    st_f(on + 8, snz);
  }
}

__global__ __launch_bounds__(256) void k_morph(MorphArgs a) {
  uint32_t v = blockIdx.x * 256 + threadIdx.x;
  if (v >= a.vertex_count)
    return;
  const uint8_t *vert = a.bind + (size_t)v * a.stride;
  uint8_t *out = a.out + (size_t)v * a.stride;
  copy_vertex(vert, out, a.stride, a.aligned);
  uint8_t *pp = out + a.pos_off;
  float p0 = ld_f(pp), p1 = ld_f(pp + 4), p2 = ld_f(pp + 8);
  for (uint32_t t = 0; t < a.target_count; t++) {
    float w = a.weights[t];
    if (w == 0.0f)
      continue;
    const float *dl = a.targets + ((size_t)t * a.vertex_count + v) * 3;
    p0 += w * dl[0];
    p1 += w * dl[1];
    p2 += w * dl[2];
  }
  st_f(pp, p0);
  st_f(pp + 4, p1);
  st_f(pp + 8, p2);
}

} // namespace

void launch_skin(const SkinArgs &a, hipStream_t s) {
  if (!a.vertex_count)
    return;
  hipLaunchKernelGGL(k_skin, dim3((a.vertex_count + 255) / 256), dim3(256), 0, s, a);
}

void launch_morph(const MorphArgs &a, hipStream_t s) {
  if (!a.vertex_count)
    return;
  hipLaunchKernelGGL(k_morph, dim3((a.vertex_count + 255) / 256), dim3(256), 0, s, a);
}
