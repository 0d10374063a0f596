/* Host check of the chunk cull (box_rejected_plane in
 * master-of-puppets_amd/csrc/cull.h, the code k_cull runs): random boxes,
 * matrices (perspective views and random ones) and depth biases; whenever a
 * box is culled by plane p, points of the box -- its corners, points on its
 * faces next to the corners, random interior points -- are transformed with
 * the vertex transform's own arithmetic (mat4_mul_vec4, then z += bias * w)
 * and must fail plane p as the raster's trivial reject tests it.  Boxes with
 * a corner inside the view volume must not be culled.
 * argv[1]: cases (default 2,000,000).  Prints the counts; exit 0 when clean. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cull.h"

static uint64_t g_rng = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return g_rng;
}
static float uni(float lo, float hi) {
  return lo + (hi - lo) * (float)((rnd() >> 40) * (1.0 / 16777216.0));
}

/* the raster's trivial-reject test of one vertex against plane p */
static bool fails(int p, float4 v) {
  switch (p) {
  case 0: return v.x < -v.w;
  case 1: return v.x > v.w;
  case 2: return v.y < -v.w;
  case 3: return v.y > v.w;
  case 4: return v.z < -v.w;
  default: return v.z > v.w;
  }
}

static float4 xform(const float *m, float bias, float x, float y, float z) {
  float4 v = mat4_mul_vec4(m, x, y, z, 1.0f);
  if (bias != 0.0f)
    v.z += bias * v.w;
  return v;
}

/* column-major perspective * look-at, as the frame driver builds them */
static void view_proj(float *m) {
  float fov = uni(0.3f, 1.6f), asp = uni(0.5f, 2.5f), n = uni(0.01f, 1.0f), f = n * uni(10.0f, 1e4f);
  float t = 1.0f / tanf(fov * 0.5f);
  float P[16] = {0};
  P[0] = t / asp;
  P[5] = t;
  P[10] = (f + n) / (n - f);
  P[11] = -1.0f;
  P[14] = 2.0f * f * n / (n - f);
  float e[3] = {uni(-50, 50), uni(-50, 50), uni(-50, 50)}, c[3] = {uni(-50, 50), uni(-50, 50), uni(-50, 50)};
  float fw[3] = {c[0] - e[0], c[1] - e[1], c[2] - e[2]};
  float l = sqrtf(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]) + 1e-6f;
  for (int i = 0; i < 3; i++)
    fw[i] /= l;
  float up[3] = {0, 1, 0};
  float s[3] = {fw[1] * up[2] - fw[2] * up[1], fw[2] * up[0] - fw[0] * up[2], fw[0] * up[1] - fw[1] * up[0]};
  l = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + 1e-6f;
  for (int i = 0; i < 3; i++)
    s[i] /= l;
  float u[3] = {s[1] * fw[2] - s[2] * fw[1], s[2] * fw[0] - s[0] * fw[2], s[0] * fw[1] - s[1] * fw[0]};
  float V[16] = {s[0], u[0], -fw[0], 0, s[1], u[1], -fw[1], 0, s[2], u[2], -fw[2], 0,
                 -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]), -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]),
                 fw[0] * e[0] + fw[1] * e[1] + fw[2] * e[2], 1};
  for (int col = 0; col < 4; col++)
    for (int row = 0; row < 4; row++) {
      float acc = 0.0f;
      for (int k = 0; k < 4; k++)
        acc += P[k * 4 + row] * V[col * 4 + k];
      m[col * 4 + row] = acc;
    }
}

int main(int argc, char **argv) {
  long cases = argc > 1 ? atol(argv[1]) : 2000000;
  long culled = 0, points = 0, wrong = 0, visible_culled = 0, visible = 0;
  for (long k = 0; k < cases; k++) {
    float m[16];
    if (rnd() & 1) {
      view_proj(m);
    } else {
      for (int i = 0; i < 16; i++)
        m[i] = uni(-2.0f, 2.0f);
    }
    float bias = (rnd() & 3) == 0 ? uni(-1e-3f, 1e-3f) : 0.0f;
    float cx = uni(-100, 100), cy = uni(-100, 100), cz = uni(-100, 100);
    float ex = uni(0, 1) < 0.2f ? 0.0f : uni(0, 20), ey = uni(0, 20), ez = uni(0, 20);
    float4 lo = make_float4(cx - ex, cy - ey, cz - ez, 0), hi = make_float4(cx + ex, cy + ey, cz + ez, 0);
    int p = box_rejected_plane(m, bias, lo, hi);
    bool corner_inside = false;
    for (int c = 0; c < 8; c++) {
      float4 v = xform(m, bias, (c & 1) ? hi.x : lo.x, (c & 2) ? hi.y : lo.y, (c & 4) ? hi.z : lo.z);
      if (v.x >= -v.w && v.x <= v.w && v.y >= -v.w && v.y <= v.w && v.z >= -v.w && v.z <= v.w)
        corner_inside = true;
    }
    if (corner_inside) {
      visible++;
      if (p >= 0)
        visible_culled++;
    }
    if (p < 0)
      continue;
    culled++;
    for (int j = 0; j < 48; j++) {
      float x, y, z;
      int c = (int)(rnd() & 7);
      float bx = (c & 1) ? hi.x : lo.x, by = (c & 2) ? hi.y : lo.y, bz = (c & 4) ? hi.z : lo.z;
      if (j < 8) { /* the corners */
        x = (j & 1) ? hi.x : lo.x;
        y = (j & 2) ? hi.y : lo.y;
        z = (j & 4) ? hi.z : lo.z;
      } else if (j < 32) { /* on a face, next to a corner */
        float t = uni(0, 1) * uni(0, 1) * uni(0, 1);
        x = bx;
        y = by;
        z = bz;
        switch (rnd() % 3) {
        case 0: x = bx + (lo.x + hi.x - 2 * bx) * t; break;
        case 1: y = by + (lo.y + hi.y - 2 * by) * t; break;
        default: z = bz + (lo.z + hi.z - 2 * bz) * t; break;
        }
      } else { /* inside */
        x = uni(lo.x, hi.x);
        y = uni(lo.y, hi.y);
        z = uni(lo.z, hi.z);
      }
      points++;
      if (!fails(p, xform(m, bias, x, y, z)))
        wrong++;
    }
  }
  printf("cases %ld culled %ld points %ld wrong %ld visible %ld visible_culled %ld\n", cases, culled,
         points, wrong, visible, visible_culled);
  return (wrong == 0 && visible_culled == 0 && culled > cases / 10) ? 0 : 1;
}
