/*
 * mop_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of the reference's
 * software-rasterizer frame path, used only by tests/, smoke() and the
 * benchmark's cpu_baseline leg as a checker.  Never linked into the product.
 *
 * It is an RHI backend (same table layout and hook set as the HIP backend,
 * symbols prefixed mop_oracle_) so the unmodified reference viewport core can
 * drive it through oracle/ref_harness.c exactly like the HIP backend.  It
 * restates the reference arithmetic but runs it in the decomposition the GPU
 * kernels use, which is what it pins:
 *
 *   per-vertex transform        cpu_prepare_triangle   cpu_backend.c:224-324
 *   classify / clip / fan       rasterize_triangle_full rasterizer.c:2458-2553
 *                               mop_sw_clip_polygon    rasterizer.c:510-626
 *   incremental half-space walk smooth_ml              rasterizer.c:1837-1957
 *   visibility                  lexicographic min of (depth, submission
 *                               index) -- equal to the reference's strict
 *                               `z < depth`, first-submitted-wins loop
 *   shading of the winners      rasterizer.c:1959-2108 with 1596-1814,
 *                               194-348 (IBL, PCF)
 *   shadow pass                 mop_sw_shadow_render_mesh rasterizer.c:357-493
 *   HDR resolve                 mop_sw_hdr_resolve     rasterizer.c:2770-2803
 *   object outline              mop_overlay_builtin_outline
 *                                                      overlay_builtin.c:1309-1400
 *   FXAA + post                 mop_sw_fxaa 2625-2761, postprocess.c:24-116
 *   environment background      pass_background_hdri_cpu viewport.c:2128-2180
 *   SSAA box downsample         mop_viewport_read_color viewport.c:3385-3411
 *
 * Checked against the reference build (oracle/_ref) by tests/test_oracle.py:
 * depth, object-ID, HDR and RGBA8 bit-identical on every scene tested.
 *
 * Draw classes: opaque smooth-lit draws (opacity >= 1, depth-tested and
 * written) go through the order-free visibility restatement above; every
 * other draw (blended, flat, unlit, wireframe, depth-test or depth-write off)
 * is rasterized in submission order by draw_ordered, as the reference's
 * single-threaded path does (cpu_draw, cpu_backend.c:588-620).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mop_hip_rhi.h"

typedef MopHipVec3 V3;

/* ---------------------------------------------------------------- math -- */

static V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 vscale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float vlen(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static V3 vnorm(V3 a) {
  float l = vlen(a);
  if (l < 1e-8f)
    return v3(0, 0, 0);
  float inv = 1.0f / l;
  return v3(a.x * inv, a.y * inv, a.z * inv);
}
/* mop_mat4_mul_vec4 row r: ((m[r]x + m[4+r]y) + m[8+r]z) + m[12+r]w */
static void mulv(const float *m, float x, float y, float z, float w, float o[4]) {
  for (int r = 0; r < 4; r++)
    o[r] = m[r] * x + m[4 + r] * y + m[8 + r] * z + m[12 + r] * w;
}
static MopHipMat4 mmul(const MopHipMat4 *a, const MopHipMat4 *b) {
  MopHipMat4 r;
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      float s = 0.0f;
      for (int k = 0; k < 4; k++)
        s += a->d[k * 4 + row] * b->d[c * 4 + k];
      r.d[c * 4 + row] = s;
    }
  return r;
}
static float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

/* ----------------------------------------------------------- resources -- */

struct MopHipBuffer {
  void *data;
  size_t size;
};
struct MopHipTexture {
  int w, h, hdr;
  void *data;
};

typedef struct ClipV {
  float p[4];
  float n[3], wp[3], c[4];
} ClipV;

typedef struct SubTri {
  float sx[3], sy[3], sz[3], iw[3];
  ClipV v[3]; /* attributes of the fan vertices */
  uint32_t draw;
} SubTri;

/* shading models (cpu_backend.c:608-619; rasterizer.c:2540-2610, 1158-1243) */
enum { SM_LIT_SMOOTH = 0, SM_LIT_FLAT = 1, SM_UNLIT_FLAT = 2, SM_UNLIT_SMOOTH = 3 };

typedef struct DrawRec {
  uint32_t object_id;
  int sm;
  V3 light_dir;
  float ambient, metallic, roughness;
  V3 eye;
  MopHipLight *lights;
  uint32_t nlights;
  int state;
} DrawRec;

typedef struct ShadeState {
  int shadow_on;
  const float *smap;
  int sw, sh;
  float svp[16];
  int ibl_on;
  const float *irr, *pf, *lut;
  int irr_w, irr_h, pf_w, pf_h, pf_levels, lut_n;
  float rot, inten;
} ShadeState;

struct MopHipFramebuffer {
  int w, h;
  uint8_t *color;
  uint8_t *ds; /* SSAA-downsampled colour */
  float *hdr, *depth;
  uint32_t *ids;
  uint64_t *vis;
  /* current segment */
  SubTri *tris;
  size_t ntris, cap;
  DrawRec *draws;
  size_t ndraws, dcap;
  int seg_open;
  float clear[4];
};

struct MopHipDevice {
  float exposure;
  ShadeState st;
  ShadeState *states;
  size_t nstates, scap;
  int state_dirty;
  /* shadow */
  float *smap;
  int ssize;
  float *ibl_irr, *ibl_pf, *ibl_lut;
  size_t skipped;
};

static uint32_t dkey(float z) {
  float zz = (z == 0.0f) ? 0.0f : z;
  uint32_t u;
  memcpy(&u, &zz, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

MopHipDevice *mop_oracle_device_create(void) {
  MopHipDevice *d = calloc(1, sizeof(*d));
  d->exposure = 1.0f;
  d->state_dirty = 1;
  return d;
}
void mop_oracle_device_destroy(MopHipDevice *d) {
  if (!d)
    return;
  free(d->smap);
  free(d->ibl_irr);
  free(d->ibl_pf);
  free(d->ibl_lut);
  free(d->states);
  free(d);
}
MopHipBuffer *mop_oracle_buffer_create(MopHipDevice *d, const MopHipBufferDesc *desc) {
  MopHipBuffer *b = calloc(1, sizeof(*b));
  b->data = malloc(desc->size ? desc->size : 1);
  memcpy(b->data, desc->data, desc->size);
  b->size = desc->size;
  return b;
}
void mop_oracle_buffer_destroy(MopHipDevice *d, MopHipBuffer *b) {
  if (b) {
    free(b->data);
    free(b);
  }
}
void mop_oracle_buffer_update(MopHipDevice *d, MopHipBuffer *b, const void *data, size_t off,
                              size_t size) {
  memcpy((uint8_t *)b->data + off, data, size);
}
const void *mop_oracle_buffer_read(MopHipBuffer *b) { return b ? b->data : NULL; }

static const MopHipVertexAttrib *fmt_attr(const MopHipVertexFormat *f, int sem) {
  for (uint32_t i = 0; i < f->attrib_count && i < 12; i++)
    if (f->attribs[i].semantic == sem)
      return &f->attribs[i];
  return NULL;
}

/* mop_skin_apply (viewport.c:1572-1680): linear-blend skinning of the bind
 * pose into the vertex buffer. */
int mop_oracle_skin_apply(MopHipDevice *d, MopHipBuffer *vb, const void *bind, uint32_t vc,
                          const MopHipVertexFormat *fmt, const MopHipMat4 *bones, uint32_t nb) {
  (void)d;
  const MopHipVertexAttrib *pa = fmt_attr(fmt, MOP_HIP_ATTRIB_POSITION);
  const MopHipVertexAttrib *na = fmt_attr(fmt, MOP_HIP_ATTRIB_NORMAL);
  const MopHipVertexAttrib *ja = fmt_attr(fmt, MOP_HIP_ATTRIB_JOINTS);
  const MopHipVertexAttrib *wa = fmt_attr(fmt, MOP_HIP_ATTRIB_WEIGHTS);
  if (!pa || !ja || !wa || (size_t)vc * fmt->stride > vb->size)
    return -1;
  uint8_t *out = vb->data;
  memcpy(out, bind, (size_t)vc * fmt->stride);
  for (uint32_t v = 0; v < vc; v++) {
    const uint8_t *src = (const uint8_t *)bind + (size_t)v * fmt->stride;
    uint8_t *dst = out + (size_t)v * fmt->stride;
    const uint8_t *j = src + ja->offset;
    float w[4], p[3];
    memcpy(w, src + wa->offset, 16);
    memcpy(p, src + pa->offset, 12);
    float acc[3] = {0, 0, 0};
    for (int b = 0; b < 4; b++) {
      if (w[b] <= 0.0f || j[b] >= nb)
        continue;
      const float *m = bones[j[b]].d;
      float r[3] = {m[0] * p[0] + m[4] * p[1] + m[8] * p[2] + m[12],
                    m[1] * p[0] + m[5] * p[1] + m[9] * p[2] + m[13],
                    m[2] * p[0] + m[6] * p[1] + m[10] * p[2] + m[14]};
      for (int k = 0; k < 3; k++)
        acc[k] += w[b] * r[k];
    }
    memcpy(dst + pa->offset, acc, 12);
    if (na) {
      float n[3], sn[3] = {0, 0, 0};
      memcpy(n, src + na->offset, 12);
      for (int b = 0; b < 4; b++) {
        if (w[b] <= 0.0f || j[b] >= nb)
          continue;
        const float *m = bones[j[b]].d;
        float r[3] = {m[0] * n[0] + m[4] * n[1] + m[8] * n[2],
                      m[1] * n[0] + m[5] * n[1] + m[9] * n[2],
                      m[2] * n[0] + m[6] * n[1] + m[10] * n[2]};
        for (int k = 0; k < 3; k++)
          sn[k] += w[b] * r[k];
      }
      float len = sqrtf(sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2]);
      if (len > 1e-8f)
        for (int k = 0; k < 3; k++)
          sn[k] /= len;
      memcpy(dst + na->offset, sn, 12);
    }
  }
  return 0;
}

/* mop_morph_apply (viewport.c:1763-1805) */
int mop_oracle_morph_apply(MopHipDevice *d, MopHipBuffer *vb, const void *bind, uint32_t vc,
                           const MopHipVertexFormat *fmt, const float *targets,
                           const float *weights, uint32_t nt) {
  (void)d;
  size_t stride = fmt ? fmt->stride : sizeof(MopHipVertex), off = 0;
  if (fmt) {
    const MopHipVertexAttrib *pa = fmt_attr(fmt, MOP_HIP_ATTRIB_POSITION);
    if (!pa)
      return -1;
    off = pa->offset;
  }
  if ((size_t)vc * stride > vb->size)
    return -1;
  uint8_t *out = vb->data;
  memcpy(out, bind, (size_t)vc * stride);
  for (uint32_t t = 0; t < nt; t++) {
    float w = weights[t];
    if (w == 0.0f)
      continue;
    const float *dl = targets + (size_t)t * vc * 3;
    for (uint32_t v = 0; v < vc; v++) {
      float p[3];
      memcpy(p, out + v * stride + off, 12);
      p[0] += w * dl[v * 3 + 0];
      p[1] += w * dl[v * 3 + 1];
      p[2] += w * dl[v * 3 + 2];
      memcpy(out + v * stride + off, p, 12);
    }
  }
  return 0;
}

static void fb_alloc(MopHipFramebuffer *fb, int w, int h) {
  size_t n = (size_t)w * h;
  fb->w = w;
  fb->h = h;
  fb->color = calloc(n * 4, 1);
  fb->hdr = calloc(n * 4, sizeof(float));
  fb->depth = malloc(n * sizeof(float));
  fb->ids = calloc(n, sizeof(uint32_t));
  fb->vis = malloc(n * sizeof(uint64_t));
}
static void fb_free(MopHipFramebuffer *fb) {
  free(fb->color);
  free(fb->hdr);
  free(fb->depth);
  free(fb->ids);
  free(fb->vis);
  free(fb->tris);
  free(fb->draws);
  free(fb->ds);
  fb->ds = NULL;
  fb->tris = NULL;
  fb->draws = NULL;
  fb->ntris = fb->cap = fb->ndraws = fb->dcap = 0;
}
MopHipFramebuffer *mop_oracle_framebuffer_create(MopHipDevice *d,
                                                 const MopHipFramebufferDesc *desc) {
  MopHipFramebuffer *fb = calloc(1, sizeof(*fb));
  fb_alloc(fb, desc->width, desc->height);
  return fb;
}
void mop_oracle_framebuffer_destroy(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (fb) {
    fb_free(fb);
    free(fb);
  }
}
void mop_oracle_framebuffer_resize(MopHipDevice *d, MopHipFramebuffer *fb, int w, int h) {
  fb_free(fb);
  fb_alloc(fb, w, h);
}

MopHipTexture *mop_oracle_texture_create(MopHipDevice *d, int w, int h, const uint8_t *rgba) {
  MopHipTexture *t = calloc(1, sizeof(*t));
  t->w = w;
  t->h = h;
  t->data = malloc((size_t)w * h * 4);
  memcpy(t->data, rgba, (size_t)w * h * 4);
  return t;
}
MopHipTexture *mop_oracle_texture_create_hdr(MopHipDevice *d, int w, int h, const float *rgba) {
  MopHipTexture *t = calloc(1, sizeof(*t));
  t->w = w;
  t->h = h;
  t->hdr = 1;
  t->data = malloc((size_t)w * h * 16);
  memcpy(t->data, rgba, (size_t)w * h * 16);
  return t;
}
void mop_oracle_texture_destroy(MopHipDevice *d, MopHipTexture *t) {
  if (t) {
    free(t->data);
    free(t);
  }
}
void mop_oracle_set_exposure(MopHipDevice *d, float e) { d->exposure = e; }
float mop_oracle_frame_gpu_time_ms(MopHipDevice *d) { return 0.0f; }

/* ------------------------------------------------------- frame / draws -- */

static void shade_segment(MopHipDevice *d, MopHipFramebuffer *fb);

void mop_oracle_frame_begin(MopHipDevice *d, MopHipFramebuffer *fb, MopHipColor c) {
  size_t n = (size_t)fb->w * fb->h;
  uint8_t r = (uint8_t)(c.r * 255.0f), g = (uint8_t)(c.g * 255.0f), b = (uint8_t)(c.b * 255.0f),
          a = (uint8_t)(c.a * 255.0f);
  for (size_t i = 0; i < n; i++) {
    fb->color[i * 4] = r;
    fb->color[i * 4 + 1] = g;
    fb->color[i * 4 + 2] = b;
    fb->color[i * 4 + 3] = a;
    fb->hdr[i * 4] = c.r;
    fb->hdr[i * 4 + 1] = c.g;
    fb->hdr[i * 4 + 2] = c.b;
    fb->hdr[i * 4 + 3] = c.a;
    fb->depth[i] = 1.0f;
    fb->ids[i] = 0;
  }
  fb->seg_open = 0;
  fb->ntris = 0;
  fb->ndraws = 0;
  d->nstates = 0;
  d->state_dirty = 1;
}

static void open_segment(MopHipFramebuffer *fb) {
  if (fb->seg_open)
    return;
  size_t n = (size_t)fb->w * fb->h;
  for (size_t i = 0; i < n; i++) {
    float z = fb->depth[i];
    uint32_t k = (z == z) ? dkey(z) : 0u;
    fb->vis[i] = (uint64_t)k << 32;
  }
  fb->ntris = 0;
  fb->seg_open = 1;
}

static void push_tri(MopHipFramebuffer *fb, const SubTri *t) {
  if (fb->ntris == fb->cap) {
    fb->cap = fb->cap ? fb->cap * 2 : 4096;
    fb->tris = realloc(fb->tris, fb->cap * sizeof(SubTri));
  }
  fb->tris[fb->ntris++] = *t;
}

static const MopHipVertexAttrib *find_attr(const MopHipVertexFormat *f, int sem) {
  for (uint32_t i = 0; i < f->attrib_count && i < 12; i++)
    if (f->attribs[i].semantic == sem)
      return &f->attribs[i];
  return NULL;
}
static void rd(const uint8_t *b, const MopHipVertexAttrib *a, int n, float *o, float dflt_w) {
  const float *f = (const float *)(b + a->offset);
  int have = a->format == MOP_HIP_FORMAT_FLOAT ? 1 : a->format == MOP_HIP_FORMAT_FLOAT2 ? 2 :
             a->format == MOP_HIP_FORMAT_FLOAT3 ? 3 : a->format == MOP_HIP_FORMAT_FLOAT4 ? 4 : 0;
  for (int k = 0; k < n; k++)
    o[k] = k < have ? f[k] : 0.0f;
  if (n == 4 && have < 4)
    o[3] = dflt_w;
}

/* per-vertex transform (cpu_prepare_triangle / _ex) */
static void xform_vertex(const MopHipDrawCall *c, const MopHipMat4 *model, const MopHipMat4 *mvp,
                         const uint8_t *vb, ClipV *o) {
  float p[3], n[3] = {0, 0, 0}, col[4], uv[2];
  int has_n = 1;
  if (!c->vertex_format) {
    const float *f = (const float *)vb;
    memcpy(p, f, 12);
    memcpy(n, f + 3, 12);
    memcpy(col, f + 6, 16);
    memcpy(uv, f + 10, 8);
  } else {
    const MopHipVertexFormat *fm = c->vertex_format;
    rd(vb, find_attr(fm, MOP_HIP_ATTRIB_POSITION), 3, p, 0);
    const MopHipVertexAttrib *na = find_attr(fm, MOP_HIP_ATTRIB_NORMAL);
    const MopHipVertexAttrib *ca = find_attr(fm, MOP_HIP_ATTRIB_COLOR);
    const MopHipVertexAttrib *ua = find_attr(fm, MOP_HIP_ATTRIB_TEXCOORD0);
    if (na) rd(vb, na, 3, n, 0); else has_n = 0;
    if (ca) rd(vb, ca, 4, col, 1.0f); else col[0] = col[1] = col[2] = col[3] = 1.0f;
    if (ua) rd(vb, ua, 2, uv, 0); else uv[0] = uv[1] = 0.0f;
  }
  mulv(mvp->d, p[0], p[1], p[2], 1.0f, o->p);
  if (c->depth_bias != 0.0f)
    o->p[2] += c->depth_bias * o->p[3];
  float w4[4];
  mulv(model->d, p[0], p[1], p[2], 1.0f, w4);
  o->wp[0] = w4[0];
  o->wp[1] = w4[1];
  o->wp[2] = w4[2];
  if (has_n) {
    mulv(model->d, n[0], n[1], n[2], 0.0f, w4);
    o->n[0] = w4[0];
    o->n[1] = w4[1];
    o->n[2] = w4[2];
  } else {
    o->n[0] = 0.0f;
    o->n[1] = 1.0f;
    o->n[2] = 0.0f;
  }
  memcpy(o->c, col, 16);
  const MopHipTexture *t = c->texture;
  if (t && !t->hdr && t->w >= 1 && t->h >= 1) {
    float tu = uv[0] - floorf(uv[0]), tv = uv[1] - floorf(uv[1]);
    int tx = (int)(tu * (float)(t->w - 1) + 0.5f), ty = (int)(tv * (float)(t->h - 1) + 0.5f);
    if (tx < 0) tx = 0;
    if (tx >= t->w) tx = t->w - 1;
    if (ty < 0) ty = 0;
    if (ty >= t->h) ty = t->h - 1;
    const uint8_t *px = (const uint8_t *)t->data + ((size_t)ty * t->w + tx) * 4;
    for (int k = 0; k < 4; k++)
      o->c[k] *= (float)px[k] / 255.0f;
  }
}

/* Sutherland-Hodgman, rasterizer.c:510-626 */
static const float PL[6][4] = {{1, 0, 0, 1}, {-1, 0, 0, 1}, {0, 1, 0, 1},
                               {0, -1, 0, 1}, {0, 0, 1, 1}, {0, 0, -1, 1}};
static float pdist(int pl, const float *p) {
  return PL[pl][0] * p[0] + PL[pl][1] * p[1] + PL[pl][2] * p[2] + PL[pl][3] * p[3];
}
static ClipV lerpv(const ClipV *a, const ClipV *b, float t) {
  ClipV r;
  for (int k = 0; k < 4; k++) r.p[k] = a->p[k] + t * (b->p[k] - a->p[k]);
  for (int k = 0; k < 3; k++) r.n[k] = a->n[k] + t * (b->n[k] - a->n[k]);
  for (int k = 0; k < 3; k++) r.wp[k] = a->wp[k] + t * (b->wp[k] - a->wp[k]);
  for (int k = 0; k < 4; k++) r.c[k] = a->c[k] + t * (b->c[k] - a->c[k]);
  return r;
}
static int clip_poly(ClipV *a, int n, ClipV *b) {
  ClipV *src = a, *dst = b;
  for (int pl = 0; pl < 6; pl++) {
    int oc = 0;
    if (n < 1)
      return 0;
    ClipV prev = src[n - 1];
    float pd = pdist(pl, prev.p);
    for (int i = 0; i < n; i++) {
      ClipV cur = src[i];
      float cd = pdist(pl, cur.p);
      if (cd >= 0.0f) {
        if (pd < 0.0f && oc < 24)
          dst[oc++] = lerpv(&prev, &cur, pd / (pd - cd));
        if (oc < 24)
          dst[oc++] = cur;
      } else if (pd >= 0.0f && oc < 24) {
        dst[oc++] = lerpv(&prev, &cur, pd / (pd - cd));
      }
      prev = cur;
      pd = cd;
    }
    n = oc;
    if (n == 0)
      return 0;
    ClipV *t = src;
    src = dst;
    dst = t;
  }
  if (src != a)
    memcpy(a, src, (size_t)n * sizeof(ClipV));
  return n;
}

typedef struct Setup {
  int x0, y0, x1, y1;
  float inv_area, edx[3], edy[3], wr[3], inv_e[3];
} Setup;

/* bbox, area, flipped edges, row-start edge values (rasterizer.c:1842-1924) */
static int setup_tri(const float *sx, const float *sy, int W, int H, Setup *s, int want_ie) {
  float fx0 = sx[0], fy0 = sy[0], fx1 = sx[0], fy1 = sy[0];
  if (sx[1] < fx0) fx0 = sx[1];
  if (sx[2] < fx0) fx0 = sx[2];
  if (sy[1] < fy0) fy0 = sy[1];
  if (sy[2] < fy0) fy0 = sy[2];
  if (sx[1] > fx1) fx1 = sx[1];
  if (sx[2] > fx1) fx1 = sx[2];
  if (sy[1] > fy1) fy1 = sy[1];
  if (sy[2] > fy1) fy1 = sy[2];
  int x0 = (int)floorf(fx0), y0 = (int)floorf(fy0), x1 = (int)ceilf(fx1), y1 = (int)ceilf(fy1);
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 >= W) x1 = W - 1;
  if (y1 >= H) y1 = H - 1;
  if (x0 > x1 || y0 > y1)
    return 0;
  float area = (sx[1] - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (sy[1] - sy[0]);
  if (fabsf(area) < 1e-6f)
    return 0;
  int flip = area < 0.0f;
  s->inv_area = 1.0f / fabsf(area);
  float edx[3] = {sy[1] - sy[2], sy[2] - sy[0], sy[0] - sy[1]};
  float edy[3] = {sx[2] - sx[1], sx[0] - sx[2], sx[1] - sx[0]};
  float px = (float)x0 + 0.5f, py = (float)y0 + 0.5f;
  float wr[3] = {(sx[2] - sx[1]) * (py - sy[1]) - (sy[2] - sy[1]) * (px - sx[1]),
                 (sx[0] - sx[2]) * (py - sy[2]) - (sy[0] - sy[2]) * (px - sx[2]),
                 (sx[1] - sx[0]) * (py - sy[0]) - (sy[1] - sy[0]) * (px - sx[0])};
  for (int k = 0; k < 3; k++) {
    if (flip) {
      edx[k] = -edx[k];
      edy[k] = -edy[k];
      wr[k] = -wr[k];
    }
    s->edx[k] = edx[k];
    s->edy[k] = edy[k];
    s->wr[k] = wr[k];
    if (want_ie) {
      float l = sqrtf(edx[k] * edx[k] + edy[k] * edy[k]);
      s->inv_e[k] = l > 1e-6f ? 1.0f / l : 0.0f;
    }
  }
  s->x0 = x0;
  s->y0 = y0;
  s->x1 = x1;
  s->y1 = y1;
  return 1;
}

/* flat path depth increments (rasterizer.c:973-976) */
static void flat_z(const Setup *s, const SubTri *t, float *zr, float *zdx, float *zdy) {
  *zdx = (s->edx[0] * t->sz[0] + s->edx[1] * t->sz[1] + s->edx[2] * t->sz[2]) * s->inv_area;
  *zdy = (s->edy[0] * t->sz[0] + s->edy[1] * t->sz[1] + s->edy[2] * t->sz[2]) * s->inv_area;
  *zr = (s->wr[0] * t->sz[0] + s->wr[1] * t->sz[1] + s->wr[2] * t->sz[2]) * s->inv_area;
}

/* rasterize_filled_triangle's opaque loop (rasterizer.c:1000-1026) as a
 * visibility update: edge test only, incremental z. */
static void walk_vis_flat(MopHipFramebuffer *fb, const SubTri *t, uint64_t seq) {
  Setup s;
  if (!setup_tri(t->sx, t->sy, fb->w, fb->h, &s, 0))
    return;
  float zr, zdx, zdy;
  flat_z(&s, t, &zr, &zdx, &zdy);
  float r0 = s.wr[0], r1 = s.wr[1], r2 = s.wr[2];
  for (int y = s.y0; y <= s.y1; y++) {
    float w0 = r0, w1 = r1, w2 = r2, z = zr;
    for (int x = s.x0; x <= s.x1; x++) {
      if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f && z == z) {
        uint64_t key = ((uint64_t)dkey(z) << 32) | seq;
        size_t i = (size_t)y * fb->w + x;
        if (key < fb->vis[i])
          fb->vis[i] = key;
      }
      w0 += s.edx[0];
      w1 += s.edx[1];
      w2 += s.edx[2];
      z += zdx;
    }
    r0 += s.edy[0];
    r1 += s.edy[1];
    r2 += s.edy[2];
    zr += zdy;
  }
}

static void walk_vis(MopHipFramebuffer *fb, const SubTri *t, uint64_t seq) {
  Setup s;
  if (!setup_tri(t->sx, t->sy, fb->w, fb->h, &s, 1))
    return;
  float r0 = s.wr[0], r1 = s.wr[1], r2 = s.wr[2];
  for (int y = s.y0; y <= s.y1; y++) {
    float w0 = r0, w1 = r1, w2 = r2;
    for (int x = s.x0; x <= s.x1; x++) {
      if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) {
        float d0 = w0 * s.inv_e[0], d1 = w1 * s.inv_e[1], d2 = w2 * s.inv_e[2];
        float md = d0;
        if (d1 < md) md = d1;
        if (d2 < md) md = d2;
        if (md >= 0.0f) {
          float b0 = w0 * s.inv_area, b1 = w1 * s.inv_area, b2 = w2 * s.inv_area;
          float z = b0 * t->sz[0] + b1 * t->sz[1] + b2 * t->sz[2];
          if (z == z) {
            uint64_t key = ((uint64_t)dkey(z) << 32) | seq;
            size_t i = (size_t)y * fb->w + x;
            if (key < fb->vis[i])
              fb->vis[i] = key;
          }
        }
      }
      w0 += s.edx[0];
      w1 += s.edx[1];
      w2 += s.edx[2];
    }
    r0 += s.edy[0];
    r1 += s.edy[1];
    r2 += s.edy[2];
  }
}

static int fan_tri(const ClipV *a, const ClipV *b, const ClipV *c, int cull, int wire, int W,
                   int H, SubTri *o) {
  if (fabsf(a->p[3]) < 1e-7f || fabsf(b->p[3]) < 1e-7f || fabsf(c->p[3]) < 1e-7f)
    return 0;
  const ClipV *v[3] = {a, b, c};
  float hw = (float)W * 0.5f, hh = (float)H * 0.5f;
  for (int k = 0; k < 3; k++) {
    float iw = 1.0f / v[k]->p[3];
    o->iw[k] = iw;
    o->sx[k] = (v[k]->p[0] * iw + 1.0f) * hw;
    o->sy[k] = (1.0f - v[k]->p[1] * iw) * hh;
    o->sz[k] = (v[k]->p[2] * iw + 1.0f) * 0.5f;
    o->v[k] = *v[k];
  }
  float sa = (o->sx[1] - o->sx[0]) * (o->sy[2] - o->sy[0]) -
             (o->sx[2] - o->sx[0]) * (o->sy[1] - o->sy[0]);
  if (cull && sa >= 0.0f)
    return 0;
  if (wire)
    return 1;
  Setup s;
  return setup_tri(o->sx, o->sy, W, H, &s, 0);
}

static int cur_state(MopHipDevice *d) {
  if (d->state_dirty || d->nstates == 0) {
    if (d->nstates == d->scap) {
      d->scap = d->scap ? d->scap * 2 : 8;
      d->states = realloc(d->states, d->scap * sizeof(ShadeState));
    }
    d->states[d->nstates++] = d->st;
    d->state_dirty = 0;
  }
  return (int)d->nstates - 1;
}

static void draw_ordered(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c,
                         const MopHipMat4 *model, const MopHipMat4 *mvp, int sm);
static void flat_color(const ShadeState *S, const DrawRec *dr, const SubTri *t, float out[3]);
static void unlit_smooth_px(const DrawRec *dr, const SubTri *t, float w0, float w1, float w2,
                            float inv_area, float out[3]);
static void shade_px(const MopHipDevice *dev, const ShadeState *S, const DrawRec *dr,
                     const SubTri *t, float w0, float w1, float w2, float inv_area,
                     float out[3]);

/* aa_plot (rasterizer.c:688-730) */
static void aa_plot(MopHipFramebuffer *fb, int x, int y, float z, const uint8_t *c8, float cov,
                    uint32_t id, int depth_test) {
  if (x < 0 || x >= fb->w || y < 0 || y >= fb->h)
    return;
  if (cov < 0.004f)
    return;
  size_t i = (size_t)y * fb->w + x;
  if (depth_test && z >= fb->depth[i])
    return;
  float *hp = fb->hdr + i * 4;
  float inv = 1.0f - cov;
  for (int q = 0; q < 3; q++)
    hp[q] = ((float)c8[q] / 255.0f) * cov + hp[q] * inv;
  hp[3] = 1.0f;
  if (cov > 0.5f) {
    fb->depth[i] = z;
    fb->ids[i] = id;
  }
}

static float fpart(float x) { return x - floorf(x); }
static float rfpart(float x) { return 1.0f - fpart(x); }

/* Xiaolin Wu line, thickness 1 (mop_sw_draw_line_aa, rasterizer.c:779-871) */
static void line_aa(MopHipFramebuffer *fb, float x0, float y0, float z0, float x1, float y1,
                    float z1, const uint8_t *c8, uint32_t id, int dt) {
  int steep = fabsf(y1 - y0) > fabsf(x1 - x0);
  float t;
  if (steep) {
    t = x0; x0 = y0; y0 = t;
    t = x1; x1 = y1; y1 = t;
  }
  if (x0 > x1) {
    t = x0; x0 = x1; x1 = t;
    t = y0; y0 = y1; y1 = t;
    t = z0; z0 = z1; z1 = t;
  }
  float dx = x1 - x0, dy = y1 - y0;
  float g = (dx < 0.001f) ? 1.0f : dy / dx;
  float xe = roundf(x0), ye = y0 + g * (xe - x0), xg = rfpart(x0 + 0.5f);
  int xp1 = (int)xe, yp1 = (int)floorf(ye);
  if (steep) {
    aa_plot(fb, yp1, xp1, z0, c8, rfpart(ye) * xg, id, dt);
    aa_plot(fb, yp1 + 1, xp1, z0, c8, fpart(ye) * xg, id, dt);
  } else {
    aa_plot(fb, xp1, yp1, z0, c8, rfpart(ye) * xg, id, dt);
    aa_plot(fb, xp1, yp1 + 1, z0, c8, fpart(ye) * xg, id, dt);
  }
  float iy_f = ye + g;
  xe = roundf(x1);
  ye = y1 + g * (xe - x1);
  xg = fpart(x1 + 0.5f);
  int xp2 = (int)xe, yp2 = (int)floorf(ye);
  if (steep) {
    aa_plot(fb, yp2, xp2, z1, c8, rfpart(ye) * xg, id, dt);
    aa_plot(fb, yp2 + 1, xp2, z1, c8, fpart(ye) * xg, id, dt);
  } else {
    aa_plot(fb, xp2, yp2, z1, c8, rfpart(ye) * xg, id, dt);
    aa_plot(fb, xp2, yp2 + 1, z1, c8, fpart(ye) * xg, id, dt);
  }
  float steps = (float)(xp2 - xp1);
  if (steps < 1.0f)
    steps = 1.0f;
  for (int x = xp1 + 1; x < xp2; x++) {
    float tt = (float)(x - xp1) / steps;
    float z = z0 + tt * (z1 - z0);
    int iy = (int)floorf(iy_f);
    if (steep) {
      aa_plot(fb, iy, x, z, c8, rfpart(iy_f), id, dt);
      aa_plot(fb, iy + 1, x, z, c8, fpart(iy_f), id, dt);
    } else {
      aa_plot(fb, x, iy, z, c8, rfpart(iy_f), id, dt);
      aa_plot(fb, x, iy + 1, z, c8, fpart(iy_f), id, dt);
    }
    iy_f += g;
  }
}

static void draw_one(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c,
                     const MopHipMat4 *model, const MopHipMat4 *mvp) {
  int lit = c->lights && c->light_count > 0;
  int smooth = c->shading_mode == MOP_HIP_SHADING_SMOOTH;
  int sm = lit ? (smooth ? SM_LIT_SMOOTH : SM_LIT_FLAT) : (smooth ? SM_UNLIT_SMOOTH : SM_UNLIT_FLAT);
  int order_free = !c->wireframe && sm != SM_UNLIT_SMOOTH && c->depth_test && c->depth_write &&
                   c->blend_mode == MOP_HIP_BLEND_OPAQUE && c->opacity >= 1.0f;
  if (!order_free) {
    /* lines ignore blend mode and opacity */
    int invisible = !c->wireframe && (c->blend_mode == MOP_HIP_BLEND_ALPHA ||
                                      c->blend_mode == MOP_HIP_BLEND_ADDITIVE) &&
                    !(c->opacity > 0.0f);
    if (invisible)
      return;
    draw_ordered(d, fb, c, model, mvp, sm);
    return;
  }
  open_segment(fb);
  if (fb->ndraws == fb->dcap) {
    fb->dcap = fb->dcap ? fb->dcap * 2 : 64;
    fb->draws = realloc(fb->draws, fb->dcap * sizeof(DrawRec));
  }
  DrawRec *dr = &fb->draws[fb->ndraws];
  dr->object_id = c->object_id;
  dr->sm = sm;
  dr->light_dir = c->light_dir;
  dr->ambient = c->ambient;
  dr->metallic = c->metallic;
  dr->roughness = c->roughness;
  dr->eye = c->cam_eye;
  dr->nlights = c->light_count;
  dr->lights = malloc(c->light_count * sizeof(MopHipLight));
  memcpy(dr->lights, c->lights, c->light_count * sizeof(MopHipLight));
  dr->state = cur_state(d);
  uint32_t di = (uint32_t)fb->ndraws++;

  uint32_t stride = c->vertex_format ? c->vertex_format->stride : sizeof(MopHipVertex);
  uint32_t nv = c->vertex_count;
  ClipV *vc = malloc((size_t)(nv ? nv : 1) * sizeof(ClipV));
  for (uint32_t i = 0; i < nv; i++)
    xform_vertex(c, model, mvp, (const uint8_t *)c->vertex_buffer->data + (size_t)i * stride,
                 &vc[i]);
  const uint32_t *ib = (const uint32_t *)c->index_buffer->data;
  int cull = c->backface_cull;
  for (uint32_t i = 0; i + 2 < c->index_count; i += 3) {
    uint32_t i0 = ib[i], i1 = ib[i + 1], i2 = ib[i + 2];
    if (i0 >= nv || i1 >= nv || i2 >= nv)
      continue;
    const float *a = vc[i0].p, *b = vc[i1].p, *cc = vc[i2].p;
    if ((a[0] < -a[3] && b[0] < -b[3] && cc[0] < -cc[3]) ||
        (a[0] > a[3] && b[0] > b[3] && cc[0] > cc[3]) ||
        (a[1] < -a[3] && b[1] < -b[3] && cc[1] < -cc[3]) ||
        (a[1] > a[3] && b[1] > b[3] && cc[1] > cc[3]) ||
        (a[2] < -a[3] && b[2] < -b[3] && cc[2] < -cc[3]) ||
        (a[2] > a[3] && b[2] > b[3] && cc[2] > cc[3]))
      continue;
    if (cull && a[3] > 0.0f && b[3] > 0.0f && cc[3] > 0.0f) {
      float ex = b[0] * a[3] - a[0] * b[3], ey = b[1] * a[3] - a[1] * b[3];
      float fx = cc[0] * a[3] - a[0] * cc[3], fy = cc[1] * a[3] - a[1] * cc[3];
      if (ex * fy - ey * fx <= 0.0f)
        continue;
    }
    ClipV poly[24], tmp[24];
    int n;
    int inside = a[3] > 0.0f && b[3] > 0.0f && cc[3] > 0.0f;
    for (int k = 0; k < 3 && inside; k++) {
      const float *q = k == 0 ? a : k == 1 ? b : cc;
      inside = q[0] >= -q[3] && q[0] <= q[3] && q[1] >= -q[3] && q[1] <= q[3] &&
               q[2] >= -q[3] && q[2] <= q[3];
    }
    poly[0] = vc[i0];
    poly[1] = vc[i1];
    poly[2] = vc[i2];
    n = inside ? 3 : clip_poly(poly, 3, tmp);
    if (n < 3)
      continue;
    for (int k = 1; k < n - 1; k++) {
      SubTri st;
      if (!fan_tri(&poly[0], &poly[k], &poly[k + 1], cull, c->wireframe, fb->w, fb->h, &st))
        continue;
      st.draw = di;
      push_tri(fb, &st);
      if (sm == SM_LIT_FLAT || sm == SM_UNLIT_FLAT)
        walk_vis_flat(fb, &fb->tris[fb->ntris - 1], (uint64_t)fb->ntris);
      else
        walk_vis(fb, &fb->tris[fb->ntris - 1], (uint64_t)fb->ntris);
    }
  }
  free(vc);
}

/* Draws whose result depends on order, restated as the reference's
 * sequential loop: per triangle, per pixel, depth test against the current
 * buffer, blend (rasterizer.c:1926-2140); depth_write = false saves and
 * restores the depth buffer around the draw (cpu_backend.c:539-545,622-628). */
static void draw_ordered(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c,
                         const MopHipMat4 *model, const MopHipMat4 *mvp, int sm) {
  shade_segment(d, fb);
  size_t npx = (size_t)fb->w * fb->h;
  float *saved = NULL;
  if (c->depth_test && !c->depth_write) {
    saved = malloc(npx * sizeof(float));
    memcpy(saved, fb->depth, npx * sizeof(float));
  }
  DrawRec dr;
  dr.object_id = c->object_id;
  dr.sm = sm;
  dr.light_dir = c->light_dir;
  dr.ambient = c->ambient;
  dr.metallic = c->metallic;
  dr.roughness = c->roughness;
  dr.eye = c->cam_eye;
  dr.nlights = c->light_count;
  dr.lights = (MopHipLight *)c->lights;
  dr.state = cur_state(d);
  float op = clamp01(c->opacity);
  uint32_t stride = c->vertex_format ? c->vertex_format->stride : sizeof(MopHipVertex);
  uint32_t nv = c->vertex_count;
  ClipV *vc = malloc((size_t)(nv ? nv : 1) * sizeof(ClipV));
  for (uint32_t i = 0; i < nv; i++)
    xform_vertex(c, model, mvp, (const uint8_t *)c->vertex_buffer->data + (size_t)i * stride,
                 &vc[i]);
  const uint32_t *ib = (const uint32_t *)c->index_buffer->data;
  int cull = c->backface_cull;
  for (uint32_t i = 0; i + 2 < c->index_count; i += 3) {
    uint32_t i0 = ib[i], i1 = ib[i + 1], i2 = ib[i + 2];
    if (i0 >= nv || i1 >= nv || i2 >= nv)
      continue;
    const float *a = vc[i0].p, *b = vc[i1].p, *cc = vc[i2].p;
    if ((a[0] < -a[3] && b[0] < -b[3] && cc[0] < -cc[3]) ||
        (a[0] > a[3] && b[0] > b[3] && cc[0] > cc[3]) ||
        (a[1] < -a[3] && b[1] < -b[3] && cc[1] < -cc[3]) ||
        (a[1] > a[3] && b[1] > b[3] && cc[1] > cc[3]) ||
        (a[2] < -a[3] && b[2] < -b[3] && cc[2] < -cc[3]) ||
        (a[2] > a[3] && b[2] > b[3] && cc[2] > cc[3]))
      continue;
    if (cull && a[3] > 0.0f && b[3] > 0.0f && cc[3] > 0.0f) {
      float ex = b[0] * a[3] - a[0] * b[3], ey = b[1] * a[3] - a[1] * b[3];
      float fx = cc[0] * a[3] - a[0] * cc[3], fy = cc[1] * a[3] - a[1] * cc[3];
      if (ex * fy - ey * fx <= 0.0f)
        continue;
    }
    ClipV poly[24], tmp[24];
    int inside = a[3] > 0.0f && b[3] > 0.0f && cc[3] > 0.0f;
    for (int k = 0; k < 3 && inside; k++) {
      const float *q = k == 0 ? a : k == 1 ? b : cc;
      inside = q[0] >= -q[3] && q[0] <= q[3] && q[1] >= -q[3] && q[1] <= q[3] &&
               q[2] >= -q[3] && q[2] <= q[3];
    }
    poly[0] = vc[i0];
    poly[1] = vc[i1];
    poly[2] = vc[i2];
    int n = inside ? 3 : clip_poly(poly, 3, tmp);
    if (n < 3)
      continue;
    for (int k = 1; k < n - 1; k++) {
      SubTri st;
      if (!fan_tri(&poly[0], &poly[k], &poly[k + 1], cull, c->wireframe, fb->w, fb->h, &st))
        continue;
      if (c->wireframe) {
        /* flat colour truncated to 8 bits, three Wu lines
         * (rasterizer.c:2590-2603, 1222-1233) */
        float col[3];
        flat_color(&d->states[dr.state], &dr, &st, col);
        uint8_t c8[3];
        for (int q = 0; q < 3; q++)
          c8[q] = (uint8_t)((col[q] > 1.0f ? 1.0f : col[q]) * 255.0f);
        for (int e = 0; e < 3; e++) {
          int a0 = e, a1 = (e + 1) % 3;
          line_aa(fb, st.sx[a0], st.sy[a0], st.sz[a0], st.sx[a1], st.sy[a1], st.sz[a1], c8,
                  c->object_id, c->depth_test);
        }
        continue;
      }
      Setup sp;
      if (!setup_tri(st.sx, st.sy, fb->w, fb->h, &sp, 1))
        continue;
      if (sm == SM_LIT_FLAT || sm == SM_UNLIT_FLAT) {
        /* rasterize_filled_triangle (rasterizer.c:890-1087) */
        float col[3];
        flat_color(&d->states[dr.state], &dr, &st, col);
        float ca = clamp01(c->opacity);
        if (sm == SM_UNLIT_FLAT)
          ca = (float)(uint8_t)(ca * 255.0f) / 255.0f;
        int opaque = c->blend_mode == MOP_HIP_BLEND_OPAQUE && ca >= 1.0f;
        float zr, zdx, zdy;
        flat_z(&sp, &st, &zr, &zdx, &zdy);
        float r0 = sp.wr[0], r1 = sp.wr[1], r2 = sp.wr[2];
        for (int y = sp.y0; y <= sp.y1; y++) {
          float w0 = r0, w1 = r1, w2 = r2, z = zr;
          for (int x = sp.x0; x <= sp.x1; x++) {
            if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) {
              size_t idx = (size_t)y * fb->w + x;
              if (!c->depth_test || z < fb->depth[idx]) {
                float *hp = fb->hdr + idx * 4;
                if (opaque) {
                  hp[0] = col[0];
                  hp[1] = col[1];
                  hp[2] = col[2];
                  hp[3] = 1.0f;
                  fb->depth[idx] = z;
                  fb->ids[idx] = c->object_id;
                } else {
                  for (int q = 0; q < 3; q++) {
                    if (c->blend_mode == MOP_HIP_BLEND_ADDITIVE)
                      hp[q] = hp[q] + col[q] * ca;
                    else if (c->blend_mode == MOP_HIP_BLEND_MULTIPLY)
                      hp[q] = hp[q] * col[q];
                    else
                      hp[q] = col[q] * ca + hp[q] * (1.0f - ca);
                  }
                  hp[3] = 1.0f;
                }
              }
            }
            w0 += sp.edx[0];
            w1 += sp.edx[1];
            w2 += sp.edx[2];
            z += zdx;
          }
          r0 += sp.edy[0];
          r1 += sp.edy[1];
          r2 += sp.edy[2];
          zr += zdy;
        }
        continue;
      }
      float r0 = sp.wr[0], r1 = sp.wr[1], r2 = sp.wr[2];
      for (int y = sp.y0; y <= sp.y1; y++) {
        float w0 = r0, w1 = r1, w2 = r2;
        for (int x = sp.x0; x <= sp.x1; x++) {
          if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) {
            float d0 = w0 * sp.inv_e[0], d1 = w1 * sp.inv_e[1], d2 = w2 * sp.inv_e[2];
            float md = d0;
            if (d1 < md) md = d1;
            if (d2 < md) md = d2;
            if (md >= 0.0f) {
              float z = w0 * sp.inv_area * st.sz[0] + w1 * sp.inv_area * st.sz[1] +
                        w2 * sp.inv_area * st.sz[2];
              size_t idx = (size_t)y * fb->w + x;
              if (!c->depth_test || z < fb->depth[idx]) {
                float rgb[3], cov;
                if (sm == SM_UNLIT_SMOOTH) {
                  /* silhouette AA (rasterizer.c:1421-1426) */
                  unlit_smooth_px(&dr, &st, w0, w1, w2, sp.inv_area, rgb);
                  int sil = (fb->depth[idx] - z) > 0.005f;
                  cov = (md < 1.0f && sil) ? md : 1.0f;
                } else {
                  shade_px(d, &d->states[dr.state], &dr, &st, w0, w1, w2, sp.inv_area, rgb);
                  cov = (md < 1.0f && c->blend_mode != MOP_HIP_BLEND_OPAQUE) ? md : 1.0f;
                }
                float fa = op * cov;
                float *hp = fb->hdr + idx * 4;
                if (c->blend_mode == MOP_HIP_BLEND_OPAQUE && fa >= 1.0f) {
                  hp[0] = rgb[0];
                  hp[1] = rgb[1];
                  hp[2] = rgb[2];
                  hp[3] = 1.0f;
                  fb->depth[idx] = z;
                  fb->ids[idx] = c->object_id;
                } else {
                  float inv = 1.0f - fa;
                  for (int q = 0; q < 3; q++) {
                    if (c->blend_mode == MOP_HIP_BLEND_ADDITIVE)
                      hp[q] = hp[q] + rgb[q] * fa;
                    else if (c->blend_mode == MOP_HIP_BLEND_MULTIPLY)
                      hp[q] = hp[q] * rgb[q];
                    else
                      hp[q] = rgb[q] * fa + hp[q] * inv;
                  }
                  hp[3] = 1.0f;
                  if (fa > 0.5f) {
                    fb->depth[idx] = z;
                    fb->ids[idx] = c->object_id;
                  }
                }
              }
            }
          }
          w0 += sp.edx[0];
          w1 += sp.edx[1];
          w2 += sp.edx[2];
        }
        r0 += sp.edy[0];
        r1 += sp.edy[1];
        r2 += sp.edy[2];
      }
    }
  }
  free(vc);
  if (saved) {
    memcpy(fb->depth, saved, npx * sizeof(float));
    free(saved);
  }
}

void mop_oracle_draw(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c) {
  if (c && c->vertex_buffer && c->index_buffer)
    draw_one(d, fb, c, &c->model, &c->mvp);
}

void mop_oracle_draw_instanced(MopHipDevice *d, MopHipFramebuffer *fb, const MopHipDrawCall *c,
                               const MopHipMat4 *xf, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) {
    MopHipMat4 vm = mmul(&c->view, &xf[i]);
    MopHipMat4 mvp = mmul(&c->projection, &vm);
    draw_one(d, fb, c, &xf[i], &mvp);
  }
}

/* ------------------------------------------------------------- shading -- */

static float pcf(const ShadeState *s, V3 wp) {
  float lp[4];
  mulv(s->svp, wp.x, wp.y, wp.z, 1.0f, lp);
  if (lp[3] <= 0.0f)
    return 1.0f;
  float iw = 1.0f / lp[3];
  float nx = lp[0] * iw, ny = lp[1] * iw, nz = (lp[2] * iw + 1.0f) * 0.5f;
  float u = (nx + 1.0f) * 0.5f * (float)s->sw, v = (1.0f - ny) * 0.5f * (float)s->sh;
  int cx = (int)u, cy = (int)v, lit = 0, tot = 0;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      int x = cx + dx, y = cy + dy;
      if (x < 0 || x >= s->sw || y < 0 || y >= s->sh) {
        lit++;
        tot++;
        continue;
      }
      if (nz <= s->smap[(size_t)y * s->sw + x] + 0.005f)
        lit++;
      tot++;
    }
  return (float)lit / (float)tot;
}

static void samp(const float *data, int w, int h, float u, float v, float out[3]) {
  u = u - floorf(u);
  v = v - floorf(v);
  float fx = u * (float)(w - 1), fy = v * (float)(h - 1);
  int x0 = (int)fx, y0 = (int)fy;
  int x1 = (x0 + 1) % w, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
  float sx = fx - (float)x0, sy = fy - (float)y0;
  const float *p00 = data + ((size_t)y0 * w + x0) * 4, *p10 = data + ((size_t)y0 * w + x1) * 4;
  const float *p01 = data + ((size_t)y1 * w + x0) * 4, *p11 = data + ((size_t)y1 * w + x1) * 4;
  for (int c = 0; c < 3; c++) {
    float top = p00[c] * (1.0f - sx) + p10[c] * sx;
    float bot = p01[c] * (1.0f - sx) + p11[c] * sx;
    out[c] = top * (1.0f - sy) + bot * sy;
  }
}

static void d2uv(float dx, float dy, float dz, float rot, float *u, float *v) {
  float phi = atan2f(dz, dx) + rot;
  float cdy = dy < -1.0f ? -1.0f : (dy > 1.0f ? 1.0f : dy);
  *u = phi / (2.0f * 3.14159265f) + 0.5f;
  *v = asinf(cdy) / 3.14159265f + 0.5f;
}

static void shade_px(const MopHipDevice *dev, const ShadeState *S, const DrawRec *dr,
                     const SubTri *t, float w0, float w1, float w2, float inv_area,
                     float out[3]) {
  float b0 = w0 * inv_area, b1 = w1 * inv_area, b2 = w2 * inv_area;
  float pc0 = b0 * t->iw[0], pc1 = b1 * t->iw[1], pc2 = b2 * t->iw[2];
  float ip = 1.0f / (pc0 + pc1 + pc2);
  pc0 *= ip;
  pc1 *= ip;
  pc2 *= ip;
  const ClipV *A = &t->v[0], *B = &t->v[1], *Cc = &t->v[2];
#define IP(f, k) (pc0 * A->f[k] + pc1 * B->f[k] + pc2 * Cc->f[k])
  V3 n = vnorm(v3(IP(n, 0), IP(n, 1), IP(n, 2)));
  float fr = IP(c, 0), fg = IP(c, 1), fb = IP(c, 2);
  V3 wp = v3(IP(wp, 0), IP(wp, 1), IP(wp, 2));
#undef IP
  V3 vd = vnorm(vsub(dr->eye, wp));
  float met = dr->metallic, rough = dr->roughness;
  /* compute_multi_light_rgb (rasterizer.c:1596-1688) */
  float lit[3] = {dr->ambient, dr->ambient, dr->ambient};
  for (uint32_t i = 0; i < dr->nlights; i++) {
    const MopHipLight *L = &dr->lights[i];
    if (!L->active)
      continue;
    float ndotl = 0.0f, att = 1.0f, spot = 1.0f;
    if (L->type == MOP_HIP_LIGHT_DIRECTIONAL) {
      ndotl = -vdot(n, vnorm(L->direction));
    } else if (L->type == MOP_HIP_LIGHT_POINT || L->type == MOP_HIP_LIGHT_SPOT) {
      V3 tl = vsub(L->position, wp);
      float dist = vlen(tl);
      if (dist < 1e-6f)
        dist = 1e-6f;
      V3 dir = vscale(tl, 1.0f / dist);
      ndotl = vdot(n, dir);
      if (L->type == MOP_HIP_LIGHT_SPOT) {
        float ca = -vdot(dir, vnorm(L->direction));
        if (ca < L->spot_outer_cos) {
          spot = 0.0f;
        } else if (ca < L->spot_inner_cos) {
          float rs = L->spot_inner_cos - L->spot_outer_cos;
          if (rs > 1e-6f) {
            float tt = (ca - L->spot_outer_cos) / rs;
            spot = tt * tt * (3.0f - 2.0f * tt);
          }
        }
      }
      if (L->range > 0.0f) {
        float r = dist / L->range;
        att = 1.0f - r;
        if (att < 0.0f)
          att = 0.0f;
        att *= att;
      } else {
        float d2 = dist * dist;
        if (d2 < 0.01f)
          d2 = 0.01f;
        att = 1.0f / d2;
      }
    }
    if (ndotl < 0.0f)
      ndotl = 0.0f;
    float sh = 1.0f;
    if (L->type == MOP_HIP_LIGHT_DIRECTIONAL && S->shadow_on)
      sh = pcf(S, wp);
    float con = ndotl * L->intensity * att * spot * sh;
    lit[0] += con * L->color.r;
    lit[1] += con * L->color.g;
    lit[2] += con * L->color.b;
  }
  /* compute_multi_specular_ggx (rasterizer.c:1701-1814) */
  float spec[3] = {0, 0, 0};
  {
    float al = rough * rough, al2 = al * al;
    if (al2 < 1e-7f)
      al2 = 1e-7f;
    float ndv = vdot(n, vd);
    if (ndv < 1e-4f)
      ndv = 1e-4f;
    float f0[3] = {0.04f * (1.0f - met) + fr * met, 0.04f * (1.0f - met) + fg * met,
                   0.04f * (1.0f - met) + fb * met};
    float k = al * 0.5f;
    float g1v = ndv / (ndv * (1.0f - k) + k);
    for (uint32_t i = 0; i < dr->nlights; i++) {
      const MopHipLight *L = &dr->lights[i];
      if (!L->active)
        continue;
      V3 ld = v3(0, 0, 0);
      float att = 1.0f;
      if (L->type == MOP_HIP_LIGHT_DIRECTIONAL) {
        V3 dd = vnorm(L->direction);
        ld = v3(-dd.x, -dd.y, -dd.z);
      } else if (L->type == MOP_HIP_LIGHT_POINT || L->type == MOP_HIP_LIGHT_SPOT) {
        V3 tl = vsub(L->position, wp);
        float dist = vlen(tl);
        if (dist < 1e-6f)
          dist = 1e-6f;
        ld = vscale(tl, 1.0f / dist);
        if (L->range > 0.0f) {
          float r = dist / L->range;
          att = 1.0f - r;
          if (att < 0.0f)
            att = 0.0f;
          att *= att;
        } else {
          float d2 = dist * dist;
          if (d2 < 0.01f)
            d2 = 0.01f;
          att = 1.0f / d2;
        }
      }
      float ndl = vdot(n, ld);
      if (ndl <= 0.0f)
        continue;
      V3 h = vnorm(vadd(ld, vd));
      float ndh = vdot(n, h), vdh = vdot(vd, h);
      if (ndh < 0.0f) ndh = 0.0f;
      if (vdh < 0.0f) vdh = 0.0f;
      float dd = ndh * ndh * (al2 - 1.0f) + 1.0f;
      float D = al2 / ((float)3.14159265358979323846 * dd * dd);
      float g1l = ndl / (ndl * (1.0f - k) + k);
      float G = g1v * g1l;
      float om = 1.0f - vdh, om2 = om * om, om5 = om2 * om2 * om;
      float sd = 4.0f * ndl * ndv;
      if (sd < 1e-6f)
        sd = 1e-6f;
      float st = D * G / sd;
      float sh = 1.0f;
      if (L->type == MOP_HIP_LIGHT_DIRECTIONAL && S->shadow_on)
        sh = pcf(S, wp);
      float wgt = st * L->intensity * att * ndl * sh * (float)3.14159265358979323846;
      float lc[3] = {L->color.r, L->color.g, L->color.b};
      for (int c = 0; c < 3; c++)
        spec[c] += wgt * (f0[c] + (1.0f - f0[c]) * om5) * lc[c];
    }
  }
  float ds = 1.0f - met;
  float env[3] = {0, 0, 0}, irr[3] = {0, 0, 0};
  int has_irr = S->ibl_on && S->irr;
  if (has_irr) {
    float len = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    if (len >= 1e-6f) {
      float u, v;
      d2uv(n.x / len, n.y / len, n.z / len, S->rot, &u, &v);
      samp(S->irr, S->irr_w, S->irr_h, u, v, irr);
      for (int c = 0; c < 3; c++)
        irr[c] *= S->inten;
    }
  }
  {
    float vdn = vdot(vd, n);
    V3 refl = v3(2.0f * vdn * n.x - vd.x, 2.0f * vdn * n.y - vd.y, 2.0f * vdn * n.z - vd.z);
    float ndv = vdot(n, vd);
    if (ndv < 0.0f)
      ndv = 0.0f;
    float f0[3] = {0.04f * (1.0f - met) + fr * met, 0.04f * (1.0f - met) + fg * met,
                   0.04f * (1.0f - met) + fb * met};
    if (S->ibl_on && S->pf && S->lut) {
      float pf[3] = {0, 0, 0};
      float len = sqrtf(refl.x * refl.x + refl.y * refl.y + refl.z * refl.z);
      if (len >= 1e-6f) {
        int lv = (int)(rough * (float)(S->pf_levels - 1));
        if (lv >= S->pf_levels)
          lv = S->pf_levels - 1;
        size_t off = 0;
        for (int l = 0; l < lv; l++) {
          int lw = S->pf_w >> l, lh = S->pf_h >> l;
          off += (size_t)(lw < 1 ? 1 : lw) * (lh < 1 ? 1 : lh);
        }
        int lw = S->pf_w >> lv, lh = S->pf_h >> lv;
        float u, v;
        d2uv(refl.x / len, refl.y / len, refl.z / len, S->rot, &u, &v);
        samp(S->pf + off * 4, lw < 1 ? 1 : lw, lh < 1 ? 1 : lh, u, v, pf);
        for (int c = 0; c < 3; c++)
          pf[c] *= S->inten;
      }
      float nv = ndv > 1.0f ? 1.0f : ndv, rr = clamp01(rough);
      int x = (int)(nv * (float)(S->lut_n - 1) + 0.5f), y = (int)(rr * (float)(S->lut_n - 1) + 0.5f);
      if (x >= S->lut_n) x = S->lut_n - 1;
      if (y >= S->lut_n) y = S->lut_n - 1;
      float bs = S->lut[((size_t)y * S->lut_n + x) * 4], bb = S->lut[((size_t)y * S->lut_n + x) * 4 + 1];
      for (int c = 0; c < 3; c++)
        env[c] = pf[c] * (f0[c] * bs + bb);
    } else if (met > 0.01f) {
      float skt = refl.y * 0.5f + 0.5f;
      if (skt < 0.0f) skt = 0.0f;
      if (skt > 1.0f) skt = 1.0f;
      float sky[3] = {0.35f, 0.38f, 0.45f}, gnd[3] = {0.08f, 0.07f, 0.06f}, he[3];
      for (int c = 0; c < 3; c++)
        he[c] = gnd[c] + skt * (sky[c] - gnd[c]);
      float rm = rough * rough, avg = (he[0] + he[1] + he[2]) * (1.0f / 3.0f);
      for (int c = 0; c < 3; c++)
        he[c] += rm * (avg - he[c]);
      float om = 1.0f - ndv, om2 = om * om, om5 = om2 * om2 * om;
      for (int c = 0; c < 3; c++)
        env[c] = he[c] * (f0[c] + (1.0f - f0[c]) * om5) * dr->ambient * 3.0f;
    }
  }
  float col[3] = {fr, fg, fb};
  for (int c = 0; c < 3; c++) {
    float p = has_irr ? col[c] * (lit[c] + irr[c]) * ds + env[c] + spec[c]
                      : col[c] * lit[c] * ds + env[c] + spec[c];
    out[c] = p < 0.0f ? 0.0f : p;
  }
}

/* compute_multi_light (rasterizer.c:1498-1589) */
static float multi_light(const ShadeState *S, const DrawRec *dr, V3 n, V3 wp) {
  float total = dr->ambient;
  for (uint32_t i = 0; i < dr->nlights; i++) {
    const MopHipLight *L = &dr->lights[i];
    if (!L->active)
      continue;
    float ndotl = 0.0f, att = 1.0f, spot = 1.0f;
    if (L->type == MOP_HIP_LIGHT_DIRECTIONAL) {
      ndotl = -vdot(n, vnorm(L->direction));
    } else if (L->type == MOP_HIP_LIGHT_POINT || L->type == MOP_HIP_LIGHT_SPOT) {
      V3 tl = vsub(L->position, wp);
      float dist = vlen(tl);
      if (dist < 1e-6f)
        dist = 1e-6f;
      V3 dir = vscale(tl, 1.0f / dist);
      ndotl = vdot(n, dir);
      if (L->type == MOP_HIP_LIGHT_SPOT) {
        float ca = -vdot(dir, vnorm(L->direction));
        if (ca < L->spot_outer_cos) {
          spot = 0.0f;
        } else if (ca < L->spot_inner_cos) {
          float rs = L->spot_inner_cos - L->spot_outer_cos;
          if (rs > 1e-6f) {
            float tt = (ca - L->spot_outer_cos) / rs;
            spot = tt * tt * (3.0f - 2.0f * tt);
          }
        }
      }
      if (L->range > 0.0f) {
        float r = dist / L->range;
        att = 1.0f - r;
        if (att < 0.0f)
          att = 0.0f;
        att *= att;
      } else if (L->type == MOP_HIP_LIGHT_POINT) {
        float d2 = dist * dist;
        if (d2 < 0.01f)
          d2 = 0.01f;
        att = 1.0f / d2;
      } else {
        att = 1.0f / (1.0f + dist * dist);
      }
    }
    if (ndotl < 0.0f)
      ndotl = 0.0f;
    float sh = 1.0f;
    if (L->type == MOP_HIP_LIGHT_DIRECTIONAL && S->shadow_on)
      sh = pcf(S, wp);
    total += ndotl * L->intensity * att * spot * sh;
  }
  return clamp01(total);
}

/* Constant colour of a flat screen triangle (rasterizer.c:2556-2596 lit,
 * 1199-1220 unlit). */
static void flat_color(const ShadeState *S, const DrawRec *dr, const SubTri *t, float out[3]) {
  const ClipV *a = &t->v[0], *b = &t->v[1], *c = &t->v[2];
  V3 fn = vnorm(v3(a->n[0] + b->n[0] + c->n[0], a->n[1] + b->n[1] + c->n[1],
                   a->n[2] + b->n[2] + c->n[2]));
  float avg[3];
  for (int k = 0; k < 3; k++)
    avg[k] = (a->c[k] + b->c[k] + c->c[k]) * (1.0f / 3.0f);
  if (dr->sm == SM_LIT_FLAT || dr->sm == SM_LIT_SMOOTH) { /* lit (wireframe: any mode) */
    V3 wp = v3((a->wp[0] + b->wp[0] + c->wp[0]) / 3.0f, (a->wp[1] + b->wp[1] + c->wp[1]) / 3.0f,
               (a->wp[2] + b->wp[2] + c->wp[2]) / 3.0f);
    float lighting = multi_light(S, dr, fn, wp);
    float ds = 1.0f - dr->metallic, envf = dr->ambient * dr->metallic * 0.3f;
    for (int k = 0; k < 3; k++)
      out[k] = avg[k] * (lighting * ds + envf);
  } else {
    float ndotl = vdot(fn, vnorm(dr->light_dir));
    if (ndotl < 0.0f)
      ndotl = 0.0f;
    float lighting = clamp01(dr->ambient + (1.0f - dr->ambient) * ndotl);
    for (int k = 0; k < 3; k++)
      out[k] = avg[k] * lighting;
  }
  for (int k = 0; k < 3; k++)
    if (out[k] < 0.0f)
      out[k] = 0.0f;
}

/* mop_sw_rasterize_triangle_smooth's colour (rasterizer.c:1368-1419) */
static void unlit_smooth_px(const DrawRec *dr, const SubTri *t, float w0, float w1, float w2,
                            float inv_area, float out[3]) {
  float b0 = w0 * inv_area, b1 = w1 * inv_area, b2 = w2 * inv_area;
  float pc0 = b0 * t->iw[0], pc1 = b1 * t->iw[1], pc2 = b2 * t->iw[2];
  float ip = 1.0f / (pc0 + pc1 + pc2);
  pc0 *= ip;
  pc1 *= ip;
  pc2 *= ip;
  const ClipV *A = &t->v[0], *B = &t->v[1], *Cc = &t->v[2];
#define IP(f, k) (pc0 * A->f[k] + pc1 * B->f[k] + pc2 * Cc->f[k])
  V3 n = vnorm(v3(IP(n, 0), IP(n, 1), IP(n, 2)));
  float col[3] = {IP(c, 0), IP(c, 1), IP(c, 2)};
#undef IP
  V3 nl = vnorm(dr->light_dir);
  float ndotl = vdot(n, nl);
  if (ndotl < 0.0f)
    ndotl = 0.0f;
  float lit = dr->ambient + (1.0f - dr->ambient) * ndotl;
  V3 hv = vnorm(v3(nl.x, nl.y + 1.0f, nl.z));
  float ndh = vdot(n, hv);
  if (ndh < 0.0f)
    ndh = 0.0f;
  float sp = ndh * ndh;
  for (int k = 0; k < 4; k++)
    sp *= sp;
  sp *= 0.25f;
  for (int k = 0; k < 3; k++) {
    float p = col[k] * lit + sp;
    out[k] = p < 0.0f ? 0.0f : p;
  }
}

/* Shade the winners of the open segment; replays each winner's incremental
 * walk to its pixel for the edge values. */
static void shade_segment(MopHipDevice *d, MopHipFramebuffer *fb) {
  if (!fb->seg_open)
    return;
  size_t n = (size_t)fb->w * fb->h;
  for (size_t i = 0; i < n; i++) {
    uint32_t s1 = (uint32_t)fb->vis[i];
    if (!s1)
      continue;
    const SubTri *t = &fb->tris[s1 - 1];
    Setup s;
    setup_tri(t->sx, t->sy, fb->w, fb->h, &s, 0);
    int px = (int)(i % fb->w), py = (int)(i / fb->w);
    const DrawRec *dr = &fb->draws[t->draw];
    float rgb[3], z;
    if (dr->sm == SM_LIT_FLAT || dr->sm == SM_UNLIT_FLAT) {
      float zdx, zdy;
      flat_z(&s, t, &z, &zdx, &zdy);
      for (int y = s.y0; y < py; y++)
        z += zdy;
      for (int x = s.x0; x < px; x++)
        z += zdx;
      flat_color(&d->states[dr->state], dr, t, rgb);
    } else {
      float w[3];
      for (int k = 0; k < 3; k++) {
        float r = s.wr[k];
        for (int y = s.y0; y < py; y++)
          r += s.edy[k];
        for (int x = s.x0; x < px; x++)
          r += s.edx[k];
        w[k] = r;
      }
      float b0 = w[0] * s.inv_area, b1 = w[1] * s.inv_area, b2 = w[2] * s.inv_area;
      z = b0 * t->sz[0] + b1 * t->sz[1] + b2 * t->sz[2];
      shade_px(d, &d->states[dr->state], dr, t, w[0], w[1], w[2], s.inv_area, rgb);
    }
    fb->hdr[i * 4] = rgb[0];
    fb->hdr[i * 4 + 1] = rgb[1];
    fb->hdr[i * 4 + 2] = rgb[2];
    fb->hdr[i * 4 + 3] = 1.0f;
    fb->depth[i] = z;
    fb->ids[i] = dr->object_id;
  }
  for (size_t k = 0; k < fb->ndraws; k++)
    free(fb->draws[k].lights);
  fb->ndraws = 0;
  fb->ntris = 0;
  fb->seg_open = 0;
}

static float aces(float x) {
  float m = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
  return m < 0.0f ? 0.0f : (m > 1.0f ? 1.0f : m);
}

static void resolve(MopHipFramebuffer *fb, float e) {
  size_t n = (size_t)fb->w * fb->h;
  for (size_t i = 0; i < n; i++) {
    float *h = fb->hdr + i * 4;
    uint8_t *c = fb->color + i * 4;
    if (fb->ids[i] == 0) {
      for (int k = 0; k < 3; k++)
        c[k] = (uint8_t)(fminf(h[k], 1.0f) * 255.0f + 0.5f);
      c[3] = 255;
    } else {
      for (int k = 0; k < 3; k++)
        c[k] = (uint8_t)(aces(h[k] * e) * 255.0f + 0.5f);
      c[3] = (uint8_t)(h[3] * 255.0f + 0.5f);
    }
  }
}

void mop_oracle_frame_end(MopHipDevice *d, MopHipFramebuffer *fb) {
  shade_segment(d, fb);
  resolve(fb, d->exposure);
}
void mop_oracle_frame_submit(MopHipDevice *d, MopHipFramebuffer *fb) {}

void mop_oracle_hdr_resolve(MopHipDevice *d, MopHipFramebuffer *fb, float e) {
  shade_segment(d, fb);
  resolve(fb, e);
}

/* ------------------------------------------------------- shadow / IBL -- */

void mop_oracle_shadow_render_mesh(MopHipDevice *d, const void *verts, uint32_t vc,
                                   const uint32_t *idx, uint32_t ic, const MopHipMat4 *model,
                                   const MopHipMat4 *lvp, int size, int first) {
  if (first || d->ssize != size) {
    free(d->smap);
    d->ssize = size;
    d->smap = malloc((size_t)size * size * sizeof(float));
    for (size_t i = 0; i < (size_t)size * size; i++)
      d->smap[i] = 1.0f;
  }
  MopHipMat4 mvp = mmul(lvp, model);
  const MopHipVertex *v = (const MopHipVertex *)verts;
  float hw = (float)size * 0.5f;
  for (uint32_t i = 0; i + 2 < ic; i += 3) {
    uint32_t id[3] = {idx[i], idx[i + 1], idx[i + 2]};
    if (id[0] >= vc || id[1] >= vc || id[2] >= vc)
      continue;
    float cp[3][4];
    for (int k = 0; k < 3; k++)
      mulv(mvp.d, v[id[k]].position.x, v[id[k]].position.y, v[id[k]].position.z, 1.0f, cp[k]);
    const float *a = cp[0], *b = cp[1], *c = cp[2];
    if ((a[0] < -a[3] && b[0] < -b[3] && c[0] < -c[3]) ||
        (a[0] > a[3] && b[0] > b[3] && c[0] > c[3]) ||
        (a[1] < -a[3] && b[1] < -b[3] && c[1] < -c[3]) ||
        (a[1] > a[3] && b[1] > b[3] && c[1] > c[3]) ||
        (a[2] < -a[3] && b[2] < -b[3] && c[2] < -c[3]) ||
        (a[2] > a[3] && b[2] > b[3] && c[2] > c[3]))
      continue;
    if (a[3] <= 1e-6f || b[3] <= 1e-6f || c[3] <= 1e-6f)
      continue;
    float sx[3], sy[3], sz[3];
    for (int k = 0; k < 3; k++) {
      float iw = 1.0f / cp[k][3];
      sx[k] = (cp[k][0] * iw + 1.0f) * hw;
      sy[k] = (1.0f - cp[k][1] * iw) * hw;
      sz[k] = (cp[k][2] * iw + 1.0f) * 0.5f;
    }
    Setup s;
    if (!setup_tri(sx, sy, size, size, &s, 0))
      continue;
    float r0 = s.wr[0], r1 = s.wr[1], r2 = s.wr[2];
    for (int y = s.y0; y <= s.y1; y++) {
      float w0 = r0, w1 = r1, w2 = r2;
      for (int x = s.x0; x <= s.x1; x++) {
        if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) {
          float z = w0 * s.inv_area * sz[0] + w1 * s.inv_area * sz[1] + w2 * s.inv_area * sz[2];
          float *dp = &d->smap[(size_t)y * size + x];
          if (z < *dp)
            *dp = z;
        }
        w0 += s.edx[0];
        w1 += s.edx[1];
        w2 += s.edx[2];
      }
      r0 += s.edy[0];
      r1 += s.edy[1];
      r2 += s.edy[2];
    }
  }
}

void mop_oracle_shadow_set(MopHipDevice *d, int w, int h, const MopHipMat4 *lvp) {
  d->st.shadow_on = d->smap != NULL;
  d->st.smap = d->smap;
  d->st.sw = w;
  d->st.sh = h;
  memcpy(d->st.svp, lvp->d, sizeof(d->st.svp));
  d->state_dirty = 1;
}
void mop_oracle_shadow_clear(MopHipDevice *d) {
  d->st.shadow_on = 0;
  d->state_dirty = 1;
}

void mop_oracle_ibl_set(MopHipDevice *d, const float *irr, int iw, int ih, const float *pf,
                        int pw, int ph, int pl, const float *lut, int ln, float rot,
                        float inten) {
  /* tables are kept by pointer, as mop_sw_ibl_set does (rasterizer.c:167-181) */
  d->st.ibl_on = irr != NULL || (pf != NULL && lut != NULL);
  d->st.irr = irr;
  d->st.pf = pf;
  d->st.lut = lut;
  d->st.irr_w = iw;
  d->st.irr_h = ih;
  d->st.pf_w = pw;
  d->st.pf_h = ph;
  d->st.pf_levels = pl;
  d->st.lut_n = ln;
  d->st.rot = rot;
  d->st.inten = inten;
  d->state_dirty = 1;
}
void mop_oracle_ibl_clear(MopHipDevice *d) {
  d->st.ibl_on = 0;
  d->st.irr = d->st.pf = d->st.lut = NULL;
  d->state_dirty = 1;
}

/* --------------------------------------------------------------- post -- */

static float luma(const uint8_t *s, int w, int x, int y) {
  const uint8_t *p = s + ((size_t)y * w + x) * 4;
  return 0.299f * (float)p[0] + 0.587f * (float)p[1] + 0.114f * (float)p[2];
}

void mop_oracle_postprocess_apply(MopHipDevice *d, MopHipFramebuffer *fb, uint32_t fx,
                                  const MopHipFogParams *fog) {
  if (fx == 0)
    return;
  int w = fb->w, h = fb->h;
  size_t n = (size_t)w * h * 4;
  if ((fx & MOP_HIP_POST_FXAA) && w >= 3 && h >= 3) {
    uint8_t *src = malloc(n);
    memcpy(src, fb->color, n);
    for (int y = 1; y < h - 1; y++)
      for (int x = 1; x < w - 1; x++) {
        float lM = luma(src, w, x, y), lN = luma(src, w, x, y - 1), lS = luma(src, w, x, y + 1);
        float lW = luma(src, w, x - 1, y), lE = luma(src, w, x + 1, y);
        float mn = lM, mx = lM;
        if (lN < mn) mn = lN;
        if (lS < mn) mn = lS;
        if (lW < mn) mn = lW;
        if (lE < mn) mn = lE;
        if (lN > mx) mx = lN;
        if (lS > mx) mx = lS;
        if (lW > mx) mx = lW;
        if (lE > mx) mx = lE;
        float range = mx - mn;
        if (range < 0.0625f * mx + 0.004f)
          continue;
        float lNW = luma(src, w, x - 1, y - 1), lNE = luma(src, w, x + 1, y - 1);
        float lSW = luma(src, w, x - 1, y + 1), lSE = luma(src, w, x + 1, y + 1);
        float avg = (lN + lS + lW + lE) * (2.0f / 12.0f) + (lNW + lNE + lSW + lSE) * (1.0f / 12.0f);
        float sp = fabsf(avg - lM) / range;
        if (sp > 1.0f)
          sp = 1.0f;
        sp = sp * sp * (3.0f - 2.0f * sp);
        sp *= 0.75f;
        float eh = fabsf(-2.0f * lW + lNW + lSW) + fabsf(-2.0f * lM + lN + lS) * 2.0f +
                   fabsf(-2.0f * lE + lNE + lSE);
        float ev = fabsf(-2.0f * lN + lNW + lNE) + fabsf(-2.0f * lM + lW + lE) * 2.0f +
                   fabsf(-2.0f * lS + lSW + lSE);
        int hz = eh >= ev;
        int stx = hz ? 0 : 1, sty = hz ? 1 : 0;
        float gp = fabsf((hz ? lS : lE) - lM), gn = fabsf((hz ? lN : lW) - lM);
        int ox = gp >= gn ? stx : -stx, oy = gp >= gn ? sty : -sty;
        int nx = x + ox, ny = y + oy;
        size_t ci = ((size_t)y * w + x) * 4, ni = ((size_t)ny * w + nx) * 4;
        float ib = 1.0f - sp;
        for (int k = 0; k < 3; k++)
          fb->color[ci + k] = (uint8_t)(src[ci + k] * ib + src[ni + k] * sp);
      }
    free(src);
  }
  float cx = (float)w * 0.5f, cy = (float)h * 0.5f, md = sqrtf(cx * cx + cy * cy);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      uint8_t *c = fb->color + ((size_t)y * w + x) * 4;
      float v[3] = {(float)c[0] / 255.0f, (float)c[1] / 255.0f, (float)c[2] / 255.0f};
      if ((fx & MOP_HIP_POST_FOG) && fog) {
        float dz = fb->depth[(size_t)y * w + x], rg = fog->far_dist - fog->near_dist, f = 0.0f;
        if (rg > 0.0f) {
          f = (dz - fog->near_dist) / rg;
          f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
        }
        float fc[3] = {fog->color.r, fog->color.g, fog->color.b};
        for (int k = 0; k < 3; k++)
          v[k] = v[k] + (fc[k] - v[k]) * f;
      }
      for (int k = 0; k < 3; k++) {
        if (fx & MOP_HIP_POST_TONEMAP)
          v[k] = v[k] / (v[k] + 1.0f);
        if (fx & MOP_HIP_POST_GAMMA) {
          float q = v[k] > 0.0f ? v[k] : 0.0f;
          v[k] = q <= 0.0031308f ? q * 12.92f : 1.055f * powf(q, 1.0f / 2.4f) - 0.055f;
        }
      }
      if (fx & MOP_HIP_POST_VIGNETTE) {
        float dx = (float)x - cx, dy = (float)y - cy;
        float dist = sqrtf(dx * dx + dy * dy) / md;
        float vg = 1.0f - dist * dist * 0.5f;
        if (vg < 0.0f)
          vg = 0.0f;
        for (int k = 0; k < 3; k++)
          v[k] *= vg;
      }
      for (int k = 0; k < 3; k++) {
        float q = v[k] > 1.0f ? 1.0f : v[k];
        if (q < 0.0f)
          q = 0.0f;
        c[k] = (uint8_t)(q * 255.0f + 0.5f);
      }
    }
}

/* ------------------------------------------------------------ readback -- */

const uint8_t *mop_oracle_framebuffer_read_color(MopHipDevice *d, MopHipFramebuffer *fb, int *w,
                                                 int *h) {
  if (w) *w = fb->w;
  if (h) *h = fb->h;
  return fb->color;
}
/* mop_overlay_builtin_outline (overlay_builtin.c:1309-1400) on the resolved
 * image: classify every pixel first (the id buffer does not change while
 * the overlay paints), then paint in row order -- 5x5 blocks around edges of
 * selected objects, single pixels at edges of other objects -- each paint
 * blending into the current colour. */
#define ORACLE_CHROME 0xFFFD0000u
void mop_oracle_draw_outline(MopHipDevice *d, MopHipFramebuffer *fb, MopHipColor accent,
                             float alpha_sel, float alpha_unsel, const uint32_t *sel,
                             uint32_t n_sel) {
  (void)d;
  int w = fb->w, h = fb->h;
  const uint32_t *ids = fb->ids;
  uint8_t acc[3] = {(uint8_t)(accent.r * 255.0f), (uint8_t)(accent.g * 255.0f),
                    (uint8_t)(accent.b * 255.0f)};
  uint8_t *kind = calloc((size_t)w * h, 1); /* 1 edge of an unselected object, 2 selected */
  if (!kind)
    return;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      uint32_t id = ids[(size_t)y * w + x];
      if (id == 0 || id >= ORACLE_CHROME)
        continue;
      const int nx[4] = {x - 1, x + 1, x, x}, ny[4] = {y, y, y - 1, y + 1};
      int edge = 0;
      for (int k = 0; k < 4 && !edge; k++) {
        if (nx[k] < 0 || nx[k] >= w || ny[k] < 0 || ny[k] >= h)
          continue;
        uint32_t o = ids[(size_t)ny[k] * w + nx[k]];
        edge = o != id && o < ORACLE_CHROME;
      }
      if (!edge)
        continue;
      int selected = 0;
      for (uint32_t i = 0; i < n_sel && !selected; i++)
        selected = sel[i] == id;
      kind[(size_t)y * w + x] = selected ? 2 : 1;
    }
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int k = kind[(size_t)y * w + x];
      if (!k)
        continue;
      int r = k == 2 ? 2 : 0;
      float a = k == 2 ? alpha_sel : alpha_unsel;
      for (int py = y - r; py <= y + r; py++)
        for (int px = x - r; px <= x + r; px++) {
          if (px < 0 || px >= w || py < 0 || py >= h || ids[(size_t)py * w + px] >= ORACLE_CHROME)
            continue;
          uint8_t *c = fb->color + ((size_t)py * w + px) * 4;
          for (int j = 0; j < 3; j++)
            c[j] = (uint8_t)(c[j] * (1.0f - a) + acc[j] * a);
        }
    }
  free(kind);
}

/* viewport.c:3385-3411: box filter of sf x sf blocks, truncating integer
 * mean per channel (the restatement's counterpart of the device kernel) */
const uint8_t *mop_oracle_framebuffer_read_color_downsampled(MopHipDevice *d,
                                                             MopHipFramebuffer *fb, int sf,
                                                             int *ow, int *oh) {
  (void)d;
  int pw = fb->w / sf, ph = fb->h / sf;
  if (ow) *ow = pw;
  if (oh) *oh = ph;
  if (sf < 1 || pw < 1 || ph < 1)
    return NULL;
  free(fb->ds);
  fb->ds = malloc((size_t)pw * ph * 4);
  if (!fb->ds)
    return NULL;
  for (int y = 0; y < ph; y++)
    for (int x = 0; x < pw; x++) {
      int acc[4] = {0, 0, 0, 0};
      for (int dy = 0; dy < sf; dy++)
        for (int dx = 0; dx < sf; dx++) {
          const uint8_t *p = &fb->color[((size_t)(y * sf + dy) * fb->w + x * sf + dx) * 4];
          for (int c = 0; c < 4; c++)
            acc[c] += p[c];
        }
      for (int c = 0; c < 4; c++)
        fb->ds[((size_t)y * pw + x) * 4 + c] = (uint8_t)(acc[c] / (sf * sf));
    }
  return fb->ds;
}

/* pass_background_hdri_cpu (viewport.c:2128-2180) restated: the camera ray
 * of every pixel mapped to the equirect environment and sampled bilinearly
 * into the HDR buffer (the restatement draws it at once: frame_begin has
 * cleared and no draw has run yet). */
void mop_oracle_draw_background_hdri(MopHipDevice *d, MopHipFramebuffer *fb, MopHipTexture *env,
                                     float fov, V3 eye, V3 target, V3 up, float rotation,
                                     float intensity) {
  (void)d;
  if (!fb || !env || !env->hdr)
    return;
  const float *data = (const float *)env->data;
  int sw = fb->w, sh = fb->h;
  float aspect = (float)sw / (float)sh;
  float half_h = tanf(fov * 0.5f);
  float half_w = half_h * aspect;
  V3 fwd = vnorm(vsub(target, eye));
  V3 c1 = v3(fwd.y * up.z - fwd.z * up.y, fwd.z * up.x - fwd.x * up.z,
             fwd.x * up.y - fwd.y * up.x);
  V3 right = vnorm(c1);
  V3 u = v3(right.y * fwd.z - right.z * fwd.y, right.z * fwd.x - right.x * fwd.z,
            right.x * fwd.y - right.y * fwd.x);
  const float pi = 3.14159265358979f;
  for (int y = 0; y < sh; y++) {
    float ndc_y = 1.0f - 2.0f * ((float)y + 0.5f) / (float)sh;
    for (int x = 0; x < sw; x++) {
      float ndc_x = 2.0f * ((float)x + 0.5f) / (float)sw - 1.0f;
      float dx = fwd.x + right.x * ndc_x * half_w + u.x * ndc_y * half_h;
      float dy = fwd.y + right.y * ndc_x * half_w + u.y * ndc_y * half_h;
      float dz = fwd.z + right.z * ndc_x * half_w + u.z * ndc_y * half_h;
      float len = sqrtf(dx * dx + dy * dy + dz * dz);
      dx /= len;
      dy /= len;
      dz /= len;
      float phi_val = atan2f(dz, dx) - rotation;
      float cdy = dy < -1.0f ? -1.0f : (dy > 1.0f ? 1.0f : dy);
      float theta_val = asinf(cdy);
      float eu = phi_val / (2.0f * pi) + 0.5f;
      float ev = theta_val / pi + 0.5f;
      eu = eu - floorf(eu);
      ev = ev - floorf(ev);
      float fx = eu * (float)(env->w - 1), fy = ev * (float)(env->h - 1);
      int x0 = (int)fx, y0 = (int)fy;
      int x1 = (x0 + 1) % env->w;
      int y1 = y0 + 1 < env->h ? y0 + 1 : env->h - 1;
      float sx = fx - (float)x0, sy = fy - (float)y0;
      const float *p00 = &data[((size_t)y0 * env->w + x0) * 4];
      const float *p10 = &data[((size_t)y0 * env->w + x1) * 4];
      const float *p01 = &data[((size_t)y1 * env->w + x0) * 4];
      const float *p11 = &data[((size_t)y1 * env->w + x1) * 4];
      float *o = &fb->hdr[((size_t)y * sw + x) * 4];
      for (int c = 0; c < 3; c++) {
        float top = p00[c] * (1.0f - sx) + p10[c] * sx;
        float bot = p01[c] * (1.0f - sx) + p11[c] * sx;
        o[c] = (top * (1.0f - sy) + bot * sy) * intensity;
      }
      o[3] = 1.0f;
    }
  }
}

const uint32_t *mop_oracle_framebuffer_read_object_id(MopHipDevice *d, MopHipFramebuffer *fb,
                                                      int *w, int *h) {
  if (w) *w = fb->w;
  if (h) *h = fb->h;
  return fb->ids;
}
const float *mop_oracle_framebuffer_read_depth(MopHipDevice *d, MopHipFramebuffer *fb, int *w,
                                               int *h) {
  if (w) *w = fb->w;
  if (h) *h = fb->h;
  return fb->depth;
}
const float *mop_oracle_framebuffer_read_hdr(MopHipDevice *d, MopHipFramebuffer *fb, int *w,
                                             int *h) {
  if (w) *w = fb->w;
  if (h) *h = fb->h;
  return fb->hdr;
}
uint32_t mop_oracle_pick_read_id(MopHipDevice *d, MopHipFramebuffer *fb, int x, int y) {
  if (x < 0 || x >= fb->w || y < 0 || y >= fb->h)
    return 0;
  return fb->ids[(size_t)y * fb->w + x];
}
float mop_oracle_pick_read_depth(MopHipDevice *d, MopHipFramebuffer *fb, int x, int y) {
  if (x < 0 || x >= fb->w || y < 0 || y >= fb->h)
    return 1.0f;
  return fb->depth[(size_t)y * fb->w + x];
}
size_t mop_oracle_skipped_draws(MopHipDevice *d) { return d->skipped; }

static const MopHipRhiBackend s_oracle = {
    .name = "oracle",
    .device_create = mop_oracle_device_create,
    .device_destroy = mop_oracle_device_destroy,
    .buffer_create = mop_oracle_buffer_create,
    .buffer_destroy = mop_oracle_buffer_destroy,
    .framebuffer_create = mop_oracle_framebuffer_create,
    .framebuffer_destroy = mop_oracle_framebuffer_destroy,
    .framebuffer_resize = mop_oracle_framebuffer_resize,
    .frame_begin = mop_oracle_frame_begin,
    .frame_end = mop_oracle_frame_end,
    .frame_submit = mop_oracle_frame_submit,
    .draw = mop_oracle_draw,
    .pick_read_id = mop_oracle_pick_read_id,
    .pick_read_depth = mop_oracle_pick_read_depth,
    .framebuffer_read_color = mop_oracle_framebuffer_read_color,
    .framebuffer_read_object_id = mop_oracle_framebuffer_read_object_id,
    .framebuffer_read_depth = mop_oracle_framebuffer_read_depth,
    .texture_create = mop_oracle_texture_create,
    .texture_create_hdr = mop_oracle_texture_create_hdr,
    .texture_destroy = mop_oracle_texture_destroy,
    .draw_instanced = mop_oracle_draw_instanced,
    .buffer_update = mop_oracle_buffer_update,
    .buffer_read = mop_oracle_buffer_read,
    .frame_gpu_time_ms = mop_oracle_frame_gpu_time_ms,
    .set_exposure = mop_oracle_set_exposure,
};

const void *mop_rhi_backend_oracle(void) { return &s_oracle; }
