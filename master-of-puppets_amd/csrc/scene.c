/*
 * scene.c -- C11 frame driver over the HIP backend's C ABI: a resident scene
 * (meshes, lights, environment, shadow map in HBM) rendered frame by frame
 * with the pass order and per-draw arguments of mop_viewport_render
 * (viewport.c:2993-3317): clear, exposure, shadow pass (pass_shadow,
 * viewport.c:2288-2442), IBL state (rg_ibl_setup, :2692-2706), opaque draws
 * (pass_scene_opaque + EMIT_DRAW, :2015-2088, :2449-2467), transparent draws
 * back to front (pass_scene_transparent, :2470-2546), HDR resolve
 * (:2727-2730) and post-processing (:2950-2961).
 *
 * It is the caller the benchmark, the batched form and the serving loop use
 * in place of the viewport core (whose render call is synchronous per
 * viewport).  Camera handling follows the core: the eye is re-derived from
 * orbit parameters every frame (mop_orbit_camera_apply, camera.c:44-63),
 * so frames match the core's bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mop_hip_scene.h"
#include "mop_hip_scene_file.h"

/* ---- host math in the reference's operation order (math.c) -------------- */

static MopHipVec3 v3(float x, float y, float z) {
  MopHipVec3 r = {x, y, z};
  return r;
}
static MopHipVec3 v3_sub(MopHipVec3 a, MopHipVec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static float v3_dot(MopHipVec3 a, MopHipVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float v3_len(MopHipVec3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
static MopHipVec3 v3_cross(MopHipVec3 a, MopHipVec3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static MopHipVec3 v3_norm(MopHipVec3 v) {
  float len = v3_len(v);
  if (len < 1e-8f)
    return v3(0.0f, 0.0f, 0.0f);
  float inv = 1.0f / len;
  return v3(v.x * inv, v.y * inv, v.z * inv);
}

#define M(m, r, c) ((m).d[(c)*4 + (r)])

static MopHipMat4 m4_identity(void) {
  MopHipMat4 m;
  memset(&m, 0, sizeof(m));
  M(m, 0, 0) = M(m, 1, 1) = M(m, 2, 2) = M(m, 3, 3) = 1.0f;
  return m;
}

static MopHipMat4 m4_mul(MopHipMat4 a, MopHipMat4 b) {
  MopHipMat4 r;
  for (int c = 0; c < 4; c++)
    for (int row = 0; row < 4; row++) {
      float sum = 0.0f;
      for (int k = 0; k < 4; k++)
        sum += M(a, row, k) * M(b, k, c);
      M(r, row, c) = sum;
    }
  return r;
}

static MopHipMat4 m4_perspective(float fov, float aspect, float n, float f) {
  float t = tanf(fov * 0.5f);
  MopHipMat4 m;
  memset(&m, 0, sizeof(m));
  M(m, 0, 0) = 1.0f / (aspect * t);
  M(m, 1, 1) = 1.0f / t;
  M(m, 2, 2) = -(f + n) / (f - n);
  M(m, 3, 2) = -1.0f;
  M(m, 2, 3) = -(2.0f * f * n) / (f - n);
  return m;
}

static MopHipMat4 m4_ortho(float l, float r, float b, float t, float n, float f) {
  MopHipMat4 m;
  memset(&m, 0, sizeof(m));
  float rl = r - l, tb = t - b, fn = f - n;
  if (fabsf(rl) < 1e-9f || fabsf(tb) < 1e-9f || fabsf(fn) < 1e-9f)
    return m4_identity();
  M(m, 0, 0) = 2.0f / rl;
  M(m, 1, 1) = 2.0f / tb;
  M(m, 2, 2) = -2.0f / fn;
  M(m, 0, 3) = -(r + l) / rl;
  M(m, 1, 3) = -(t + b) / tb;
  M(m, 2, 3) = -(f + n) / fn;
  M(m, 3, 3) = 1.0f;
  return m;
}

static MopHipMat4 m4_look_at(MopHipVec3 eye, MopHipVec3 center, MopHipVec3 up) {
  MopHipVec3 f = v3_norm(v3_sub(center, eye));
  MopHipVec3 s = v3_norm(v3_cross(f, up));
  MopHipVec3 u = v3_cross(s, f);
  MopHipMat4 m = m4_identity();
  M(m, 0, 0) = s.x;
  M(m, 0, 1) = s.y;
  M(m, 0, 2) = s.z;
  M(m, 1, 0) = u.x;
  M(m, 1, 1) = u.y;
  M(m, 1, 2) = u.z;
  M(m, 2, 0) = -f.x;
  M(m, 2, 1) = -f.y;
  M(m, 2, 2) = -f.z;
  M(m, 0, 3) = -v3_dot(s, eye);
  M(m, 1, 3) = -v3_dot(u, eye);
  M(m, 2, 3) = v3_dot(f, eye);
  return m;
}

/* ---- scene ------------------------------------------------------------- */

typedef struct SceneMesh {
  MopHipBuffer *vb, *ib;
  uint32_t vertex_count, index_count, object_id;
  MopHipMat4 model;
  int has_material;
  float metallic, roughness;
  MopHipVec3 aabb_min, aabb_max; /* local */
  uint32_t blend;                 /* MopHipBlendMode */
  float opacity;
} SceneMesh;

struct MopHipScene {
  MopHipDevice *dev;
  SceneMesh *meshes;
  uint32_t mesh_count, mesh_cap;
  MopHipLight *lights;
  uint32_t light_count, light_cap;
  MopHipIbl *ibl;
  float env_rotation, env_intensity;
  float exposure;
  float ambient;
  MopHipColor clear;
  uint32_t post_effects;
  int shadow_size;
  MopHipShadowMap *shadow;
  int shadow_cache;  /* reuse the rendered map while the scene is unchanged */
  int shadow_valid;
  MopHipMat4 shadow_vp;
  uint32_t last_triangles;
};

MopHipScene *mop_hip_scene_create(MopHipDevice *dev) {
  if (!dev)
    return NULL;
  MopHipScene *s = calloc(1, sizeof(*s));
  if (!s)
    return NULL;
  s->dev = dev;
  s->exposure = 1.0f;
  s->ambient = 0.25f;                                  /* viewport.c:643 */
  s->clear = (MopHipColor){0.10f, 0.10f, 0.11f, 1.0f}; /* theme.c:17,31 */
  s->post_effects = MOP_HIP_POST_GAMMA | MOP_HIP_POST_FXAA; /* viewport.c:646 */
  s->shadow_size = 1024;                               /* viewport.c:2286 */
  s->env_intensity = 1.0f;
  return s;
}

void mop_hip_scene_destroy(MopHipScene *s) {
  if (!s)
    return;
  for (uint32_t i = 0; i < s->mesh_count; i++) {
    mop_hip_buffer_destroy(s->dev, s->meshes[i].vb);
    mop_hip_buffer_destroy(s->dev, s->meshes[i].ib);
  }
  free(s->meshes);
  free(s->lights);
  mop_hip_ibl_destroy(s->dev, s->ibl);
  mop_hip_shadow_map_destroy(s->dev, s->shadow);
  free(s);
}

int mop_hip_scene_add_mesh(MopHipScene *s, const MopHipVertex *v, uint32_t nv,
                           const uint32_t *idx, uint32_t ni, uint32_t object_id,
                           const float *model16, int has_material, float metallic,
                           float roughness) {
  if (!s || !v || !idx || nv == 0 || ni < 3)
    return -1;
  if (s->mesh_count == s->mesh_cap) {
    uint32_t nc = s->mesh_cap ? s->mesh_cap * 2 : 64;
    SceneMesh *nm = realloc(s->meshes, nc * sizeof(SceneMesh));
    if (!nm)
      return -1;
    s->meshes = nm;
    s->mesh_cap = nc;
  }
  SceneMesh *m = &s->meshes[s->mesh_count];
  memset(m, 0, sizeof(*m));
  MopHipBufferDesc vd = {v, (size_t)nv * sizeof(MopHipVertex)};
  MopHipBufferDesc id = {idx, (size_t)ni * sizeof(uint32_t)};
  m->vb = mop_hip_buffer_create(s->dev, &vd);
  m->ib = mop_hip_buffer_create(s->dev, &id);
  if (!m->vb || !m->ib) {
    mop_hip_buffer_destroy(s->dev, m->vb);
    mop_hip_buffer_destroy(s->dev, m->ib);
    return -1;
  }
  m->vertex_count = nv;
  m->index_count = ni;
  m->object_id = object_id;
  if (model16)
    memcpy(m->model.d, model16, sizeof(m->model.d));
  else
    m->model = m4_identity();
  m->has_material = has_material;
  m->metallic = metallic;
  m->roughness = roughness;
  m->blend = MOP_HIP_BLEND_OPAQUE; /* viewport.c mesh defaults */
  m->opacity = 1.0f;
  m->aabb_min = m->aabb_max = v[0].position;
  for (uint32_t i = 1; i < nv; i++) {
    MopHipVec3 p = v[i].position;
    if (p.x < m->aabb_min.x) m->aabb_min.x = p.x;
    if (p.y < m->aabb_min.y) m->aabb_min.y = p.y;
    if (p.z < m->aabb_min.z) m->aabb_min.z = p.z;
    if (p.x > m->aabb_max.x) m->aabb_max.x = p.x;
    if (p.y > m->aabb_max.y) m->aabb_max.y = p.y;
    if (p.z > m->aabb_max.z) m->aabb_max.z = p.z;
  }
  s->mesh_count++;
  s->shadow_valid = 0;
  return (int)s->mesh_count - 1;
}

int mop_hip_scene_update_mesh_vertices(MopHipScene *s, uint32_t mesh, const MopHipVertex *v,
                                       uint32_t nv) {
  if (!s || !v || mesh >= s->mesh_count || nv != s->meshes[mesh].vertex_count)
    return -1;
  SceneMesh *m = &s->meshes[mesh];
  mop_hip_buffer_update(s->dev, m->vb, v, 0, (size_t)nv * sizeof(MopHipVertex));
  m->aabb_min = m->aabb_max = v[0].position;
  for (uint32_t i = 1; i < nv; i++) {
    MopHipVec3 p = v[i].position;
    if (p.x < m->aabb_min.x) m->aabb_min.x = p.x;
    if (p.y < m->aabb_min.y) m->aabb_min.y = p.y;
    if (p.z < m->aabb_min.z) m->aabb_min.z = p.z;
    if (p.x > m->aabb_max.x) m->aabb_max.x = p.x;
    if (p.y > m->aabb_max.y) m->aabb_max.y = p.y;
    if (p.z > m->aabb_max.z) m->aabb_max.z = p.z;
  }
  s->shadow_valid = 0;
  return 0;
}

int mop_hip_scene_add_loaded_mesh(MopHipScene *s, MopHipLoadedMesh *lm, uint32_t object_id,
                                  const float *model16, int has_material, float metallic,
                                  float roughness) {
  if (!s || !lm || !lm->vertex_buffer || !lm->index_buffer || lm->vertex_count == 0 ||
      lm->index_count < 3)
    return -1;
  if (s->mesh_count == s->mesh_cap) {
    uint32_t nc = s->mesh_cap ? s->mesh_cap * 2 : 64;
    SceneMesh *nm = realloc(s->meshes, nc * sizeof(SceneMesh));
    if (!nm)
      return -1;
    s->meshes = nm;
    s->mesh_cap = nc;
  }
  SceneMesh *m = &s->meshes[s->mesh_count];
  memset(m, 0, sizeof(*m));
  m->vb = lm->vertex_buffer;
  m->ib = lm->index_buffer;
  m->vertex_count = lm->vertex_count;
  m->index_count = lm->index_count;
  m->object_id = object_id;
  if (model16)
    memcpy(m->model.d, model16, sizeof(m->model.d));
  else
    m->model = m4_identity();
  m->has_material = has_material;
  m->metallic = metallic;
  m->roughness = roughness;
  m->blend = MOP_HIP_BLEND_OPAQUE; /* viewport.c mesh defaults */
  m->opacity = 1.0f;
  m->aabb_min = lm->aabb_min;
  m->aabb_max = lm->aabb_max;
  memset(lm, 0, sizeof(*lm));
  s->mesh_count++;
  s->shadow_valid = 0;
  return (int)s->mesh_count - 1;
}

int mop_hip_scene_add_light(MopHipScene *s, const MopHipLight *l) {
  if (!s || !l)
    return -1;
  if (s->light_count == s->light_cap) {
    uint32_t nc = s->light_cap ? s->light_cap * 2 : 8;
    MopHipLight *nl = realloc(s->lights, nc * sizeof(MopHipLight));
    if (!nl)
      return -1;
    s->lights = nl;
    s->light_cap = nc;
  }
  s->lights[s->light_count] = *l;
  s->lights[s->light_count].viewport = NULL;
  s->light_count++;
  s->shadow_valid = 0;
  return (int)s->light_count - 1;
}

/* Auto-exposure from the log-average luminance (environment.c:806-835). */
static float auto_exposure(const float *rgba, int w, int h, float cur) {
  size_t px = (size_t)w * h;
  double log_sum = 0.0;
  size_t count = 0;
  for (size_t i = 0; i < px; i++) {
    float lum = 0.2126f * rgba[i * 4] + 0.7152f * rgba[i * 4 + 1] + 0.0722f * rgba[i * 4 + 2];
    if (lum > 1e-6f) {
      log_sum += log(lum);
      count++;
    }
  }
  if (!count)
    return cur;
  float avg = (float)exp(log_sum / (double)count);
  float e = 0.18f / avg;
  if (e < 0.01f)
    e = 0.01f;
  if (e > 16.0f)
    e = 16.0f;
  return e;
}

int mop_hip_scene_set_environment(MopHipScene *s, const float *rgba, int w, int h,
                                  float rotation, float intensity) {
  if (!s)
    return -1;
  mop_hip_ibl_destroy(s->dev, s->ibl);
  s->ibl = NULL;
  if (!rgba)
    return 0;
  s->ibl = mop_hip_ibl_create(s->dev, rgba, w, h);
  if (!s->ibl)
    return -2;
  s->env_rotation = rotation;
  s->env_intensity = intensity > 0.0f ? intensity : 1.0f;
  s->exposure = auto_exposure(rgba, w, h, s->exposure);
  return 0;
}

void mop_hip_scene_set_shadow_size(MopHipScene *s, int size) {
  if (s && size > 0 && size != s->shadow_size) {
    s->shadow_size = size;
    s->shadow_valid = 0;
  }
}
void mop_hip_scene_set_shadow_cache(MopHipScene *s, int enable) {
  if (s)
    s->shadow_cache = enable;
}
void mop_hip_scene_set_ambient(MopHipScene *s, float a) {
  if (s)
    s->ambient = a;
}
void mop_hip_scene_set_exposure(MopHipScene *s, float e) {
  if (s)
    s->exposure = e;
}
float mop_hip_scene_exposure(const MopHipScene *s) { return s ? s->exposure : 1.0f; }
void mop_hip_scene_set_clear_color(MopHipScene *s, MopHipColor c) {
  if (s)
    s->clear = c;
}
int mop_hip_scene_set_mesh_blend(MopHipScene *s, uint32_t mesh, uint32_t blend_mode,
                                 float opacity) {
  if (!s || mesh >= s->mesh_count || blend_mode > MOP_HIP_BLEND_MULTIPLY)
    return -1;
  s->meshes[mesh].blend = blend_mode;
  s->meshes[mesh].opacity = opacity < 0.0f ? 0.0f : (opacity > 1.0f ? 1.0f : opacity);
  s->shadow_valid = 0; /* the shadow pass draws opaque meshes only */
  return 0;
}

void mop_hip_scene_set_post_effects(MopHipScene *s, uint32_t effects) {
  if (s)
    s->post_effects = effects;
}
uint32_t mop_hip_scene_triangles(const MopHipScene *s) { return s ? s->last_triangles : 0; }

/* Light view-projection of pass_shadow (viewport.c:2291-2415); returns 0
 * when no shadow pass runs. */
static int shadow_light_vp(const MopHipScene *s, MopHipMat4 *out) {
  int li = -1;
  for (uint32_t i = 0; i < s->light_count; i++)
    if (s->lights[i].active && s->lights[i].type == MOP_HIP_LIGHT_DIRECTIONAL &&
        s->lights[i].cast_shadows) {
      li = (int)i;
      break;
    }
  if (li < 0)
    return 0;
  float mn[3] = {1e9f, 1e9f, 1e9f}, mx[3] = {-1e9f, -1e9f, -1e9f};
  int any = 0;
  for (uint32_t i = 0; i < s->mesh_count; i++) {
    if (s->meshes[i].blend != MOP_HIP_BLEND_OPAQUE)
      continue;
    any = 1;
    const float *d = s->meshes[i].model.d;
    float cx = d[12], cy = d[13], cz = d[14];
    float sx = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float sy = sqrtf(d[4] * d[4] + d[5] * d[5] + d[6] * d[6]);
    float sz = sqrtf(d[8] * d[8] + d[9] * d[9] + d[10] * d[10]);
    float r = sx > sy ? sx : sy;
    if (sz > r)
      r = sz;
    r *= 1.5f;
    if (cx - r < mn[0]) mn[0] = cx - r;
    if (cy - r < mn[1]) mn[1] = cy - r;
    if (cz - r < mn[2]) mn[2] = cz - r;
    if (cx + r > mx[0]) mx[0] = cx + r;
    if (cy + r > mx[1]) mx[1] = cy + r;
    if (cz + r > mx[2]) mx[2] = cz + r;
  }
  if (!any)
    return 0;
  MopHipVec3 nd = v3_norm(s->lights[li].direction);
  float ccx = (mn[0] + mx[0]) * 0.5f, ccy = (mn[1] + mx[1]) * 0.5f, ccz = (mn[2] + mx[2]) * 0.5f;
  float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
  float radius = sqrtf(dx * dx + dy * dy + dz * dz) * 0.5f;
  if (radius < 1.0f)
    radius = 1.0f;
  MopHipVec3 eye = v3(ccx - nd.x * radius * 2.0f, ccy - nd.y * radius * 2.0f,
                      ccz - nd.z * radius * 2.0f);
  MopHipVec3 up = v3(0, 1, 0);
  if (fabsf(v3_dot(nd, up)) > 0.99f)
    up = v3(1, 0, 0);
  MopHipMat4 view = m4_look_at(eye, v3(ccx, ccy, ccz), up);
  float os = radius * 1.2f;
  MopHipMat4 proj = m4_ortho(-os, os, -os, os, 0.1f, radius * 4.0f);
  *out = m4_mul(proj, view);
  return 1;
}

/* View matrices as the core derives them after mop_viewport_set_camera:
 * orbit parameters (viewport.c:1081-1091), then the per-frame eye/up
 * (camera.c:35-63) and mop_viewport_set_camera_orbit (viewport.c:1095-1118). */
static void camera_matrices(const MopHipCamera *cam, int w, int h, MopHipMat4 *view,
                            MopHipMat4 *proj, MopHipVec3 *eye_out) {
  MopHipVec3 eye = cam->eye, up = cam->up;
  MopHipVec3 off = v3_sub(cam->eye, cam->target);
  float dist = v3_len(off);
  if (dist > 1e-6f) {
    float yaw = atan2f(off.x, off.z);
    float pitch = asinf(off.y / dist);
    float cp = cosf(pitch);
    eye = v3(cam->target.x + dist * cp * sinf(yaw), cam->target.y + dist * sinf(pitch),
             cam->target.z + dist * cp * cosf(yaw));
    float sp = sinf(pitch), sy = sinf(yaw), cy = cosf(yaw);
    up = v3(-sp * sy, cp, -sp * cy);
  }
  float fov = cam->fov_deg * (3.14159265358979323846f / 180.0f);
  *view = m4_look_at(eye, cam->target, up);
  *proj = m4_perspective(fov, (float)w / (float)h, cam->znear, cam->zfar);
  *eye_out = eye;
}

/* Conservative mesh-vs-frustum test (pass_scene_opaque, viewport.c:2450-2464):
 * world AABB of the local box's 8 corners against the 6 planes of P*V. */
static int mesh_outside(const SceneMesh *m, const MopHipMat4 *vpm) {
  MopHipVec3 c[8];
  for (int k = 0; k < 8; k++) {
    MopHipVec3 p = v3((k & 1) ? m->aabb_max.x : m->aabb_min.x,
                      (k & 2) ? m->aabb_max.y : m->aabb_min.y,
                      (k & 4) ? m->aabb_max.z : m->aabb_min.z);
    const float *d = m->model.d;
    c[k] = v3(d[0] * p.x + d[4] * p.y + d[8] * p.z + d[12],
              d[1] * p.x + d[5] * p.y + d[9] * p.z + d[13],
              d[2] * p.x + d[6] * p.y + d[10] * p.z + d[14]);
  }
  MopHipVec3 mn = c[0], mx = c[0];
  for (int k = 1; k < 8; k++) {
    if (c[k].x < mn.x) mn.x = c[k].x;
    if (c[k].y < mn.y) mn.y = c[k].y;
    if (c[k].z < mn.z) mn.z = c[k].z;
    if (c[k].x > mx.x) mx.x = c[k].x;
    if (c[k].y > mx.y) mx.y = c[k].y;
    if (c[k].z > mx.z) mx.z = c[k].z;
  }
  const float *R = vpm->d;
#define RR(r, col) R[(col)*4 + (r)]
  for (int i = 0; i < 6; i++) {
    int row = i / 2;
    float sg = (i & 1) ? -1.0f : 1.0f;
    float nx = RR(3, 0) + sg * RR(row, 0), ny = RR(3, 1) + sg * RR(row, 1);
    float nz = RR(3, 2) + sg * RR(row, 2), dd = RR(3, 3) + sg * RR(row, 3);
    MopHipVec3 pv = v3(nx >= 0 ? mx.x : mn.x, ny >= 0 ? mx.y : mn.y, nz >= 0 ? mx.z : mn.z);
    /* a small slack keeps the test conservative under rounding */
    float sc = sqrtf(nx * nx + ny * ny + nz * nz);
    if (nx * pv.x + ny * pv.y + nz * pv.z + dd < -1e-4f * (sc > 1.0f ? sc : 1.0f))
      return 1;
  }
#undef RR
  return 0;
}

/* EMIT_DRAW (viewport.c:2015-2088) of one scene mesh; returns its triangles. */
static uint32_t emit_draw(MopHipScene *s, MopHipFramebuffer *fb, const SceneMesh *m,
                          const MopHipMat4 *view, const MopHipMat4 *proj, MopHipVec3 eye,
                          bool cast) {
  MopHipDrawCall c;
  memset(&c, 0, sizeof(c));
  c.vertex_buffer = m->vb;
  c.index_buffer = m->ib;
  c.vertex_count = m->vertex_count;
  c.index_count = m->index_count;
  c.object_id = m->object_id;
  c.model = m->model;
  c.view = *view;
  c.projection = *proj;
  c.mvp = m4_mul(*proj, m4_mul(*view, m->model));
  c.base_color = (MopHipColor){1, 1, 1, 1};
  c.opacity = m->opacity;
  c.light_dir = v3(0.3f, 1.0f, 0.5f);
  c.ambient = s->ambient;
  c.shading_mode = MOP_HIP_SHADING_SMOOTH;
  c.depth_test = true;
  c.depth_write = true;
  c.backface_cull = true;
  c.blend_mode = (int32_t)m->blend;
  c.metallic = m->has_material ? m->metallic : 0.0f;
  c.roughness = m->has_material ? m->roughness : 0.5f;
  c.lights = s->lights;
  c.light_count = s->light_count;
  c.cam_eye = eye;
  c.cast_shadows = cast;
  mop_hip_draw(s->dev, fb, &c);
  return m->index_count / 3;
}

uint32_t mop_hip_scene_record(MopHipScene *s, MopHipFramebuffer *fb, const MopHipCamera *cam,
                              int width, int height) {
  if (!s || !fb || !cam)
    return 0;
  MopHipDevice *dev = s->dev;
  MopHipMat4 view, proj;
  MopHipVec3 eye;
  camera_matrices(cam, width, height, &view, &proj, &eye);
  MopHipMat4 vpm = m4_mul(proj, view);

  mop_hip_frame_begin(dev, fb, s->clear);
  mop_hip_set_exposure(dev, s->exposure);

  /* shadow pass */
  int shadowed = 0;
  MopHipMat4 lvp;
  if (shadow_light_vp(s, &lvp)) {
    if (!s->shadow) {
      s->shadow = mop_hip_shadow_map_create(dev, s->shadow_size);
      s->shadow_valid = 0;
    } else if (s->shadow_valid && memcmp(&lvp, &s->shadow_vp, sizeof(lvp)) != 0) {
      s->shadow_valid = 0;
    }
    if (s->shadow) {
      if (s->shadow_cache && s->shadow_valid) {
        mop_hip_shadow_bind(dev, s->shadow, &lvp);
      } else {
        mop_hip_shadow_begin(dev, s->shadow);
        for (uint32_t i = 0; i < s->mesh_count; i++) {
          const SceneMesh *m = &s->meshes[i];
          if (m->blend != MOP_HIP_BLEND_OPAQUE)
            continue;
          mop_hip_shadow_draw(dev, m->vb, m->vertex_count, m->ib, m->index_count, &m->model,
                              &lvp);
        }
        mop_hip_shadow_end(dev, &lvp);
        s->shadow_valid = 1;
        s->shadow_vp = lvp;
      }
      shadowed = 1;
    }
  }
  if (s->ibl)
    mop_hip_ibl_bind(dev, s->ibl, s->env_rotation, s->env_intensity);

  /* opaque pass: EMIT_DRAW per visible mesh */
  bool cast = false;
  for (uint32_t i = 0; i < s->light_count; i++)
    if (s->lights[i].active && s->lights[i].type == MOP_HIP_LIGHT_DIRECTIONAL &&
        s->lights[i].cast_shadows)
      cast = true;
  uint32_t tris = 0;
  for (uint32_t i = 0; i < s->mesh_count; i++) {
    const SceneMesh *m = &s->meshes[i];
    if (m->blend != MOP_HIP_BLEND_OPAQUE || mesh_outside(m, &vpm))
      continue;
    tris += emit_draw(s, fb, m, &view, &proj, eye, cast);
  }
  /* transparent pass (pass_scene_transparent, viewport.c:2470-2546): the
   * visible blended meshes, farthest first by the squared distance of their
   * model translation to the eye; the insertion sort keeps submission order
   * among equal distances */
  uint32_t nt = 0;
  for (uint32_t i = 0; i < s->mesh_count; i++)
    if (s->meshes[i].blend != MOP_HIP_BLEND_OPAQUE)
      nt++;
  if (nt) {
    uint32_t *ord = malloc(nt * sizeof(uint32_t));
    float *dist = malloc(nt * sizeof(float));
    if (ord && dist) {
      uint32_t k = 0;
      for (uint32_t i = 0; i < s->mesh_count; i++) {
        const SceneMesh *m = &s->meshes[i];
        if (m->blend == MOP_HIP_BLEND_OPAQUE || mesh_outside(m, &vpm))
          continue;
        float dx = m->model.d[12] - eye.x, dy = m->model.d[13] - eye.y,
              dz = m->model.d[14] - eye.z;
        ord[k] = i;
        dist[k] = dx * dx + dy * dy + dz * dz;
        k++;
      }
      for (uint32_t j = 1; j < k; j++) {
        uint32_t oi = ord[j];
        float od = dist[j];
        int q = (int)j - 1;
        while (q >= 0 && dist[q] < od) {
          ord[q + 1] = ord[q];
          dist[q + 1] = dist[q];
          q--;
        }
        ord[q + 1] = oi;
        dist[q + 1] = od;
      }
      for (uint32_t j = 0; j < k; j++)
        tris += emit_draw(s, fb, &s->meshes[ord[j]], &view, &proj, eye, cast);
    }
    free(ord);
    free(dist);
  }
  /* rg_frame_end + rg_cleanup: resolve with the exposure, then post */
  mop_hip_frame_request_post(dev, fb, s->exposure, s->post_effects, NULL);
  if (shadowed)
    mop_hip_shadow_clear(dev);
  if (s->ibl)
    mop_hip_ibl_clear(dev);
  s->last_triangles = tris;
  return tris;
}

uint32_t mop_hip_scene_render(MopHipScene *s, MopHipFramebuffer *fb, const MopHipCamera *cam,
                              int width, int height) {
  uint32_t tris = mop_hip_scene_record(s, fb, cam, width, height);
  if (s && fb)
    mop_hip_frames_execute(s->dev, &fb, 1);
  return tris;
}
