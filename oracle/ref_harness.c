/*
 * ref_harness.c -- TEST INFRASTRUCTURE, never part of the product.
 *
 * Drives the *reference* MOP viewport core (compiled from the reference's own
 * sources into oracle/_ref/libmopref.a) through its public API, so the parity
 * tests and the CPU baseline use exactly the reference behaviour:
 *
 *   mode 0  reference CPU backend, multithreaded as shipped (timing baseline;
 *           cpu_backend.c:55-75 sizes its pool to nproc-1 workers)
 *   mode 1  reference CPU backend forced single-threaded: the wrapped
 *           mop_sw_threadpool_create returns NULL, so cpu_draw always takes
 *           the submission-order loop (cpu_backend.c:547,588-620).  Goldens.
 *   mode 2  the same unmodified viewport core driving the HIP backend.  The
 *           harness hands the HIP RHI table out for MOP_BACKEND_CPU and routes
 *           the core's CPU-gated software calls (shadow pass, IBL state, HDR
 *           resolve, post-process: viewport.c:2433,2699,2729,2953-2959) to the
 *           HIP backend's hooks.  This is the drop-in exercised end to end
 *           with no edit of the reference sources.
 *
 * Interposition uses the linker's --wrap (see oracle/Makefile).
 */
#include "core/viewport_internal.h"
#include "rasterizer/rasterizer.h"
#include "rasterizer/rasterizer_mt.h"
#include "rhi/rhi.h"

#include <dlfcn.h>
#include <mop/mop.h>
#include <stddef.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* tinyexr (linked from the reference build) */
extern int SaveEXR(const float *data, const int width, const int height,
                   const int components, const int save_as_fp16,
                   const char *filename, const char **err);

typedef const MopRhiBackend *(*fn_backend)(void);

/* ---- mode + backend entry points (resolved with dlsym) ------------------ */

static int g_mode = 1;

/* One routed backend: the RHI table and the hooks standing in for the
 * core's CPU-gated software calls.  Slot 0 serves mode 2 (HIP backend,
 * prefix "mop_hip_"), slot 1 serves mode 3 (CPU restatement in
 * oracle/mop_oracle.c, prefix "mop_oracle_"). */
typedef struct Routed {
  const MopRhiBackend *rhi;
  void (*shadow_mesh)(MopRhiDevice *, const void *, uint32_t, const uint32_t *, uint32_t,
                      const MopMat4 *, const MopMat4 *, int, int);
  void (*shadow_set)(MopRhiDevice *, int, int, const MopMat4 *);
  void (*shadow_clear)(MopRhiDevice *);
  void (*ibl_set)(MopRhiDevice *, const float *, int, int, const float *, int, int, int,
                  const float *, int, float, float);
  void (*ibl_clear)(MopRhiDevice *);
  void (*hdr_resolve)(MopRhiDevice *, MopRhiFramebuffer *, float);
  void (*postprocess)(MopRhiDevice *, MopRhiFramebuffer *, uint32_t, const MopFogParams *);
  const float *(*read_hdr)(MopRhiDevice *, MopRhiFramebuffer *, int *, int *);
  /* optional: the SSAA downsample of mop_viewport_read_color on the backend */
  const uint8_t *(*read_color_ds)(MopRhiDevice *, MopRhiFramebuffer *, int, int *, int *);
  /* optional: the CPU skybox (viewport.c:2128-2180) on the backend, and the
   * table with frame_begin wrapped to issue it */
  void (*background)(MopRhiDevice *, MopRhiFramebuffer *, MopRhiTexture *, float, MopVec3,
                     MopVec3, MopVec3, float, float);
  void (*frame_begin)(MopRhiDevice *, MopRhiFramebuffer *, MopColor);
  /* optional: the object outline (overlay_builtin.c:1309-1400) on the backend */
  void (*outline)(MopRhiDevice *, MopRhiFramebuffer *, MopColor, float, float, const uint32_t *,
                  uint32_t);
  MopRhiBackend table;
} Routed;

static Routed g_routed[2];

static Routed *routed(void) { return (g_mode >= 2 && g_mode <= 3) ? &g_routed[g_mode - 2] : NULL; }

/* viewport whose frame is being rendered (for hooks without a vp argument) */
static MopViewport *g_cur_vp;
static int g_shadow_started;

/* Shadow-map size override (0: the core's fixed MOP_SHADOW_MAP_SIZE,
 * viewport.c:2286).  The core's pass_shadow still computes the light
 * view-projection and the caster list; the wrappers below render those
 * casters into a map of this size instead: in modes 0/1 the reference's own
 * mop_sw_shadow_render_mesh draws into the harness's map (cleared to 1.0
 * before the first caster, as pass_shadow clears its own, viewport.c:2365-2368)
 * and mop_sw_shadow_set points the shading at it; in modes 2/3 the routed
 * hooks get the size.  This is how the benchmarked 2048^2 and 4096^2 maps
 * are compared with the reference's arithmetic. */
static int g_shadow_size;
static MopSwFramebuffer g_shadow_fb;

void mh_set_shadow_size(int size) { g_shadow_size = size > 0 ? size : 0; }

static int override_map_ready(void) {
  if (g_shadow_fb.width != g_shadow_size || g_shadow_fb.height != g_shadow_size) {
    mop_sw_framebuffer_free(&g_shadow_fb);
    memset(&g_shadow_fb, 0, sizeof(g_shadow_fb));
    if (!mop_sw_framebuffer_alloc(&g_shadow_fb, g_shadow_size, g_shadow_size))
      return 0;
  }
  return 1;
}

static void override_map_clear(void) {
  size_t n = (size_t)g_shadow_size * (size_t)g_shadow_size;
  for (size_t p = 0; p < n; p++)
    g_shadow_fb.depth[p] = 1.0f;
}

int mh_set_mode(int mode) {
  if ((mode == 2 || mode == 3) && !g_routed[mode - 2].rhi)
    return -1;
  g_mode = mode;
  return 0;
}

static void *sym(void *h, const char *prefix, const char *name) {
  char buf[128];
  snprintf(buf, sizeof(buf), "%s%s", prefix, name);
  return dlsym(h, buf);
}

/* The viewport whose environment background the routed backend draws in
 * this frame (mh_render sets it; the core's own background pass is skipped
 * for it, since it would write the CPU backend's framebuffer layout). */
static MopViewport *g_bg_vp;

static void routed_frame_begin(MopRhiDevice *dev, MopRhiFramebuffer *fb, MopColor clear) {
  Routed *r = routed();
  if (!r)
    return;
  r->frame_begin(dev, fb, clear);
  MopViewport *vp = g_bg_vp;
  if (vp && vp->device == dev && vp->framebuffer == fb && vp->env_texture)
    r->background(dev, fb, vp->env_texture, vp->cam_fov_radians, vp->cam_eye, vp->cam_target,
                  vp->cam_up, vp->env_rotation, vp->env_intensity);
}

/* Loads a backend library for mode 2 (prefix "mop_hip_") or mode 3
 * (prefix "mop_oracle_"); its table comes from mop_rhi_backend_<table>(). */
int mh_load_backend(int mode, const char *so_path, const char *prefix, const char *table) {
  if (mode != 2 && mode != 3)
    return -4;
  void *h = dlopen(so_path, RTLD_NOW | RTLD_GLOBAL);
  if (!h) {
    fprintf(stderr, "mh_load_backend: %s\n", dlerror());
    return -1;
  }
  Routed *r = &g_routed[mode - 2];
  char tb[64];
  snprintf(tb, sizeof(tb), "mop_rhi_backend_%s", table);
  fn_backend fb = (fn_backend)dlsym(h, tb);
  r->shadow_mesh = sym(h, prefix, "shadow_render_mesh");
  r->shadow_set = sym(h, prefix, "shadow_set");
  r->shadow_clear = sym(h, prefix, "shadow_clear");
  r->ibl_set = sym(h, prefix, "ibl_set");
  r->ibl_clear = sym(h, prefix, "ibl_clear");
  r->hdr_resolve = sym(h, prefix, "hdr_resolve");
  r->postprocess = sym(h, prefix, "postprocess_apply");
  r->read_hdr = sym(h, prefix, "framebuffer_read_hdr");
  r->read_color_ds = sym(h, prefix, "framebuffer_read_color_downsampled");
  r->outline = sym(h, prefix, "draw_outline");
  if (!fb || !r->shadow_mesh || !r->shadow_set || !r->shadow_clear || !r->ibl_set ||
      !r->ibl_clear || !r->hdr_resolve || !r->postprocess || !r->read_hdr) {
    fprintf(stderr, "mh_load_backend: missing symbol in %s\n", so_path);
    return -2;
  }
  r->rhi = fb();
  if (!r->rhi)
    return -3;
  r->background = sym(h, prefix, "draw_background_hdri");
  if (r->background) {
    /* frame_begin followed by the environment background, which the core
     * would draw next (viewport.c:3211-3219) */
    r->table = *r->rhi;
    r->frame_begin = r->rhi->frame_begin;
    r->table.frame_begin = routed_frame_begin;
    r->rhi = &r->table;
  }
  return 0;
}

int mh_load_hip(const char *so_path) { return mh_load_backend(2, so_path, "mop_hip_", "hip"); }

/* ---- wrapped reference symbols ----------------------------------------- */

MopSwThreadPool *__real_mop_sw_threadpool_create(int num_threads);
/* Mode 1: no pool, so cpu_draw takes its single-threaded submission-order
 * path.  Otherwise the reference's own rule (cores - 1 workers plus the
 * calling thread, cpu_backend.c:60-71) applied to the cores this process may
 * run on: sysconf counts every online CPU, which oversubscribes a process
 * pinned to fewer cores. */
MopSwThreadPool *__wrap_mop_sw_threadpool_create(int num_threads) {
  if (g_mode == 1)
    return NULL;
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) {
    int n = CPU_COUNT(&set);
    int workers = n > 1 ? n - 1 : 1;
    if (workers < num_threads)
      num_threads = workers;
  }
  return __real_mop_sw_threadpool_create(num_threads);
}

const MopRhiBackend *__real_mop_rhi_get_backend(MopBackendType type);
const MopRhiBackend *__wrap_mop_rhi_get_backend(MopBackendType type) {
  Routed *r = routed();
  if (r && (type == MOP_BACKEND_CPU || type == MOP_BACKEND_AUTO))
    return r->rhi;
  return __real_mop_rhi_get_backend(type);
}

void __real_mop_sw_shadow_render_mesh(const MopVertex *, uint32_t,
                                      const uint32_t *, uint32_t, MopMat4,
                                      MopMat4, MopSwFramebuffer *);
void __wrap_mop_sw_shadow_render_mesh(const MopVertex *v, uint32_t vc,
                                      const uint32_t *idx, uint32_t ic,
                                      MopMat4 model, MopMat4 light_vp,
                                      MopSwFramebuffer *sfb) {
  Routed *r = routed();
  if (!r) {
    if (g_shadow_size && override_map_ready()) {
      if (!g_shadow_started)
        override_map_clear();
      g_shadow_started = 1;
      __real_mop_sw_shadow_render_mesh(v, vc, idx, ic, model, light_vp, &g_shadow_fb);
      return;
    }
    __real_mop_sw_shadow_render_mesh(v, vc, idx, ic, model, light_vp, sfb);
    return;
  }
  /* &vp->shadow_fb -> vp (pass_shadow passes the viewport's own member) */
  MopViewport *vp = (MopViewport *)((char *)sfb - offsetof(MopViewport, shadow_fb));
  r->shadow_mesh(vp->device, v, vc, idx, ic, &model, &light_vp,
                 g_shadow_size ? g_shadow_size : sfb->width, g_shadow_started ? 0 : 1);
  g_shadow_started = 1;
}

void __real_mop_sw_shadow_set(const float *, int, int, MopMat4);
void __wrap_mop_sw_shadow_set(const float *depth, int w, int h, MopMat4 lvp) {
  Routed *r = routed();
  if (!r) {
    if (g_shadow_size && override_map_ready()) {
      if (!g_shadow_started)
        override_map_clear(); /* a pass with no caster */
      g_shadow_started = 0;
      __real_mop_sw_shadow_set(g_shadow_fb.depth, g_shadow_size, g_shadow_size, lvp);
      return;
    }
    __real_mop_sw_shadow_set(depth, w, h, lvp);
    return;
  }
  g_shadow_started = 0;
  if (g_shadow_size)
    w = h = g_shadow_size;
  if (g_cur_vp)
    r->shadow_set(g_cur_vp->device, w, h, &lvp);
}

void __real_mop_sw_shadow_clear(void);
void __wrap_mop_sw_shadow_clear(void) {
  Routed *r = routed();
  g_shadow_started = 0;
  if (!r) {
    __real_mop_sw_shadow_clear();
    return;
  }
  if (g_cur_vp)
    r->shadow_clear(g_cur_vp->device);
}

void __real_mop_sw_ibl_set(const float *, int, int, const float *, int, int,
                           int, const float *, int, float, float);
void __wrap_mop_sw_ibl_set(const float *irr, int iw, int ih, const float *pf,
                           int pw, int ph, int pl, const float *lut, int ls,
                           float rot, float inten) {
  Routed *r = routed();
  if (!r) {
    __real_mop_sw_ibl_set(irr, iw, ih, pf, pw, ph, pl, lut, ls, rot, inten);
    return;
  }
  if (g_cur_vp)
    r->ibl_set(g_cur_vp->device, irr, iw, ih, pf, pw, ph, pl, lut, ls, rot,
              inten);
}

void __real_mop_sw_ibl_clear(void);
void __wrap_mop_sw_ibl_clear(void) {
  Routed *r = routed();
  if (!r) {
    __real_mop_sw_ibl_clear();
    return;
  }
  if (g_cur_vp)
    r->ibl_clear(g_cur_vp->device);
}

void __real_mop_sw_hdr_resolve(MopSwFramebuffer *fb, float exposure);
void __wrap_mop_sw_hdr_resolve(MopSwFramebuffer *fb, float exposure) {
  Routed *r = routed();
  if (!r) {
    __real_mop_sw_hdr_resolve(fb, exposure);
    return;
  }
  /* the core passes vp->framebuffer cast to MopSwFramebuffer* */
  r->hdr_resolve(g_cur_vp->device, (MopRhiFramebuffer *)fb, exposure);
}

void __real_mop_postprocess_apply(MopSwFramebuffer *, uint32_t,
                                  const MopFogParams *);
void __wrap_mop_postprocess_apply(MopSwFramebuffer *fb, uint32_t effects,
                                  const MopFogParams *fog) {
  Routed *r = routed();
  if (!r) {
    __real_mop_postprocess_apply(fb, effects, fog);
    return;
  }
  r->postprocess(g_cur_vp->device, (MopRhiFramebuffer *)fb, effects, fog);
}

/* mop_viewport_read_color (viewport.c:3364-3418): with a routed backend that
 * has its own SSAA downsample, the box filter runs there (as the core would
 * call it for that backend) instead of on the host. */
const uint8_t *__real_mop_viewport_read_color(MopViewport *, int *, int *);
const uint8_t *__wrap_mop_viewport_read_color(MopViewport *vp, int *w, int *h) {
  Routed *r = routed();
  if (!r || !r->read_color_ds || !vp || vp->ssaa_factor <= 1)
    return __real_mop_viewport_read_color(vp, w, h);
  return r->read_color_ds(vp->device, vp->framebuffer, vp->ssaa_factor, w, h);
}

/* mop_overlay_builtin_outline (overlay_builtin.c:1309-1400), called after the
 * HDR resolve and before the post effects (viewport.c:2733-2739): with a
 * routed backend that has its own outline, the viewport's theme and
 * selection go to it instead of the host loop over the read-back image. */
void __real_mop_overlay_builtin_outline(MopViewport *vp, void *user_data);
void __wrap_mop_overlay_builtin_outline(MopViewport *vp, void *user_data) {
  Routed *r = routed();
  if (!r || !r->outline || !vp) {
    __real_mop_overlay_builtin_outline(vp, user_data);
    return;
  }
  r->outline(vp->device, vp->framebuffer, vp->theme.accent, vp->theme.outline_opacity_selected,
             vp->theme.outline_opacity_unselected, vp->selected_ids, vp->selected_count);
}

/* ---- scene construction through the public API ------------------------- */

MopViewport *mh_viewport_ssaa(int w, int h, int ssaa) {
  MopViewportDesc d = {.width = w,
                       .height = h,
                       .backend = MOP_BACKEND_CPU,
                       .ssaa_factor = ssaa};
  MopViewport *vp = mop_viewport_create(&d);
  if (!vp)
    return NULL;
  mop_viewport_set_chrome(vp, false);
  mop_viewport_set_overlay_enabled(vp, MOP_OVERLAY_OUTLINE, false);
  mop_viewport_clear_lights(vp);
  return vp;
}

MopViewport *mh_viewport(int w, int h) { return mh_viewport_ssaa(w, h, 1); }

void mh_destroy(MopViewport *vp) {
  if (g_cur_vp == vp)
    g_cur_vp = NULL;
  mop_viewport_destroy(vp);
}

/* model: 16 floats column-major (NULL = identity); material sets metallic /
 * roughness through mop_mesh_set_material like the configs need. */
MopMesh *mh_add_mesh(MopViewport *vp, const MopVertex *verts, uint32_t nv,
                     const uint32_t *idx, uint32_t ni, uint32_t object_id,
                     const float *model, int has_material, float metallic,
                     float roughness, int blend_mode, float opacity) {
  MopMeshDesc md = {.vertices = verts,
                    .vertex_count = nv,
                    .indices = idx,
                    .index_count = ni,
                    .object_id = object_id};
  MopMesh *m = mop_viewport_add_mesh(vp, &md);
  if (!m)
    return NULL;
  if (model) {
    MopMat4 t;
    memcpy(t.d, model, sizeof(t.d));
    mop_mesh_set_transform(m, &t);
  }
  if (has_material) {
    MopMaterial mat = mop_material_default();
    mat.metallic = metallic;
    mat.roughness = roughness;
    mop_mesh_set_material(m, &mat);
  }
  if (blend_mode != 0)
    mop_mesh_set_blend_mode(m, (MopBlendMode)blend_mode);
  if (opacity < 1.0f) {
    mop_mesh_set_opacity(m, opacity);
    if (blend_mode == 0)
      mop_mesh_set_blend_mode(m, MOP_BLEND_OPAQUE);
  }
  return m;
}

/* Instanced mesh (one geometry, many transforms; pass_instanced,
 * viewport.c:2602-2646).  transforms: count x 16 floats column-major. */
int mh_add_instanced(MopViewport *vp, const MopVertex *verts, uint32_t nv,
                     const uint32_t *idx, uint32_t ni, uint32_t object_id,
                     const float *transforms, uint32_t count) {
  MopMeshDesc md = {.vertices = verts,
                    .vertex_count = nv,
                    .indices = idx,
                    .index_count = ni,
                    .object_id = object_id};
  return mop_viewport_add_instanced_mesh(vp, &md, (const MopMat4 *)transforms, count) ? 0
                                                                                       : -1;
}

int mh_add_light(MopViewport *vp, int type, const float *pos, const float *dir,
                 const float *color, float intensity, float range,
                 float inner_cos, float outer_cos, int cast_shadows) {
  MopLight l;
  memset(&l, 0, sizeof(l));
  l.type = (MopLightType)type;
  l.position = (MopVec3){pos[0], pos[1], pos[2]};
  l.direction = (MopVec3){dir[0], dir[1], dir[2]};
  l.color = (MopColor){color[0], color[1], color[2], color[3]};
  l.intensity = intensity;
  l.range = range;
  l.spot_inner_cos = inner_cos;
  l.spot_outer_cos = outer_cos;
  l.active = true;
  l.cast_shadows = cast_shadows != 0;
  return mop_viewport_add_light(vp, &l) ? 0 : -1;
}

void mh_set_camera(MopViewport *vp, const float *eye, const float *target,
                   const float *up, float fov_deg, float znear, float zfar) {
  mop_viewport_set_camera(vp, (MopVec3){eye[0], eye[1], eye[2]},
                          (MopVec3){target[0], target[1], target[2]},
                          (MopVec3){up[0], up[1], up[2]}, fov_deg, znear, zfar);
}

/* mop_viewport_set_environment_background (environment.c:868-874): the
 * environment map as the skybox behind the scene */
void mh_set_env_background(MopViewport *vp, int on) {
  mop_viewport_set_environment_background(vp, on != 0);
}

/* The object outline overlay (overlay.h:68) with the theme's accent and
 * outline opacities (theme.h:26-69), and the selection it outlines
 * (selection.h:48; additive toggles). */
void mh_set_outline(MopViewport *vp, int on, const float *accent, float alpha_selected,
                    float alpha_unselected) {
  MopTheme t = *mop_viewport_get_theme(vp);
  t.accent = (MopColor){accent[0], accent[1], accent[2], 1.0f};
  t.outline_opacity_selected = alpha_selected;
  t.outline_opacity_unselected = alpha_unselected;
  mop_viewport_set_theme(vp, &t);
  mop_viewport_set_overlay_enabled(vp, MOP_OVERLAY_OUTLINE, on != 0);
}

void mh_select(MopViewport *vp, uint32_t id, int additive) {
  mop_viewport_select_object(vp, id, additive != 0);
}

/* mop_scene_save (mop_scene.c:544-768) of the viewport's meshes, lights and
 * camera; flags MOP_SAVE_QUANTIZE for 18-byte quantized vertices */
int mh_save_scene(MopViewport *vp, const char *path, uint32_t flags) {
  return mop_scene_save(vp, path, flags);
}

/* mop_scene_load and its counts and camera (mop_scene.c:225-519); cam gets
 * eye, target, up, fov (radians), near, far.  0 on success. */
int mh_scene_file_info(const char *path, uint32_t *nmesh, uint32_t *nlight, float *cam) {
  MopSceneFile *f = mop_scene_load(path);
  if (!f)
    return -1;
  *nmesh = mop_scene_mesh_count(f);
  *nlight = mop_scene_light_count(f);
  MopVec3 e, t, u;
  float fov = 0, n = 0, fa = 0;
  if (cam && mop_scene_get_camera(f, &e, &t, &u, &fov, &n, &fa)) {
    float c[12] = {e.x, e.y, e.z, t.x, t.y, t.z, u.x, u.y, u.z, fov, n, fa};
    memcpy(cam, c, sizeof(c));
  }
  mop_scene_free(f);
  return 0;
}

/* mop_scene_get_mesh (mop_scene.c:397-487): the reference's host decode of
 * mesh idx.  Sizes go to *vc / *ic; the arrays are copied when the
 * capacities suffice (vertices as 48-byte MopVertex).  bbox6 gets the
 * header's box.  0 on success. */
int mh_scene_file_mesh(const char *path, uint32_t idx, void *verts, uint32_t vcap,
                       uint32_t *indices, uint32_t icap, uint32_t *vc, uint32_t *ic,
                       float *bbox6) {
  MopSceneFile *f = mop_scene_load(path);
  if (!f)
    return -1;
  MopLoadedMesh m;
  bool ok = mop_scene_get_mesh(f, idx, &m);
  mop_scene_free(f);
  if (!ok)
    return -2;
  *vc = m.vertex_count;
  *ic = m.index_count;
  if (verts && vcap >= m.vertex_count)
    memcpy(verts, m.vertices, (size_t)m.vertex_count * sizeof(MopVertex));
  if (indices && icap >= m.index_count)
    memcpy(indices, m.indices, (size_t)m.index_count * sizeof(uint32_t));
  if (bbox6) {
    float b[6] = {m.bbox_min.x, m.bbox_min.y, m.bbox_min.z,
                  m.bbox_max.x, m.bbox_max.y, m.bbox_max.z};
    memcpy(bbox6, b, sizeof(b));
  }
  mop_load_free(&m);
  return 0;
}

/* mop_viewport_resize (viewport.h:86): the serving checks render many
 * requests of one resident scene at different resolutions */
void mh_resize(MopViewport *vp, int w, int h) { mop_viewport_resize(vp, w, h); }

void mh_set_clear_color(MopViewport *vp, const float *c) {
  mop_viewport_set_clear_color(vp, (MopColor){c[0], c[1], c[2], c[3]});
}

void mh_set_post_effects(MopViewport *vp, uint32_t effects) {
  mop_viewport_set_post_effects(vp, effects);
}

/* mop_viewport_set_fog (postprocess.h:42): colour rgba, near, far */
void mh_set_fog(MopViewport *vp, const float *rgba, float near_dist, float far_dist) {
  MopFogParams f = {.color = {rgba[0], rgba[1], rgba[2], rgba[3]},
                    .near_dist = near_dist, .far_dist = far_dist};
  mop_viewport_set_fog(vp, &f);
}

void mh_set_ambient(MopViewport *vp, float a) { mop_viewport_set_ambient(vp, a); }

void mh_set_shading(MopViewport *vp, int mode) {
  mop_viewport_set_shading(vp, (MopShadingMode)mode);
}

void mh_set_exposure(MopViewport *vp, float e) { mop_viewport_set_exposure(vp, e); }

/* per-mesh shading override (mop_mesh_set_shading); mode < 0 keeps the
 * viewport's mode */
void mh_mesh_set_shading(MopMesh *m, int mode) {
  if (m && mode >= 0)
    mop_mesh_set_shading(m, (MopShadingMode)mode);
}

/* Mesh with a flexible vertex format (mop_viewport_add_mesh_ex): raw
 * interleaved bytes, nattr attributes (semantic, format, offset). */
MopMesh *mh_add_mesh_flex(MopViewport *vp, const void *data, uint32_t nv, uint32_t stride,
                          const uint32_t *idx, uint32_t ni, uint32_t object_id,
                          uint32_t nattr, const int *sem, const int *fmt, const uint32_t *off,
                          const float *model) {
  MopVertexFormat vf;
  memset(&vf, 0, sizeof(vf));
  if (nattr > MOP_MAX_VERTEX_ATTRIBS)
    return NULL;
  for (uint32_t i = 0; i < nattr; i++) {
    vf.attribs[i].semantic = (MopAttribSemantic)sem[i];
    vf.attribs[i].format = (MopAttribFormat)fmt[i];
    vf.attribs[i].offset = off[i];
  }
  vf.attrib_count = nattr;
  vf.stride = stride;
  MopMeshDescEx md = {.vertex_data = data,
                      .vertex_count = nv,
                      .indices = idx,
                      .index_count = ni,
                      .object_id = object_id,
                      .vertex_format = &vf};
  MopMesh *m = mop_viewport_add_mesh_ex(vp, &md);
  if (m && model) {
    MopMat4 t;
    memcpy(t.d, model, sizeof(t.d));
    mop_mesh_set_transform(m, &t);
  }
  return m;
}

void mh_mesh_set_bones(MopViewport *vp, MopMesh *m, const float *mats, uint32_t n) {
  mop_mesh_set_bone_matrices(m, vp, (const MopMat4 *)mats, n);
}

void mh_mesh_set_morph(MopViewport *vp, MopMesh *m, const float *targets,
                       const float *weights, uint32_t n) {
  mop_mesh_set_morph_targets(m, vp, targets, weights, n);
}

/* The mesh's current vertex bytes as the backend returns them (after a
 * render these are the core's skinned / morphed vertices) and its bind-pose
 * snapshot. */
int mh_mesh_vertex_bytes(MopViewport *vp, MopMesh *m, void *out, size_t bytes) {
  const void *p = vp->rhi->buffer_read(m->vertex_buffer);
  if (!p)
    return -1;
  memcpy(out, p, bytes);
  return 0;
}

const void *mh_mesh_bind_pose(MopMesh *m) { return m->bind_pose_data; }

void mh_set_render_mode(MopViewport *vp, int mode) {
  mop_viewport_set_render_mode(vp, (MopRenderMode)mode);
}

void mh_set_light_dir(MopViewport *vp, const float *d) {
  mop_viewport_set_light_dir(vp, (MopVec3){d[0], d[1], d[2]});
}

/* Writes the RGBA32F equirect to `path` as float EXR (lossless) and loads it
 * as the viewport environment (MOP_ENV_HDRI, environment.c:730-846). */
int mh_set_environment(MopViewport *vp, const float *rgba, int w, int h,
                       const char *path, float rotation, float intensity) {
  const char *err = NULL;
  if (SaveEXR(rgba, w, h, 4, 0, path, &err) != 0) {
    fprintf(stderr, "SaveEXR failed: %s\n", err ? err : "?");
    return -1;
  }
  MopEnvironmentDesc d = {.type = MOP_ENV_HDRI,
                          .hdr_path = path,
                          .rotation = rotation,
                          .intensity = intensity};
  return mop_viewport_set_environment(vp, &d) ? 0 : -2;
}

float mh_get_exposure(MopViewport *vp) { return vp->exposure; }

/* IBL tables as the core holds them (inputs to mop_sw_ibl_set). */
int mh_get_ibl(MopViewport *vp, const float **irr, int *iw, int *ih,
               const float **pf, int *pw, int *ph, int *pl, const float **lut) {
  if (!vp->env_irradiance_data)
    return -1;
  *irr = vp->env_irradiance_data;
  *iw = vp->env_irradiance_w;
  *ih = vp->env_irradiance_h;
  *pf = vp->env_prefiltered_data;
  *pw = vp->env_prefiltered_w;
  *ph = vp->env_prefiltered_h;
  *pl = vp->env_prefiltered_levels;
  *lut = vp->env_brdf_lut_data;
  return 0;
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* One frame: mop_viewport_render + mop_viewport_read_color (the frame path the
 * CPU baseline times).  Returns wall milliseconds. */
double mh_render(MopViewport *vp) {
  g_cur_vp = vp;
  Routed *r = routed();
  bool bg = r && r->background && vp->show_env_background && vp->env_hdr_data &&
            (vp->env_type == MOP_ENV_HDRI || vp->env_type == MOP_ENV_PROCEDURAL_SKY);
  if (bg) {
    vp->show_env_background = false;
    g_bg_vp = vp;
  }
  double t0 = now_ms();
  mop_viewport_render(vp);
  if (bg) {
    vp->show_env_background = true;
    g_bg_vp = NULL;
  }
  int w, h;
  (void)mop_viewport_read_color(vp, &w, &h);
  return now_ms() - t0;
}

uint32_t mh_triangle_count(MopViewport *vp) {
  return mop_viewport_get_stats(vp).triangle_count;
}

/* Full-buffer readbacks through the RHI table the core holds. */
int mh_read(MopViewport *vp, uint8_t *color, float *depth, uint32_t *ids,
            float *hdr) {
  int w = 0, h = 0;
  size_t n;
  if (color) {
    const uint8_t *c = mop_viewport_read_color(vp, &w, &h);
    if (!c)
      return -1;
    memcpy(color, c, (size_t)w * h * 4);
  }
  if (depth) {
    const float *d = vp->rhi->framebuffer_read_depth(vp->device, vp->framebuffer,
                                                      &w, &h);
    if (!d)
      return -2;
    memcpy(depth, d, (size_t)w * h * sizeof(float));
  }
  if (ids) {
    const uint32_t *o = vp->rhi->framebuffer_read_object_id(
        vp->device, vp->framebuffer, &w, &h);
    if (!o)
      return -3;
    memcpy(ids, o, (size_t)w * h * sizeof(uint32_t));
  }
  if (hdr) {
    const float *src;
    Routed *r = routed();
    if (r) {
      src = r->read_hdr(vp->device, vp->framebuffer, &w, &h);
    } else {
      MopSwFramebuffer *sw = (MopSwFramebuffer *)vp->framebuffer;
      src = sw->color_hdr;
      w = sw->width;
      h = sw->height;
    }
    if (!src)
      return -4;
    n = (size_t)w * h * 4;
    memcpy(hdr, src, n * sizeof(float));
  }
  return 0;
}

int mh_pick(MopViewport *vp, int x, int y, uint32_t *id, float *depth) {
  MopPickResult r = mop_viewport_pick(vp, x, y);
  *id = r.object_id;
  *depth = r.depth;
  return r.hit ? 1 : 0;
}
