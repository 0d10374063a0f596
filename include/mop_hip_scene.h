/*
 * mop_hip_scene.h -- frame driver over the HIP backend (C11, no counterpart
 * in the reference's RHI): resident scenes rendered with the pass order and
 * draw arguments of mop_viewport_render (src/core/viewport.c:2993-3317).
 * Used by the benchmark and the batched / serving forms, where many frames
 * of many scenes are in flight and the synchronous viewport core cannot be.
 */
#ifndef MOP_HIP_SCENE_H
#define MOP_HIP_SCENE_H

#include "mop_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct MopHipScene MopHipScene;

typedef struct MopHipCamera {
  MopHipVec3 eye, target, up;
  float fov_deg, znear, zfar;
} MopHipCamera;

MopHipScene *mop_hip_scene_create(MopHipDevice *dev);
void mop_hip_scene_destroy(MopHipScene *s);
/* Copies the arrays into HBM (mop_viewport_add_mesh + set_transform +
 * set_material); model16 NULL = identity. Returns the mesh index or < 0. */
int mop_hip_scene_add_mesh(MopHipScene *s, const MopHipVertex *v, uint32_t nv,
                           const uint32_t *idx, uint32_t ni, uint32_t object_id,
                           const float *model16, int has_material, float metallic,
                           float roughness);
/* Rewrite the nv vertices of mesh `mesh` in place (nv must equal its vertex
 * count): the in-place path of mop_mesh_update_geometry
 * (viewport.c:1430-1452) through the backend's buffer_update.  Returns 0 or
 * -1. */
int mop_hip_scene_update_mesh_vertices(MopHipScene *s, uint32_t mesh, const MopHipVertex *v,
                                       uint32_t nv);
int mop_hip_scene_add_light(MopHipScene *s, const MopHipLight *l);
/* Equirect RGBA32F environment: IBL tables computed on the device and the
 * auto-exposure of environment.c:806-835. NULL removes it. */
int mop_hip_scene_set_environment(MopHipScene *s, const float *rgba, int w, int h,
                                  float rotation, float intensity);
void mop_hip_scene_set_shadow_size(MopHipScene *s, int size);
/* Reuse the rendered shadow map while meshes, lights and its light matrix
 * are unchanged (the map does not depend on the camera). */
void mop_hip_scene_set_shadow_cache(MopHipScene *s, int enable);
void mop_hip_scene_set_ambient(MopHipScene *s, float ambient);
void mop_hip_scene_set_exposure(MopHipScene *s, float exposure);
float mop_hip_scene_exposure(const MopHipScene *s);
void mop_hip_scene_set_clear_color(MopHipScene *s, MopHipColor c);
void mop_hip_scene_set_post_effects(MopHipScene *s, uint32_t effects);
/* Blend mode (MOP_HIP_BLEND_*) and opacity (clamped to [0, 1]) of mesh
 * `mesh`: mop_mesh_set_blend_mode and mop_mesh_set_opacity
 * (scene.h:111-125, viewport.c:1815-1837) in one call, the mode given being
 * the mode drawn (no promotion of an opaque mesh to alpha).  Blended meshes
 * cast no shadow and are drawn after the opaque ones, farthest first
 * (pass_scene_transparent, viewport.c:2470-2546).  Returns 0 or -1. */
int mop_hip_scene_set_mesh_blend(MopHipScene *s, uint32_t mesh, uint32_t blend_mode,
                                 float opacity);
/* Record one frame of s into fb without executing it (execute a set of
 * recorded frames with mop_hip_frames_execute).  Returns triangles submitted. */
uint32_t mop_hip_scene_record(MopHipScene *s, MopHipFramebuffer *fb, const MopHipCamera *cam,
                              int width, int height);
/* Record and execute one frame into fb (width x height); readbacks through
 * the mop_hip_framebuffer_read_* calls.  Returns triangles submitted. */
uint32_t mop_hip_scene_render(MopHipScene *s, MopHipFramebuffer *fb, const MopHipCamera *cam,
                              int width, int height);
uint32_t mop_hip_scene_triangles(const MopHipScene *s);

#ifdef __cplusplus
}
#endif

#endif
