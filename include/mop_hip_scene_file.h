/*
 * mop_hip_scene_file.h -- the .mop v2 scene file (src/loader/mop_scene.c,
 * include/mop/loader/mop_scene.h) read into HBM.
 *
 * The file is memory-mapped and parsed on the host (header, table of
 * contents: mop_scene_load, mop_scene.c:225-337).  A mesh section is copied
 * to the device as it lies in the file -- 18-byte quantized vertices and
 * 16-bit indices when the file holds them -- and decoded there into the
 * 48-byte MopHipVertex records and 32-bit indices the frame path draws
 * (mop_scene_get_mesh, mop_scene.c:397-487, with dequantize_vertex and
 * oct_decode, :142-219).  The decoded vertices equal the reference's host
 * decode bit for bit.  A quantized vertex crosses PCIe as 18 bytes instead
 * of 48 and a 16-bit index as 2 instead of 4, and no host thread decodes.
 */
#ifndef MOP_HIP_SCENE_FILE_H
#define MOP_HIP_SCENE_FILE_H

#include "mop_hip.h"
#include "mop_hip_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct MopHipSceneFile MopHipSceneFile;

/* One mesh section decoded into device buffers (MopLoadedMesh,
 * loader.h:85-97, with RHI buffers in place of host arrays). */
typedef struct MopHipLoadedMesh {
  MopHipBuffer *vertex_buffer; /* vertex_count MopHipVertex records */
  MopHipBuffer *index_buffer;  /* index_count uint32 indices */
  uint32_t vertex_count, index_count;
  MopHipVec3 bbox_min, bbox_max; /* the section header's box, as MopLoadedMesh */
  MopHipVec3 aabb_min, aabb_max; /* box of the decoded positions */
} MopHipLoadedMesh;

/* mop_scene_load (mop_scene.c:225-337): NULL when the file cannot be opened
 * or mapped, is shorter than the 64-byte header, has another magic
 * (0x4D4F5002) or version (2), more than 256 sections, or a table of
 * contents past its end; the reason is in mop_hip_last_error(). */
MopHipSceneFile *mop_hip_scene_file_load(const char *path);
/* mop_scene_free (mop_scene.c:343-357) */
void mop_hip_scene_file_free(MopHipSceneFile *f);
/* mop_scene_mesh_count / mop_scene_light_count (mop_scene.c:385-395) */
uint32_t mop_hip_scene_file_mesh_count(const MopHipSceneFile *f);
uint32_t mop_hip_scene_file_light_count(const MopHipSceneFile *f);
/* mop_scene_get_camera (mop_scene.c:489-519); any output may be NULL */
bool mop_hip_scene_file_get_camera(const MopHipSceneFile *f, MopHipVec3 *eye, MopHipVec3 *target,
                                   MopHipVec3 *up, float *fov_radians, float *near_p,
                                   float *far_p);
/* The idx-th light section (MopLightSection, mop_scene.c:79-89, as
 * mop_scene_save writes it, :619-649).  The reference offers no reader for
 * it; the fields the section stores are filled, active is true and
 * cast_shadows false (the section does not store it). */
bool mop_hip_scene_file_get_light(const MopHipSceneFile *f, uint32_t idx, MopHipLight *out);
/* mop_scene_get_mesh (mop_scene.c:397-487), decoded on the device.  False
 * (out zeroed) when the section is missing, lies outside the file, or is
 * shorter than its header says -- the last check is this loader's own: the
 * reference reads vertex and index bytes without it. */
bool mop_hip_scene_file_get_mesh(MopHipDevice *dev, const MopHipSceneFile *f, uint32_t idx,
                                 MopHipLoadedMesh *out);
/* mop_load_free (loader.h:103-104): the mesh's buffers */
void mop_hip_loaded_mesh_free(MopHipDevice *dev, MopHipLoadedMesh *m);

/* A decoded mesh added to a resident scene without leaving the device (the
 * scene takes the buffers; m is zeroed).  Same meaning of object_id, model16
 * and the material as mop_hip_scene_add_mesh.  Returns the mesh index or
 * < 0. */
int mop_hip_scene_add_loaded_mesh(MopHipScene *s, MopHipLoadedMesh *m, uint32_t object_id,
                                  const float *model16, int has_material, float metallic,
                                  float roughness);

#ifdef __cplusplus
}
#endif

#endif
