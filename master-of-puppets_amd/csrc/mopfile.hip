/*
 * mopfile.hip -- the .mop v2 scene file (src/loader/mop_scene.c) read into
 * HBM: host parsing of the memory-mapped file, device decoding of its mesh
 * sections.
 *
 * On-disk layout (little-endian, every section 8-byte aligned; the layout
 * mop_scene_save writes, mop_scene.c:544-768):
 *   [0, 64)        header: magic 0x4D4F5002, version 2, section count, flags
 *   [64, ...)      table of contents, 24 bytes per section:
 *                  type (1 mesh, 2 camera, 3 light), flags, offset, size
 *   mesh section   48-byte header (vertex count, index count, vertex
 *                  encoding 0 raw / 1 quantized, index size 2 / 4, bbox min
 *                  and max), the vertices (48-byte MopVertex or 18-byte
 *                  quantized records), padding to 8 bytes, the indices
 *   camera         eye, target, up, fov (radians), near, far, aspect (56 B)
 *   light          type, position, direction, rgba, intensity, range, spot
 *                  inner and outer cosines (64 B)
 * A quantized vertex is 18 bytes (the struct in mop_scene.h is commented
 * "exactly 20 bytes"; its sizeof is 18): position u16 x 3 in the mesh's box,
 * octahedral normal i16 x 2, RGBA u8 x 4, uv u16 x 2.
 *
 * Device decoding (k_dequantize): a workgroup copies the 256 records it
 * decodes (4608 bytes, 16-byte aligned) into LDS with 16-byte loads, each
 * lane decodes one record and writes its 48-byte vertex as three 16-byte
 * stores; the box of the decoded positions is reduced per wave and merged
 * with six atomics.  The arithmetic is dequantize_vertex / oct_decode
 * (mop_scene.c:142-219) operation for operation: IEEE division by 65535,
 * 32767 and 255, the octahedral fold, sqrtf and the three divisions by the
 * length, no fused multiply-add (the build uses -ffp-contract=off).
 * Traffic per vertex: 18 bytes read, 48 written; per 16-bit index 2 read, 4
 * written -- HBM-bound, and small next to the PCIe copy of the section.
 */
#include <fcntl.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <vector>

#include "mop_hip_internal.h"
#include "mop_hip_scene_file.h"

namespace {

constexpr uint32_t kMagic = 0x4D4F5002u;
constexpr uint32_t kVersion = 2u;
constexpr uint32_t kMaxSections = 256;
constexpr size_t kHeaderBytes = 64;
constexpr size_t kTocBytes = 24;
constexpr size_t kMeshHeaderBytes = 48;
constexpr size_t kCameraBytes = 56;
constexpr size_t kLightBytes = 64;
constexpr size_t kQuantBytes = 18;
constexpr size_t kVertexBytes = 48;
constexpr uint32_t kSectionMesh = 1, kSectionCamera = 2, kSectionLight = 3;
constexpr uint32_t kVtxQuantized = 1;
constexpr int kDecodeWG = 256;

struct TocEntry {
  uint32_t type, flags;
  uint64_t offset, size;
};

uint32_t rd_u32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}
uint64_t rd_u64(const uint8_t *p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}
float rd_f32(const uint8_t *p) {
  float v;
  memcpy(&v, p, 4);
  return v;
}
size_t up8(size_t v) { return (v + 7) & ~(size_t)7; }
size_t up16(size_t v) { return (v + 15) & ~(size_t)15; }

/* order-preserving map of a float onto uint32 (for atomic min / max) */
__device__ __forceinline__ uint32_t fkey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
float fkey_inv(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct DecodeArgs {
  const uint8_t *src; /* quantized records (16-byte aligned, padded to 16) */
  float *dst;         /* MopHipVertex records */
  uint32_t n;
  float bmin[3], range[3];
  uint32_t *box; /* min keys x, y, z then max keys x, y, z */
};

__device__ void box_merge(float x, float y, float z, bool valid, uint32_t *box) {
  uint32_t k[6] = {valid ? fkey(x) : 0xffffffffu, valid ? fkey(y) : 0xffffffffu,
                   valid ? fkey(z) : 0xffffffffu, valid ? fkey(x) : 0u,
                   valid ? fkey(y) : 0u,          valid ? fkey(z) : 0u};
  for (int off = 32; off >= 1; off >>= 1)
    for (int c = 0; c < 6; c++) {
      uint32_t o = __shfl_xor(k[c], off, 64);
      k[c] = c < 3 ? min(k[c], o) : max(k[c], o);
    }
  if ((threadIdx.x & 63) == 0) {
    for (int c = 0; c < 3; c++) {
      atomicMin(&box[c], k[c]);
      atomicMax(&box[3 + c], k[3 + c]);
    }
  }
}

/* dequantize_vertex + oct_decode (mop_scene.c:142-163, 198-219) */
__global__ __launch_bounds__(kDecodeWG) void k_dequantize(DecodeArgs a) {
  __shared__ uint4 stage[kDecodeWG * kQuantBytes / 16];
  uint32_t base = blockIdx.x * kDecodeWG;
  uint32_t nv = a.n - base < (uint32_t)kDecodeWG ? a.n - base : (uint32_t)kDecodeWG;
  uint32_t n16 = (uint32_t)((nv * kQuantBytes + 15) / 16);
  const uint4 *src = reinterpret_cast<const uint4 *>(a.src + (size_t)base * kQuantBytes);
  for (uint32_t i = threadIdx.x; i < n16; i += kDecodeWG)
    stage[i] = src[i];
  __syncthreads();
  uint32_t t = threadIdx.x;
  bool valid = t < nv;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  if (valid) {
    const uint8_t *r = reinterpret_cast<const uint8_t *>(stage) + (size_t)t * kQuantBytes;
    const uint16_t *q = reinterpret_cast<const uint16_t *>(r);
    px = a.bmin[0] + ((float)q[0] / 65535.0f) * a.range[0];
    py = a.bmin[1] + ((float)q[1] / 65535.0f) * a.range[1];
    pz = a.bmin[2] + ((float)q[2] / 65535.0f) * a.range[2];
    float ox = (float)(int16_t)q[3] / 32767.0f;
    float oy = (float)(int16_t)q[4] / 32767.0f;
    float nz = 1.0f - fabsf(ox) - fabsf(oy);
    if (nz < 0.0f) {
      float fx = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
      float fy = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
      ox = fx;
      oy = fy;
    }
    float nx = ox, ny = oy;
    float len = sqrtf(nx * nx + ny * ny + nz * nz);
    if (len > 1e-8f) {
      nx /= len;
      ny /= len;
      nz /= len;
    }
    float cr = (float)r[10] / 255.0f, cg = (float)r[11] / 255.0f;
    float cb = (float)r[12] / 255.0f, ca = (float)r[13] / 255.0f;
    float u = (float)q[7] / 65535.0f, v = (float)q[8] / 65535.0f;
    float4 *o = reinterpret_cast<float4 *>(a.dst + (size_t)(base + t) * 12);
    o[0] = make_float4(px, py, pz, nx);
    o[1] = make_float4(ny, nz, cr, cg);
    o[2] = make_float4(cb, ca, u, v);
  }
  box_merge(px, py, pz, valid, a.box);
}

/* box of raw (48-byte) vertices already in place */
__global__ __launch_bounds__(kDecodeWG) void k_vertex_box(const float *v, uint32_t n,
                                                          uint32_t *box) {
  uint32_t i = blockIdx.x * kDecodeWG + threadIdx.x;
  bool valid = i < n;
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (valid) {
    x = v[(size_t)i * 12];
    y = v[(size_t)i * 12 + 1];
    z = v[(size_t)i * 12 + 2];
  }
  box_merge(x, y, z, valid, box);
}

/* 16-bit indices widened to 32 bits (mop_scene.c:467-471), eight per lane */
__global__ __launch_bounds__(kDecodeWG) void k_widen_indices(const uint16_t *src, uint32_t n,
                                                             uint32_t *dst) {
  size_t i = ((size_t)blockIdx.x * kDecodeWG + threadIdx.x) * 8;
  if (i + 8 <= n) {
    uint4 w = *reinterpret_cast<const uint4 *>(src + i);
    uint4 lo = make_uint4(w.x & 0xffffu, w.x >> 16, w.y & 0xffffu, w.y >> 16);
    uint4 hi = make_uint4(w.z & 0xffffu, w.z >> 16, w.w & 0xffffu, w.w >> 16);
    reinterpret_cast<uint4 *>(dst + i)[0] = lo;
    reinterpret_cast<uint4 *>(dst + i)[1] = hi;
  } else {
    for (; i < n; i++)
      dst[i] = src[i];
  }
}

} // namespace

struct MopHipSceneFile {
  const uint8_t *data = nullptr;
  size_t size = 0;
  std::vector<TocEntry> toc;
};

extern "C" MopHipSceneFile *mop_hip_scene_file_load(const char *path) {
  if (!path) {
    runtime_set_err("mop_hip_scene_file_load: NULL path");
    return nullptr;
  }
  int fd = open(path, O_RDONLY);
  if (fd < 0) {
    runtime_set_err("mop_hip_scene_file_load: failed to open '%s'", path);
    return nullptr;
  }
  struct stat st;
  if (fstat(fd, &st) < 0 || st.st_size < (off_t)kHeaderBytes) {
    close(fd);
    runtime_set_err("mop_hip_scene_file_load: file too small '%s'", path);
    return nullptr;
  }
  size_t size = (size_t)st.st_size;
  void *map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (map == MAP_FAILED) {
    runtime_set_err("mop_hip_scene_file_load: mmap failed for '%s'", path);
    return nullptr;
  }
  const uint8_t *p = (const uint8_t *)map;
  const char *why = nullptr;
  uint32_t nsec = rd_u32(p + 8);
  if (rd_u32(p) != kMagic)
    why = "bad magic";
  else if (rd_u32(p + 4) != kVersion)
    why = "unsupported version";
  else if (nsec > kMaxSections)
    why = "too many sections";
  else if (kHeaderBytes + (size_t)nsec * kTocBytes > size)
    why = "table of contents exceeds the file";
  if (why) {
    munmap(map, size);
    runtime_set_err("mop_hip_scene_file_load: %s in '%s'", why, path);
    return nullptr;
  }
  MopHipSceneFile *f = new MopHipSceneFile();
  f->data = p;
  f->size = size;
  f->toc.resize(nsec);
  for (uint32_t i = 0; i < nsec; i++) {
    const uint8_t *e = p + kHeaderBytes + (size_t)i * kTocBytes;
    f->toc[i] = TocEntry{rd_u32(e), rd_u32(e + 4), rd_u64(e + 8), rd_u64(e + 16)};
  }
  return f;
}

extern "C" void mop_hip_scene_file_free(MopHipSceneFile *f) {
  if (!f)
    return;
  munmap((void *)f->data, f->size);
  delete f;
}

/* the nth section of a type (find_section, mop_scene.c:363-374) */
static const TocEntry *nth_section(const MopHipSceneFile *f, uint32_t type, uint32_t nth) {
  for (const TocEntry &e : f->toc)
    if (e.type == type && nth-- == 0)
      return &e;
  return nullptr;
}

static uint32_t count_of(const MopHipSceneFile *f, uint32_t type) {
  uint32_t n = 0;
  for (const TocEntry &e : f->toc)
    n += e.type == type;
  return n;
}

/* section bytes [offset, offset + need) inside the file */
static bool in_file(const MopHipSceneFile *f, uint64_t offset, uint64_t need) {
  return offset <= f->size && need <= f->size - offset;
}

extern "C" uint32_t mop_hip_scene_file_mesh_count(const MopHipSceneFile *f) {
  return f ? count_of(f, kSectionMesh) : 0;
}

extern "C" uint32_t mop_hip_scene_file_light_count(const MopHipSceneFile *f) {
  return f ? count_of(f, kSectionLight) : 0;
}

extern "C" bool mop_hip_scene_file_get_camera(const MopHipSceneFile *f, MopHipVec3 *eye,
                                              MopHipVec3 *target, MopHipVec3 *up,
                                              float *fov_radians, float *near_p, float *far_p) {
  if (!f)
    return false;
  const TocEntry *e = nth_section(f, kSectionCamera, 0);
  if (!e || !in_file(f, e->offset, kCameraBytes))
    return false;
  const uint8_t *c = f->data + e->offset;
  auto vec = [c](int k) { return MopHipVec3{rd_f32(c + k), rd_f32(c + k + 4), rd_f32(c + k + 8)}; };
  if (eye)
    *eye = vec(0);
  if (target)
    *target = vec(12);
  if (up)
    *up = vec(24);
  if (fov_radians)
    *fov_radians = rd_f32(c + 36);
  if (near_p)
    *near_p = rd_f32(c + 40);
  if (far_p)
    *far_p = rd_f32(c + 44);
  return true;
}

extern "C" bool mop_hip_scene_file_get_light(const MopHipSceneFile *f, uint32_t idx,
                                             MopHipLight *out) {
  if (!f || !out)
    return false;
  const TocEntry *e = nth_section(f, kSectionLight, idx);
  if (!e || !in_file(f, e->offset, kLightBytes))
    return false;
  const uint8_t *c = f->data + e->offset;
  memset(out, 0, sizeof(*out));
  out->type = (int32_t)rd_u32(c);
  out->position = MopHipVec3{rd_f32(c + 4), rd_f32(c + 8), rd_f32(c + 12)};
  out->direction = MopHipVec3{rd_f32(c + 16), rd_f32(c + 20), rd_f32(c + 24)};
  out->color = MopHipColor{rd_f32(c + 28), rd_f32(c + 32), rd_f32(c + 36), rd_f32(c + 40)};
  out->intensity = rd_f32(c + 44);
  out->range = rd_f32(c + 48);
  out->spot_inner_cos = rd_f32(c + 52);
  out->spot_outer_cos = rd_f32(c + 56);
  out->active = true;
  out->cast_shadows = false;
  return true;
}

extern "C" void mop_hip_loaded_mesh_free(MopHipDevice *dev, MopHipLoadedMesh *m) {
  if (!m)
    return;
  mop_hip_buffer_destroy(dev, m->vertex_buffer);
  mop_hip_buffer_destroy(dev, m->index_buffer);
  memset(m, 0, sizeof(*m));
}

extern "C" bool mop_hip_scene_file_get_mesh(MopHipDevice *dev, const MopHipSceneFile *f,
                                            uint32_t idx, MopHipLoadedMesh *out) {
  if (!out)
    return false;
  memset(out, 0, sizeof(*out));
  if (!f)
    return false;
  const TocEntry *e = nth_section(f, kSectionMesh, idx);
  if (!e)
    return false;
  if (!in_file(f, e->offset, e->size)) {
    runtime_set_err("mop_hip_scene_file_get_mesh: section data out of bounds");
    return false;
  }
  if (e->size < kMeshHeaderBytes)
    return false;
  const uint8_t *sec = f->data + e->offset;
  uint32_t nv = rd_u32(sec), ni = rd_u32(sec + 4);
  bool quantized = rd_u32(sec + 8) == kVtxQuantized;
  bool idx16 = rd_u32(sec + 12) == 2;
  float bmin[3] = {rd_f32(sec + 16), rd_f32(sec + 20), rd_f32(sec + 24)};
  float bmax[3] = {rd_f32(sec + 28), rd_f32(sec + 32), rd_f32(sec + 36)};
  size_t vtx_bytes = (size_t)nv * (quantized ? kQuantBytes : kVertexBytes);
  size_t idx_bytes = (size_t)ni * (idx16 ? 2 : 4);
  if (kMeshHeaderBytes + up8(vtx_bytes) + idx_bytes > e->size) {
    runtime_set_err("mop_hip_scene_file_get_mesh: section %u shorter than its header says", idx);
    return false;
  }
  if (!dev) {
    runtime_set_err("mop_hip_scene_file_get_mesh: no device");
    return false;
  }
  const uint8_t *vsrc = sec + kMeshHeaderBytes;
  const uint8_t *isrc = vsrc + up8(vtx_bytes);
  hipStream_t s = device_stream(dev);

  void *vdev = nullptr, *idev = nullptr, *stage = nullptr;
  uint32_t *box = nullptr;
  size_t stage_v = quantized ? up16(vtx_bytes) : 0, stage_i = idx16 ? up16(idx_bytes) : 0;
  bool ok = hipMalloc(&vdev, nv ? (size_t)nv * kVertexBytes : 4) == hipSuccess &&
            hipMalloc(&idev, ni ? (size_t)ni * 4 : 4) == hipSuccess &&
            hipMalloc((void **)&box, 6 * sizeof(uint32_t)) == hipSuccess &&
            (stage_v + stage_i == 0 || hipMalloc(&stage, stage_v + stage_i) == hipSuccess);
  const uint32_t box_init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  uint32_t box_keys[6] = {0, 0, 0, 0, 0, 0};
  /* the section's bytes as they lie in the file: straight into the vertex
   * and index buffers when no decoding is needed, else into the staging
   * area the kernels read */
  ok = ok && hipMemcpy(box, box_init, sizeof(box_init), hipMemcpyHostToDevice) == hipSuccess;
  if (ok && vtx_bytes)
    ok = hipMemcpy(quantized ? stage : vdev, vsrc, vtx_bytes, hipMemcpyHostToDevice) ==
         hipSuccess;
  if (ok && idx_bytes)
    ok = hipMemcpy(idx16 ? (uint8_t *)stage + stage_v : idev, isrc, idx_bytes,
                   hipMemcpyHostToDevice) == hipSuccess;
  if (ok && nv) {
    uint32_t blocks = (nv + kDecodeWG - 1) / kDecodeWG;
    if (quantized) {
      DecodeArgs a;
      a.src = (const uint8_t *)stage;
      a.dst = (float *)vdev;
      a.n = nv;
      for (int c = 0; c < 3; c++) {
        a.bmin[c] = bmin[c];
        a.range[c] = bmax[c] - bmin[c]; /* dequantize_vertex's range, mop_scene.c:204 */
      }
      a.box = box;
      hipLaunchKernelGGL(k_dequantize, dim3(blocks), dim3(kDecodeWG), 0, s, a);
    } else {
      hipLaunchKernelGGL(k_vertex_box, dim3(blocks), dim3(kDecodeWG), 0, s,
                         (const float *)vdev, nv, box);
    }
    ok = hipGetLastError() == hipSuccess;
  }
  if (ok && idx16 && ni) {
    uint32_t lanes = (ni + 7) / 8;
    hipLaunchKernelGGL(k_widen_indices, dim3((lanes + kDecodeWG - 1) / kDecodeWG),
                       dim3(kDecodeWG), 0, s, (const uint16_t *)((uint8_t *)stage + stage_v), ni,
                       (uint32_t *)idev);
    ok = hipGetLastError() == hipSuccess;
  }
  ok = ok && hipStreamSynchronize(s) == hipSuccess &&
       hipMemcpy(box_keys, box, sizeof(box_keys), hipMemcpyDeviceToHost) == hipSuccess;
  if (stage)
    (void)hipFree(stage);
  if (box)
    (void)hipFree(box);
  if (ok) {
    out->vertex_buffer = buffer_adopt_device(dev, vdev, (size_t)nv * kVertexBytes);
    if (out->vertex_buffer)
      vdev = nullptr;
    out->index_buffer = buffer_adopt_device(dev, idev, (size_t)ni * 4);
    if (out->index_buffer)
      idev = nullptr;
    ok = out->vertex_buffer && out->index_buffer;
  }
  if (!ok) {
    runtime_set_err("mop_hip_scene_file_get_mesh: device decode of mesh %u failed", idx);
    if (vdev)
      (void)hipFree(vdev);
    if (idev)
      (void)hipFree(idev);
    mop_hip_loaded_mesh_free(dev, out);
    return false;
  }
  out->vertex_count = nv;
  out->index_count = ni;
  out->bbox_min = MopHipVec3{bmin[0], bmin[1], bmin[2]};
  out->bbox_max = MopHipVec3{bmax[0], bmax[1], bmax[2]};
  if (nv) {
    out->aabb_min = MopHipVec3{fkey_inv(box_keys[0]), fkey_inv(box_keys[1]), fkey_inv(box_keys[2])};
    out->aabb_max = MopHipVec3{fkey_inv(box_keys[3]), fkey_inv(box_keys[4]), fkey_inv(box_keys[5])};
  }
  return true;
}
