/* Host harness for the shading pass's key checks (shade_key_fetch in
 * master-of-puppets_amd/csrc/shade_key.h, the code k_shade runs): random
 * and stale visibility keys against tables filled with garbage -- the state
 * a key of an earlier frame would meet (triangle records of culled
 * triangles keep old contents, the checkpoint store old offsets) -- built
 * with AddressSanitizer, every table an allocation of its exact size.  A
 * read outside any table aborts the run; the program also checks what an
 * accepted key hands to the shading (draw and attribute slots in range).
 * argv[1]: keys per path (default 2,000,000).  Exit status 0 when clean. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "shade_key.h"

static uint64_t g_rng = 0x243F6A8885A308D3ull;
static uint64_t rnd() {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return g_rng;
}
static uint32_t rnd_below(uint32_t n) { return (uint32_t)(rnd() % n); }
static float rnd_float(float lo, float hi) {
  return lo + (hi - lo) * (float)((rnd() >> 40) * (1.0 / 16777216.0));
}
/* garbage float: a random bit pattern, or a plausible value */
static float junk(float lo, float hi) {
  uint32_t u = (uint32_t)rnd();
  float f;
  memcpy(&f, &u, 4);
  return (rnd() & 3) ? rnd_float(lo, hi) : f;
}
static uint32_t junk_index(uint32_t n) {
  switch (rnd() & 3) {
  case 0: return (uint32_t)rnd();
  case 1: return n + rnd_below(64);
  default: return rnd_below(n);
  }
}
template <class T> static T *table(size_t n) { return (T *)malloc(n * sizeof(T)); }

int main(int argc, char **argv) {
  long keys = argc > 1 ? atol(argv[1]) : 2000000;
  const uint32_t n_verts = 6000, n_tris = 10000, clip_cap = 96, ndraws = 7, nviews = 3;
  const uint32_t chk_cap = 30000, subtri_cap = 15000;
  const int vw[nviews] = {200, 96, 333}, vh[nviews] = {150, 64, 41};

  DrawGPU *draws = table<DrawGPU>(ndraws);
  memset(draws, 0, ndraws * sizeof(DrawGPU));
  for (uint32_t d = 0; d < ndraws; d++) {
    draws[d].view = d % nviews;
    draws[d].flags = (uint32_t)rnd_below(4) << MOPH_DRAW_SM_SHIFT;
  }
  ViewGPU *views = table<ViewGPU>(nviews);
  memset(views, 0, nviews * sizeof(ViewGPU));
  for (uint32_t v = 0; v < nviews; v++) {
    views[v].w = vw[v];
    views[v].h = vh[v];
    views[v].wbuf = table<float4>((size_t)vw[v] * vh[v]);
    for (int i = 0; i < vw[v] * vh[v]; i++)
      views[v].wbuf[i] = make_float4(junk(-1e3f, 1e3f), junk(-1e3f, 1e3f), junk(-1e3f, 1e3f),
                                     junk(0.0f, 1.0f));
  }
  /* records: mostly small grid triangles (v, v+1, v+41) as a mesh gives,
   * some clipped, some of garbage */
  uint4 *tri_rec = table<uint4>(n_tris);
  for (uint32_t t = 0; t < n_tris; t++) {
    uint32_t r = rnd_below(8), v = rnd_below(n_verts - 42);
    tri_rec[t] = r == 0   ? make_uint4(junk_index(clip_cap), (uint32_t)rnd(), (uint32_t)rnd(),
                                       junk_index(ndraws))
                 : r == 1 ? make_uint4(junk_index(n_verts), junk_index(n_verts),
                                       junk_index(n_verts), junk_index(ndraws))
                          : make_uint4(v, v + 1 + (r & 1) * 40, v + 41 - (r & 1) * 40,
                                       rnd_below(ndraws));
  }
  /* screen positions: a 40-column grid of 8 x 6 px cells (some jitter, some
   * far outside the views), some garbage */
  float4 *vscr = table<float4>(n_verts);
  for (uint32_t i = 0; i < n_verts; i++) {
    float gx = (float)(i % 40) * 8.0f - 8.0f + rnd_float(-2.0f, 2.0f);
    float gy = (float)((i / 40) % 40) * 6.0f - 6.0f + rnd_float(-2.0f, 2.0f);
    if ((rnd() & 15) == 0)
      gx *= 40.0f; /* a big triangle */
    vscr[i] = (rnd() & 7) ? make_float4(gx, gy, rnd_float(0.0f, 1.0f), rnd_float(0.05f, 5.0f))
                          : make_float4(junk(-40.0f, 380.0f), junk(-40.0f, 200.0f),
                                        junk(0.0f, 1.0f), junk(0.01f, 10.0f));
  }
  float4 *cpos = table<float4>((size_t)clip_cap * MOPH_MAX_POLY);
  for (uint32_t i = 0; i < clip_cap * MOPH_MAX_POLY; i++)
    cpos[i] = make_float4(junk(-2.0f, 2.0f), junk(-2.0f, 2.0f), junk(-2.0f, 2.0f),
                          junk(0.05f, 3.0f));
  float4 *vattr = table<float4>((size_t)n_verts * 3);
  for (uint32_t i = 0; i < n_verts * 3; i++)
    vattr[i] = make_float4(junk(-1.0f, 1.0f), junk(-1.0f, 1.0f), junk(-1.0f, 1.0f), junk(0.0f, 1.0f));
  float4 *cattr = table<float4>((size_t)clip_cap * MOPH_MAX_POLY * 3);
  for (uint32_t i = 0; i < clip_cap * MOPH_MAX_POLY * 3; i++)
    cattr[i] = make_float4(junk(-1.0f, 1.0f), junk(-1.0f, 1.0f), junk(-1.0f, 1.0f), junk(0.0f, 1.0f));
  uint32_t nbo = n_tris + clip_cap * 16 + 1;
  uint32_t *big_off = table<uint32_t>(nbo);
  for (uint32_t i = 0; i < nbo; i++)
    big_off[i] = (rnd() & 7) == 0 ? ~0u : junk_index(chk_cap);
  float4 *chk = table<float4>(chk_cap);
  for (uint32_t i = 0; i < chk_cap; i++)
    chk[i] = make_float4(junk(-1e3f, 1e3f), junk(-1e3f, 1e3f), junk(-1e3f, 1e3f), junk(0.0f, 1.0f));
  SubTri *subtris = table<SubTri>(subtri_cap);
  for (uint32_t i = 0; i < subtri_cap; i++) {
    for (int k = 0; k < 3; k++) {
      subtris[i].sx[k] = junk(-40.0f, 380.0f);
      subtris[i].sy[k] = junk(-40.0f, 200.0f);
      subtris[i].sz[k] = junk(0.0f, 1.0f);
      subtris[i].iw[k] = junk(0.01f, 10.0f);
      subtris[i].a[k] = junk_index(n_verts + clip_cap * MOPH_MAX_POLY);
    }
    subtris[i].draw = junk_index(ndraws);
  }

  long accepted[2] = {0, 0}, big[2] = {0, 0}, bad = 0;
  double sum = 0.0; /* consumes every fetched value, so no load is optimized away */
  for (int fused = 1; fused >= 0; fused--) {
    FrameArgs fa;
    memset(&fa, 0, sizeof(fa));
    fa.draws = draws;
    fa.views = views;
    fa.ndraws = ndraws;
    fa.nviews = nviews;
    fa.n_verts = n_verts;
    fa.n_tris = n_tris;
    fa.clip_cap = clip_cap;
    fa.small_area = 64;
    fa.fused = (uint32_t)fused;
    fa.vscr = vscr;
    fa.tri_rec = tri_rec;
    fa.chk = chk;
    fa.chk_cap = chk_cap;
    fa.big_off = big_off;
    fa.cpos = cpos;
    fa.vattr = vattr;
    fa.cattr = cattr;
    fa.subtris = subtris;
    fa.subtri_cap = subtri_cap;
    for (long i = 0; i < keys; i++) {
      uint32_t s;
      switch (rnd() & 3) {
      case 0: s = (uint32_t)rnd(); break;                                 /* any value */
      case 1: s = fused ? n_tris * 16u + rnd_below(4096) : subtri_cap + rnd_below(4096); break;
      default: s = fused ? (rnd_below(n_tris) << 4) | ((rnd() & 1) ? 15u : rnd_below(16))
                         : rnd_below(subtri_cap);                         /* in range */
      }
      uint32_t v = rnd_below(nviews);
      int px = (int)rnd_below((uint32_t)vw[v]), py = (int)rnd_below((uint32_t)vh[v]);
      if (fused && (rnd() & 1) && (s >> 4) < n_tris && tri_rec[s >> 4].x < n_verts) {
        /* a pixel near the keyed triangle's first vertex */
        float4 p = vscr[tri_rec[s >> 4].x];
        if (p.x > -1e6f && p.x < 1e6f && p.y > -1e6f && p.y < 1e6f) {
          px = (int)p.x + (int)rnd_below(9) - 4;
          py = (int)p.y + (int)rnd_below(9) - 4;
          px = px < 0 ? 0 : (px >= vw[v] ? vw[v] - 1 : px);
          py = py < 0 ? 0 : (py >= vh[v] ? vh[v] - 1 : py);
        }
      }
      size_t idx = (size_t)py * (size_t)vw[v] + (size_t)px;
      KeyFetch k;
      if (shade_key_fetch(fa, views[v], s, px, py, idx, k)) {
        accepted[fused]++;
        sum += (double)k.w0 + (double)k.w1 + (double)k.w2 + (double)k.zf + (double)k.sm +
               (double)k.st.sx[0] + (double)k.st.sy[2] + (double)k.st.iw[1] + (double)k.st.sz[2];
        for (int j = 0; j < 3; j++)
          sum += (double)k.A0[j].x + (double)k.A1[j].w + (double)k.A2[j].y;
        big[fused] += bbox_area(k.r.es) > (long long)fa.small_area;
        uint32_t na = n_verts + clip_cap * MOPH_MAX_POLY;
        if (k.st.draw >= ndraws || k.st.a[0] >= na || k.st.a[1] >= na || k.st.a[2] >= na ||
            px < k.r.es.min_x || px > k.r.es.max_x || py < k.r.es.min_y || py > k.r.es.max_y)
          bad++;
      }
    }
  }
  printf("order-free: %ld of %ld keys accepted (%ld via checkpoints); store path: %ld of %ld "
         "(%ld big); inconsistent %ld; checksum %g\n",
         accepted[1], keys, big[1], accepted[0], keys, big[0], bad, sum);
  fflush(stdout);
  for (uint32_t v = 0; v < nviews; v++)
    free(views[v].wbuf);
  free(views);
  free(draws);
  free(tri_rec);
  free(vscr);
  free(cpos);
  free(vattr);
  free(cattr);
  free(big_off);
  free(chk);
  free(subtris);
  return bad ? 1 : 0;
}
