/* Host check of master-of-puppets_amd/csrc/libmf.h against the C library:
 * bit-identical results on a stride of every float in [-1, 1] (asinf), of
 * every finite float (atanf), on random argument pairs (atan2f), and on the
 * special values.  Prints the mismatch counts; exit status 1 on a mismatch.
 * argv[1]: stride (1 = exhaustive). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "libmf.h"

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return rng;
}

static int same(float a, float b) { return lmf_bits(a) == lmf_bits(b) || (a != a && b != b); }

int main(int argc, char **argv) {
  uint32_t stride = argc > 1 ? (uint32_t)strtoul(argv[1], 0, 10) : 97;
  long bad_as = 0, bad_at = 0, bad_a2 = 0, n_as = 0, n_at = 0, n_a2 = 0;
  for (uint64_t u = 0; u <= 0x3f800000u; u += stride)
    for (int sg = 0; sg < 2; sg++) {
      float x = lmf_float((uint32_t)u | (sg ? 0x80000000u : 0u));
      bad_as += !same(asinf(x), mop_asinf(x));
      n_as++;
    }
  for (uint64_t u = 0; u <= 0x7f800000u; u += stride)
    for (int sg = 0; sg < 2; sg++) {
      float x = lmf_float((uint32_t)u | (sg ? 0x80000000u : 0u));
      bad_at += !same(atanf(x), mop_atanf(x));
      n_at++;
    }
  long pairs = 40000000L / (long)stride + 100000;
  for (long i = 0; i < pairs; i++) {
    uint64_t r = next();
    float y, x;
    if (i & 1) { /* any bit patterns */
      y = lmf_float((uint32_t)r);
      x = lmf_float((uint32_t)(r >> 32));
    } else {     /* unit-range directions, as the shading calls it */
      y = (float)((r >> 40) * (1.0 / 16777216.0)) * 2.0f - 1.0f;
      x = (float)(((r >> 16) & 0xffffff) * (1.0 / 16777216.0)) * 2.0f - 1.0f;
    }
    bad_a2 += !same(atan2f(y, x), mop_atan2f(y, x));
    n_a2++;
  }
  const float sp[] = {0.0f, -0.0f, 1.0f, -1.0f, INFINITY, -INFINITY, NAN, 0.5f, -0.5f,
                      1e-30f, -1e-30f, 3.0f, -3.0f, 1e30f, -1e30f};
  for (unsigned i = 0; i < sizeof(sp) / sizeof(sp[0]); i++)
    for (unsigned j = 0; j < sizeof(sp) / sizeof(sp[0]); j++) {
      bad_a2 += !same(atan2f(sp[i], sp[j]), mop_atan2f(sp[i], sp[j]));
      n_a2++;
    }
  printf("asinf %ld/%ld atanf %ld/%ld atan2f %ld/%ld\n", bad_as, n_as, bad_at, n_at, bad_a2,
         n_a2);
  return (bad_as || bad_at || bad_a2) ? 1 : 0;
}
