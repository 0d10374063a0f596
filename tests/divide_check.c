/* Host check of mop_div_recip (master-of-puppets_amd/csrc/libmf.h) against
 * IEEE division: every float numerator (on a stride; 1 = exhaustive) for
 * the constant divisors the shading uses (2 pi and pi as float, 9), the
 * numerators 0..9 over 9 exactly, and random (a, b) pairs over the whole
 * normal range.  Prints the mismatch counts; exit status 1 on a mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "libmf.h"

static uint64_t rng = 0x2545F4914F6CDD1Dull;
static uint64_t next(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return rng;
}

static int same(float a, float b) { return lmf_bits(a) == lmf_bits(b) || (a != a && b != b); }

int main(int argc, char **argv) {
  uint32_t stride = argc > 1 ? (uint32_t)strtoul(argv[1], 0, 10) : 101;
  const float divs[3] = {2.0f * 3.14159265f, 3.14159265f, 9.0f};
  long bad_c = 0, n_c = 0, bad_r = 0, n_r = 0;
  for (int k = 0; k < 3; k++) {
    volatile float c = divs[k];
    float y = 1.0f / c;
    for (uint64_t u = 0; u <= 0xffffffffull; u += stride) {
      float a = lmf_float((uint32_t)u);
      bad_c += !same(a / c, mop_div_recip(a, c, y));
      n_c++;
    }
    for (int i = 0; i <= 9; i++) {
      bad_c += !same((float)i / c, mop_div_recip((float)i, c, y));
      n_c++;
    }
  }
  long pairs = 2000000000L / (long)stride + 1000000;
  for (long i = 0; i < pairs; i++) {
    uint64_t r = next();
    float a = lmf_float((uint32_t)r), b = lmf_float((uint32_t)(r >> 32));
    volatile float vb = b;
    float y = 1.0f / vb;
    bad_r += !same(a / vb, mop_div_recip(a, vb, y));
    n_r++;
  }
  printf("constant divisors %ld/%ld random pairs %ld/%ld\n", bad_c, n_c, bad_r, n_r);
  return (bad_c || bad_r) ? 1 : 0;
}
