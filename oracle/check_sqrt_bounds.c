/* Exhaustive check of the two squared-distance bounds the GPU nearest-higher
 * searches compare against instead of taking square roots (not_farther_than and
 * strictly_nearer in clue_device.hpp, restated here operation by operation), on
 * every finite float s >= 0 with sqrtf as the judge:
 *   up(s) = s * (1 + 2^-20) + 2^-126:  every x > up(s) has sqrtf(x) > sqrtf(s)
 *     (it suffices to test x = the next float above up(s), sqrtf being monotone);
 *   nn(s) = fma(s, 1 - 2^-20, -2^-126): every x >= 0 below nn(s) has
 *     sqrtf(x) < sqrtf(s) (tested on the float just below nn(s) when nn(s) > 0).
 * Test infrastructure only.  Build and run: make -C oracle check-sqrt (about 10 s). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

int main(void) {
  long bad_up = 0, bad_nn = 0, nn_positive = 0;
#pragma omp parallel for reduction(+ : bad_up, bad_nn, nn_positive) schedule(static)
  for (int64_t i = 0; i < 0x7f800000LL; ++i) {
    const float s = from_bits((uint32_t)i);
    const float r = sqrtf(s);
    /* the product and the sum rounded separately (no contraction) */
    volatile float prod = s * (1.f + 0x1p-20f);
    const float up = prod + 0x1p-126f;
    if (up < INFINITY) {
      const float above = from_bits(to_bits(up) + 1u);
      if (!(sqrtf(above) > r)) ++bad_up;
    }
    const float nn = fmaf(s, 1.f - 0x1p-20f, -0x1p-126f);
    if (nn > 0.f) {
      ++nn_positive;
      const float below = from_bits(to_bits(nn) - 1u);
      if (!(sqrtf(below) < r)) ++bad_nn;
    }
  }
  printf("exhaustive: not_farther_than over 2139095040 floats, %ld violations\n", bad_up);
  printf("exhaustive: strictly_nearer over %ld floats with a positive bound, %ld violations\n", nn_positive,
         bad_nn);
  return (bad_up || bad_nn) ? 1 : 0;
}
