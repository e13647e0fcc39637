/* Exhaustive check that the engine exponential (oracle_expf_engine, the same
 * definition the GPU kernels evaluate) is the correctly rounded float exp on
 * every float in [-110, 89]: compared with expl() in long double (64-bit
 * significand) rounded once to float.  Outside that range the engine returns
 * +inf / 0, which are the correctly rounded values there.
 * Build and run: make -C oracle check-exp   (about 20 s on 8 cores). */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

float oracle_expf_engine(float x);
float oracle_expf_longdouble(float x);

int main(void) {
  long long bad = 0, total = 0;
#pragma omp parallel for reduction(+ : bad, total) schedule(dynamic, 1 << 20)
  for (long long u = 0; u < (1ll << 32); ++u) {
    const uint32_t b = (uint32_t)u;
    float x;
    memcpy(&x, &b, 4);
    if (!(x >= -110.0f && x <= 89.0f)) continue;
    ++total;
    const float e = oracle_expf_engine(x), l = oracle_expf_longdouble(x);
    if (memcmp(&e, &l, 4) != 0) ++bad;
  }
  printf("exhaustive: %lld inputs in [-110, 89], %lld mismatches\n", total, bad);
  return bad != 0;
}
