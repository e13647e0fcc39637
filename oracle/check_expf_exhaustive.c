/* Exhaustive check that the engine exponential (oracle_expf_engine, the same
 * definition the GPU kernels evaluate) is the correctly rounded float exp on
 * every float in [-110, 89]: compared with expl() in long double (64-bit
 * significand) rounded once to float.  Outside that range the engine returns
 * +inf / 0, which are the correctly rounded values there.
 * Build and run: make -C oracle check-exp   (about 20 s on 8 cores). */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <math.h>

float oracle_expf_engine(float x);
float oracle_expf_longdouble(float x);

/* The GPU density pass's evaluation for arguments in [-14, 0] (engine_expf_unit
 * in clue_device.hpp), restated operation by operation: k = x*64/ln2 rounded to
 * an integer by adding 1.5*2^52 in one fused multiply-add (k is the low word),
 * one-constant reduction, exp(r) - 1 = r + r^2 (1/2 + r/6 + r^2 (1 + r/5) / 24). */
static double unit_table[64];
static float expf_unit(float x) {
  const double xd = (double)x;
  const double sh = fma(xd, 0x1.71547652b82fep+6, 0x1.8p+52);
  uint64_t bits;
  memcpy(&bits, &sh, 8);
  const int k = (int)(uint32_t)bits;
  const double kd = sh - 0x1.8p+52;
  const double r = fma(-kd, 0x1.62e42fefa39efp-7, xd);
  const double h = fma(r, 1.0 / 6.0, 0.5);
  const double q = fma(r, 0.2, 1.0) * (1.0 / 24.0);
  const double r2 = r * r;
  const double p = fma(r2, fma(r2, q, h), r);
  const double t = unit_table[k & 63];
  return (float)ldexp(fma(t, p, t), k >> 6);
}

/* The same evaluation with 2^(k/64) read whole from the wide table the density
 * pass builds in LDS (entry k = unit_table[k & 63] * 2^(k >> 6), k in
 * [-1293, 0]): no mask, no shift, no final ldexp.  Its bits must be expf_unit's,
 * and k must stay inside the table on every argument. */
#define WIDE 1294
static double wide_table[WIDE];
static long long wide_out_of_range;
static float expf_unit_wide(float x) {
  const double xd = (double)x;
  const double sh = fma(xd, 0x1.71547652b82fep+6, 0x1.8p+52);
  uint64_t bits;
  memcpy(&bits, &sh, 8);
  const int k = (int)(uint32_t)bits;
  if (k < -(WIDE - 1) || k > 0) {
#pragma omp atomic
    ++wide_out_of_range;
    return 0.0f;
  }
  const double kd = sh - 0x1.8p+52;
  const double r = fma(-kd, 0x1.62e42fefa39efp-7, xd);
  const double h = fma(r, 1.0 / 6.0, 0.5);
  const double q = fma(r, 0.2, 1.0) * (1.0 / 24.0);
  const double r2 = r * r;
  const double p = fma(r2, fma(r2, q, h), r);
  const double t = wide_table[k + WIDE - 1];
  return (float)fma(t, p, t);
}

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
  for (int i = 0; i < 64; ++i) unit_table[i] = (double)exp2l((long double)i / 64.0L);
  for (int k = -(WIDE - 1); k <= 0; ++k) wide_table[k + WIDE - 1] = ldexp(unit_table[k & 63], k >> 6);
  long long bad_u = 0, total_u = 0, bad_w = 0;
#pragma omp parallel for reduction(+ : bad_u, total_u, bad_w) schedule(dynamic, 1 << 20)
  for (long long u = 0x80000000ll; u < (1ll << 32); ++u) {
    const uint32_t b = (uint32_t)u;
    float x;
    memcpy(&x, &b, 4);
    if (!(x >= -14.0f)) continue;
    ++total_u;
    const float e = expf_unit(x), l = oracle_expf_longdouble(x), w = expf_unit_wide(x);
    if (memcmp(&e, &l, 4) != 0) ++bad_u;
    if (memcmp(&w, &e, 4) != 0) ++bad_w;
  }
  printf("exhaustive: %lld inputs in [-14, 0], unit-amplitude evaluation, %lld mismatches\n", total_u, bad_u);
  printf("exhaustive: %lld inputs in [-14, 0], wide-table evaluation against the unit one, %lld mismatches"
         " (%lld indexes outside the table)\n", total_u, bad_w, wide_out_of_range);
  return bad != 0 || bad_u != 0 || bad_w != 0 || wide_out_of_range != 0;
}
