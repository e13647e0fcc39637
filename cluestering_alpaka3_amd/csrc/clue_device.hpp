// clue_device.hpp -- device arithmetic of the MI355X CLUE engine.
//
// Every helper reproduces the reference's float arithmetic operation by
// operation (the library is compiled with -ffp-contract=off, so no multiply-add
// is fused), which is what makes the engine's tile ids, densities,
// nearest-highers and labels bit-identical to the serial backend.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <climits>

namespace clue_hip {

constexpr int kMaxDim = 10;
constexpr int kWave = 64;

// Constant address space view of read-only device memory: a wave-uniform index
// into it compiles to scalar (SMEM) loads that feed VALU operands directly.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)(p);
}

// float -> int32 as static_cast<int> behaves on the reference's x86-64 host
// (cvttss2si): truncation toward zero, INT_MIN for NaN or out-of-range values.
__device__ __forceinline__ int32_t f2i(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? static_cast<int32_t>(v) : INT32_MIN;
}

// One event of a batch: an independent clustering problem with its own tile
// grid.  The host fills the shape (SetupTiles.hpp:23-26 evaluated with the
// host's double pow, as the reference does); the bounding-box kernel fills the
// extremes and tile sizes (CoordinateExtremes, ComputeTiles.hpp:21-26).
struct EventDesc {
  int32_t begin;        // first point of the event in the batch arrays
  int32_t n;            // points
  int32_t npd;          // tiles per dimension
  int32_t tile_base;    // first global tile id of the event
  int32_t ntiles;       // npd^D
  int32_t chunk_begin;  // chunks of the event (bounding-box partials)
  int32_t chunk_end;
  int32_t pad;
  int32_t stride[kMaxDim];  // npd^(D-1-d): integer form of powf(npd, D-d-1), TilesView.hpp:56,67
  float mn[kMaxDim];
  float mx[kMaxDim];
  float ts[kMaxDim];
};

// At most kChunk consecutive points of one event: the unit of the per-point
// kernels that need their event's grid uniformly.
struct Chunk {
  int32_t event;
  int32_t begin;
  int32_t end;
  int32_t pad;
};

// Wrapped (periodic) coordinates, Clusterer::setWrappedCoordinates.
struct Wrap {
  uint8_t w[kMaxDim];
};

// The grid of one event as the kernels use it (wave-uniform values).
template <int D>
struct EvGrid {
  int32_t npd;
  int32_t tile_base;
  int32_t ntiles;
  int32_t begin;
  int32_t stride[D];
  float mn[D], mx[D], ts[D];
  bool wrapped[D];
};

// TilesView::normalizeCoordinate (TilesView.hpp:93-101).
__device__ __forceinline__ float normalize_coord(float coord, float mn, float mx) {
  const float range = mx - mn;
  float rem = coord - static_cast<float>(f2i(coord / range)) * range;
  if (rem >= mx)
    rem -= range;
  else if (rem < mn)
    rem += range;
  return rem;
}

// TilesView::getBin (TilesView.hpp:37-51).
__device__ __forceinline__ int32_t tile_bin(float coord, float mn, float mx, float ts, int32_t npd,
                                            bool wrapped) {
  int32_t b = wrapped ? f2i((normalize_coord(coord, mn, mx) - mn) / ts) : f2i((coord - mn) / ts);
  b = b < npd - 1 ? b : npd - 1;
  b = b > 0 ? b : 0;
  return b;
}

// ---------------------------------------------------------------------------
// Engine exponential: the correctly rounded float exp, evaluated with IEEE
// double operations only (__builtin_fma is one correctly rounded fused
// multiply-add).  exp(x) = 2^(k/64) * exp(r) with k = rint(x*64/ln2),
// r = x - k*ln2/64 (two-constant Cody-Waite split; hi has 33 significant bits so
// k*hi is exact), exp(r)-1 by a degree-5 Taylor polynomial in Horner form
// (|r| <= ln2/128, truncation below 4e-17), 2^(j/64) from a table of correctly
// rounded doubles; one final rounding double -> float.  The oracle evaluates
// the same sequence.  Written without branches: out-of-range and NaN inputs
// are resolved by selects after a clamped evaluation.
// ---------------------------------------------------------------------------
static __device__ __constant__ double kExp2Table[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0};

// Inputs are clamped to [-110, 89] (fmaxf/fminf map NaN to a bound): the
// double-precision evaluation then rounds to 0 below -103.97 and to +inf above
// 88.72 by itself, and NaN is restored at the end.
__device__ __forceinline__ float engine_expf(float x) {
  const float xc = __builtin_fminf(__builtin_fmaxf(x, -110.0f), 89.0f);
  const double xd = static_cast<double>(xc);
  const double kd = __builtin_rint(xd * 0x1.71547652b82fep+6);
  const int k = static_cast<int>(kd);
  double r = __builtin_fma(-kd, 0x1.62e42fefp-7, xd);
  r = __builtin_fma(-kd, 0x1.473de6af278edp-40, r);
  double p = __builtin_fma(r, 1.0 / 120.0, 1.0 / 24.0);
  p = __builtin_fma(r, p, 1.0 / 6.0);
  p = __builtin_fma(r, p, 0.5);
  p = __builtin_fma(r * r, p, r);
  const int j = k & 63;
  const int m = (k - j) / 64;
  const double t = kExp2Table[j];
  const float y = static_cast<float>(__builtin_ldexp(__builtin_fma(t, p, t), m));
  return x != x ? x : y;
}

// engine_expf for arguments known to be <= 0 and not NaN (the Gaussian and
// exponential kernels with finite positive width / decay): the same sequence
// without the upper clamp and the NaN pass-through, hence the same bits.
// table: kExp2Table or a copy of it (in LDS for the density pass).
__device__ __forceinline__ float engine_expf_nonpos(float x, const double* table = kExp2Table) {
  const double xd = static_cast<double>(__builtin_fmaxf(x, -110.0f));
  const double kd = __builtin_rint(xd * 0x1.71547652b82fep+6);
  const int k = static_cast<int>(kd);
  double r = __builtin_fma(-kd, 0x1.62e42fefp-7, xd);
  r = __builtin_fma(-kd, 0x1.473de6af278edp-40, r);
  double p = __builtin_fma(r, 1.0 / 120.0, 1.0 / 24.0);
  p = __builtin_fma(r, p, 1.0 / 6.0);
  p = __builtin_fma(r, p, 0.5);
  p = __builtin_fma(r * r, p, r);
  const int j = k & 63;
  const int m = (k - j) / 64;
  const double t = table[j];
  return static_cast<float>(__builtin_ldexp(__builtin_fma(t, p, t), m));
}

// The correctly rounded float exp of x in [-14, 0], a shorter evaluation for
// the kernels' in-radius terms once the host has bounded their argument: k is
// x*64/ln2 rounded to an integer by adding 1.5*2^52 in the fused multiply-add
// (k is then the low word of the sum, no conversion), the reduction uses one
// constant (ln2/64 rounded to double: |k| <= 1293 keeps its error below 2^-49
// of the result) and exp(r) - 1 is r + r^2 (1/2 + r/6 + r^2 (1 + r/5) / 24),
// whose additive constants 0.5 and 1.0 are inline operands (no register
// copies).  Checked on every float in [-14, 0] against long double rounded once
// to float: the same bits as engine_expf, the correctly rounded value.
__device__ __forceinline__ float engine_expf_unit(float x, const double* table = kExp2Table) {
  const double xd = static_cast<double>(x);
  const double sh = __builtin_fma(xd, 0x1.71547652b82fep+6, 0x1.8p+52);
  const int k = static_cast<int>(static_cast<uint32_t>(__builtin_bit_cast(uint64_t, sh)));
  const double kd = sh - 0x1.8p+52;
  const double r = __builtin_fma(-kd, 0x1.62e42fefa39efp-7, xd);
  const double h = __builtin_fma(r, 1.0 / 6.0, 0.5);
  const double q = __builtin_fma(r, 0.2, 1.0) * (1.0 / 24.0);
  const double r2 = r * r;
  const double p = __builtin_fma(r2, __builtin_fma(r2, q, h), r);
  const double t = table[k & 63];
  return static_cast<float>(__builtin_ldexp(__builtin_fma(t, p, t), k >> 6));
}

// The unit exponential reads 2^(k/64) whole from a wide table: x in [-14, 0]
// rounds x*64/ln2 to k in [-1293, 0], and entry k = kExp2Table[k & 63] * 2^(k >> 6)
// is an exact scaling of the 64-entry table.
constexpr int kExpWide = 1294;
__device__ __forceinline__ double exp2_wide_entry(int k) {
  return __builtin_ldexp(kExp2Table[k & 63], k >> 6);
}

// engine_expf_unit with 2^(k/64) from the wide table (wide0: its entry k = 0, so
// that k indexes it directly): no index mask, no shift, no final ldexp.  The
// fused multiply-add of a table entry scaled by 2^m is 2^m times that of the
// unscaled entry (no result lies near the subnormal range), so the bits are
// engine_expf_unit's (checked on every float in [-14, 0], together with the
// range of k, by oracle/check_expf_exhaustive.c).
__device__ __forceinline__ float engine_expf_unit_wide(float x, const double* wide0) {
  const double xd = static_cast<double>(x);
  const double sh = __builtin_fma(xd, 0x1.71547652b82fep+6, 0x1.8p+52);
  const int k = static_cast<int>(static_cast<uint32_t>(__builtin_bit_cast(uint64_t, sh)));
  const double kd = sh - 0x1.8p+52;
  const double r = __builtin_fma(-kd, 0x1.62e42fefa39efp-7, xd);
  const double h = __builtin_fma(r, 1.0 / 6.0, 0.5);
  const double q = __builtin_fma(r, 0.2, 1.0) * (1.0 / 24.0);
  const double r2 = r * r;
  const double p = __builtin_fma(r2, __builtin_fma(r2, q, h), r);
  const double t = wide0[k];
  return static_cast<float>(__builtin_fma(t, p, t));
}

// Correctly rounded sqrt of x in [2^-96, FLT_MAX]: the hardware estimate (within
// one ulp) corrected by the signs of the residuals of its two neighbours -- the
// IEEE sqrtf lowering without its small-input scaling and special-value paths.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
  const float r = __builtin_amdgcn_sqrtf(x);
  const float rm = __int_as_float(__float_as_int(r) - 1);
  const float rp = __int_as_float(__float_as_int(r) + 1);
  const float em = __builtin_fmaf(-rm, r, x);
  const float ep = __builtin_fmaf(-rp, r, x);
  float y = em <= 0.f ? rm : r;
  y = ep > 0.f ? rp : y;
  return y;
}

// Correctly rounded sqrtf of a finite s >= +0: below 2^-96 the fast evaluation
// runs on s * 2^64, whose square root is exactly 2^32 times sqrtf(s) (both are
// normal floats there, and scaling by a power of two commutes with rounding).
__device__ __forceinline__ float sqrt_rn_finite(float s) {
  const bool tiny = s < 0x1p-96f;
  const float r = sqrt_rn_normal(tiny ? s * 0x1p64f : s);
  return tiny ? r * 0x1p-32f : r;
}

// Gaussian kernel argument -(a*a) / den as the correctly rounded float quotient,
// computed as (float)((double)num * (1/den)) with 1/den rounded to double on the
// host: a quotient of two floats lies at least ~2^-48 (relative) from any float
// rounding boundary, and the double product is within 2^-52 of it, so the
// final rounding lands on the same float as the IEEE division.
__device__ __forceinline__ float quotient_via_rcp(float num, double rcp_den) {
  return static_cast<float>(static_cast<double>(num) * rcp_den);
}

// Convolutional kernels (core/detail/ConvolutionalKernel.hpp:16-65), j != i.
enum { KFLAT = 0, KGAUSS = 1, KEXP = 2 };
// rcp_den: 1 / ((2*sigma)*sigma) rounded to double (Gaussian only).
// NONPOS: the exp argument is known to be <= 0 and not NaN (checked on the host).
// UNIT: also the amplitude is exactly 1 (amplitude * e == e) and the argument is
// known to lie in [-14, 0] for every in-radius distance (checked on the host);
// 2 (Gaussian): also the mean is 0 and (2 sigma) sigma a power of two.
// table: with UNIT, the wide table of 2^(k/64) at its entry k = 0 (every caller
// passes it; the default is for the other evaluations); otherwise kExp2Table or
// a copy of it.
template <int KERNEL, bool NONPOS = false, int UNIT = 0>
__device__ __forceinline__ float kernel_value(float dist, float p0, float p1, float p2, double rcp_den,
                                              const double* table = kExp2Table) {
  if constexpr (KERNEL == KFLAT) {
    return p0;
  } else if constexpr (KERNEL == KGAUSS && UNIT == 2) {
    // mean 0 and (2 sigma) sigma = 2^e: the quotient -(d*d) / 2^e is the exact
    // scaling of d*d by -2^-e (rcp_den), so one float product gives its bits
    return engine_expf_unit_wide((dist * dist) * static_cast<float>(-rcp_den), table);
  } else if constexpr (KERNEL == KGAUSS) {
    const float a = dist - p0;
    const float num = (-a) * a;  // -(d-mu)*(d-mu), divided by (2*sigma)*sigma
    const float x = quotient_via_rcp(num, rcp_den);
    if constexpr (UNIT) return engine_expf_unit_wide(x, table);
    return p2 * (NONPOS ? engine_expf_nonpos(x, table) : engine_expf(x));
  } else {
    const float x = (-p0) * dist;
    if constexpr (UNIT) return engine_expf_unit_wide(x, table);
    return p1 * (NONPOS ? engine_expf_nonpos(x, table) : engine_expf(x));
  }
}

// Unit evaluations that are exactly 1 at every squared distance below 2^-96,
// where the fast square root is not exact: the host takes UNIT 2 for the
// Gaussian only with 1/((2 sigma) sigma) <= 2^70, and UNIT for the exponential
// only with a rate <= 2^22, so the argument stays above -2^-25 there, whose
// correctly rounded exp is 1.  They need no test of the distance against
// 2^-96, and no test of j == i (the self term's kernel value is 1).
template <int KERNEL, int UNIT>
constexpr bool unit_tiny_exact() {
  return (UNIT != 0 && KERNEL == KEXP) || (UNIT == 2 && KERNEL == KGAUSS);
}

// Distance metrics (core/DistanceMetrics.hpp).  The squared-term sum is the
// right fold of meta::accumulate (internal/meta/accumulate.hpp:14-15).
enum { MEUCLID = 0, MWEUCLID = 1, MPERIODIC = 2, MMANHATTAN = 3, MCHEBYSHEV = 4, MWCHEBYSHEV = 5 };

template <int D, int METRIC>
__device__ __forceinline__ float metric_distance(const float* a, const float* b, const float* mp) {
  if constexpr (METRIC == MEUCLID || METRIC == MWEUCLID || METRIC == MPERIODIC) {
    float s = 0.f;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const float df = a[d] - b[d];
      float t;
      if constexpr (METRIC == MEUCLID) {
        t = df * df;
      } else if constexpr (METRIC == MWEUCLID) {
        t = (mp[d] * df) * df;
      } else {
        const float diff = __builtin_fabsf(df);
        const float per = mp[d] - diff;
        const float pd = per < diff ? per : diff;  // std::min(diff, per)
        t = pd * pd;
      }
      s = t + s;  // t_{D-1} + 0.f first, as the fold's init
    }
    return __builtin_sqrtf(s);
  } else if constexpr (METRIC == MMANHATTAN) {
    float s = 0.f;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) s = __builtin_fabsf(a[d] - b[d]) + s;
    return s;
  } else {
    float m = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float v = __builtin_fabsf(a[d] - b[d]);
      if constexpr (METRIC == MWCHEBYSHEV) v = mp[d] * v;
      m = (m < v) ? v : m;  // std::max(m, v)
    }
    return m;
  }
}

// Squared Euclidean distance with the reference's right fold; the candidate
// test d <= r becomes s2 <= thr with thr = max{s : sqrtf(s) <= r} (exact,
// sqrtf is correctly rounded and monotone).
template <int D>
__device__ __forceinline__ float squared_distance(const float* a, const float* b) {
  float s = 0.f;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    const float df = a[d] - b[d];
    const float t = df * df;
    s = (d == D - 1) ? t : t + s;
  }
  return s;
}

// Largest float s with sqrtf(s) < delta (-1 when none): an update of the
// nearest-higher (d < delta) is possible only for candidates with s2 <= it.
__device__ __forceinline__ float sqrt_lt_bound(float delta) {
  if (!(delta > 0.f)) return -1.f;
  float t = delta * delta;
  if (!(t <= FLT_MAX)) t = FLT_MAX;
  while (t > 0.f && !(__builtin_sqrtf(t) < delta)) t = __int_as_float(__float_as_int(t) - 1);
  if (!(__builtin_sqrtf(t) < delta)) return -1.f;
  while (t < FLT_MAX) {
    const float nt = __int_as_float(__float_as_int(t) + 1);
    if (!(__builtin_sqrtf(nt) < delta)) break;
    t = nt;
  }
  return t;
}

// The floats s >= 0 with sqrtf(s) == d, for d the sqrtf of a finite float s2 >= 0,
// form an interval [lo, hi].  sqrtf rounds to nearest and the square root of a
// float is never half-way between two floats, so s lies in it exactly when
// (d - g_lo/2)^2 < s < (d + g_hi/2)^2, with g_lo and g_hi the gaps from d to the
// floats beside it.  Both midpoints have at most 25 significant bits and their
// squares at most 50, so the squares are exact in double precision and are not
// floats themselves.
__device__ __forceinline__ void sqrt_class(float d, float& lo, float& hi) {
  const double dd = d;
  const double mh = 0.5 * (dd + static_cast<double>(__uint_as_float(__float_as_uint(d) + 1u)));
  const double th = mh * mh;
  float h = static_cast<float>(th);
  if (static_cast<double>(h) > th) h = __uint_as_float(__float_as_uint(h) - 1u);
  hi = h;
  float l = 0.f;
  if (d > 0.f) {
    const double ml = 0.5 * (dd + static_cast<double>(__uint_as_float(__float_as_uint(d) - 1u)));
    const double tl = ml * ml;
    l = static_cast<float>(tl);
    if (static_cast<double>(l) < tl) l = __uint_as_float(__float_as_uint(l) + 1u);
  }
  lo = l;
}

// Every squared distance x with sqrtf(x) <= sqrtf(s) has x <= not_farther_than(s):
// a float's sqrtf class [lo, hi] spans less than 2^-21 of it (the midpoints of
// sqrtf(s) and its neighbours, squared); the margin of 2^-20 (and of the least
// normal float, for tiny s) holds it with room for the rounding of the product.
// The nearest-higher searches bound their candidates' squared distances with it
// and take a square root only for those within.
__device__ __forceinline__ float not_farther_than(float s) { return s * (1.f + 0x1p-20f) + 0x1p-126f; }

// Every squared distance x >= 0 with x < strictly_nearer(s) has sqrtf(x) < sqrtf(s)
// (s finite, >= 0): the class of sqrtf(s) reaches less than 2^-21 of s below it,
// and the margin of 2^-20 minus the least normal float (one rounding) stays below
// that; for s below 2^-126 the result is negative, for +inf it is +inf.  Checked
// on every finite float.
// The nearest-higher searches replace their best by a candidate below it without
// taking a square root; only candidates between it and not_farther_than(s) need
// the square roots compared.
__device__ __forceinline__ float strictly_nearer(float s) { return __builtin_fmaf(s, 1.f - 0x1p-20f, -0x1p-126f); }

// (sqrtf(s2), key) < (sqrtf(s2b), best) in lexicographic order (s2, s2b finite, >= 0)
__device__ __forceinline__ bool sqrt_nearer_or_earlier(float s2, int64_t key, float s2b, int64_t best) {
  const float dc = sqrt_rn_finite(s2), db = sqrt_rn_finite(s2b);
  return (dc < db) | ((dc == db) & (key < best));
}

// Order key of the nearest-higher relation (ClusteringKernels.hpp:132-134): point j
// is higher than point i when rho_j > rho_i, or rho_j == rho_i > 0 and j > i.  With
//   rank_key(rho, id) = ord(rho) << 32 | id,
// ord monotone in rho, equal for -0 and +0, and 0 (below every number) for NaN,
// j is higher than i exactly when rank_key(rho_j, j) > higher_threshold(rank_key(rho_i, i)).
__device__ __forceinline__ uint64_t rank_key(float rho, int32_t id) {
  uint32_t b = __float_as_uint(rho);
  if (b == 0x80000000u) b = 0u;
  uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  if (rho != rho) ord = 0u;
  return (static_cast<uint64_t>(ord) << 32) | static_cast<uint32_t>(id);
}
// rho > 0: the key itself (equal densities are ordered by id); rho <= 0: above
// every key of that density (no tie-break); NaN: above every key.
__device__ __forceinline__ uint64_t higher_threshold(uint64_t key) {
  const uint32_t ord = static_cast<uint32_t>(key >> 32);
  if (ord == 0u) return ~uint64_t{0};
  return ord > 0x80000000u ? key : (key | 0xffffffffull);
}

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v)));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v >> 32)));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

}  // namespace clue_hip
