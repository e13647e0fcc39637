// clue_oracle.cpp -- CPU restatement of CLUEstering 2.9.0's serial backend.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Never linked into the product.
// Built with g++ -O2 -ffp-contract=off (no fast-math): every float expression
// below is evaluated as written, one rounding per operation.
//
// Reference files restated (paths relative to the CLUEstering tree):
//   core/detail/SetupTiles.hpp:16-49, core/detail/ComputeTiles.hpp:13-27  grid sizing
//   data_structures/internal/TilesView.hpp:37-101                          bins, box, global bin
//   data_structures/AssociationMap.hpp:122-251, detail/AssociationMap.hpp  tile / follower CSR
//   core/detail/ClusteringKernels.hpp:18-281                               K4..K7
//   core/detail/ConvolutionalKernel.hpp:10-65                              kernels
//   core/DistanceMetrics.hpp:33-230, internal/meta/accumulate.hpp:10-18     metrics, right fold
//   data_structures/internal/PointsCommon.hpp:75-86                        index -1 sentinel
//   data_structures/internal/SeedArray.hpp:33-41, core/detail/SetupSeeds.hpp:10-20  seed capacity
//   core/detail/Clusterer.hpp:134-171                                      stage order
#include "clue_oracle.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int g_mutation = 0;  // planted mistakes for the pinning tests (0 = faithful)
enum {
  MUT_LEFT_FOLD = 1,        // distance squares summed left to right
  MUT_NH_LAST_TIE = 2,      // nearest-higher keeps the last minimum instead of the first
  MUT_SELF_KERNEL = 3,      // self term weighted by the kernel instead of 1
  MUT_SEED_STRICT_RHOC = 4, // seed needs rho > rhoc instead of rho >= rhoc
  MUT_NO_CAPACITY = 5,      // seed capacity ignored
  MUT_DENSITY_STRICT = 6,   // density counts d < dc instead of d <= dc
  MUT_NH_BOX_DC = 7,        // nearest-higher search box built with dc instead of dm
  MUT_TILE_DESCENDING = 8,  // points inside a tile in descending id order
  MUT_NO_EQUAL_RHO_TIE = 9, // equal-rho tie clause dropped from the "higher" predicate
  MUT_GRID_PLUS_ONE = 10,   // one more tile per dimension than the reference rule
};

// float -> int32 as the reference's static_cast<int> behaves on x86-64
// (cvttss2si): truncation toward zero, and INT_MIN for NaN / out of range.
inline int32_t f2i(float v) {
  if (v >= -2147483648.0f && v < 2147483648.0f) return static_cast<int32_t>(v);
  return INT32_MIN;
}

// ----------------------------------------------------------------------------
// Engine exponential: the correctly rounded float exp, evaluated with IEEE
// double operations only (each std::fma is one correctly rounded fused
// multiply-add).  exp(x) = 2^(k/64) * exp(r), k = rint(x*64/ln2),
// r = x - k*ln2/64 (two-constant Cody-Waite split; hi has 33 significant bits so
// k*hi is exact), exp(r) - 1 by a degree-5 Taylor polynomial in Horner form
// (|r| <= ln2/128, truncation < 4e-17), 2^(j/64) from a 64-entry table rounded
// to double, one final rounding double -> float.  The GPU engine evaluates the
// identical sequence, so both produce identical bits.
// ----------------------------------------------------------------------------
double g_exp_table[64];
bool g_exp_table_ready = false;

void init_exp_table() {
  if (g_exp_table_ready) return;
  for (int i = 0; i < 64; ++i)
    g_exp_table[i] = static_cast<double>(exp2l(static_cast<long double>(i) / 64.0L));
  g_exp_table_ready = true;
}

constexpr double kInvLn2x64 = 0x1.71547652b82fep+6;  // 64/ln2 rounded to double
constexpr double kLn2by64Hi = 0x1.62e42fefp-7;       // 33 significant bits
constexpr double kLn2by64Lo = 0x1.473de6af278edp-40; // ln2/64 - hi, rounded

inline float expf_engine(float x) {
  if (x != x) return x;
  if (x > 89.0f) return std::numeric_limits<float>::infinity();
  if (x < -110.0f) return 0.0f;
  const double xd = static_cast<double>(x);
  const double kd = std::rint(xd * kInvLn2x64);
  const int k = static_cast<int>(kd);
  double r = std::fma(-kd, kLn2by64Hi, xd);
  r = std::fma(-kd, kLn2by64Lo, r);
  double p = std::fma(r, 1.0 / 120.0, 1.0 / 24.0);
  p = std::fma(r, p, 1.0 / 6.0);
  p = std::fma(r, p, 0.5);
  p = std::fma(r * r, p, r);
  const int j = k & 63;
  const int m = (k - j) / 64;
  const double t = g_exp_table[j];
  return static_cast<float>(std::ldexp(std::fma(t, p, t), m));
}

inline float expf_sel(float x, int mode) {
  return mode == ORACLE_EXP_LIBM ? ::expf(x) : expf_engine(x);
}

// ----------------------------------------------------------------------------
// Metrics (core/DistanceMetrics.hpp).  meta::accumulate is a right fold:
// t_0 + (t_1 + (... + (t_{D-1} + 0))) (internal/meta/accumulate.hpp:14-15).
// ----------------------------------------------------------------------------
template <int D>
inline float metric_distance(const float* a, const float* b, const oracle_params& p) {
  switch (p.metric_type) {
    case ORACLE_METRIC_EUCLIDEAN:
    case ORACLE_METRIC_WEIGHTED_EUCLIDEAN:
    case ORACLE_METRIC_PERIODIC_EUCLIDEAN: {
      float t[D];
      for (int d = 0; d < D; ++d) {
        const float df = a[d] - b[d];
        if (p.metric_type == ORACLE_METRIC_EUCLIDEAN) {
          t[d] = df * df;  // DistanceMetrics.hpp:49
        } else if (p.metric_type == ORACLE_METRIC_WEIGHTED_EUCLIDEAN) {
          t[d] = (p.metric_params[d] * df) * df;  // :93 w*(a-b)*(a-b)
        } else {
          const float diff = std::fabs(df);  // :136-138
          const float per = p.metric_params[d] - diff;
          const float pd = std::min(diff, per);
          t[d] = pd * pd;
        }
      }
      float s = 0.f;
      if (g_mutation == MUT_LEFT_FOLD) {
        for (int d = 0; d < D; ++d) s = s + t[d];
      } else {
        for (int d = D - 1; d >= 0; --d) s = t[d] + s;
      }
      return std::sqrt(s);
    }
    case ORACLE_METRIC_MANHATTAN: {  // :161-162, right fold of |a-b|
      float s = 0.f;
      for (int d = D - 1; d >= 0; --d) s = std::fabs(a[d] - b[d]) + s;
      return s;
    }
    case ORACLE_METRIC_CHEBYSHEV:  // :184-185 via meta::maximum (maximum.hpp:9-19)
    case ORACLE_METRIC_WEIGHTED_CHEBYSHEV: {
      float mx = 0.f;
      for (int d = 0; d < D; ++d) {
        float v = std::fabs(a[d] - b[d]);
        if (p.metric_type == ORACLE_METRIC_WEIGHTED_CHEBYSHEV) v = p.metric_params[d] * v;
        mx = std::max(mx, v);
      }
      return mx;
    }
  }
  return 0.f;
}

// Convolutional kernels, core/detail/ConvolutionalKernel.hpp:16-65.
inline float kernel_value(const oracle_params& p, float dist, int32_t point_id, int32_t j) {
  if (point_id == j && g_mutation != MUT_SELF_KERNEL) return 1.f;
  switch (p.kernel_type) {
    case ORACLE_KERNEL_FLAT:
      return p.kernel_params[0];
    case ORACLE_KERNEL_GAUSSIAN: {
      const float mean = p.kernel_params[0], sd = p.kernel_params[1], amp = p.kernel_params[2];
      const float a = dist - mean;
      const float num = (-a) * a;           // -(d-mu)*(d-mu)
      const float den = (2.f * sd) * sd;    // 2*sigma*sigma
      return amp * expf_sel(num / den, p.exp_mode);
    }
    case ORACLE_KERNEL_EXPONENTIAL: {
      const float avg = p.kernel_params[0], amp = p.kernel_params[1];
      return amp * expf_sel((-avg) * dist, p.exp_mode);
    }
  }
  return 0.f;
}

// ----------------------------------------------------------------------------
// Tiles (TilesView.hpp:37-101)
// ----------------------------------------------------------------------------
struct Tiles {
  int D;
  int32_t npd, ntiles;
  float mn[ORACLE_MAX_DIM], mx[ORACLE_MAX_DIM], ts[ORACLE_MAX_DIM];
  uint8_t wrapped[ORACLE_MAX_DIM];
  int32_t stride[ORACLE_MAX_DIM];
  std::vector<int32_t> offsets, indexes;

  float normalize(float coord, int d) const {  // TilesView.hpp:93-101
    const float range = mx[d] - mn[d];
    float rem = coord - static_cast<float>(f2i(coord / range)) * range;
    if (rem >= mx[d])
      rem -= range;
    else if (rem < mn[d])
      rem += range;
    return rem;
  }
  int32_t bin(float coord, int d) const {  // TilesView.hpp:37-51
    int32_t b;
    if (wrapped[d])
      b = f2i((normalize(coord, d) - mn[d]) / ts[d]);
    else
      b = f2i((coord - mn[d]) / ts[d]);
    b = std::min(b, npd - 1);
    b = std::max(b, 0);
    return b;
  }
  int32_t global_bin(const float* c) const {  // TilesView.hpp:53-61 (integer strides)
    int32_t g = 0;
    for (int d = 0; d < D; ++d) g += stride[d] * bin(c[d], d);
    return g;
  }
  int32_t global_bin_by_bin(const int32_t* b) const {  // TilesView.hpp:63-70
    int32_t g = 0;
    for (int d = 0; d < D; ++d) g += stride[d] * (wrapped[d] ? (b[d] % npd) : b[d]);
    return g;
  }
  void search_box(const float* c, float r, int32_t* lo, int32_t* hi) const {  // :72-82
    for (int d = 0; d < D; ++d) {
      int32_t inf = bin(c[d] - r, d);
      int32_t sup = bin(c[d] + r, d);
      if (wrapped[d] && inf > sup) sup += npd;
      lo[d] = inf;
      hi[d] = sup;
    }
  }
};

// SetupTiles.hpp:16-49 + ComputeTiles.hpp:13-27
void grid_setup(int D, int32_t n, const float* coords, int32_t ppb, oracle_grid* g) {
  int32_t nt = static_cast<int32_t>(std::ceil(static_cast<float>(n) / static_cast<float>(ppb)));
  int32_t npd = static_cast<int32_t>(std::ceil(std::pow(static_cast<double>(nt), 1. / D)));
  if (g_mutation == MUT_GRID_PLUS_ONE) npd += 1;
  nt = static_cast<int32_t>(std::pow(static_cast<double>(npd), static_cast<double>(D)));
  g->npd = npd;
  g->ntiles = nt;
  for (int d = 0; d < D; ++d) {
    const float* c = coords + static_cast<size_t>(d) * n;
    float lo = c[0], hi = c[0];
    for (int32_t i = 1; i < n; ++i) {  // std::ranges::min_element / max_element
      if (c[i] < lo) lo = c[i];
      if (hi < c[i]) hi = c[i];
    }
    g->minmax[2 * d] = lo;
    g->minmax[2 * d + 1] = hi;
    g->tile_size[d] = (hi - lo) / static_cast<float>(npd);
  }
}

void tiles_init(Tiles& t, int D, const oracle_grid& g, const oracle_params& p) {
  t.D = D;
  t.npd = g.npd;
  t.ntiles = g.ntiles;
  for (int d = 0; d < D; ++d) {
    t.mn[d] = g.minmax[2 * d];
    t.mx[d] = g.minmax[2 * d + 1];
    t.ts[d] = g.tile_size[d];
    t.wrapped[d] = p.wrapped[d];
  }
  int32_t s = 1;
  for (int d = D - 1; d >= 0; --d) {
    t.stride[d] = s;
    s *= g.npd;
  }
}

// Tile CSR: counting sort by global bin, ascending point id inside a tile
// (serial execution of KernelFillAssociator, detail/AssociationMap.hpp:50-67).
void tiles_fill(Tiles& t, int32_t n, const float* coords) {
  const int D = t.D;
  std::vector<int32_t> key(n);
  t.offsets.assign(static_cast<size_t>(t.ntiles) + 1, 0);
  float c[ORACLE_MAX_DIM];
  for (int32_t i = 0; i < n; ++i) {
    for (int d = 0; d < D; ++d) c[d] = coords[static_cast<size_t>(d) * n + i];
    key[i] = t.global_bin(c);
    t.offsets[key[i] + 1] += 1;
  }
  for (int32_t k = 0; k < t.ntiles; ++k) t.offsets[k + 1] += t.offsets[k];
  std::vector<int32_t> cursor(t.offsets.begin(), t.offsets.end() - 1);
  t.indexes.assign(n, 0);
  for (int32_t i = 0; i < n; ++i) t.indexes[cursor[key[i]]++] = i;
  if (g_mutation == MUT_TILE_DESCENDING)
    for (int32_t k = 0; k < t.ntiles; ++k)
      std::reverse(t.indexes.begin() + t.offsets[k], t.indexes.begin() + t.offsets[k + 1]);
}

template <int D>
inline void load_point(const float* coords, int32_t n, int32_t i, float* out) {
  if (i == -1) {  // PointsCommon.hpp:75-81
    for (int d = 0; d < D; ++d) out[d] = FLT_MAX;
    return;
  }
  for (int d = 0; d < D; ++d) out[d] = coords[static_cast<size_t>(d) * n + i];
}

// Visit the tiles of a search box in the reference order: lexicographic over
// (b_0, ..., b_{D-1}) with dim 0 outermost (ClusteringKernels.hpp:47-62).
template <int D, typename F>
inline void for_each_box_tile(const Tiles& t, const int32_t* lo, const int32_t* hi, F&& f) {
  int32_t b[D];
  for (int d = 0; d < D; ++d) {
    if (lo[d] > hi[d]) return;
    b[d] = lo[d];
  }
  while (true) {
    f(t.global_bin_by_bin(b));
    int d = D - 1;
    while (d >= 0) {
      if (++b[d] <= hi[d]) break;
      b[d] = lo[d];
      --d;
    }
    if (d < 0) return;
  }
}

struct Counters {
  int64_t cand_rho = 0, cand_nh = 0, in_rho = 0, in_nh = 0;
};

// The points in tile order (the CSR's slot order): coordinates and weights of
// slot s are those of point indexes[s].  The loops below read candidates from
// this copy -- the values and their visit order are the reference's; only the
// memory they are read from is contiguous per tile.
struct TileOrdered {
  std::vector<float> x;  // D x n, x[d * n + s]
  std::vector<float> w;
};

template <int D>
void tile_order(const Tiles& t, int32_t n, const float* coords, const float* w, TileOrdered& to) {
  to.x.resize(static_cast<size_t>(D) * n);
  to.w.resize(n);
  for (int32_t s = 0; s < n; ++s) {
    const int32_t j = t.indexes[s];
    for (int d = 0; d < D; ++d) to.x[static_cast<size_t>(d) * n + s] = coords[static_cast<size_t>(d) * n + j];
    to.w[s] = w[j];
  }
}

template <int D>
inline void slot_point(const TileOrdered& to, int32_t n, int32_t s, float* out) {
  for (int d = 0; d < D; ++d) out[d] = to.x[static_cast<size_t>(d) * n + s];
}

// K4: KernelCalculateLocalDensity (ClusteringKernels.hpp:18-107), one point.
// Every candidate of the search box contributes ((float)(d <= dc) * k) * w_j
// (:38-44).  For a candidate outside dc whose distance is a number, whose weight
// is finite and whose kernel value is finite for every such distance
// (`bounded`), that term is a signed zero and adding it leaves rho_i unchanged
// (rho_i starts at +0 and a float sum never becomes -0 from such terms), so the
// kernel is not evaluated for it.
template <int D>
inline float density_point(const Tiles& t, const TileOrdered& to, int32_t n, const float* coords,
                           const oracle_params& p, bool bounded, int32_t i, Counters& cnt) {
  float ci[D], cj[D];
  load_point<D>(coords, n, i, ci);
  float rho_i = 0.f;
  int32_t lo[D], hi[D];
  t.search_box(ci, p.dc, lo, hi);
  for_each_box_tile<D>(t, lo, hi, [&](int32_t tile) {
    for (int32_t s = t.offsets[tile]; s < t.offsets[tile + 1]; ++s) {
      const int32_t j = t.indexes[s];
      slot_point<D>(to, n, s, cj);
      const float wj = to.w[s];
      const float dist = metric_distance<D>(ci, cj, p);
      const bool inside = (g_mutation == MUT_DENSITY_STRICT) ? (dist < p.dc) : (dist <= p.dc);
      cnt.cand_rho += 1;
      cnt.in_rho += inside ? 1 : 0;
      if (!inside && bounded && dist == dist && std::isfinite(wj)) continue;
      const float k = kernel_value(p, dist, i, j);
      rho_i += static_cast<float>(static_cast<int>(inside)) * k * wj;
    }
  });
  return rho_i;
}

// K5: KernelCalculateNearestHigher (ClusteringKernels.hpp:109-212), one point.
// rho_s: the densities in tile order.
template <int D>
inline int32_t nh_point(const Tiles& t, const TileOrdered& to, int32_t n, const float* coords, const float* rho,
                        const float* rho_s, const oracle_params& p, int32_t i, float* delta_out, Counters& cnt) {
  float ci[D], cj[D];
  load_point<D>(coords, n, i, ci);
  float delta_i = std::numeric_limits<float>::max();
  int32_t nh_i = -1;
  const float rho_i = rho[i];
  int32_t lo[D], hi[D];
  t.search_box(ci, g_mutation == MUT_NH_BOX_DC ? p.dc : p.dm, lo, hi);
  for_each_box_tile<D>(t, lo, hi, [&](int32_t tile) {
    for (int32_t s = t.offsets[tile]; s < t.offsets[tile + 1]; ++s) {
      const int32_t j = t.indexes[s];
      const float rho_j = rho_s[s];
      bool higher = (rho_j > rho_i);
      if (g_mutation != MUT_NO_EQUAL_RHO_TIE)
        higher = higher || ((rho_j == rho_i) && (rho_j > 0.f) && (j > i));
      slot_point<D>(to, n, s, cj);
      const float dist = metric_distance<D>(ci, cj, p);
      cnt.cand_nh += 1;
      cnt.in_nh += (dist <= p.dm) ? 1 : 0;
      if (higher && dist <= p.dm) {
        const bool better = (g_mutation == MUT_NH_LAST_TIE) ? (dist <= delta_i) : (dist < delta_i);
        if (better) {
          delta_i = dist;
          nh_i = j;
        }
      }
    }
  });
  if (delta_out) *delta_out = delta_i;
  return nh_i;
}

// Kernel values finite for every distance in [0, +inf] (see density_point).
bool kernel_bounded(const oracle_params& p) {
  const float* k = p.kernel_params;
  switch (p.kernel_type) {
    case ORACLE_KERNEL_FLAT:
      return std::isfinite(k[0]);
    case ORACLE_KERNEL_GAUSSIAN: {
      const float den = (2.f * k[1]) * k[1];
      return std::isfinite(k[0]) && std::isfinite(k[2]) && std::isfinite(den) && den > 0.f;
    }
    case ORACLE_KERNEL_EXPONENTIAL:
      return std::isfinite(k[0]) && k[0] > 0.f && std::isfinite(k[1]);
  }
  return false;
}

// K6 + K7 (ClusteringKernels.hpp:214-281) with SeedArray capacity semantics.
template <int D>
void seeds_and_labels(int32_t n, const float* coords, const float* rho, std::vector<int32_t>& nh,
                      const oracle_params& p, int64_t seed_candidates, int32_t* is_seed,
                      int32_t* cluster, oracle_stats* st, bool parallel) {
  // SetupSeeds.hpp:14-18: grow-only capacity.
  int64_t cap = seed_candidates;
  if (p.seed_capacity_floor > cap) cap = p.seed_capacity_floor;
  const bool unbounded = (p.seed_mode == ORACLE_SEEDS_UNBOUNDED) || g_mutation == MUT_NO_CAPACITY;
  std::vector<int32_t> seeds;
  int64_t n_seeds = 0;
  float ci[D], cn[D];
  for (int32_t i = 0; i < n; ++i) {  // serial order: ascending i
    cluster[i] = -1;
    load_point<D>(coords, n, i, ci);
    load_point<D>(coords, n, nh[i], cn);
    const float dist = metric_distance<D>(ci, cn, p);
    const bool seed = (dist > p.seed_dc) &&
                      ((g_mutation == MUT_SEED_STRICT_RHOC) ? (rho[i] > p.rhoc) : (rho[i] >= p.rhoc));
    if (seed) {
      is_seed[i] = 1;
      nh[i] = -1;
      ++n_seeds;
      if (unbounded || static_cast<int64_t>(seeds.size()) < cap) seeds.push_back(i);
    } else {
      is_seed[i] = 0;
    }
  }
  // Followers CSR (AssociationMap.hpp:193-251): children of each point, ascending id.
  std::vector<int32_t> off(static_cast<size_t>(n) + 1, 0), child(n);
  for (int32_t i = 0; i < n; ++i)
    if (nh[i] >= 0) off[nh[i] + 1] += 1;
  for (int32_t i = 0; i < n; ++i) off[i + 1] += off[i];
  {
    std::vector<int32_t> cur(off.begin(), off.end() - 1);
    for (int32_t i = 0; i < n; ++i)
      if (nh[i] >= 0) child[cur[nh[i]]++] = i;
  }
  // KernelAssignClusters: per-seed DFS with an explicit stack.
  int64_t max_stack = 0;
  const int64_t ns = static_cast<int64_t>(seeds.size());
#pragma omp parallel for schedule(dynamic, 16) reduction(max : max_stack) if (parallel)
  for (int64_t k = 0; k < ns; ++k) {
    std::vector<int32_t> stack;
    stack.push_back(seeds[k]);
    cluster[seeds[k]] = static_cast<int32_t>(k);
    int64_t local_max = 1;
    while (!stack.empty()) {
      const int32_t top = stack.back();
      stack.pop_back();
      const int32_t lab = cluster[top];
      for (int32_t c = off[top]; c < off[top + 1]; ++c) {
        cluster[child[c]] = lab;
        stack.push_back(child[c]);
      }
      local_max = std::max<int64_t>(local_max, static_cast<int64_t>(stack.size()));
    }
    max_stack = std::max(max_stack, local_max);
  }
  if (st) {
    st->seed_candidates = seed_candidates;
    st->seed_capacity = unbounded ? n_seeds : cap;
    st->n_seeds = n_seeds;
    st->n_stored_seeds = ns;
    st->n_dropped_seeds = n_seeds - ns;
    st->max_dfs_stack = max_stack;
    int64_t nc = 0;
    for (int32_t i = 0; i < n; ++i) nc += cluster[i] != -1;
    st->n_clustered = nc;
  }
}

template <int D>
int run_pipeline(int32_t n, const float* coords, const float* w, const oracle_params& p,
                 bool parallel, int nthreads, float* rho_out, int32_t* nh_out, float* delta_out,
                 int32_t* is_seed, int32_t* cluster, oracle_stats* st, int32_t* tile_offsets,
                 int32_t* tile_indexes, oracle_grid* grid_out) {
  init_exp_table();
  oracle_grid g;
  grid_setup(D, n, coords, p.points_per_tile, &g);
  Tiles t;
  tiles_init(t, D, g, p);
  tiles_fill(t, n, coords);
  if (grid_out) *grid_out = g;
  if (tile_offsets) std::copy(t.offsets.begin(), t.offsets.end(), tile_offsets);
  if (tile_indexes) std::copy(t.indexes.begin(), t.indexes.end(), tile_indexes);

  std::vector<float> rho(n), delta(n), rho_s(n);
  std::vector<int32_t> nh(n);
  TileOrdered to;
  tile_order<D>(t, n, coords, w, to);
  const bool bounded = kernel_bounded(p);
  int64_t cand_rho = 0, cand_nh = 0, in_rho = 0, in_nh = 0, seed_candidates = 0;
#ifdef _OPENMP
  if (parallel && nthreads > 0) omp_set_num_threads(nthreads);
#endif
  // every point's density and nearest-higher are independent of the others';
  // the queries run in tile order (neighbouring queries read the same tiles)
#pragma omp parallel if (parallel)
  {
    Counters c;
#pragma omp for schedule(dynamic, 256)
    for (int32_t q = 0; q < n; ++q) {
      const int32_t i = t.indexes[q];
      rho[i] = density_point<D>(t, to, n, coords, p, bounded, i, c);
    }
#pragma omp for schedule(static)
    for (int32_t q = 0; q < n; ++q) rho_s[q] = rho[t.indexes[q]];
#pragma omp for schedule(dynamic, 256) reduction(+ : seed_candidates)
    for (int32_t q = 0; q < n; ++q) {
      const int32_t i = t.indexes[q];
      nh[i] = nh_point<D>(t, to, n, coords, rho.data(), rho_s.data(), p, i, &delta[i], c);
      seed_candidates += (nh[i] == -1);
    }
#pragma omp atomic
    cand_rho += c.cand_rho;
#pragma omp atomic
    cand_nh += c.cand_nh;
#pragma omp atomic
    in_rho += c.in_rho;
#pragma omp atomic
    in_nh += c.in_nh;
  }
  if (delta_out) std::copy(delta.begin(), delta.end(), delta_out);
  std::vector<int32_t> seed_flag(n);
  seeds_and_labels<D>(n, coords, rho.data(), nh, p, seed_candidates,
                      is_seed ? is_seed : seed_flag.data(), cluster, st, parallel);
  if (st) {
    st->candidate_pairs_density = cand_rho;
    st->candidate_pairs_nh = cand_nh;
    st->inradius_pairs_density = in_rho;
    st->inradius_pairs_nh = in_nh;
  }
  if (rho_out) std::copy(rho.begin(), rho.end(), rho_out);
  if (nh_out) std::copy(nh.begin(), nh.end(), nh_out);
  return 0;
}

bool params_valid(int32_t ndim, int32_t n, const oracle_params* p) {
  if (!p || ndim < 1 || ndim > ORACLE_MAX_DIM || n <= 0) return false;
  // Clusterer.hpp:32-35
  if (p->dc <= 0.f || p->rhoc < 0.f || p->dm <= 0.f || p->seed_dc <= 0.f ||
      p->points_per_tile <= 0)
    return false;
  return true;
}

template <typename F>
int dispatch_dim(int32_t ndim, F&& f) {
  switch (ndim) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 7: return f(std::integral_constant<int, 7>{});
    case 8: return f(std::integral_constant<int, 8>{});
    case 9: return f(std::integral_constant<int, 9>{});
    case 10: return f(std::integral_constant<int, 10>{});
  }
  return -1;
}

}  // namespace

extern "C" {

int oracle_grid_setup(int32_t ndim, int32_t n, const float* coords, int32_t ppb,
                      oracle_grid* grid) {
  if (ndim < 1 || ndim > ORACLE_MAX_DIM || n <= 0 || ppb <= 0 || !grid) return -1;
  grid_setup(ndim, n, coords, ppb, grid);
  return 0;
}

int oracle_tiles(int32_t ndim, int32_t n, const float* coords, const oracle_params* p,
                 const oracle_grid* grid, int32_t* offsets, int32_t* indexes) {
  if (ndim < 1 || ndim > ORACLE_MAX_DIM || n <= 0 || !p || !grid) return -1;
  Tiles t;
  tiles_init(t, ndim, *grid, *p);
  tiles_fill(t, n, coords);
  std::copy(t.offsets.begin(), t.offsets.end(), offsets);
  std::copy(t.indexes.begin(), t.indexes.end(), indexes);
  return 0;
}

int oracle_make_clusters(int32_t ndim, int32_t n, const float* coords, const float* weights,
                         const oracle_params* p, float* rho, int32_t* nh, float* delta,
                         int32_t* is_seed, int32_t* cluster, oracle_stats* stats,
                         int32_t* tile_offsets, int32_t* tile_indexes, oracle_grid* grid_out) {
  if (!params_valid(ndim, n, p) || !cluster) return -1;
  return dispatch_dim(ndim, [&](auto dimc) {
    return run_pipeline<decltype(dimc)::value>(n, coords, weights, *p, false, 0, rho, nh, delta,
                                               is_seed, cluster, stats, tile_offsets,
                                               tile_indexes, grid_out);
  });
}

int oracle_make_clusters_omp(int32_t ndim, int32_t n, const float* coords,
                             const float* weights, const oracle_params* p, int32_t nthreads,
                             float* rho, int32_t* nh, int32_t* is_seed, int32_t* cluster,
                             oracle_stats* stats) {
  if (!params_valid(ndim, n, p) || !cluster) return -1;
  return dispatch_dim(ndim, [&](auto dimc) {
    return run_pipeline<decltype(dimc)::value>(n, coords, weights, *p, true, nthreads, rho, nh,
                                               nullptr, is_seed, cluster, stats, nullptr,
                                               nullptr, nullptr);
  });
}

int oracle_make_clusters_par(int32_t ndim, int32_t n, const float* coords, const float* weights,
                             const oracle_params* p, int32_t nthreads, float* rho, int32_t* nh, float* delta,
                             int32_t* is_seed, int32_t* cluster, oracle_stats* stats, int32_t* tile_offsets,
                             int32_t* tile_indexes, oracle_grid* grid_out) {
  if (!params_valid(ndim, n, p) || !cluster) return -1;
  return dispatch_dim(ndim, [&](auto dimc) {
    return run_pipeline<decltype(dimc)::value>(n, coords, weights, *p, true, nthreads, rho, nh, delta,
                                               is_seed, cluster, stats, tile_offsets, tile_indexes,
                                               grid_out);
  });
}

int oracle_bruteforce_density_fp64(int32_t ndim, int32_t n, const float* coords,
                                   const float* weights, const oracle_params* p, double* rho) {
  if (!params_valid(ndim, n, p)) return -1;
  if (p->metric_type != ORACLE_METRIC_EUCLIDEAN) return -2;
  for (int32_t i = 0; i < n; ++i) {
    double acc = 0.0;
    for (int32_t j = 0; j < n; ++j) {
      double s = 0.0;
      for (int d = 0; d < ndim; ++d) {
        const double df = static_cast<double>(coords[static_cast<size_t>(d) * n + i]) -
                          static_cast<double>(coords[static_cast<size_t>(d) * n + j]);
        s += df * df;
      }
      const double dist = std::sqrt(s);
      if (dist > static_cast<double>(p->dc)) continue;
      double k;
      if (i == j) {
        k = 1.0;
      } else if (p->kernel_type == ORACLE_KERNEL_FLAT) {
        k = p->kernel_params[0];
      } else if (p->kernel_type == ORACLE_KERNEL_GAUSSIAN) {
        const double a = dist - p->kernel_params[0];
        const double sd = p->kernel_params[1];
        k = p->kernel_params[2] * std::exp(-(a * a) / (2.0 * sd * sd));
      } else {
        k = p->kernel_params[1] * std::exp(-static_cast<double>(p->kernel_params[0]) * dist);
      }
      acc += k * static_cast<double>(weights[j]);
    }
    rho[i] = acc;
  }
  return 0;
}

int oracle_bruteforce_nh(int32_t ndim, int32_t n, const float* coords, const float* rho,
                         const oracle_params* p, int32_t* nh, float* delta) {
  if (!params_valid(ndim, n, p)) return -1;
  return dispatch_dim(ndim, [&](auto dimc) {
    constexpr int D = decltype(dimc)::value;
    oracle_grid g;
    grid_setup(D, n, coords, p->points_per_tile, &g);
    Tiles t;
    tiles_init(t, D, g, *p);
    // Slot of each point in the visit order: (global bin, id).
    std::vector<int64_t> slot_key(n);
    float c[D], ci[D];
    for (int32_t i = 0; i < n; ++i) {
      load_point<D>(coords, n, i, c);
      slot_key[i] = static_cast<int64_t>(t.global_bin(c)) * (static_cast<int64_t>(n) + 1) + i;
    }
    for (int32_t i = 0; i < n; ++i) {
      load_point<D>(coords, n, i, ci);
      int32_t lo[D], hi[D];
      t.search_box(ci, p->dm, lo, hi);
      float best = std::numeric_limits<float>::max();
      int32_t best_j = -1;
      for (int32_t j = 0; j < n; ++j) {
        load_point<D>(coords, n, j, c);
        bool in_box = true;
        for (int d = 0; d < D; ++d) {
          const int32_t b = t.bin(c[d], d);
          if (b < lo[d] || b > hi[d]) in_box = false;
        }
        if (!in_box) continue;
        const bool higher =
            rho[j] > rho[i] || (rho[j] == rho[i] && rho[j] > 0.f && j > i);
        if (!higher) continue;
        const float dist = metric_distance<D>(ci, c, *p);
        if (!(dist <= p->dm)) continue;
        if (dist < best || (dist == best && best_j >= 0 && slot_key[j] < slot_key[best_j])) {
          best = dist;
          best_j = j;
        }
      }
      nh[i] = best_j;
      if (delta) delta[i] = best;
    }
    return 0;
  });
}

// Margins of a result (rho as computed by the pipeline, float32) against the
// decisions that depend on it; see oracle_margins in clue_oracle.h.  The search
// boxes, visit order and predicates are those of K4/K5
// (ClusteringKernels.hpp:18-212); the FP64 density keeps the reference's
// membership test (float distance <= dc) and evaluates distance, kernel and sum
// in double from the float inputs.
int oracle_parity_margins(int32_t ndim, int32_t n, const float* coords, const float* weights,
                          const oracle_params* p, const float* rho, int32_t nthreads, double band,
                          oracle_margins* out) {
  if (!params_valid(ndim, n, p) || !rho || !out) return -1;
  if (p->metric_type != ORACLE_METRIC_EUCLIDEAN) return -2;
  init_exp_table();
  return dispatch_dim(ndim, [&](auto dimc) {
    constexpr int D = decltype(dimc)::value;
    oracle_grid g;
    grid_setup(D, n, coords, p->points_per_tile, &g);
    Tiles t;
    tiles_init(t, D, g, *p);
    tiles_fill(t, n, coords);
    TileOrdered to;
    tile_order<D>(t, n, coords, weights, to);
    std::vector<float> rho_s(n);
    for (int32_t q = 0; q < n; ++q) rho_s[q] = rho[t.indexes[q]];
    int64_t rhoc_band = 0, dist_ties = 0, dist_near = 0, dens_near = 0, above = 0, boundary = 0;
    double max_rel = 0.0;
    const float rhoc = p->rhoc;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : rhoc_band, dist_ties, dist_near, dens_near, above, boundary) reduction(max : max_rel)
    for (int32_t q = 0; q < n; ++q) {  // queries in tile order
      const int32_t i = t.indexes[q];
      float ci[D], cj[D];
      load_point<D>(coords, n, i, ci);
      const float rho_i = rho[i];
      if (std::fabs(static_cast<double>(rho_i) - rhoc) <= band * rhoc) ++rhoc_band;
      // FP64 density with the reference's membership
      int32_t lo[D], hi[D];
      t.search_box(ci, p->dc, lo, hi);
      double acc = 0.0;
      for_each_box_tile<D>(t, lo, hi, [&](int32_t tile) {
        for (int32_t s = t.offsets[tile]; s < t.offsets[tile + 1]; ++s) {
          const int32_t j = t.indexes[s];
          slot_point<D>(to, n, s, cj);
          const float df = metric_distance<D>(ci, cj, *p);
          double s2 = 0.0;
          for (int d = 0; d < D; ++d) {
            const double dd = static_cast<double>(ci[d]) - static_cast<double>(cj[d]);
            s2 += dd * dd;
          }
          const double dist = std::sqrt(s2);
          const bool in_f = df <= p->dc;
          if (in_f != (dist <= static_cast<double>(p->dc))) ++boundary;
          if (!in_f) continue;
          double k;
          if (i == j) {
            k = 1.0;
          } else if (p->kernel_type == ORACLE_KERNEL_FLAT) {
            k = p->kernel_params[0];
          } else if (p->kernel_type == ORACLE_KERNEL_GAUSSIAN) {
            const double a = dist - p->kernel_params[0];
            const double sd = p->kernel_params[1];
            k = p->kernel_params[2] * std::exp(-(a * a) / (2.0 * sd * sd));
          } else {
            k = p->kernel_params[1] * std::exp(-static_cast<double>(p->kernel_params[0]) * dist);
          }
          acc += k * static_cast<double>(to.w[s]);
        }
      });
      const double rel = acc != 0.0 ? std::fabs(static_cast<double>(rho_i) - acc) / std::fabs(acc)
                                    : std::fabs(static_cast<double>(rho_i));
      if (rel > max_rel) max_rel = rel;
      if (rel > band) ++above;
      // nearest-higher: best and runner-up higher candidates, and candidates at least
      // as near as the best whose density lies in the band around rho_i
      t.search_box(ci, p->dm, lo, hi);
      float d1 = std::numeric_limits<float>::max(), d2 = std::numeric_limits<float>::max();
      for_each_box_tile<D>(t, lo, hi, [&](int32_t tile) {
        for (int32_t s = t.offsets[tile]; s < t.offsets[tile + 1]; ++s) {
          const int32_t j = t.indexes[s];
          const float rho_j = rho_s[s];
          const bool higher = rho_j > rho_i || (rho_j == rho_i && rho_j > 0.f && j > i);
          if (!higher) continue;
          slot_point<D>(to, n, s, cj);
          const float dist = metric_distance<D>(ci, cj, *p);
          if (!(dist <= p->dm)) continue;
          if (dist < d1) {
            d2 = d1;
            d1 = dist;
          } else if (dist < d2) {
            d2 = dist;
          }
        }
      });
      if (d2 <= p->dm) {
        if (d2 == d1) ++dist_ties;
        if (static_cast<double>(d2) - d1 <= band * d1) ++dist_near;
      }
      const float reach = d1 <= p->dm ? d1 : p->dm;
      bool fragile = false;
      for_each_box_tile<D>(t, lo, hi, [&](int32_t tile) {
        for (int32_t s = t.offsets[tile]; s < t.offsets[tile + 1] && !fragile; ++s) {
          const int32_t j = t.indexes[s];
          if (j == i) continue;
          const double a = rho_s[s], b = rho_i;
          const double m = std::max(std::fabs(a), std::fabs(b));
          if (!(std::fabs(a - b) <= band * m)) continue;
          slot_point<D>(to, n, s, cj);
          if (metric_distance<D>(ci, cj, *p) <= reach) fragile = true;
        }
      });
      if (fragile) ++dens_near;
    }
    out->rhoc_band = rhoc_band;
    out->nh_distance_ties = dist_ties;
    out->nh_distance_near_ties = dist_near;
    out->nh_density_near_ties = dens_near;
    out->density_max_rel_err_fp64 = max_rel;
    out->density_above_band_fp64 = above;
    out->dc_boundary_pairs = boundary;
    return 0;
  });
}

float oracle_expf_engine(float x) {
  init_exp_table();
  return expf_engine(x);
}

float oracle_expf_longdouble(float x) {
  return static_cast<float>(expl(static_cast<long double>(x)));
}

void oracle_expf_engine_batch(const float* x, float* y, int64_t n) {
  init_exp_table();
  for (int64_t i = 0; i < n; ++i) y[i] = expf_engine(x[i]);
}

void oracle_expf_libm_batch(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = ::expf(x[i]);
}

float oracle_distance(int32_t ndim, const float* a, const float* b, const oracle_params* p) {
  float r = 0.f;
  dispatch_dim(ndim, [&](auto dimc) {
    r = metric_distance<decltype(dimc)::value>(a, b, *p);
    return 0;
  });
  return r;
}

int64_t oracle_read_csv(const char* path, int32_t ndim, float* coords, float* weights,
                        int64_t capacity) {
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  std::vector<char> line(1 << 16);
  if (!std::fgets(line.data(), static_cast<int>(line.size()), f)) {  // header
    std::fclose(f);
    return -1;
  }
  std::vector<std::vector<float>> cols(ndim + 1);
  while (std::fgets(line.data(), static_cast<int>(line.size()), f)) {
    char* s = line.data();
    if (*s == '\n' || *s == '\0' || *s == '\r') continue;
    for (int d = 0; d <= ndim; ++d) {
      char* end = nullptr;
      const float v = std::strtof(s, &end);  // std::stof == strtof
      cols[d].push_back(v);
      s = end;
      if (d < ndim) {
        while (*s && *s != ',') ++s;
        if (*s == ',') ++s;
      }
    }
  }
  std::fclose(f);
  const int64_t n = static_cast<int64_t>(cols[0].size());
  if (coords) {
    if (n > capacity) return -2;
    for (int d = 0; d < ndim; ++d)
      std::copy(cols[d].begin(), cols[d].end(), coords + static_cast<size_t>(d) * n);
    std::copy(cols[ndim].begin(), cols[ndim].end(), weights);
  }
  return n;
}

void oracle_set_mutation(int32_t mutation) { g_mutation = mutation; }

// Cluster -> points map of a label array: internal/MakeAssociator.hpp:36-42
// (bins = max label + 1) filled as the host AssociationMap fill does, a counting
// sort in ascending point order (data_structures/AssociationMap.hpp).  offsets:
// n+1 entries, indexes: n.  Returns the number of clusters.
int32_t oracle_cluster_map(int32_t n, const int32_t* cluster, int32_t* offsets, int32_t* indexes,
                           int32_t* clustered) {
  int32_t mx = -1, cnt = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (cluster[i] > mx) mx = cluster[i];
    cnt += cluster[i] > -1;  // non_negative, utils/detail/get_clusters.hpp:14-17
  }
  const int32_t bins = mx + 1;
  std::vector<int32_t> sizes(static_cast<size_t>(bins), 0);
  for (int32_t i = 0; i < n; ++i)
    if (cluster[i] >= 0) ++sizes[cluster[i]];
  offsets[0] = 0;
  for (int32_t b = 0; b < bins; ++b) offsets[b + 1] = offsets[b] + sizes[b];
  std::vector<int32_t> cursor(offsets, offsets + bins);
  for (int32_t i = 0; i < n; ++i)
    if (cluster[i] >= 0) indexes[cursor[cluster[i]]++] = i;
  *clustered = cnt;
  return bins;
}

// cluster_centroids (utils/detail/cluster_centroid.hpp:38-65): per dimension,
// centroids[cluster][dim] += coord over the points in index order (from 0.f),
// then divided by (float)count(cluster).  centroids: n_clusters * ndim.
void oracle_cluster_centroids(int32_t ndim, int32_t n, const float* coords, const int32_t* cluster,
                              int32_t n_clusters, const int32_t* offsets, float* centroids) {
  for (int32_t c = 0; c < n_clusters * ndim; ++c) centroids[c] = 0.f;
  for (int32_t dim = 0; dim < ndim; ++dim) {
    const float* x = coords + static_cast<size_t>(dim) * n;
    for (int32_t i = 0; i < n; ++i)
      if (cluster[i] >= 0) centroids[static_cast<size_t>(cluster[i]) * ndim + dim] += x[i];
    for (int32_t c = 0; c < n_clusters; ++c)
      centroids[static_cast<size_t>(c) * ndim + dim] /= static_cast<float>(offsets[c + 1] - offsets[c]);
  }
}

}  // extern "C"
