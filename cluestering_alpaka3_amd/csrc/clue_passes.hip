// clue_passes.hip -- the templated kernels of one dimension D = CLUE_DIM: binning,
// gathers, the density pass (K4, ClusteringKernels.hpp:18-107), the nearest-higher
// searches (K5, ClusteringKernels.hpp:109-212) and the seed classification (K6,
// ClusteringKernels.hpp:214-244).  The build compiles this file once per dimension
// 1..10 (-DCLUE_DIM=D) so that the instantiations compile in parallel.
#include "clue_internal.hpp"

#ifndef CLUE_DIM
#error "compile with -DCLUE_DIM=<1..10>"
#endif

namespace clue_hip {

// ===========================================================================
// Tile id per point (getGlobalBin, TilesView.hpp:53-61), its event, its Morton
// key inside the tile, and a packed copy of the point for the tile-order gather.
// ===========================================================================

template <int D>
__global__ __launch_bounds__(kThreads) void k_bin(Coords c, const float* w, const Chunk* chunks,
                                                   const EventDesc* events, Wrap wrap, int32_t* key,
                                                   int morton, int sub_bits, int32_t* pev, float* packed) {
  constexpr int S = packed_stride<D>();
  const Chunk ch = chunks[blockIdx.x];
  const EvGrid<D> g = load_event<D>(events, ch.event, wrap);
  for (int i = ch.begin + threadIdx.x; i < ch.end; i += kThreads) {
    int32_t gb = g.tile_base;
    uint32_t code = 0;
    float v[S];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float x = c.p[d][i];
      v[d] = x;
      const int32_t b = tile_bin(x, g.mn[d], g.mx[d], g.ts[d], g.npd, g.wrapped[d]);
      gb += g.stride[d] * b;
      if (morton) {
        // position inside the tile on a 2^sub_bits grid, interleaved into a Morton code
        const float f = (x - g.mn[d]) / g.ts[d] - static_cast<float>(b);
        int32_t sub = f2i(f * static_cast<float>(1 << sub_bits));
        sub = sub < 0 ? 0 : (sub > (1 << sub_bits) - 1 ? (1 << sub_bits) - 1 : sub);
        for (int bit = 0; bit < sub_bits; ++bit) code |= ((sub >> bit) & 1u) << (bit * D + d);
      }
    }
    key[i] = gb;
    pev[i] = ch.event;
    v[D] = w[i];
    v[D + 1] = __int_as_float(morton ? static_cast<int32_t>((static_cast<uint32_t>(gb) << (sub_bits * D)) | code) : 0);
#pragma unroll
    for (int k = D + 2; k < S; ++k) v[k] = 0.f;
    float4* out = reinterpret_cast<float4*>(packed + static_cast<int64_t>(i) * S);
#pragma unroll
    for (int k = 0; k < S / 4; ++k) out[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
  }
}

// Tile-ordered SoA copy: xs[d][slot] = x[d][perm[slot]] (the coalesced tile
// array), read from the packed copy: one random 16-byte read per 4 values.
// Also flags the tiles holding a weight that is infinite or NaN (tnf, cleared by
// k_tile_offsets).  The reference adds ((float)(d <= dc) * k) * w_j for every
// candidate of the search box (ClusteringKernels.hpp:42); the density pass skips
// the out-of-radius terms, which are +-0 -- except that 0 * w_j is NaN when w_j
// is infinite or NaN, and the density pass applies those terms in flagged tiles.
template <int D>
__global__ __launch_bounds__(kThreads) void k_gather(const float* packed, const int32_t* perm,
                                                      const int32_t* skey, int32_t n, int64_t ld, float* xs,
                                                      float* ws, float* hw, float h, int32_t* key_nh_s,
                                                      int32_t* tnf) {
  constexpr int S = packed_stride<D>();
  const int s = logical_block(1) * kThreads + threadIdx.x;
  if (s >= n) return;
  const int32_t id = perm[s];
  const float4* in = reinterpret_cast<const float4*>(packed + static_cast<int64_t>(id) * S);
  float v[S];
#pragma unroll
  for (int k = 0; k < S / 4; ++k) {
    const float4 q = in[k];
    v[4 * k] = q.x;
    v[4 * k + 1] = q.y;
    v[4 * k + 2] = q.z;
    v[4 * k + 3] = q.w;
  }
#pragma unroll
  for (int d = 0; d < D; ++d) xs[d * ld + s] = v[d];
  ws[s] = v[D];
  hw[s] = h * v[D];  // FlatKernel term k*w with k = h, precomputed once per point
  if (!__builtin_isfinite(v[D])) tnf[skey[s]] = 1;
  if (key_nh_s) key_nh_s[s] = __float_as_int(v[D + 1]);  // (tile, Morton code) key, in tile order
}

// Size class of each tile-ordered point's search box of radius r (ceil(log2) of
// its tile count): the launch order of the tile-walk passes in more than three
// dimensions, where one lane walks its own box and a wave lasts as long as its
// largest box -- queries sorted by class keep a wave's boxes alike.
template <int D>
__global__ __launch_bounds__(kThreads) void k_box_class(const PassArgs a, float r, int tile_bits, int32_t* key) {
  const int s = blockIdx.x * kThreads + threadIdx.x;
  const bool valid = s < a.n;
  const int sq = valid ? s : a.n - 1;
  float q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = a.xs[d][sq];
  const int my_e = a.pev[a.perm[sq]];
  for_each_event(valid, my_e, [&](int e, bool act) {
    const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
    int lo[D], hi[D];
    search_box<D>(q, r, g, lo, hi);
    int64_t count = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) count *= (hi[d] >= lo[d] ? hi[d] - lo[d] + 1 : 0);
    int c = 0;
    while (c < 31 && (int64_t{1} << c) < count) ++c;
    // then by the box's first tile: queries with equal boxes walk them in step
    if (act) key[s] = tile_bits > 0 ? (c << tile_bits) | tile_index<D>(lo, g) : c;
  });
}

// QNH: the lanes take their queries in Morton order (spatially coherent waves,
// so the lanes of a wave share most of their search box); otherwise in tile order.
// UNIT: Euclidean metric, exp argument in [-14, 0] and amplitude 1 (exp_unit);
// 2: also a Gaussian of mean 0 and power-of-two (2 sigma) sigma.
// SPLIT (filtered walk only): the workgroup's four waves share one wave's 64
// queries.  Each builds the same ring of survivors; every batch of 32 is
// evaluated a quarter per wave into LDS, and wave 0 adds the 32 terms in ring
// order -- the same sums, with the chain of a query's terms cut to the adds.
// For calls whose waves do not fill the GPU, where a pass lasts as long as its
// heaviest wave.
//
// Occupancy: in 2D with the flat kernel, eight waves per SIMD (64 VGPRs and 78
// SGPRs, no spills; seven at 68 VGPRs and 102 SGPRs without the request): the
// 4096-event calorimeter batch's pass 4.15 -> 3.84 ms.
template <int D, int METRIC, int KERNEL, bool UNION, bool QNH, int UNIT = 0, bool FILT = false, bool SPLIT = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(D == 2 && KERNEL == KFLAT ? 8 : 1)))
void k_density(const PassArgs a) {
  // the exponential's table in LDS: 2^(j/64), or for UNIT the wide table of
  // 2^(k/64) over the unit evaluation's whole range of k
  __shared__ double s_exp2[KERNEL == KFLAT ? 1 : (UNIT != 0 ? kExpWide : 64)];
  if constexpr (KERNEL != KFLAT) {
    if constexpr (UNIT != 0) {
      for (int i = threadIdx.x; i < kExpWide; i += kThreads) s_exp2[i] = exp2_wide_entry(i - (kExpWide - 1));
    } else {
      if (threadIdx.x < 64) s_exp2[threadIdx.x] = kExp2Table[threadIdx.x];
    }
    __syncthreads();
  }
  const double* const exp_table = UNIT != 0 ? s_exp2 + (kExpWide - 1) : s_exp2;
  // the filtered union walk's per-wave ring of candidates (coordinates, value, slot)
  constexpr bool kFilter = FILT && UNION && METRIC == MEUCLID;
  static_assert(!SPLIT || (kFilter && QNH), "the split walk is a variant of the filtered walk");
  constexpr int kRing = 128;         // entries per wave; a batch never wraps (kRing % kFB == 0)
  constexpr int kFB = SPLIT ? 32 : 8;  // candidates per evaluation batch
  constexpr int kWaves = kThreads / kWave;
  constexpr int kShare = kFB / kWaves;  // split: candidates of a batch per wave
  // split: the terms of a batch, [batch parity][candidate][lane]
  __shared__ float s_terms[SPLIT ? 2 : 1][SPLIT ? kFB : 1][SPLIT ? kWave : 1];
  // kernel value 1 at distance 0, the self term's value: the self test is not needed
  constexpr bool kSelfFree = unit_tiny_exact<KERNEL, UNIT>();
  __shared__ float4 s_ring[kFilter ? kThreads / kWave : 1][kFilter ? kRing : 1];
  __shared__ int32_t s_ringj[kFilter && D == 3 ? kThreads / kWave : 1][kFilter && D == 3 ? kRing : 1];
  const int qb = a.qorder ? a.qorder[blockIdx.x] : logical_block(a.xcd & 1);
  const int lane = __lane_id();
  const int wv = threadIdx.x / kWave;
  const int pos = SPLIT ? qb * kWave + lane : qb * kThreads + static_cast<int>(threadIdx.x);
  const bool valid = pos < a.n;
  const int pq = valid ? pos : a.n - 1;
  float q[D];
  float w_self;  // the point's own term is 1 * w_i (kernel value 1 for j == i)
  int s;         // the query's slot in the tile-ordered candidate layout
  int pid;       // the query's point id
  if constexpr (QNH) {
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = a.xn[d][pq];
    w_self = a.wn[pq];
    s = valid ? a.sn[pq] : -1;
    pid = a.pn[pq];
  } else {
    const int sq = a.qperm ? a.qperm[pq] : pq;
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = a.xs[d][sq];
    w_self = a.ws[sq];
    s = valid ? sq : -1;
    pid = a.perm[sq];
  }
  const int my_e = a.pev[pid];
  cptr<float> X[D];
#pragma unroll
  for (int d = 0; d < D; ++d) X[d] = as_const(a.xs[d]);
  cptr<float> W = as_const(a.ws);
  cptr<float> HW = as_const(a.hw);
  float rho = 0.f;
  bool poisoned = false;
  // The Euclidean flat-kernel loop reads the precomputed h*w_j; every other
  // combination reads w_j and evaluates the kernel.
  constexpr bool kUsesHW = (METRIC == MEUCLID) && (KERNEL == KFLAT);

  // One candidate: rho_i += ((float)(d <= dc) * k) * w_j in the visit order
  // (ClusteringKernels.hpp:38-44); out-of-radius terms are +0 and skipped (the
  // density starts at +0 and an add never makes it -0, so +0 terms change
  // nothing).  sink receives the in-radius term.
  auto add = [&](float t) { rho += t; };
  auto eval = [&](auto self_c, int j, const float* c, float wj, float hwj, auto&& sink) {
    constexpr bool SELF = decltype(self_c)::value;  // may this tile hold the lane's own point?
    if constexpr (METRIC == MEUCLID) {
      const float s2 = squared_distance<D>(q, c);
      if constexpr (KERNEL == KFLAT) {
        float t = hwj;
        if constexpr (SELF) t = (j == s) ? w_self : hwj;
        if (s2 <= a.thr) sink(t);
      } else {
        if (s2 <= a.thr) {
          float k = 1.f;
          if (!SELF || j != s) {
            if (unit_tiny_exact<KERNEL, UNIT>() || ((UNIT || a.exp_nonpos) && s2 >= 0x1p-96f)) {
              k = kernel_value<KERNEL, true, UNIT>(sqrt_rn_normal(s2), a.kp[0], a.kp[1], a.kp[2], a.kden_rcp,
                                                   exp_table);
            } else {
              k = kernel_value<KERNEL>(__builtin_sqrtf(s2), a.kp[0], a.kp[1], a.kp[2], a.kden_rcp);
            }
          }
          sink(k * wj);
        }
      }
    } else {
      const float dist = metric_distance<D, METRIC>(q, c, a.mp);
      if (dist <= a.radius) {
        const float k = (j == s) ? 1.f : kernel_value<KERNEL>(dist, a.kp[0], a.kp[1], a.kp[2], a.kden_rcp);
        sink(k * wj);
      }
    }
  };
  // uni: every lane on this tile (always so in the union walk), so candidates
  // sit at wave-uniform addresses and are read through the scalar cache
  auto run = [&](auto self_c, int beg, int end, bool need, bool uni) {
    int j = beg;
    if (UNION || uni) {
      constexpr int U = D <= 3 ? 16 : 8;  // candidates loaded ahead of their use
      const int ue = __builtin_amdgcn_readfirstlane(end);
      for (j = __builtin_amdgcn_readfirstlane(beg); j + U <= ue; j += U) {
        float c[U][D], wv4[U], hv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
          for (int d = 0; d < D; ++d) c[u][d] = X[d][j + u];
          if constexpr (kUsesHW) hv[u] = HW[j + u];
          else wv4[u] = W[j + u];
        }
        if (need) {
#pragma unroll
          for (int u = 0; u < U; ++u) eval(self_c, j + u, c[u], wv4[u], hv[u], add);
        }
      }
    }
    for (; j < end; ++j) {
      float c[D];
#pragma unroll
      for (int d = 0; d < D; ++d) c[d] = X[d][j];
      float wj = 0.f, hwj = 0.f;
      if constexpr (kUsesHW) hwj = HW[j];
      else wj = W[j];
      if (need) eval(self_c, j, c, wj, hwj, add);
    }
  };
  // out-of-radius candidates with an infinite or NaN weight add 0 * w = NaN
  auto poison_check = [&](int t, int beg, int end, bool need) {
    if (a.tnf[t] != 0) {
      for (int j = beg; j < end; ++j) {
        const float wj = W[j];
        if (need && !__builtin_isfinite(wj)) {
          float c[D];
#pragma unroll
          for (int d = 0; d < D; ++d) c[d] = X[d][j];
          bool in;
          if constexpr (METRIC == MEUCLID) in = squared_distance<D>(q, c) <= a.thr;
          else in = metric_distance<D, METRIC>(q, c, a.mp) <= a.radius;
          poisoned = poisoned || !in;
        }
      }
    }
  };
  for_each_event(valid, my_e, [&](int e, bool act) {
    const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
    int lo[D], hi[D];
    search_box<D>(q, a.radius, g, lo, hi);
    if constexpr (kFilter) {
      // Filtered union walk.  The wave drops every candidate farther than R (in
      // some dimension) from the box of its queries: no lane has it within the
      // radius.  The survivors are compacted in the visit order into a per-wave
      // ring in LDS (lane order = CSR order inside a chunk of a tile, tiles
      // ascending) and evaluated kFB at a time, every lane reading the same
      // candidate at one address.  When every lane passes box_exact, each lane's
      // in-radius candidates all lie in its own box, so the lane adds every
      // in-radius survivor, whatever tile it came from: exactly its box's terms,
      // in its box's order.  Otherwise the ring is emptied at the end of each
      // tile and only the lanes whose box holds the tile add its survivors.
      const float R = __builtin_fmaxf(a.radius, 0x1p-40f);
      const bool exact = __ballot(act && !box_exact<D>(q, a.radius, R, g, lo, hi)) == 0;
      float qlo[D], qhi[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        qlo[d] = uniform_f32(wave_min_f32(act ? q[d] : __builtin_inff()));
        qhi[d] = uniform_f32(wave_max_f32(act ? q[d] : -__builtin_inff()));
      }
      float4* ring = s_ring[wv];
      int32_t* ringj = s_ringj[D <= 2 ? 0 : wv];
      int head = 0, tail = 0;  // wave-uniform ring positions (head a multiple of kFB)
      auto put = [&](int at, const float* c, float v, int j) {
        float4 r4;
        if constexpr (D == 1) r4 = make_float4(c[0], v, __int_as_float(j), 0.f);
        else if constexpr (D == 2) r4 = make_float4(c[0], c[1], v, __int_as_float(j));
        else r4 = make_float4(c[0], c[1], c[2], v);
        ring[at] = r4;
        if constexpr (D == 3) ringj[at] = j;
      };
      int tb = 0;  // split: the parity of the batch (its term buffer)
      auto drain = [&](bool take) {
        __builtin_amdgcn_wave_barrier();
        constexpr int kOwn = SPLIT ? kShare : kFB;  // candidates this wave evaluates
        const int first = head + (SPLIT ? wv * kShare : 0);
        float4 r4[kOwn];
        int jj[kOwn];
#pragma unroll
        for (int u = 0; u < kOwn; ++u) {
          r4[u] = ring[first + u];
          if constexpr (D == 3) jj[u] = ringj[first + u];
        }
        auto cand = [&](int u, auto&& sink) {
          float c[D];
          float v;
          int j;
          if constexpr (D == 1) {
            c[0] = r4[u].x; v = r4[u].y; j = __float_as_int(r4[u].z);
          } else if constexpr (D == 2) {
            c[0] = r4[u].x; c[1] = r4[u].y; v = r4[u].z; j = __float_as_int(r4[u].w);
          } else {
            c[0] = r4[u].x; c[1] = r4[u].y; c[2] = r4[u].z; v = r4[u].w; j = jj[u];
          }
          if constexpr (kSelfFree) eval(std::false_type{}, j, c, v, v, sink);
          else eval(std::true_type{}, j, c, v, v, sink);
        };
        if constexpr (SPLIT) {
          // every wave sees the same ring, batches and flags (same queries, same
          // candidates), so all reach each barrier
#pragma unroll
          for (int u = 0; u < kOwn; ++u) {
            float t = 0.f;
            if (take) cand(u, [&](float x) { t = x; });
            s_terms[tb][wv * kShare + u][lane] = t;
          }
          __syncthreads();
          if (wv == 0) {
#pragma unroll
            for (int u = 0; u < kFB; ++u) rho += s_terms[tb][u][lane];
          }
          tb ^= 1;
        } else if (take) {
#pragma unroll
          for (int u = 0; u < kFB; ++u) cand(u, add);
        }
        __builtin_amdgcn_wave_barrier();
        head = (head + kFB) & (kRing - 1);
      };
      // the last batch is padded with candidates no lane adds (NaN distance, slot -2)
      auto flush = [&](bool take) {
        const int left = (tail - head) & (kRing - 1);
        if (left > 0) {
          if (lane < kFB - left) {
            float c[D];
#pragma unroll
            for (int d = 0; d < D; ++d) c[d] = __builtin_nanf("");
            put((tail + lane) & (kRing - 1), c, 0.f, -2);
          }
          tail = (head + kFB) & (kRing - 1);
          drain(take);
        }
      };
      visit_box<D, true>(act, lo, hi, g, as_const(a.offsets), [&](int t, int beg, int end, bool need) {
        // lanes of the wave's other events take no part
        const bool take = exact ? act : need;
        for (int base = beg; base < end; base += kWave) {
          const int j = base + lane;
          bool pass = j < end;
          const int jl = pass ? j : end - 1;  // loads stay inside the tile
          float c[D];
#pragma unroll
          for (int d = 0; d < D; ++d) c[d] = a.xs[d][jl];
          const float v = kUsesHW ? a.hw[jl] : a.ws[jl];
#pragma unroll
          for (int d = 0; d < D; ++d) pass = pass && (qlo[d] - c[d] <= R) && (c[d] - qhi[d] <= R);
          const uint64_t m = __ballot(pass);
          if (pass) put((tail + mbcnt64(m)) & (kRing - 1), c, v, j);
          tail = (tail + __builtin_popcountll(m)) & (kRing - 1);
          while (((tail - head) & (kRing - 1)) >= kFB) drain(take);
        }
        if (!exact) flush(take);
        poison_check(t, beg, end, need);
      });
      flush(act);
      return;
    }
    visit_box<D, UNION>(act, lo, hi, g, as_const(a.offsets), [&](int t, int beg, int end, bool need) {
      // Only the tile holding a lane's own point needs the j == i test.
      const bool self = !UNION || __ballot(s >= beg && s < end) != 0;
      bool uni = UNION;
      if constexpr (!UNION) uni = __ballot(t != __builtin_amdgcn_readfirstlane(t)) == 0;
      if (self) run(std::true_type{}, beg, end, need, uni);
      else run(std::false_type{}, beg, end, need, uni);
      poison_check(t, beg, end, need);
    });
  });
  if (poisoned) rho = __builtin_nanf("");
  if (valid && (!SPLIT || wv == 0)) {
    a.rho_s_out[s] = rho;
    if constexpr (QNH) {
      const int32_t id = a.pn[pos];
      a.hk_out[pos] = rank_key(rho, id);
      a.rho_out[id] = rho;
    } else {
      const int32_t id = a.perm[s];
      a.rho_out[id] = rho;
      a.hk_out[s] = rank_key(rho, id);  // tile order: the tile walk's candidates
    }
  }
}

// Density for small calls: one wave per query.  The 64 lanes take 64
// consecutive candidates of the query's visit order (the tiles of its search
// box in lexicographic order, each tile's points in CSR order) and evaluate
// their terms together; the in-radius terms are then added to the density one
// lane after another, in lane order -- the reference's order of the sum
// (ClusteringKernels.hpp:38-44).  With few queries the lane-per-query pass
// leaves most of the GPU idle and its time is the longest query's chain of
// terms; here that chain is one add per in-radius term.
template <int D, int METRIC, int KERNEL, bool QNH, int UNIT>
__global__ __launch_bounds__(kThreads) void k_density_wave(const PassArgs a) {
  __shared__ double s_exp2[KERNEL == KFLAT ? 1 : (UNIT != 0 ? kExpWide : 64)];
  if constexpr (KERNEL != KFLAT) {
    if constexpr (UNIT != 0) {
      for (int i = threadIdx.x; i < kExpWide; i += kThreads) s_exp2[i] = exp2_wide_entry(i - (kExpWide - 1));
    } else {
      if (threadIdx.x < 64) s_exp2[threadIdx.x] = kExp2Table[threadIdx.x];
    }
    __syncthreads();
  }
  const double* const exp_table = UNIT != 0 ? s_exp2 + (kExpWide - 1) : s_exp2;
  const int pos = uniform(static_cast<int>(blockIdx.x) * (kThreads / kWave) + static_cast<int>(threadIdx.x / kWave));
  if (pos >= a.n) return;  // the whole wave
  const int lane = __lane_id();
  float q[D];
  float w_self;
  int s, pid;
  if constexpr (QNH) {
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = as_const(a.xn[d])[pos];
    w_self = as_const(a.wn)[pos];
    s = as_const(a.sn)[pos];
    pid = as_const(a.pn)[pos];
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = as_const(a.xs[d])[pos];
    w_self = as_const(a.ws)[pos];
    s = pos;
    pid = as_const(a.perm)[pos];
  }
  const int e = as_const(a.pev)[pid];
  const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
  int lo[D], hi[D];
  search_box<D>(q, a.radius, g, lo, hi);
  constexpr bool kUsesHW = (METRIC == MEUCLID) && (KERNEL == KFLAT);
  float rho = 0.f;
  bool poisoned = false;
  cptr<int32_t> OFF = as_const(a.offsets);
  int b[D];
  bool more = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    more = more && lo[d] <= hi[d];
    b[d] = lo[d];
  }
  while (more) {
    const int t = uniform(tile_index<D>(b, g));
    const int beg = OFF[t], end = OFF[t + 1];
    const bool nonfinite = as_const(a.tnf)[t] != 0;
    for (int j0 = beg; j0 < end; j0 += kWave) {
      const int j = j0 + lane;
      const bool in = j < end;
      float term = 0.f;
      bool inr = false, poison = false;
      if (in) {
        float c[D];
#pragma unroll
        for (int d = 0; d < D; ++d) c[d] = a.xs[d][j];
        const float wj = kUsesHW ? 0.f : a.ws[j];
        if constexpr (METRIC == MEUCLID) {
          const float s2 = squared_distance<D>(q, c);
          inr = s2 <= a.thr;
          if (inr) {
            if constexpr (KERNEL == KFLAT) {
              term = (j == s) ? w_self : a.hw[j];
            } else {
              float k = 1.f;
              if (j != s) {
                if (unit_tiny_exact<KERNEL, UNIT>() || ((UNIT || a.exp_nonpos) && s2 >= 0x1p-96f)) {
                  k = kernel_value<KERNEL, true, UNIT>(sqrt_rn_normal(s2), a.kp[0], a.kp[1], a.kp[2], a.kden_rcp,
                                                       exp_table);
                } else {
                  k = kernel_value<KERNEL>(__builtin_sqrtf(s2), a.kp[0], a.kp[1], a.kp[2], a.kden_rcp);
                }
              }
              term = k * wj;
            }
          }
        } else {
          const float dist = metric_distance<D, METRIC>(q, c, a.mp);
          inr = dist <= a.radius;
          if (inr) {
            const float k = (j == s) ? 1.f : kernel_value<KERNEL>(dist, a.kp[0], a.kp[1], a.kp[2], a.kden_rcp);
            term = k * wj;
          }
        }
        // out-of-radius candidates with an infinite or NaN weight add 0 * w = NaN
        if (nonfinite && !inr) poison = !__builtin_isfinite(a.ws[j]);
      }
      uint64_t m = __ballot(inr);
      while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        rho = rho + readlane_f32(term, l);
      }
      poisoned = poisoned || __ballot(poison) != 0;
    }
    int d = D - 1;
    for (; d >= 0; --d) {
      if (++b[d] <= hi[d]) break;
      b[d] = lo[d];
    }
    more = d >= 0;
  }
  if (poisoned) rho = __builtin_nanf("");
  if (lane == 0) {
    a.rho_s_out[s] = rho;
    a.rho_out[pid] = rho;
    a.hk_out[QNH ? pos : s] = rank_key(rho, pid);
  }
}

template <int D, int METRIC, bool UNION>
__global__ __launch_bounds__(kThreads) void k_nearest_higher(const PassArgs a) {
  const int s = logical_block((a.xcd >> 1) & 1) * kThreads + threadIdx.x;
  const bool valid = s < a.n;
  const int sq = a.qperm ? a.qperm[valid ? s : a.n - 1] : (valid ? s : a.n - 1);
  float q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = a.xs[d][sq];
  const uint64_t key_i = a.hk[sq];
  const uint64_t thr_i = higher_threshold(key_i);
  const int32_t id_i = static_cast<int32_t>(key_i & 0xffffffffu);
  const int my_e = a.pev[id_i];
  cptr<float> X[D];
#pragma unroll
  for (int d = 0; d < D; ++d) X[d] = as_const(a.xs[d]);
  cptr<uint64_t> HK = as_const(a.hk);
  float delta = FLT_MAX;
  int32_t nh = -1;
  // Euclidean: updates need s2 <= bound = min(thr(dm), the best's squared
  // distance) -- sqrtf is monotone -- and then a strictly smaller sqrtf, whose
  // square root is taken only for such candidates
  float bound = a.thr;
  // One candidate (ClusteringKernels.hpp:129-144): j is "higher" if rho_j > rho_i,
  // or equal, positive and with a larger id (one comparison of rank keys); the
  // first strict minimum of the distance in the visit order wins.
  auto eval = [&](const float* c, uint64_t key_j) {
    const bool higher = key_j > thr_i;
    if constexpr (METRIC == MEUCLID) {
      const float s2 = squared_distance<D>(q, c);
      if (higher && s2 <= bound) {
        const float dc = __builtin_sqrtf(s2);
        if (dc < delta) {
          delta = dc;
          nh = static_cast<int32_t>(key_j & 0xffffffffu);
          bound = s2;
        }
      }
    } else {
      const float dist = metric_distance<D, METRIC>(q, c, a.mp);
      if (higher && dist <= a.radius && dist < delta) {
        delta = dist;
        nh = static_cast<int32_t>(key_j & 0xffffffffu);
      }
    }
  };
  for_each_event(valid, my_e, [&](int e, bool act) {
  const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
  int lo[D], hi[D];
  search_box<D>(q, a.radius, g, lo, hi);
  visit_box<D, UNION>(act, lo, hi, g, as_const(a.offsets), [&](int t, int beg, int end, bool need) {
    if (!need) return;
    int j = beg;
    // every lane on one tile (always so in the union walk): candidates at
    // wave-uniform addresses, read through the scalar cache
    bool uni = UNION;
    if constexpr (!UNION) uni = __ballot(t != __builtin_amdgcn_readfirstlane(t)) == 0;
    if (uni) {
      const int ue = __builtin_amdgcn_readfirstlane(end);
      for (j = __builtin_amdgcn_readfirstlane(beg); j + 4 <= ue; j += 4) {
        float c[4][D];
        uint64_t kv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
          for (int d = 0; d < D; ++d) c[u][d] = X[d][j + u];
          kv[u] = HK[j + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) eval(c[u], kv[u]);
      }
    }
    for (; j < end; ++j) {
      float c[D];
#pragma unroll
      for (int d = 0; d < D; ++d) c[d] = X[d][j];
      eval(c, HK[j]);
    }
  });
  });
  if (valid) {
    a.nh_out[id_i] = nh;
    a.delta_out[id_i] = delta;
    if (nh == -1) atomicAdd(&a.sc[my_e], 1);
  }
}

// Nearest-higher by the tile walk with one wave per query (more than three
// dimensions, other metrics, wrapped grids).  The lanes take 64 consecutive
// candidates of the query's visit order; the batch's least distance among its
// higher candidates within dm replaces the query's best only when strictly
// smaller, and among the lanes at that distance the first in lane order wins --
// together the first strict minimum in the visit order of
// for_recursion_nearest_higher (ClusteringKernels.hpp:109-166).  Each lane's
// distance is the metric's float value, as the reference compares it.
template <int D, int METRIC>
__global__ __launch_bounds__(kThreads) void k_nearest_higher_wave(const PassArgs a) {
  const int pos = uniform(static_cast<int>(blockIdx.x) * (kThreads / kWave) + static_cast<int>(threadIdx.x / kWave));
  if (pos >= a.n) return;  // the whole wave
  const int lane = __lane_id();
  float q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = as_const(a.xs[d])[pos];
  const uint64_t key_i = uniform(as_const(a.hk)[pos]);
  const uint64_t thr_i = higher_threshold(key_i);
  const int32_t id_i = static_cast<int32_t>(key_i & 0xffffffffu);
  const int e = as_const(a.pev)[id_i];
  const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
  int lo[D], hi[D];
  search_box<D>(q, a.radius, g, lo, hi);
  float delta = FLT_MAX;
  int32_t nh = -1;
  cptr<int32_t> OFF = as_const(a.offsets);
  int b[D];
  bool more = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    more = more && lo[d] <= hi[d];
    b[d] = lo[d];
  }
  while (more) {
    const int t = uniform(tile_index<D>(b, g));
    const int beg = OFF[t], end = OFF[t + 1];
    for (int j0 = beg; j0 < end; j0 += kWave) {
      const int j = j0 + lane;
      float dist = __builtin_inff();
      uint64_t key_j = 0;
      if (j < end) {
        key_j = a.hk[j];
        if (key_j > thr_i) {
          float c[D];
#pragma unroll
          for (int d = 0; d < D; ++d) c[d] = a.xs[d][j];
          float dj;
          if constexpr (METRIC == MEUCLID) {
            // squared distances: sqrtf is monotone, the square root is taken once per batch
            const float s2 = squared_distance<D>(q, c);
            dj = s2 <= a.thr ? s2 : __builtin_inff();
          } else {
            dj = metric_distance<D, METRIC>(q, c, a.mp);
            dj = dj <= a.radius ? dj : __builtin_inff();
          }
          dist = dj;
        }
      }
      const float dmin = wave_min_f32(dist);
      if constexpr (METRIC == MEUCLID) {
        if (dmin <= a.thr) {  // uniform; some higher candidate within dm
          const float dd = __builtin_sqrtf(dmin);
          if (dd < delta) {
            // the first lane whose sqrtf equals the least one (squared distances
            // within the class of dmin, all at least dmin)
            float lo2, hi2;
            sqrt_class(dd, lo2, hi2);
            const int l = __builtin_ctzll(__ballot(dist <= hi2));
            delta = dd;
            nh = static_cast<int32_t>(readlane_u64(key_j, l) & 0xffffffffu);
          }
        }
      } else if (dmin < delta) {  // uniform
        const int l = __builtin_ctzll(__ballot(dist == dmin));
        delta = dmin;
        nh = static_cast<int32_t>(readlane_u64(key_j, l) & 0xffffffffu);
      }
    }
    int d = D - 1;
    for (; d >= 0; --d) {
      if (++b[d] <= hi[d]) break;
      b[d] = lo[d];
    }
    more = d >= 0;
  }
  if (lane == 0) {
    a.nh_out[id_i] = nh;
    a.delta_out[id_i] = delta;
    if (nh == -1) atomicAdd(&a.sc[e], 1);
  }
}

// Squared distance from q to the nearest point of a box, with the fold order of
// squared_distance.  Rounding is monotone, so for every point c of the box
// lower_bound2 <= squared_distance(q, c) exactly: a box whose bound exceeds a
// lane's threshold holds no candidate that lane can accept.
template <int D>
__device__ __forceinline__ float lower_bound2(const float* q, const float* lo, const float* hi) {
  float s = 0.f;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    float g = lo[d] - q[d];
    const float g2 = q[d] - hi[d];
    g = g2 > g ? g2 : g;
    g = g > 0.f ? g : 0.f;
    const float t = g * g;
    s = (d == D - 1) ? t : t + s;
  }
  return s;
}

// Squared distance from q to the farthest point of a box, with the fold order of
// squared_distance: an upper bound of squared_distance(q, c) for every c in the box.
template <int D>
__device__ __forceinline__ float upper_bound2(const float* q, const float* lo, const float* hi) {
  float s = 0.f;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    const float g1 = q[d] - lo[d];
    const float g2 = hi[d] - q[d];
    const float g = g1 > g2 ? g1 : g2;
    const float t = g * g;
    s = (d == D - 1) ? t : t + s;
  }
  return s;
}

// Nearest-higher (K5) as a block search over the Morton-ordered layout
// (Euclidean metric, no wrapped dimension).  The result is the serial one: the
// candidate minimising (distance, visit position) where the visit position is
// (tile id, point id) -- ascending tile ids and ascending ids inside a tile are
// the order of for_recursion_nearest_higher (ClusteringKernels.hpp:109-166) on a
// grid without wrapping -- so candidates may be examined in any order, and a
// block whose bounding box cannot beat a lane's current distance is skipped.
//
// This first phase (one lane per query, a wave per 64 Morton-consecutive
// queries) limits the search to the radius thr1: the blocks of its lanes'
// search boxes, the wave's own tile first.  (A first scan of the wave's own 64
// points repeated part of that tile's scan and was 1-5 % slower.)  A lane that
// finds a higher point within it is final (every candidate at or below its distance was
// examined); the others are queued for the second phase (k_nh_pending), the
// full search within dm.
//
// SPLIT: the workgroup's four waves share one wave's 64 queries.  Block k of a
// tile belongs to wave k mod 4; each wave prunes its blocks with its own best,
// which is never better than the combined one, so the combined minimum's block
// is scanned by its owner, and the four minima are combined at the end (the
// minimum of (distance, visit position) does not depend on the order).
//
// Occupancy: in 2D the request of at least seven waves per SIMD keeps the SGPRs
// at 94 (106 without it, which allows seven waves), so that with 63 VGPRs the
// kernel runs eight (the batch's first phase about 5 % faster, the 2D 1M-point
// one 2-3 %); in 3D it would spill.
template <int D, bool SPLIT = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(D == 2 ? 7 : 1)))
void k_nearest_higher_blocks(const PassArgs a) {
  const int wv = threadIdx.x / kWave;
  const int gid = SPLIT ? logical_block((a.xcd >> 1) & 1) * kWave + __lane_id()
                        : logical_block((a.xcd >> 1) & 1) * kThreads + static_cast<int>(threadIdx.x);
  const bool valid = gid < a.n;
  const int pos = valid ? gid : a.n - 1;
  float q[D];
#pragma unroll
  for (int d = 0; d < D; ++d) q[d] = a.xn[d][pos];
  const uint64_t key_i = a.hk[pos];
  const uint64_t thr_i = higher_threshold(key_i);
  const int32_t id_i = static_cast<int32_t>(key_i & 0xffffffffu);
  const int my_e = a.pev[id_i];
  cptr<float> X[D], TMIN[D], TMAX[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    X[d] = as_const(a.xn[d]);
    TMIN[d] = as_const(a.tmin[d]);
    TMAX[d] = as_const(a.tmax[d]);
  }
  cptr<uint64_t> HK = as_const(a.hk);
  cptr<uint64_t> BMK = as_const(a.bmaxk);
  cptr<uint64_t> TMK = as_const(a.tmaxk);
  cptr<int32_t> OFF = as_const(a.offsets);
  cptr<int32_t> BOFF = as_const(a.blk_off);
  // a block or tile whose largest rank key is at most a lane's threshold holds
  // no point higher than that lane's; at most the wave's least threshold, none
  // higher than any of its lanes'
  const uint64_t wthr = uniform(wave_min_u64(valid ? thr_i : ~uint64_t{0}));
  int64_t best = INT64_MAX;  // visit key (tile, id) of the best candidate
  // The best candidate so far is at squared distance s2b (its distance, the
  // output, is sqrtf(s2b)).  A candidate replaces it when its sqrtf is smaller,
  // or equal and it is earlier in the visit order (key < best); it cannot when
  // s2 > not_farther_than(s2b), so candidates need s2 <= bound = min(search
  // radius, not_farther_than(s2b)).  Below strictly_nearer(s2b) (+inf while
  // there is no best) a candidate replaces the best without a square root;
  // between that and the bound the two square roots decide.
  const float cap = a.thr1 < a.thr ? a.thr1 : a.thr;
  float bound = cap;
  float s2b = __builtin_inff();
  // the wave's query box, for the block prefilter
  float qlo[D], qhi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    qlo[d] = wave_min_f32(valid ? q[d] : __builtin_inff());
    qhi[d] = wave_max_f32(valid ? q[d] : -__builtin_inff());
  }

  // One candidate: a higher point within the bound replaces the best when it is
  // nearer, or as near and earlier in the visit order (the update by selects).
  auto cand = [&](int t, const float* c, uint64_t key_j) {
    const float s2 = squared_distance<D>(q, c);
    if (key_j > thr_i && s2 <= bound) {
      const int64_t key = (static_cast<int64_t>(t) << 32) | static_cast<uint32_t>(key_j);
      bool better = s2 < strictly_nearer(s2b);
      if (!better) better = sqrt_nearer_or_earlier(s2, key, s2b, best);
      const float up = not_farther_than(s2);
      s2b = better ? s2 : s2b;
      best = better ? key : best;
      bound = better ? (up < cap ? up : cap) : bound;
    }
  };
  // candidates [beg, end) of tile t, wave-uniform: eight at a time, all loads
  // issued before the first use
  auto scan = [&](int t, int beg, int end) {
    constexpr int U = 8;
    int j = beg;
    for (; j + U <= end; j += U) {
      float c[U][D];
      uint64_t k[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int d = 0; d < D; ++d) c[u][d] = X[d][j + u];
        k[u] = HK[j + u];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) cand(t, c[u], k[u]);
    }
    for (; j < end; ++j) {
      float c[D];
#pragma unroll
      for (int d = 0; d < D; ++d) c[d] = X[d][j];
      cand(t, c, HK[j]);
    }
  };
  // A block [beg, end) of tile t that only the lanes of nm need (called by the
  // whole wave): the block's candidates are spread over a half-wave per needing
  // lane, two needing lanes at a time (lane c of a half takes candidate beg + c),
  // each half keeps its least (distance, id) among the candidates that lane
  // accepts (higher, squared distance within its bound), and the needing lane
  // merges it into its best.  The minimum of (distance, visit position) does not
  // depend on the order of examination, and inside one tile the visit position
  // orders as the id, so the result is the point-by-point scan's.  The lane of
  // each half that holds its half's least passes its squared distance along.
  auto scan_across = [&](int t, int beg, int end, uint64_t nm) {
    constexpr int kHalf = kWave / 2;
    static_assert(kBlock <= kHalf, "a block fits a half-wave");
    const int lane = __lane_id();
    const bool upper = lane >= kHalf;
    const int j = beg + (lane & (kHalf - 1));
    const bool jv = j < end;
    const int jl = jv ? j : beg;
    float c[D];
#pragma unroll
    for (int d = 0; d < D; ++d) c[d] = X[d][jl];
    const uint64_t kj = HK[jl];
    while (nm) {
      const int l1 = __builtin_ctzll(nm);
      nm &= nm - 1;
      const int l2 = nm ? __builtin_ctzll(nm) : l1;
      nm &= nm - 1;
      float hq[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float q1 = readlane_f32(q[d], l1);
        const float q2 = readlane_f32(q[d], l2);
        hq[d] = upper ? q2 : q1;
      }
      const uint64_t t1 = readlane_u64(thr_i, l1);
      const uint64_t t2 = readlane_u64(thr_i, l2);
      const uint64_t hthr = upper ? t2 : t1;
      const float b1 = readlane_f32(bound, l1);
      const float b2 = readlane_f32(bound, l2);
      const float hb = upper ? b2 : b1;
      const float s2 = squared_distance<D>(hq, c);
      const bool ok = jv && kj > hthr && s2 <= hb;
      uint64_t pk =
          ok ? (static_cast<uint64_t>(__float_as_uint(sqrt_rn_finite(s2))) << 32) | static_cast<uint32_t>(kj)
             : ~uint64_t{0};
#pragma unroll
      for (int off = kHalf / 2; off > 0; off >>= 1) {
        const uint64_t o = shfl_xor_u64(pk, off);
        pk = o < pk ? o : pk;
      }
      const uint64_t r1 = readlane_u64(pk, 0);
      const uint64_t r2 = readlane_u64(pk, kHalf);
      // the winners' lanes (ids differ inside a half: at most one lane each)
      const uint64_t wm = __ballot(ok && static_cast<uint32_t>(kj) == static_cast<uint32_t>(upper ? r2 : r1));
      const uint32_t wlo = static_cast<uint32_t>(wm), whi = static_cast<uint32_t>(wm >> 32);
      const float sw1 = readlane_f32(s2, wlo ? __builtin_ctz(wlo) : 0);
      const float sw2 = readlane_f32(s2, whi ? kHalf + __builtin_ctz(whi) : 0);
      const uint64_t r = lane == l1 ? r1 : r2;
      if ((lane == l1 || lane == l2) && r != ~uint64_t{0}) {
        const float sw = lane == l1 ? sw1 : sw2;
        const int64_t key = (static_cast<int64_t>(t) << 32) | static_cast<uint32_t>(r);
        bool better = sw < strictly_nearer(s2b);
        if (!better) {
          const float dr = __uint_as_float(static_cast<uint32_t>(r >> 32));  // sqrtf(sw)
          const float db = sqrt_rn_finite(s2b);
          better = (dr < db) | ((dr == db) & (key < best));
        }
        const float up = not_farther_than(sw);
        s2b = better ? sw : s2b;
        best = better ? key : best;
        bound = better ? (up < cap ? up : cap) : bound;
      }
    }
  };

  for_each_event(valid, my_e, [&](int e, bool act) {
    const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
    // only candidates within thr1: a box of a slightly larger radius (1/1024 of
    // it, plus 2^-20 of the coordinate for the rounding of q -+ r) holds every
    // such candidate and lies inside the reference's box of dm
    int lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float r = a.box1 + __builtin_fabsf(q[d]) * 0x1p-20f;
      r = r < a.radius ? r : a.radius;
      lo[d] = tile_bin(q[d] - r, g.mn[d], g.mx[d], g.ts[d], g.npd, false);
      hi[d] = tile_bin(q[d] + r, g.mn[d], g.mx[d], g.ts[d], g.npd, false);
    }
    auto in_box = [&](int t) {
      const int tl = t - g.tile_base;
      bool in = act && tl >= 0 && tl < g.ntiles;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int b = (tl / g.stride[d]) % g.npd;
        in = in && b >= lo[d] && b <= hi[d];
      }
      return in;
    };

    // the wave's own tile (its first point's), walked first
    int t0 = -1;
    {
      const int base = uniform(gid & ~(kWave - 1));
      if (base < a.n) t0 = uniform(as_const(a.kn)[base] >> a.mshift);
    }

    int ulo[D], uhi[D];
    bool any = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ulo[d] = uniform(wave_min_i32(act ? lo[d] : INT_MAX));
      uhi[d] = uniform(wave_max_i32(act ? hi[d] : INT_MIN));
      any = any && ulo[d] <= uhi[d];
    }
    if (!any) return;
    const int lane = __lane_id();
    // one tile of the union: the blocks whose box and largest key can still
    // improve a lane that needs the tile (need: the tile is in the lane's box)
    // (pfirst: a Morton position inside tile t whose block is walked first, the
    // tile's blocks following in rotated order; -1 for the tile's own order)
    auto tile_pass = [&](int t, bool need, int pfirst) {
      float tl[D], th[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        tl[d] = TMIN[d][t];
        th[d] = TMAX[d][t];
      }
      need = need && TMK[t] > thr_i && lower_bound2<D>(q, tl, th) <= bound;
      const float maxb = wave_max_f32(need ? bound : -1.f);
      if (maxb >= 0.f) {
        const int k0 = BOFF[t], k1 = BOFF[t + 1];
        const int off0 = OFF[t], off1 = OFF[t + 1];
        const int nb = k1 - k0;
        const int rel = (pfirst >= off0 && pfirst < off1) ? (pfirst - off0) / kBlock : 0;
        for (int ci = 0; ci < nb; ci += kWave) {
          // one block per lane against the wave's query box: lane l takes the
          // tile's block (rel + ci + l) mod nb
          const bool inb = ci + lane < nb;
          int r = rel + ci + lane;
          r = r >= nb ? r - nb : r;
          const int k = k0 + (inb ? r : 0);
          float bl[D], bh[D];
#pragma unroll
          for (int d = 0; d < D; ++d) {
            bl[d] = inb ? a.bmin[d][k] : __builtin_inff();
            bh[d] = inb ? a.bmax[d][k] : -__builtin_inff();
          }
          const uint64_t bk = inb ? BMK[k] : 0;
          float lbw = 0.f;
#pragma unroll
          for (int d = D - 1; d >= 0; --d) {
            float gg = bl[d] - qhi[d];
            const float g2 = qlo[d] - bh[d];
            gg = g2 > gg ? g2 : gg;
            gg = gg > 0.f ? gg : 0.f;
            const float tt = gg * gg;
            lbw = (d == D - 1) ? tt : tt + lbw;
          }
          const bool own = !SPLIT || (k & (kThreads / kWave - 1)) == wv;
          uint64_t m = __ballot(inb && own && lbw <= maxb && bk > wthr);
          while (m) {
            const int kl = __builtin_ctzll(m);
            m &= m - 1;
            float ul[D], uh[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
              ul[d] = readlane_f32(bl[d], kl);
              uh[d] = readlane_f32(bh[d], kl);
            }
            const bool nk = need && readlane_u64(bk, kl) > thr_i && lower_bound2<D>(q, ul, uh) <= bound;
            const uint64_t nm = __ballot(nk);
            if (nm != 0) {
              int rk = rel + ci + kl;
              rk = rk >= nb ? rk - nb : rk;
              const int start = off0 + rk * kBlock;
              const int end = min(start + kBlock, off1);
              if (__builtin_popcountll(nm) <= a.nh_lone) scan_across(t, start, end, nm);
              else if (nk) scan(t, start, end);
            }
          }
        }
      }
    };
    // the wave's own tile first, from the block of the wave's first query on
    // (the nearest candidates, so that the lanes' bounds shrink before the rest
    // is walked), then the rest of the union's tiles in lexicographic order
    if (t0 >= 0) {
      const bool need0 = in_box(t0);
      if (__ballot(need0) != 0) tile_pass(t0, need0, uniform(gid & ~(kWave - 1)));
    }
    int b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = ulo[d];
    while (true) {
      bool need = act;
#pragma unroll
      for (int d = 0; d < D; ++d) need = need && b[d] >= lo[d] && b[d] <= hi[d];
      if (__ballot(need) != 0) {
        const int t = uniform(tile_index<D>(b, g));
        if (t != t0) tile_pass(t, need, -1);
      }
      int d = D - 1;
      for (; d >= 0; --d) {
        if (++b[d] <= uhi[d]) break;
        b[d] = ulo[d];
      }
      if (d < 0) break;
    }
  });

  float delta = best == INT64_MAX ? FLT_MAX : sqrt_rn_finite(s2b);
  if constexpr (SPLIT) {
    // the four waves' minima, combined by wave 0
    __shared__ float s_delta[kThreads / kWave][kWave];
    __shared__ int64_t s_best[kThreads / kWave][kWave];
    const int lane = __lane_id();
    s_delta[wv][lane] = delta;
    s_best[wv][lane] = best;
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int w = 1; w < kThreads / kWave; ++w) {
      const float dw = s_delta[w][lane];
      const int64_t bw = s_best[w][lane];
      if (dw < delta || (dw == delta && bw < best)) {
        delta = dw;
        best = bw;
      }
    }
  }
  // unresolved positions are appended to the pending list, a wave's together in
  // lane (Morton) order, so that the full search runs on spatially coherent
  // waves (the order of the waves does not change any result)
  const int32_t nh = best == INT64_MAX ? -1 : static_cast<int32_t>(best & 0xffffffff);
  const uint64_t pm = __ballot(valid && nh == -1);
  if (pm != 0) {  // uniform; every lane of the wave is active here
    int base = 0;
    if (__lane_id() == 0) base = atomicAdd(a.pending_count, __builtin_popcountll(pm));
    base = __builtin_amdgcn_readfirstlane(base);
    if (valid && nh == -1) a.pending[base + mbcnt64(pm)] = pos;
  }
  if (valid && nh != -1) {
    a.nh_out[id_i] = nh;
    a.delta_out[id_i] = delta;
  }
}

// Second nearest-higher phase: a group of G lanes per unresolved point
// (pending[0..npend), positions in the Morton layout), kWave / G points per
// wave.  The group's lanes split the candidates of the point's search box
// (tiles, then their 32-point Morton blocks, pruned by their boxes against the
// group's best bound), each keeps its best (distance, tile id, point id), and a
// group reduction takes the least:
// the serial backend's first strict minimum on a grid without wrapping.  The
// grid is persistent: every wave strides over the pending list and leaves when
// it is exhausted.  SINGLE: one event, whose grid is then wave-uniform.
template <int G>
__device__ __forceinline__ float group_min_f32(float v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    const float o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

//
// Occupancy: a single 2D event asks for six waves per SIMD (80 VGPRs, no
// spills; five at 83 without the request); batched 2D events and 3D would spill.
template <int D, int G, bool SINGLE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(D == 2 && SINGLE ? 6 : 1)))
void k_nh_pending(const PassArgs a) {
  static_assert(kWave % G == 0, "groups must tile the wave");
  const int npend = uniform(*a.pending_count);
  const int lane = __lane_id();
  const int gl = lane % G;
  const int wave = static_cast<int>((blockIdx.x * kThreads + threadIdx.x) / kWave);
  const int per_wave = kWave / G;
  const int stride = gridDim.x * (kThreads / kWave) * per_wave;
  const int grp_shift = (lane / G) * G;  // the group's bits in a wave ballot
  for (int base = uniform(wave * per_wave); base < npend; base += stride) {
    const int w = base + lane / G;
    const bool active = w < npend;
    const int pos = a.pending[active ? w : npend - 1];
    float q[D];
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = a.xn[d][pos];
    const uint64_t key_i = a.hk[pos];
    const uint64_t thr_i = higher_threshold(key_i);
    const int32_t id_i = static_cast<int32_t>(key_i & 0xffffffffu);
    const int e = SINGLE ? 0 : a.pev[id_i];
    const EvGrid<D> g = load_event<D>(a.ev, e, a.wrap);
    int lo[D], hi[D];
    search_box<D>(q, a.radius, g, lo, hi);
    int64_t best = INT64_MAX;
    // candidates need s2 <= bound = min(thr(dm), not_farther_than(s2b)) (s2b: the
    // lane's best squared distance), and replace the lane's best when their sqrtf
    // is smaller (certainly so below strictly_nearer(s2b); s2b is +inf while
    // there is no best), or equal with key < best
    float bound = a.thr;
    float s2b = __builtin_inff();
    float group_bound = a.thr;
    // Levels of growing radius order the visits nearest first, so that the
    // bound shrinks early.  Each block is visited once, at the first level whose
    // radius reaches its box, and then every candidate of it within the bound is
    // examined; blocks inside the first phase's radius hold no higher point.
    // After a level the search is complete when the group's bound lies within
    // the level's radius: every block not yet visited is farther than that.
    const float prev0 = a.lvl_prev0;
    float prev = -1.f;  // the previous level's threshold (none before the first)
    bool done = false;
    // The box's tiles, one per lane of the group when they fit (their boxes and
    // ranges are then loaded once, together, and shuffled to the group at each
    // level); otherwise each level walks them and loads them one by one.
    int dims[D];
    int nt = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      dims[d] = hi[d] - lo[d] + 1;
      nt *= dims[d];
    }
    const bool cached = nt <= G;
    auto tile_of = [&](int ti) {
      int bb[D];
      int rem = ti;
#pragma unroll
      for (int d = D - 1; d >= 0; --d) {
        bb[d] = lo[d] + rem % dims[d];
        rem /= dims[d];
      }
      return tile_index<D>(bb, g);
    };
    int my_t = 0, my_k0 = 0, my_k1 = 0, my_o0 = 0, my_o1 = 0;
    uint64_t my_tk = 0;
    float my_tl[D], my_th[D];
    if (cached) {
      const bool tv = active && gl < nt;
      my_t = tv ? tile_of(gl) : 0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        my_tl[d] = tv ? a.tmin[d][my_t] : __builtin_inff();
        my_th[d] = tv ? a.tmax[d][my_t] : -__builtin_inff();
      }
      my_k0 = tv ? a.blk_off[my_t] : 0;
      my_k1 = tv ? a.blk_off[my_t + 1] : 0;
      my_o0 = tv ? a.offsets[my_t] : 0;
      my_o1 = tv ? a.offsets[my_t + 1] : 0;
      my_tk = tv ? a.tmaxk[my_t] : 0;
    }
    auto cand = [&](int t, const float* c, uint64_t key_j) {
      const float s2 = squared_distance<D>(q, c);
      if (key_j > thr_i && s2 <= bound) {
        const int64_t key = (static_cast<int64_t>(t) << 32) | static_cast<uint32_t>(key_j);
        bool better = s2 < strictly_nearer(s2b);
        if (!better) better = sqrt_nearer_or_earlier(s2, key, s2b, best);
        const float up = not_farther_than(s2);
        s2b = better ? s2 : s2b;
        best = better ? key : best;
        bound = better ? (up < a.thr ? up : a.thr) : bound;
      }
    };
    // at every level the point's own tile is walked first (box index own_ti
    // swapped with 0): its blocks hold the nearest candidates, so the bound
    // shrinks before the other tiles are walked (any order gives the same result)
    int own_ti = 0;
    {
      int idx = 0;
#pragma unroll
      for (int d = 0; d < D; ++d)
        idx = idx * dims[d] + (tile_bin(q[d], g.mn[d], g.mx[d], g.ts[d], g.npd, false) - lo[d]);
      own_ti = (idx >= 0 && idx < nt) ? idx : 0;
    }
    for (int lv = 0; lv < a.nlvl; ++lv) {
      const float cap = a.lvl[lv];
      int ti = 0;
      bool more = active && !done;
      while (__ballot(more) != 0) {
        if (more) {  // uniform within a group
          const int tv = ti == 0 ? own_ti : (ti == own_ti ? 0 : ti);
          int t, k0, k1, off0, off1;
          uint64_t tk;
          float tl[D], th[D];
          if (cached) {
            const int src = grp_shift + tv;
            t = __shfl(my_t, src);
            tk = (static_cast<uint64_t>(__shfl(static_cast<uint32_t>(my_tk >> 32), src)) << 32) |
                 __shfl(static_cast<uint32_t>(my_tk), src);
            k0 = __shfl(my_k0, src);
            k1 = __shfl(my_k1, src);
            off0 = __shfl(my_o0, src);
            off1 = __shfl(my_o1, src);
#pragma unroll
            for (int d = 0; d < D; ++d) {
              tl[d] = __shfl(my_tl[d], src);
              th[d] = __shfl(my_th[d], src);
            }
          } else {
            t = tile_of(tv);
#pragma unroll
            for (int d = 0; d < D; ++d) {
              tl[d] = a.tmin[d][t];
              th[d] = a.tmax[d][t];
            }
            k0 = a.blk_off[t];
            k1 = a.blk_off[t + 1];
            off0 = a.offsets[t];
            off1 = a.offsets[t + 1];
            tk = a.tmaxk[t];
          }
          const float lim_t = cap < group_bound ? cap : group_bound;
          // tiles without a point higher than this one, beyond the level's
          // radius, inside the first phase's radius or whose every block an
          // earlier level reached are skipped
          const float ub_t = upper_bound2<D>(q, tl, th);
          if (tk > thr_i && lower_bound2<D>(q, tl, th) <= lim_t && ub_t > prev0 && ub_t > prev) {
            // the tile's Morton blocks: G at a time against the group's bound,
            // then the points of each block visited at this level
            for (int kb = k0; kb < k1; kb += G) {
              const int k = kb + gl;
              const bool inb = k < k1;
              float bl[D], bh[D];
#pragma unroll
              for (int d = 0; d < D; ++d) {
                bl[d] = inb ? a.bmin[d][k] : __builtin_inff();
                bh[d] = inb ? a.bmax[d][k] : -__builtin_inff();
              }
              const uint64_t bk = inb ? a.bmaxk[k] : 0;
              const float lim_b = cap < group_bound ? cap : group_bound;
              const float lb = lower_bound2<D>(q, bl, bh);
              const bool pass = inb && bk > thr_i && lb <= lim_b && lb > prev &&
                                upper_bound2<D>(q, bl, bh) > prev0;
              uint64_t gm = __ballot(pass) >> grp_shift;
              if constexpr (G < kWave) gm &= (uint64_t{1} << G) - 1;
              while (gm) {
                const int kl = __builtin_ctzll(gm);
                gm &= gm - 1;
                const int start = off0 + (kb + kl - k0) * kBlock;
                const int end = min(start + kBlock, off1);
                // two candidates per lane, both loads issued before either is used
                for (int j = start + gl; j < end; j += 2 * G) {
                  const int j2 = j + G;
                  const bool v2 = j2 < end;
                  const int jj = v2 ? j2 : j;
                  float c0[D], c1[D];
#pragma unroll
                  for (int d = 0; d < D; ++d) {
                    c0[d] = a.xn[d][j];
                    c1[d] = a.xn[d][jj];
                  }
                  const uint64_t h0 = a.hk[j];
                  const uint64_t h1 = a.hk[jj];
                  cand(t, c0, h0);
                  if (v2) cand(t, c1, h1);
                }
              }
              group_bound = group_min_f32<G>(bound);  // the point's best so far over its lanes
            }
          }
          more = ++ti < nt;
        }
      }
      group_bound = group_min_f32<G>(bound);
      done = group_bound <= cap;
      prev = cap;
    }
    // lexicographic (distance, key) minimum over the group
    float delta = best == INT64_MAX ? FLT_MAX : sqrt_rn_finite(s2b);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      const float od = __shfl_xor(delta, off);
      const int32_t klo = __shfl_xor(static_cast<int32_t>(best & 0xffffffff), off);
      const int32_t khi = __shfl_xor(static_cast<int32_t>(best >> 32), off);
      const int64_t ok = (static_cast<int64_t>(khi) << 32) | static_cast<uint32_t>(klo);
      if (od < delta || (od == delta && ok < best)) {
        delta = od;
        best = ok;
      }
    }
    if (active && gl == 0) {
      const int32_t nh = best == INT64_MAX ? -1 : static_cast<int32_t>(best & 0xffffffff);
      a.nh_out[id_i] = nh;
      a.delta_out[id_i] = nh == -1 ? FLT_MAX : delta;
      if (nh == -1) atomicAdd(&a.sc[e], 1);
    }
  }
}

template <int D, int METRIC>
__global__ __launch_bounds__(kThreads) void k_classify(const ClassifyArgs a) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  // the call's totals start at zero; [1], the grow-only capacity, carries over
  // unless cleared
  if (i < 11 && (i != 1 || a.clear_cap)) a.counters[i] = 0;
  if (i >= a.n) return;
  const int32_t nhv = a.nh[i];
  float dist;
  if (nhv == -1) {
    // metric(point_i, PointsView[-1]) with the all-FLT_MAX sentinel (PointsCommon.hpp:75-81)
    float x[D], m[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      x[d] = a.c.p[d][i];
      m[d] = FLT_MAX;
    }
    dist = metric_distance<D, METRIC>(x, m, a.mp);
  } else {
    dist = a.delta[i];  // the nearest-higher kernel's metric(point_i, point_nh)
  }
  const bool seed = (dist > a.seed_dc) && (a.rho[i] >= a.rhoc);
  a.is_seed[i] = seed ? 1 : 0;
  a.flag[i] = seed ? 1 : 0;
  const int32_t parent = seed ? -1 : nhv;
  if (seed) a.nh[i] = -1;
  a.root[i] = parent == -1 ? i : parent;
}

// ===========================================================================
// Small calls fused per event (one workgroup each).
// ===========================================================================
// Layout of one event: bounding box and tile sizes (k_bbox_partial,
// k_bbox_final), tile ids (k_bin), the stable counting sort by tile id into the
// CSR -- ascending point id inside a tile, the serial fill order -- and the
// tile-ordered copies (k_radix_*, k_tile_offsets, k_gather, k_tile_nonfinite).
// Each wave owns a contiguous range of the event's points: it counts them per
// tile, the bases of a tile's waves follow in wave order after the scanned tile
// offsets, and each wave ranks its points in order, 64 at a time, equal tiles by
// lane order (ballot match).
template <int D>
__global__ __launch_bounds__(kFusedThreads) void k_fused_layout(const FusedArgs a) {
  __shared__ uint16_t s_key[kFusedMaxPoints];
  __shared__ uint16_t s_cnt[kFusedWaves][kFusedMaxTiles];
  __shared__ float s_lo[kFusedWaves][D], s_hi[kFusedWaves][D];
  __shared__ float s_mn[D], s_mx[D], s_ts[D];
  __shared__ int s_scan[kFusedWaves];
  const int e = blockIdx.x;
  const int tid = threadIdx.x, lane = __lane_id(), w = tid / kWave;
  const int begin = a.ev[e].begin, n = a.ev[e].n, T = a.ev[e].ntiles, npd = a.ev[e].npd;
  const int tb = a.ev[e].tile_base;
  int stride[D];
#pragma unroll
  for (int d = 0; d < D; ++d) stride[d] = a.ev[e].stride[d];
  if (e == 0 && tid < 11 && (tid != 1 || a.clear_cap)) a.counters[tid] = 0;
  if (tid == 0) a.sc[e] = 0;
  // bounding box: NaN is never selected
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float lo = __builtin_inff(), hi = -__builtin_inff();
    for (int i = tid; i < n; i += kFusedThreads) {
      const float v = a.c.p[d][begin + i];
      lo = v < lo ? v : lo;
      hi = hi < v ? v : hi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float olo = __shfl_xor(lo, off), ohi = __shfl_xor(hi, off);
      lo = olo < lo ? olo : lo;
      hi = hi < ohi ? ohi : hi;
    }
    if (lane == 0) {
      s_lo[w][d] = lo;
      s_hi[w][d] = hi;
    }
  }
  __syncthreads();
  if (tid < D) {
    float lo = s_lo[0][tid], hi = s_hi[0][tid];
    for (int k = 1; k < kFusedWaves; ++k) {
      lo = s_lo[k][tid] < lo ? s_lo[k][tid] : lo;
      hi = hi < s_hi[k][tid] ? s_hi[k][tid] : hi;
    }
    const float ts = (hi - lo) / static_cast<float>(npd);  // ComputeTiles.hpp:21-26
    s_mn[tid] = lo;
    s_mx[tid] = hi;
    s_ts[tid] = ts;
    a.ev[e].mn[tid] = lo;
    a.ev[e].mx[tid] = hi;
    a.ev[e].ts[tid] = ts;
  }
  for (int t = tid; t < kFusedWaves * T; t += kFusedThreads) s_cnt[t / T][t % T] = 0;
  __syncthreads();
  // tile ids (local to the event)
  for (int i = tid; i < n; i += kFusedThreads) {
    int32_t t = 0;
#pragma unroll
    for (int d = 0; d < D; ++d)
      t += stride[d] * tile_bin(a.c.p[d][begin + i], s_mn[d], s_mx[d], s_ts[d], npd, a.wrap.w[d] != 0);
    s_key[i] = static_cast<uint16_t>(t);
    a.pev[begin + i] = e;
  }
  __syncthreads();
  int kbits = 0;
  while ((1 << kbits) < T) ++kbits;
  const int span = (n + kFusedWaves - 1) / kFusedWaves;
  const int r0 = w * span, r1 = min(r0 + span, n);
  // per-wave counts: one lane per distinct tile of each 64 points adds their number
  for (int base = r0; base < r1; base += kWave) {
    const int i = base + lane;
    const bool in = i < r1;
    const int k = in ? s_key[i] : 0;
    const uint64_t peers = match_peers(k, in, kbits);
    if (in && mbcnt64(peers) == 0) s_cnt[w][k] = static_cast<uint16_t>(s_cnt[w][k] + __builtin_popcountll(peers));
  }
  __syncthreads();
  // tile offsets (two tiles per thread: tid and tid + kFusedThreads), then the
  // base of each wave inside each tile
  const int t0 = tid, t1 = tid + kFusedThreads;
  int c0 = 0, c1 = 0;
  if (t0 < T)
    for (int k = 0; k < kFusedWaves; ++k) c0 += s_cnt[k][t0];
  if (t1 < T)
    for (int k = 0; k < kFusedWaves; ++k) c1 += s_cnt[k][t1];
  int total0, total1;
  const int x0 = block_incl_scan<kFusedThreads>(c0, s_scan, total0) - c0;
  const int x1 = total0 + block_incl_scan<kFusedThreads>(c1, s_scan, total1) - c1;
  if (t0 < T) {
    int run = x0;
    for (int k = 0; k < kFusedWaves; ++k) {
      const int c = s_cnt[k][t0];
      s_cnt[k][t0] = static_cast<uint16_t>(run);
      run += c;
    }
    a.offsets[tb + t0] = begin + x0;
    a.tnf[tb + t0] = 0;
  }
  if (t1 < T) {
    int run = x1;
    for (int k = 0; k < kFusedWaves; ++k) {
      const int c = s_cnt[k][t1];
      s_cnt[k][t1] = static_cast<uint16_t>(run);
      run += c;
    }
    a.offsets[tb + t1] = begin + x1;
    a.tnf[tb + t1] = 0;
  }
  if (tid == 0) a.offsets[tb + T] = begin + n;
  __syncthreads();
  // stable scatter: every lane reads its tile's base before the lowest lane of
  // the tile's peers moves it on
  for (int base = r0; base < r1; base += kWave) {
    const int i = base + lane;
    const bool in = i < r1;
    const int k = in ? s_key[i] : 0;
    const uint64_t peers = match_peers(k, in, kbits);
    const int before = mbcnt64(peers);
    const int rb = in ? s_cnt[w][k] : 0;
    __builtin_amdgcn_wave_barrier();
    if (in && before == 0) s_cnt[w][k] = static_cast<uint16_t>(rb + __builtin_popcountll(peers));
    __builtin_amdgcn_wave_barrier();
    if (in) {
      const int slot = begin + rb + before;
      const int gi = begin + i;
      a.perm[slot] = gi;
#pragma unroll
      for (int d = 0; d < D; ++d) a.xs[d * a.ld + slot] = a.c.p[d][gi];
      const float wv = a.w[gi];
      a.ws[slot] = wv;
      a.hw[slot] = a.h * wv;  // FlatKernel term k*w with k = h
      if (!__builtin_isfinite(wv)) a.tnf[tb + k] = 1;
    }
  }
}

// Seeds, seed ranks, capacity and labels of one event (k_classify,
// scan_exclusive of the flags, k_capacity, k_jump, k_labels): roots by pointer
// jumping in LDS, every step replacing a parent by one of its ancestors.
template <int D, int METRIC>
__global__ __launch_bounds__(kFusedThreads) void k_fused_labels(const FusedArgs a) {
  __shared__ int16_t s_root[kFusedMaxPoints];
  __shared__ int16_t s_rank[kFusedMaxPoints];  // local seed rank, -1 for other points
  __shared__ int s_scan[kFusedWaves];
  __shared__ int s_used;
  const int e = blockIdx.x;
  const int tid = threadIdx.x;
  const int begin = a.ev[e].begin, n = a.ev[e].n;
  for (int i = tid; i < n; i += kFusedThreads) {
    const int gi = begin + i;
    const int32_t nhv = a.nh[gi];
    float dist;
    if (nhv == -1) {
      // metric(point_i, PointsView[-1]) with the all-FLT_MAX sentinel (PointsCommon.hpp:75-81)
      float x[D], m[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        x[d] = a.c.p[d][gi];
        m[d] = FLT_MAX;
      }
      dist = metric_distance<D, METRIC>(x, m, a.mp);
    } else {
      dist = a.delta[gi];
    }
    const bool seed = (dist > a.seed_dc) && (a.rho[gi] >= a.rhoc);  // ClusteringKernels.hpp:233
    a.is_seed[gi] = seed ? 1 : 0;
    s_root[i] = static_cast<int16_t>((seed || nhv == -1) ? i : nhv - begin);
    s_rank[i] = seed ? 1 : 0;
    a.nh[gi] = (seed || nhv < 0) ? -1 : nhv - begin;
  }
  __syncthreads();
  // seed ranks in point order (each thread 32 consecutive points)
  const int c0 = tid * (kFusedMaxPoints / kFusedThreads);
  int cnt = 0;
  for (int k = 0; k < kFusedMaxPoints / kFusedThreads; ++k)
    if (c0 + k < n) cnt += s_rank[c0 + k];
  int ns;
  int r = block_incl_scan<kFusedThreads>(cnt, s_scan, ns) - cnt;
  for (int k = 0; k < kFusedMaxPoints / kFusedThreads; ++k) {
    if (c0 + k < n) {
      const bool sd = s_rank[c0 + k] != 0;
      s_rank[c0 + k] = static_cast<int16_t>(sd ? r : -1);
      r += sd ? 1 : 0;
    }
  }
  if (tid == 0) {
    const int32_t cand = a.sc[e];
    int32_t cap = cand;  // SetupSeeds.hpp:14-18
    if (a.persistent) {
      cap = a.counters[1] > cand ? a.counters[1] : cand;
      a.counters[1] = cap;
    }
    const int32_t used = a.unbounded ? ns : cap;
    const int32_t stored = ns < used ? ns : used;
    s_used = used;
    atomicAdd(&a.counters[0], cand);
    atomicAdd(&a.counters[2], used);
    atomicAdd(&a.counters[3], ns);
    atomicAdd(&a.counters[4], stored);
    atomicAdd(&a.counters[5], ns - stored);
  }
  __syncthreads();
  // each round halves every distance to a root: 16 rounds reach it from any
  // depth below 2^15 (the nearest-higher graph is acyclic, a strict order on
  // (density, id)); the bound only keeps a corrupt input from looping
  for (int round = 0; round < 20; ++round) {
    bool changed = false;
    for (int i = tid; i < n; i += kFusedThreads) {
      const int rt = s_root[i];
      const int up = s_root[rt];
      if (up != rt) {
        s_root[i] = static_cast<int16_t>(up);
        changed = true;
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
  const int used = s_used;
  for (int i = tid; i < n; i += kFusedThreads) {
    const int rk = s_rank[s_root[i]];
    a.cluster[begin + i] = (rk >= 0 && rk < used) ? rk : -1;  // ClusteringKernels.hpp:252-279
  }
}

// ===========================================================================
// Launches
// ===========================================================================
template <int D>
clue_status launch_bin(const Coords& c, const float* weights, int32_t n_chunks, clue_plan_s* p, const Wrap& wrap,
                       int sub_bits, hipStream_t st) {
  k_bin<D><<<n_chunks, kThreads, 0, st>>>(c, weights, p->chunks.as<Chunk>(), p->events.as<EventDesc>(), wrap,
                                          p->key.as<int32_t>(), sub_bits >= 0 ? 1 : 0, sub_bits < 0 ? 0 : sub_bits,
                                          p->pev.as<int32_t>(), p->packed.as<float>());
  CLUE_LAUNCH_CHECK("bin");
  return CLUE_OK;
}

template <int D>
clue_status launch_gather(clue_plan_s* p, const int32_t* perm, const int32_t* skey, int32_t n, float h, bool blocks,
                          hipStream_t st) {
  k_gather<D><<<div_up(n, kThreads), kThreads, 0, st>>>(p->packed.as<float>(), perm, skey, n, p->ld,
                                                        p->xs.as<float>(), p->ws.as<float>(), p->hw.as<float>(), h,
                                                        blocks ? p->kn2.as<int32_t>() : nullptr,
                                                        p->tnf.as<int32_t>());
  CLUE_LAUNCH_CHECK("gather");
  return CLUE_OK;
}

template <int D, int METRIC>
clue_status launch_passes(const PassArgs& base_args, int kernel_type, float dc, float dm,
                          hipStream_t st, clue_plan_s* p, bool blocks) {
  constexpr bool U = (D <= 3);
  const int grid = static_cast<int>(div_up(base_args.n, kThreads));
  PassArgs a = base_args;
  a.radius = dc;
  a.thr = sqrt_le_bound(dc);
  clue_status rc = mark(p, 4, st);
  if (rc != CLUE_OK) return rc;
  a.qorder = nullptr;
  a.qperm = nullptr;
  // more than three dimensions: each lane walks its own box; queries sorted by
  // box size class (5 bits) -- for the density pass then by the box's first
  // tile -- keep a wave's boxes alike (a stable sort)
  auto order_by_box = [&](float r, bool by_tile) -> clue_status {
    const int tile_bits = by_tile ? std::min(bits_for(a.ntiles), 26) : 0;
    k_box_class<D><<<grid, kThreads, 0, st>>>(a, r, tile_bits, p->qk.as<int32_t>());
    CLUE_LAUNCH_CHECK("box_class");
    int32_t* keys = nullptr;
    int32_t* vals = nullptr;
    const clue_status r2 = radix_sort(p, st, a.n, 5 + tile_bits, p->qk.as<int32_t>(), p->qk2.as<int32_t>(),
                                      p->qv.as<int32_t>(), p->qv2.as<int32_t>(), false, &keys, &vals);
    a.qperm = vals;
    return r2;
  };
  // the nearest-higher pass (after either density pass)
  // one wave per query (k_density_wave, k_nearest_higher_wave) for small calls
  const bool wave_mode = D > p->tuning.wave_dims || (p->tuning.small_call_points > 0 && a.n <= p->tuning.small_call_points);
  // four waves per 64 queries in the density pass (the filtered walk's split
  // variant) and the first nearest-higher phase, for calls of at most
  // split_call_points points with the Morton layout; the density's query blocks
  // are then 64 queries, unordered, and the split density takes precedence over
  // the wave-per-query one
  const bool split = U && METRIC == MEUCLID && blocks && p->tuning.split_call_points > 0 &&
                     a.n <= p->tuning.split_call_points;
  const bool dsplit = split && a.dfilter;
  auto nearest_higher = [&]() -> clue_status {
    a.radius = dm;
    if constexpr (!U) {
      if (!wave_mode) {
        rc = order_by_box(dm, false);  // the first tile does not pay here (measured)
        if (rc != CLUE_OK) return rc;
      }
    }
    a.thr = sqrt_le_bound(dm);
    rc = mark(p, 5, st);
    if (rc != CLUE_OK) return rc;
    if constexpr (METRIC == MEUCLID && D <= 3) {
      if (blocks) {
        // automatic first radius and group: measured per dimension (3D 4M points:
        // 3.76 ms at 0.15 dm and 16 lanes, 3.0 ms at 0.2 dm and 8 lanes, and with
        // the first phase walking its own tile from the wave's block, 2.70 ms at
        // 0.2 dm, 2.61 at 0.3 and 0.4; 2D 1M points: 0.47 ms at 0.15 dm and 16
        // lanes, 0.50 ms at 0.2 dm and 8 lanes, now 0.360 ms at 0.15 dm and 16 lanes,
        // 0.350 at 0.3 dm and 8, 0.356 at 0.25 and 8, 0.371 at 0.15 and 8)
        const bool d23 = D == 2 || D == 3;
        const float r1 = p->tuning.nh_first_radius > 0.f ? p->tuning.nh_first_radius : (d23 ? 0.3f : 0.15f);
        const int group = p->tuning.pending_group > 0 ? p->tuning.pending_group : (d23 ? 8 : 16);
        a.thr1 = sqrt_le_bound(dm * r1);
        a.box1 = dm * r1 * (1.f + 1.f / 1024.f);
        // second phase levels: twice the radius each time, dm last
        a.nlvl = 0;
        a.lvl_prev0 = std::min(a.thr1, a.thr);
        float r = dm * r1;
        if (r < dm) {
          while (a.nlvl < 7) {
            r *= 2.f;
            if (!(r < dm)) break;
            a.lvl[a.nlvl++] = sqrt_le_bound(r);
          }
          a.lvl[a.nlvl++] = a.thr;
        }
        rc = launch_nh_maxkeys(p, a, st);
        if (rc != CLUE_OK) return rc;
        if (split) k_nearest_higher_blocks<D, true><<<div_up(a.n, kWave), kThreads, 0, st>>>(a);
        else k_nearest_higher_blocks<D><<<grid, kThreads, 0, st>>>(a);
        CLUE_LAUNCH_CHECK("nearest_higher_blocks");
        // enough waves to fill the GPU; surplus waves find the pending list exhausted
        // default grid of the persistent second phase: 8192 workgroups for calls
        // of 2^20 points or more (2D 1M points: 0.51 -> 0.49 ms, 3D 4M: 7.43 ->
        // 7.31 ms; the 4096-event batch's nearest-higher stage 5.28 -> 4.93 ms
        // against 2048), 2048 otherwise (small served requests lose more than
        // they gain)
        const int pblocks = p->tuning.pending_blocks > 0 ? p->tuning.pending_blocks
                                                         : (a.n >= (1 << 20) ? 8192 : 2048);
        const int nblk = static_cast<int>(std::min<int64_t>(pblocks, div_up(a.n, kThreads / kWave)));
        const bool single = a.n_events == 1;
        switch (group) {
          case 8:
            if (single) k_nh_pending<D, 8, true><<<nblk, kThreads, 0, st>>>(a);
            else k_nh_pending<D, 8, false><<<nblk, kThreads, 0, st>>>(a);
            break;
          case 16:
            if (single) k_nh_pending<D, 16, true><<<nblk, kThreads, 0, st>>>(a);
            else k_nh_pending<D, 16, false><<<nblk, kThreads, 0, st>>>(a);
            break;
          case 32:
            if (single) k_nh_pending<D, 32, true><<<nblk, kThreads, 0, st>>>(a);
            else k_nh_pending<D, 32, false><<<nblk, kThreads, 0, st>>>(a);
            break;
          default:
            if (single) k_nh_pending<D, 64, true><<<nblk, kThreads, 0, st>>>(a);
            else k_nh_pending<D, 64, false><<<nblk, kThreads, 0, st>>>(a);
            break;
        }
        CLUE_LAUNCH_CHECK("nearest_higher_pending");
        return CLUE_OK;
      }
    }
    if (wave_mode) {
      k_nearest_higher_wave<D, METRIC><<<div_up(a.n, kThreads / kWave), kThreads, 0, st>>>(a);
      CLUE_LAUNCH_CHECK("nearest_higher_wave");
      return CLUE_OK;
    }
    k_nearest_higher<D, METRIC, U><<<grid, kThreads, 0, st>>>(a);
    CLUE_LAUNCH_CHECK("nearest_higher");
    return CLUE_OK;
  };
  // small calls: one wave per query (k_density_wave), the same sums in the same order,
  // unless the call runs the split density
  if (wave_mode && !dsplit) {
    const int wgrid = static_cast<int>(div_up(a.n, kThreads / kWave));
    if constexpr (METRIC == MEUCLID) {
      if (blocks) {
        const int unit = a.exp_unit;
        switch (kernel_type) {
          case KFLAT: k_density_wave<D, METRIC, KFLAT, true, 0><<<wgrid, kThreads, 0, st>>>(a); break;
          case KGAUSS:
            if (unit == 2) k_density_wave<D, METRIC, KGAUSS, true, 2><<<wgrid, kThreads, 0, st>>>(a);
            else if (unit) k_density_wave<D, METRIC, KGAUSS, true, 1><<<wgrid, kThreads, 0, st>>>(a);
            else k_density_wave<D, METRIC, KGAUSS, true, 0><<<wgrid, kThreads, 0, st>>>(a);
            break;
          case KEXP:
            if (unit) k_density_wave<D, METRIC, KEXP, true, 1><<<wgrid, kThreads, 0, st>>>(a);
            else k_density_wave<D, METRIC, KEXP, true, 0><<<wgrid, kThreads, 0, st>>>(a);
            break;
          default: return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown convolutional kernel type");
        }
        CLUE_LAUNCH_CHECK("density_wave");
        return nearest_higher();
      }
    }
    switch (kernel_type) {
      case KFLAT: k_density_wave<D, METRIC, KFLAT, false, 0><<<wgrid, kThreads, 0, st>>>(a); break;
      case KGAUSS: k_density_wave<D, METRIC, KGAUSS, false, 0><<<wgrid, kThreads, 0, st>>>(a); break;
      case KEXP: k_density_wave<D, METRIC, KEXP, false, 0><<<wgrid, kThreads, 0, st>>>(a); break;
      default: return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown convolutional kernel type");
    }
    CLUE_LAUNCH_CHECK("density_wave");
    return nearest_higher();
  }
  if constexpr (!U) {
    rc = order_by_box(dc, true);
    if (rc != CLUE_OK) return rc;
  }
  // heaviest-first pays in 2D (a shorter tail); in 3D the lost spatial locality
  // of concurrent blocks costs more than it saves (measured), hence the default
  const bool order = p->tuning.density_order > 0 || (p->tuning.density_order < 0 && D <= 2);
  if (blocks && order && grid <= kOrderMax && !dsplit) {
    rc = launch_density_order(p, a, grid, st);
    if (rc != CLUE_OK) return rc;
    a.qorder = p->qorder.as<int32_t>();
  }
  if (blocks) {
    int unit = 0;
    if constexpr (METRIC == MEUCLID) unit = a.exp_unit;
    constexpr int kU1 = METRIC == MEUCLID ? 1 : 0, kU2 = METRIC == MEUCLID ? 2 : 0;
    auto launch = [&](auto filt_c, auto split_c) -> bool {
      constexpr bool F = decltype(filt_c)::value;
      constexpr bool S = decltype(split_c)::value;
      const int g = S ? static_cast<int>(div_up(a.n, kWave)) : grid;
      switch (kernel_type) {
        case KFLAT: k_density<D, METRIC, KFLAT, U, true, 0, F, S><<<g, kThreads, 0, st>>>(a); return true;
        case KGAUSS:
          if (unit == 2) k_density<D, METRIC, KGAUSS, U, true, kU2, F, S><<<g, kThreads, 0, st>>>(a);
          else if (unit) k_density<D, METRIC, KGAUSS, U, true, kU1, F, S><<<g, kThreads, 0, st>>>(a);
          else k_density<D, METRIC, KGAUSS, U, true, 0, F, S><<<g, kThreads, 0, st>>>(a);
          return true;
        case KEXP:
          if (unit) k_density<D, METRIC, KEXP, U, true, kU1, F, S><<<g, kThreads, 0, st>>>(a);
          else k_density<D, METRIC, KEXP, U, true, 0, F, S><<<g, kThreads, 0, st>>>(a);
          return true;
      }
      return false;
    };
    constexpr std::false_type no{};
    constexpr std::true_type yes{};
    bool ok;
    if constexpr (U && METRIC == MEUCLID) ok = !a.dfilter ? launch(no, no) : dsplit ? launch(yes, yes) : launch(yes, no);
    else ok = launch(no, no);
    if (!ok) return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown convolutional kernel type");
  } else {
    switch (kernel_type) {
      case KFLAT: k_density<D, METRIC, KFLAT, U, false><<<grid, kThreads, 0, st>>>(a); break;
      case KGAUSS: k_density<D, METRIC, KGAUSS, U, false><<<grid, kThreads, 0, st>>>(a); break;
      case KEXP: k_density<D, METRIC, KEXP, U, false><<<grid, kThreads, 0, st>>>(a); break;
      default: return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown convolutional kernel type");
    }
  }
  CLUE_LAUNCH_CHECK("density");
  return nearest_higher();
}

template <int D, int METRIC>
clue_status launch_classify(const ClassifyArgs& a, hipStream_t st) {
  k_classify<D, METRIC><<<div_up(a.n, kThreads), kThreads, 0, st>>>(a);
  CLUE_LAUNCH_CHECK("classify");
  return CLUE_OK;
}

template <int D>
clue_status launch_fused_layout(const FusedArgs& a, int32_t n_events, hipStream_t st) {
  k_fused_layout<D><<<n_events, kFusedThreads, 0, st>>>(a);
  CLUE_LAUNCH_CHECK("fused_layout");
  return CLUE_OK;
}

template <int D>
clue_status dispatch_fused_labels(int metric, const FusedArgs& a, int32_t n_events, hipStream_t st) {
  switch (metric) {
    case MEUCLID: k_fused_labels<D, MEUCLID><<<n_events, kFusedThreads, 0, st>>>(a); break;
    case MWEUCLID: k_fused_labels<D, MWEUCLID><<<n_events, kFusedThreads, 0, st>>>(a); break;
    case MPERIODIC: k_fused_labels<D, MPERIODIC><<<n_events, kFusedThreads, 0, st>>>(a); break;
    case MMANHATTAN: k_fused_labels<D, MMANHATTAN><<<n_events, kFusedThreads, 0, st>>>(a); break;
    case MCHEBYSHEV: k_fused_labels<D, MCHEBYSHEV><<<n_events, kFusedThreads, 0, st>>>(a); break;
    case MWCHEBYSHEV: k_fused_labels<D, MWCHEBYSHEV><<<n_events, kFusedThreads, 0, st>>>(a); break;
    default: return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown distance metric type");
  }
  CLUE_LAUNCH_CHECK("fused_labels");
  return CLUE_OK;
}

template <int D>
clue_status dispatch_metric_passes(int metric, const PassArgs& a, int kt, float dc, float dm,
                                   hipStream_t st, clue_plan_s* p, bool blocks) {
  switch (metric) {
    case MEUCLID: return launch_passes<D, MEUCLID>(a, kt, dc, dm, st, p, blocks);
    case MWEUCLID: return launch_passes<D, MWEUCLID>(a, kt, dc, dm, st, p, false);
    case MPERIODIC: return launch_passes<D, MPERIODIC>(a, kt, dc, dm, st, p, false);
    case MMANHATTAN: return launch_passes<D, MMANHATTAN>(a, kt, dc, dm, st, p, false);
    case MCHEBYSHEV: return launch_passes<D, MCHEBYSHEV>(a, kt, dc, dm, st, p, false);
    case MWCHEBYSHEV: return launch_passes<D, MWCHEBYSHEV>(a, kt, dc, dm, st, p, false);
  }
  return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown distance metric type");
}

template <int D>
clue_status dispatch_metric_classify(int metric, const ClassifyArgs& a, hipStream_t st) {
  switch (metric) {
    case MEUCLID: return launch_classify<D, MEUCLID>(a, st);
    case MWEUCLID: return launch_classify<D, MWEUCLID>(a, st);
    case MPERIODIC: return launch_classify<D, MPERIODIC>(a, st);
    case MMANHATTAN: return launch_classify<D, MMANHATTAN>(a, st);
    case MCHEBYSHEV: return launch_classify<D, MCHEBYSHEV>(a, st);
    case MWCHEBYSHEV: return launch_classify<D, MWCHEBYSHEV>(a, st);
  }
  return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown distance metric type");
}

template clue_status launch_bin<CLUE_DIM>(const Coords&, const float*, int32_t, clue_plan_s*, const Wrap&, int,
                                          hipStream_t);
template clue_status launch_gather<CLUE_DIM>(clue_plan_s*, const int32_t*, const int32_t*, int32_t, float, bool,
                                             hipStream_t);
template clue_status dispatch_metric_passes<CLUE_DIM>(int, const PassArgs&, int, float, float, hipStream_t,
                                                      clue_plan_s*, bool);
template clue_status launch_fused_layout<CLUE_DIM>(const FusedArgs&, int32_t, hipStream_t);
template clue_status dispatch_fused_labels<CLUE_DIM>(int, const FusedArgs&, int32_t, hipStream_t);
template clue_status dispatch_metric_classify<CLUE_DIM>(int, const ClassifyArgs&, hipStream_t);

}  // namespace clue_hip
