// clue_internal.hpp -- shared declarations of the engine's translation units:
// launch constants, per-call argument structs, wave-level device helpers, the
// plan state and the host helpers.  The templated neighbour kernels live in
// clue_passes.hip, compiled once per dimension; everything else in clue_engine.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "clue_abi.h"
#include "clue_device.hpp"

namespace clue_hip {

constexpr int kThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kThreads * kScanItems;  // 4096
constexpr int kScanRawBlocks = 4 * kThreads;
constexpr int kSortItems = 16;
constexpr int kSortTile = kThreads * kSortItems;  // 4096 keys per block, 1024 per wave
constexpr int kBBoxBlocks = 256;
// Stages timed by clue_plan_stage_times, in pipeline order.
constexpr int kStages = 8;

struct Coords {
  const float* p[kMaxDim];
};

// ===========================================================================
// Block-level helpers
// ===========================================================================
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

// Inclusive scan across a block of NT threads; `total` receives the block sum.
template <int NT>
__device__ __forceinline__ int block_incl_scan(int v, int* smem, int& total) {
  constexpr int NW = NT / kWave;
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  int incl = wave_incl_scan(v);
  if (lane == kWave - 1) smem[w] = incl;
  __syncthreads();
  if (w == 0) {
    int t = lane < NW ? smem[lane] : 0;
    t = wave_incl_scan(t);
    if (lane < NW) smem[lane] = t;
  }
  __syncthreads();
  if (w > 0) incl += smem[w - 1];
  total = smem[NW - 1];
  __syncthreads();
  return incl;
}

// Workgroups are dispatched to the 8 XCDs round robin (block b to XCD b % 8).
// With xcd set, XCD x takes the contiguous range x of the logical blocks, so
// neighbouring query blocks -- which read the same tiles -- share one L2.
__device__ __forceinline__ int logical_block(int xcd) {
  const int b = blockIdx.x;
  if (!xcd) return b;
  const int nb = gridDim.x;
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, i = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

constexpr int kChunk = 4096;  // points per chunk

template <int D>
__device__ __forceinline__ EvGrid<D> load_event(const EventDesc* events, int e, const Wrap& wrap) {
  cptr<EventDesc> ev = as_const(events) + e;
  EvGrid<D> g;
  g.npd = ev->npd;
  g.tile_base = ev->tile_base;
  g.ntiles = ev->ntiles;
  g.begin = ev->begin;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    g.stride[d] = ev->stride[d];
    g.mn[d] = ev->mn[d];
    g.mx[d] = ev->mx[d];
    g.ts[d] = ev->ts[d];
    g.wrapped[d] = wrap.w[d] != 0;
  }
  return g;
}

// Floats per point of the packed copy written by k_bin: D coordinates, the
// weight and the (tile, Morton code) key, rounded up to whole 16-byte vectors.
template <int D>
constexpr int packed_stride() { return (D + 2 + 3) / 4 * 4; }

constexpr int kBlock = 32;       // points per block of the Morton layout
constexpr int kBoxLanes = 8;    // lanes per block of the block boxes
constexpr int kTileLanes = 16;  // lanes per tile of the tile boxes
constexpr int kOrderMax = 16384;  // query blocks the in-LDS ordering handles

// ===========================================================================
// Density (K4) and nearest-higher (K5).
// ===========================================================================
struct PassArgs {
  const float* xs[kMaxDim];  // tile-ordered coordinates
  const float* ws;           // tile-ordered weights
  const float* hw;           // tile-ordered h*w (flat kernel)
  const float* rho_s;        // tile-ordered densities (nearest-higher input)
  const int32_t* perm;       // slot -> point id
  const int32_t* offsets;    // tile CSR offsets (global tile ids)
  const EventDesc* ev;       // per-event grids
  const int32_t* pev;        // point id -> event
  int32_t* sc;               // per-event count of points without a nearest-higher
  Wrap wrap;
  int32_t n;
  int32_t n_events;
  float radius;              // dc (density) or dm (nearest-higher)
  float thr;                 // max{s : sqrtf(s) <= radius}
  float kp[3];
  double kden_rcp;           // Gaussian: 1 / ((2*sigma)*sigma) rounded to double
  float mp[kMaxDim];
  // Morton-ordered layout (queries of both passes; candidates of the block search)
  const float* xn[kMaxDim];
  const float* wn;
  const int32_t* sn;         // Morton position -> tile-order slot
  const int32_t* pn;         // Morton position -> point id
  const int32_t* kn;         // per Morton position: its tile id << mshift (with the Morton code below when mshift > 0)
  int32_t mshift;
  const int32_t* blk_off;    // first block of each tile
  const int32_t* btile;      // tile of each block
  int32_t ntiles;
  const float* bmin[kMaxDim];
  const float* bmax[kMaxDim];
  const float* tmin[kMaxDim];
  const float* tmax[kMaxDim];
  int32_t* pending;          // nearest-higher: Morton positions left for the full search
  int32_t* pending_count;    // its length (cleared by k_block_maxkey, then atomic appends)
  float thr1;                // first-phase search radius threshold
  float box1;                // first-phase search box radius (<= dm)
  float lvl[8];              // second phase: squared radius thresholds of its levels, the last one dm's
  int nlvl;
  float lvl_prev0;           // threshold every first-phase candidate was examined within
  uint64_t* hk_out;          // density: rank_key of each Morton position
  const uint64_t* hk;        // nearest-higher input: rank_key, Morton order
  const uint64_t* bmaxk;     // largest rank_key of each block
  const uint64_t* tmaxk;     // largest rank_key of each tile
  float* rho_s_out;          // density: tile order
  float* rho_out;            // density: original order
  int32_t* nh_out;           // nearest-higher: original order
  float* delta_out;          // nearest-higher distance: original order
  int xcd;                   // bit 0: XCD-aware block order of density, bit 1: of nearest-higher
  int exp_nonpos;            // the kernel's exp argument is always <= 0 and never NaN
  int exp_unit;              // and >= -14 for in-radius distances, with amplitude exactly 1;
                             // 2: a Gaussian of mean 0 and power-of-two (2 sigma) sigma besides
  const int32_t* qorder;     // density: query blocks in decreasing estimated cost (or null)
  const int32_t* tnf;        // density: 1 for tiles holding an infinite or NaN weight
  const int32_t* qperm;      // tile-order walks: query slots by growing search-box size (or null)
  int dfilter;               // density: filtered union walk (tuning density_filter)
  int nh_lone;               // first nearest-higher phase: blocks that at most this many lanes need
                             // are scanned across the wave (tuning nh_lone_lanes)
};


// Search box of a point: [bin(x - r), bin(x + r)] per dimension, wrapped
// dimensions unrolled past npd (TilesView::searchBox, TilesView.hpp:72-82).
template <int D>
__device__ __forceinline__ void search_box(const float* q, float r, const EvGrid<D>& g, int* lo,
                                           int* hi) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int inf = tile_bin(q[d] - r, g.mn[d], g.mx[d], g.ts[d], g.npd, g.wrapped[d]);
    int sup = tile_bin(q[d] + r, g.mn[d], g.mx[d], g.ts[d], g.npd, g.wrapped[d]);
    if (g.wrapped[d] && inf > sup) sup += g.npd;
    lo[d] = inf;
    hi[d] = sup;
  }
}

// Global tile id of a box position (TilesView::getGlobalBinByBin, TilesView.hpp:63-70).
template <int D>
__device__ __forceinline__ int tile_index(const int* b, const EvGrid<D>& g) {
  int t = g.tile_base;
#pragma unroll
  for (int d = 0; d < D; ++d) t += g.stride[d] * (g.wrapped[d] ? b[d] % g.npd : b[d]);
  return t;
}

// Runs body(e, act) once per event present among the wave's valid lanes, with
// act true on the lanes of that event (wave-uniform e).
template <typename F>
__device__ __forceinline__ void for_each_event(bool valid, int my_e, F&& body) {
  uint64_t todo = __ballot(valid);
  while (todo) {
    const int e = __builtin_amdgcn_readlane(my_e, __builtin_ctzll(todo));
    const bool act = valid && my_e == e;
    todo &= ~__ballot(act);
    body(e, act);
  }
}

// Visitor over the tiles a lane needs, in the reference order (lexicographic,
// dimension 0 outermost; within a tile, CSR order).  UNION: the wave walks the
// union of its lanes' boxes with wave-uniform tile and candidate indices (the
// candidates arrive through scalar loads); body(tile, begin, end, need) runs
// with the whole wave and `need` tells a lane whether the tile is in its box.
// Otherwise each lane walks its own box.
template <int D, bool UNION, typename F>
__device__ __forceinline__ void visit_box(bool valid, const int* lo, const int* hi,
                                          const EvGrid<D>& gs, cptr<int32_t> offsets, F&& body) {
  if constexpr (UNION) {
    int ulo[D], uhi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ulo[d] = uniform(wave_min_i32(valid ? lo[d] : INT_MAX));
      uhi[d] = uniform(wave_max_i32(valid ? hi[d] : INT_MIN));
      if (ulo[d] > uhi[d]) return;
    }
    int b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = ulo[d];
    while (true) {
      bool need = valid;
#pragma unroll
      for (int d = 0; d < D; ++d) need = need && b[d] >= lo[d] && b[d] <= hi[d];
      if (__ballot(need) != 0) {
        const int g = uniform(tile_index<D>(b, gs));
        body(g, offsets[g], offsets[g + 1], need);
      }
      int d = D - 1;
      for (; d >= 0; --d) {
        if (++b[d] <= uhi[d]) break;
        b[d] = ulo[d];
      }
      if (d < 0) break;
    }
  } else {
    if (!valid) return;
    int b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = lo[d];
    while (true) {
      const int g = tile_index<D>(b, gs);
      body(g, offsets[g], offsets[g + 1], true);
      int d = D - 1;
      for (; d >= 0; --d) {
        if (++b[d] <= hi[d]) break;
        b[d] = lo[d];
      }
      if (d < 0) break;
    }
  }
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ float uniform_f32(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// Number of set bits of m below this lane.
__device__ __forceinline__ int mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0));
}
// Neighbouring floats (nextafterf towards -inf / +inf); NaN stays NaN.
__device__ __forceinline__ float next_down(float x) {
  if (x != x || x == -__builtin_inff()) return x;
  if (x == 0.f) return -0x1p-149f;
  const int b = __float_as_int(x);
  return __int_as_float(x > 0.f ? b - 1 : b + 1);
}
__device__ __forceinline__ float next_up(float x) { return -next_down(-x); }

// Whether every candidate within radius r of q (Euclidean, squared distance
// <= thr) lies in q's search box [lo, hi], so that a walk may ignore which
// tiles a lane's box holds.  A candidate c outside the box in dimension d has
// c < m = fl(q - r) (or c > M = fl(q + r)): the bins are monotone while
// (x - min)/tile_size stays inside the int range, as it does for every
// candidate.  Then fl(q - c) >= fl(q - next_down(m)), and once that exceeds
// R = max(r, 2^-40), fl((q-c)^2) and with it the squared distance exceed thr
// (sqrtf(fl(x*x)) = |x| for |x| >= 2^-40).  Boxes touching the grid's edge, and
// dimensions of zero extent (every bin 0, the box [0, 0]), exclude nothing.
// Non-finite queries, wrapped dimensions and overflowing bins answer false.
template <int D>
__device__ __forceinline__ bool box_exact(const float* q, float r, float R, const EvGrid<D>& g, const int* lo,
                                          const int* hi) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    ok = ok && !g.wrapped[d] && __builtin_isfinite(q[d]);
    if (g.ts[d] == 0.f) continue;
    ok = ok && __builtin_isfinite(g.ts[d]) && g.ts[d] > 0.f;
    if (lo[d] > 0) {
      const float m = q[d] - r;
      const float v = (m - g.mn[d]) / g.ts[d];
      ok = ok && v >= -2147483648.0f && v < 2147483648.0f && (q[d] - next_down(m)) > R;
    }
    if (hi[d] < g.npd - 1) {
      const float mm = q[d] + r;
      const float v = (mm - g.mn[d]) / g.ts[d];
      ok = ok && v >= -2147483648.0f && v < 2147483648.0f && (next_up(mm) - q[d]) > R;
    }
  }
  return ok;
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int off) {
  const uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), off);
  const uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), off);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = shfl_xor_u64(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// ===========================================================================
// Seeds (K6), seed ranks and labels (K7).
// ===========================================================================
struct ClassifyArgs {
  Coords c;
  const float* rho;
  const float* delta;
  int32_t* nh;        // in: nearest-higher; out: -1 for seeds
  int32_t* is_seed;
  int32_t* flag;      // seed flags to be scanned into ranks
  int32_t* root;      // parent pointer, roots point to themselves
  int32_t n;
  float seed_dc;
  float rhoc;
  float mp[kMaxDim];
  int32_t* counters;  // call totals, cleared here: [0], [1] when clear_cap, [2, 11)
  int clear_cap;
};

// ===========================================================================
// Small calls fused per event: one workgroup of kFusedThreads builds an event's
// layout (bounding box, tile ids, stable counting sort into the tile CSR,
// tile-ordered copies), another its seeds, seed ranks, capacity and labels.
// ===========================================================================
constexpr int kFusedThreads = 1024;
constexpr int kFusedWaves = kFusedThreads / kWave;
constexpr int kFusedMaxPoints = 32768;  // per event: local indices fit 16 bits
constexpr int kFusedMaxTiles = 2048;    // per event: 16 per-wave counters per tile in LDS

struct FusedArgs {
  Coords c;            // caller coordinates, original order
  const float* w;      // caller weights
  EventDesc* ev;       // per-event grids: extremes and tile sizes are written here
  Wrap wrap;
  int32_t* pev;        // point -> event
  int32_t* perm;       // slot -> point id
  int32_t* offsets;    // tile CSR over global tile ids
  float* xs;           // tile-ordered coordinates, leading dimension ld
  int64_t ld;
  float* ws;           // tile-ordered weights
  float* hw;           // tile-ordered h * w (flat kernel)
  float h;
  int32_t* tnf;        // tiles holding an infinite or NaN weight
  int32_t* sc;         // per event: points without a nearest-higher (cleared by the layout)
  int32_t* counters;   // call totals (cleared by the layout: [0], [1] when clear_cap, [2, 11))
  int clear_cap;
  // seeds and labels
  const float* rho;    // original order
  const float* delta;  // nearest-higher distance, original order
  int32_t* nh;         // in: nearest-higher (call-wide ids); out: local to the event, -1 for seeds
  int32_t* is_seed;
  int32_t* cluster;
  float seed_dc;
  float rhoc;
  float mp[kMaxDim];
  int unbounded;       // every seed stored (no capacity truncation)
  int persistent;      // the plan's grow-only capacity applies (plain calls)
};

// Lanes of the wave holding the same key (< 2^bits) as this one, among the
// lanes with `in` set (none for a lane without).
__device__ __forceinline__ uint64_t match_peers(int key, bool in, int bits) {
  uint64_t peers = __ballot(in);
  for (int b = 0; b < bits; ++b) {
    const bool bit = (key >> b) & 1;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return in ? peers : 0;
}

// ===========================================================================
// Host side
// ===========================================================================
extern thread_local std::string g_error;
clue_status fail(clue_status st, const std::string& msg);
void set_error(const std::string& msg);

#define CLUE_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(e_ == hipErrorOutOfMemory ? CLUE_ERR_OUT_OF_MEMORY : CLUE_ERR_RUNTIME,   \
                  std::string(#call) + ": " + hipGetErrorString(e_));                      \
  } while (0)

#define CLUE_LAUNCH_CHECK(what)                                                            \
  do {                                                                                     \
    hipError_t e_ = hipGetLastError();                                                     \
    if (e_ != hipSuccess)                                                                  \
      return fail(CLUE_ERR_RUNTIME, std::string("launch ") + what + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      bytes = 0;
    }
    const size_t alloc = need < 256 ? 256 : need;
    hipError_t e = hipMalloc(&p, alloc);
    if (e == hipSuccess) bytes = alloc;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace clue_hip

using namespace clue_hip;

struct clue_plan_s {
  int device = 0;
  DevBuf key, key2, perm, perm2, offsets, hist, part, xs, ws, hw, rho_s, delta, flag, root,
      nh_tmp, rho_tmp, seed_tmp, bbox_part, counters;
  // events of the call: descriptors, chunks, point -> event, per-event counters
  DevBuf events, chunks, pev, sc, cap_used, rank_base;
  // Morton-ordered layout and the block boxes of the nearest-higher search
  DevBuf btile;
  DevBuf packed;  // per point: coordinates, weight, Morton key (k_bin -> k_gather)
  DevBuf kn, kn2, pn, pn2, mid, xn, wn, hk, bcnt, blk_off, bbox_lo, bbox_hi, tbox_lo, tbox_hi;
  DevBuf bmaxk, tmaxk;  // largest rank key of each block / tile
  DevBuf qk, qk2, qv, qv2;  // more than three dimensions: sort of the queries by box size
  DevBuf tnf;           // tiles holding an infinite or NaN weight
  DevBuf pending, pending_count, qcost, qorder;
  // cluster -> points map (post-processing): sort keys and values
  DevBuf mkey, mkey2, mval, mval2;
  clue_tuning tuning{0.f, -1, 7, 0, 0, -1, 65536, 3, 32768, 1, 12288, 262144, 4};
  int64_t ldb = 0;  // leading dimension of the block boxes
  int64_t ld = 0;   // leading dimension of xs / xn
  // pinned staging ring of the event and chunk tables (a slot is reused once its
  // copy has run, so the host can enqueue a few calls ahead of the device)
  struct StageSlot {
    void* host = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
  };
  static constexpr int kStageSlots = 4;
  StageSlot stage[kStageSlots];
  int stage_next = 0;
  // the next plain call starts with a fresh grow-only seed capacity
  bool reset_capacity = false;
  // last call
  bool has_run = false;
  hipEvent_t done = nullptr;  // recorded after each call's last kernel (introspection waits on it)
  int32_t last_n = 0, last_d = 0, last_T = 0, last_events = 0;
  int32_t* last_perm = nullptr;
  // a plan over a region of device memory it does not own (the no-allocation
  // entry point): reserve() carves every buffer from it and allocates nothing
  unsigned char* arena = nullptr;
  size_t arena_bytes = 0, arena_used = 0;
  // optional per-stage HIP event timing of the last call
  bool timing = false;
  hipEvent_t ev[kStages + 1] = {};

  DevBuf* all_buffers(int k) {
    DevBuf* b[] = {&key,     &key2,   &perm,     &perm2,   &offsets,  &hist,    &part,
                   &xs,      &ws,     &hw,       &rho_s,   &delta,   &flag,     &root,    &nh_tmp,
                   &rho_tmp, &seed_tmp, &bbox_part, &counters, &events, &chunks, &pev,    &sc,
                   &cap_used, &rank_base, &kn,   &kn2,     &pn,      &pn2,      &mid,     &xn,
                   &wn,      &hk,       &bcnt,    &blk_off, &bbox_lo,  &bbox_hi, &tbox_lo,
                   &tbox_hi, &bmaxk, &tmaxk, &tnf, &qk, &qk2, &qv, &qv2, &pending, &pending_count, &qcost, &qorder, &mkey,  &mkey2,
                   &mval,    &mval2,  &btile, &packed};
    return k < static_cast<int>(sizeof(b) / sizeof(b[0])) ? b[k] : nullptr;
  }
};

namespace clue_hip {

void grid_shape(int32_t D, int32_t n, int32_t ppb, int32_t& npd, int32_t& T);
float sqrt_le_bound(float r);
clue_status scan_exclusive(const int32_t* in, int32_t* out, int64_t n, int32_t* part, hipStream_t st);
clue_status mark(clue_plan_s* p, int k, hipStream_t st);
clue_status radix_sort(clue_plan_s* p, hipStream_t st, int32_t n, int total_bits, int32_t* k0,
                       int32_t* k1, int32_t* v0, int32_t* v1, bool keep_keys,
                       int32_t** keys_sorted, int32_t** vals_sorted);
int bits_for(int64_t count);
// launches of the non-templated kernels the neighbour passes need (clue_engine.hip)
clue_status launch_density_order(clue_plan_s* p, const PassArgs& a, int grid, hipStream_t st);
clue_status launch_nh_maxkeys(clue_plan_s* p, const PassArgs& a, hipStream_t st);

// per-dimension entry points (clue_passes.hip, one translation unit per D)
template <int D>
clue_status launch_bin(const Coords& c, const float* weights, int32_t n_chunks, clue_plan_s* p, const Wrap& wrap,
                       int sub_bits, hipStream_t st);
template <int D>
clue_status launch_gather(clue_plan_s* p, const int32_t* perm, const int32_t* skey, int32_t n, float h, bool blocks,
                          hipStream_t st);
template <int D>
clue_status dispatch_metric_passes(int metric, const PassArgs& a, int kt, float dc, float dm,
                                   hipStream_t st, clue_plan_s* p, bool blocks);
template <int D>
clue_status dispatch_metric_classify(int metric, const ClassifyArgs& a, hipStream_t st);
template <int D>
clue_status launch_fused_layout(const FusedArgs& a, int32_t n_events, hipStream_t st);
template <int D>
clue_status dispatch_fused_labels(int metric, const FusedArgs& a, int32_t n_events, hipStream_t st);

template <typename F>
clue_status dispatch_dim(int D, F&& f) {
  switch (D) {
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
  return fail(CLUE_ERR_INVALID_ARGUMENT, "number of dimensions must be in [1, 10]");
}

}  // namespace clue_hip
