// clue_engine.hip -- MI355X (gfx950) CLUE engine: kernels, plan state and the C ABI.
//
// Pipeline of one clustering call (all on one HIP stream, no host round trip):
//   bbox        per-dimension min/max, tile sizes      ComputeTiles.hpp:13-27
//   bin         tile id per point + tile histogram     Tiles.hpp:76-89, detail/AssociationMap.hpp:13-48
//   sort        stable LSD counting sort of (tile, id) AssociationMap.hpp:122-191 (serial order)
//   offsets     exclusive scan of the histogram        AssociationMap.hpp:170-174
//   gather      tile-ordered SoA copy of coords/weight
//   density     K4, rho in the reference visit order   ClusteringKernels.hpp:18-107
//   nearest     K5, (distance, visit order) argmin     ClusteringKernels.hpp:109-212
//   classify    K6, seed flags                         ClusteringKernels.hpp:214-244
//   seed rank   exclusive scan of the seed flags       (serial push_back order)
//   labels      K7 by pointer jumping to the root      ClusteringKernels.hpp:246-281
#include <mutex>
#include <new>

#include "clue_internal.hpp"

namespace clue_hip {

// Points per tile the autotuner tries (with CLUE_TUNE_POINTS_PER_TILE).
constexpr int32_t kTunePointsPerTile[] = {16, 32, 64, 128, 256, 512};

// Stages timed by clue_plan_stage_times, in pipeline order.
const char* const kStageNames[kStages] = {"bbox",    "bin",       "sort",     "offsets_gather",
                                          "density", "nearest_higher", "classify_rank", "labels"};


__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ===========================================================================
// Exclusive scan of int32 (3 phases).  out has n+1 entries; out[n] = total.
// in and out may alias.
// ===========================================================================
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const int32_t* in, int64_t n,
                                                           int32_t* part) {
  __shared__ int smem[kThreads / kWave];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t idx = base + k * kThreads + threadIdx.x;
    if (idx < n) sum += in[idx];
  }
  int total;
  block_incl_scan<kThreads>(sum, smem, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Short arrays (at most kScanSmall entries): one workgroup scans them and writes
// out[n] (one launch instead of two; for longer arrays one workgroup is slower
// than the two launches over many blocks).
constexpr int kScanSmallItems = 16;
constexpr int kScanSmallTile = 1024 * kScanSmallItems;
constexpr int64_t kScanSmall = kScanSmallTile;
__global__ __launch_bounds__(1024) void k_scan_small(const int32_t* in, int64_t n, int32_t* out) {
  __shared__ int32_t tile[kScanSmallTile];
  __shared__ int smem[1024 / kWave];
  const int tid = threadIdx.x;
  int carry = 0;
  for (int64_t base = 0; base < n; base += kScanSmallTile) {
#pragma unroll
    for (int k = 0; k < kScanSmallItems; ++k) {
      const int64_t idx = base + k * 1024 + tid;
      tile[k * 1024 + tid] = idx < n ? in[idx] : 0;
    }
    __syncthreads();
    int local = 0;
#pragma unroll
    for (int k = 0; k < kScanSmallItems; ++k) local += tile[tid * kScanSmallItems + k];
    int total;
    const int incl = block_incl_scan<1024>(local, smem, total);
    int prefix = carry + incl - local;
#pragma unroll
    for (int k = 0; k < kScanSmallItems; ++k) {
      const int v = tile[tid * kScanSmallItems + k];
      tile[tid * kScanSmallItems + k] = prefix;
      prefix += v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanSmallItems; ++k) {
      const int64_t idx = base + k * 1024 + tid;
      if (idx < n) out[idx] = tile[k * 1024 + tid];
    }
    carry += total;
    __syncthreads();
  }
  if (tid == 0) out[n] = carry;
}

__global__ __launch_bounds__(1024) void k_scan_partials(int32_t* part, int nb, int32_t* total_out) {
  __shared__ int smem[1024 / kWave];
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int idx = base + threadIdx.x;
    const int v = idx < nb ? part[idx] : 0;
    int total;
    const int incl = block_incl_scan<1024>(v, smem, total);
    if (idx < nb) part[idx] = carry + incl - v;
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// raw_partials: part holds the blocks' sums (not yet scanned); each block then
// adds up those before it itself, and the last block writes out[n] -- the
// middle phase's launch saved where the blocks are few.
__global__ __launch_bounds__(kThreads) void k_scan_down(const int32_t* in, int64_t n,
                                                         const int32_t* part, int32_t* out,
                                                         int raw_partials) {
  __shared__ int32_t tile[kScanTile];
  __shared__ int smem[kThreads / kWave];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t idx = base + k * kThreads + threadIdx.x;
    tile[k * kThreads + threadIdx.x] = idx < n ? in[idx] : 0;
  }
  int carry;
  if (raw_partials) {
    int v = 0;
    for (int b = threadIdx.x; b < static_cast<int>(blockIdx.x); b += kThreads) v += part[b];
    block_incl_scan<kThreads>(v, smem, carry);
  } else {
    carry = part[blockIdx.x];
  }
  __syncthreads();
  int local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) local += tile[threadIdx.x * kScanItems + k];
  int total;
  const int incl = block_incl_scan<kThreads>(local, smem, total);
  if (raw_partials && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = carry + total;
  int prefix = carry + incl - local;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int v = tile[threadIdx.x * kScanItems + k];
    tile[threadIdx.x * kScanItems + k] = prefix;
    prefix += v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t idx = base + k * kThreads + threadIdx.x;
    if (idx < n) out[idx] = tile[k * kThreads + threadIdx.x];
  }
}

// ===========================================================================
// Bounding box of every event (per-dimension min / max), then its tile sizes.
// ===========================================================================

__global__ __launch_bounds__(kThreads) void k_bbox_partial(Coords c, const Chunk* chunks, int D,
                                                            float* part) {
  __shared__ float slo[kThreads / kWave], shi[kThreads / kWave];
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  const Chunk ch = chunks[blockIdx.x];
  for (int d = 0; d < D; ++d) {
    float lo = __builtin_inff(), hi = -__builtin_inff();
    for (int i = ch.begin + threadIdx.x; i < ch.end; i += kThreads) {
      const float v = c.p[d][i];
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
      slo[w] = lo;
      shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < kThreads / kWave; ++k) {
        lo = slo[k] < lo ? slo[k] : lo;
        hi = hi < shi[k] ? shi[k] : hi;
      }
      part[(blockIdx.x * D + d) * 2 + 0] = lo;
      part[(blockIdx.x * D + d) * 2 + 1] = hi;
    }
    __syncthreads();
  }
}

// ComputeTiles.hpp:21-26: tile size = (max - min) / nPerDim, one float division.
// One wave per event and dimension.
__global__ void k_bbox_final(const float* part, int D, int n_events, EventDesc* ev) {
  const int e = blockIdx.x;
  const int d = threadIdx.x / kWave;
  const int lane = __lane_id();
  if (e >= n_events || d >= D) return;
  float lo = __builtin_inff(), hi = -__builtin_inff();
  for (int b = ev[e].chunk_begin + lane; b < ev[e].chunk_end; b += kWave) {
    const float plo = part[(b * D + d) * 2 + 0], phi = part[(b * D + d) * 2 + 1];
    lo = plo < lo ? plo : lo;
    hi = hi < phi ? phi : hi;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float olo = __shfl_xor(lo, off), ohi = __shfl_xor(hi, off);
    lo = olo < lo ? olo : lo;
    hi = hi < ohi ? ohi : hi;
  }
  if (lane == 0) {
    ev[e].mn[d] = lo;
    ev[e].mx[d] = hi;
    ev[e].ts[d] = (hi - lo) / static_cast<float>(ev[e].npd);
  }
}

// ===========================================================================
// Stable LSD counting sort of (tile id, point id), digits of `bits` <= 8 bits.
// Block = 4 waves x 1024 consecutive keys; inside a wave, keys with equal
// digits are ranked by lane order (ballot match), waves by index order, blocks
// by the scanned digit-major histogram: the order of equal tile ids is the
// input order, i.e. ascending point id after the first pass -- exactly the
// serial fill order of KernelFillAssociator.
// ===========================================================================
__global__ __launch_bounds__(kThreads) void k_radix_hist(const int32_t* keys, int32_t n, int shift,
                                                          int bits, int nblocks, int32_t* hist) {
  __shared__ int32_t h[256];
  const int radix = 1 << bits;
  for (int t = threadIdx.x; t < 256; t += kThreads) h[t] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
  const int lane = __lane_id();
  int32_t kk[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t idx = base + r * kThreads + threadIdx.x;
    kk[r] = idx < n ? keys[idx] : -1;
  }
  // lanes holding the same digit are matched by one ballot per digit bit and
  // counted by their lowest lane: one LDS atomic per distinct digit of a wave
  // (keys of later passes arrive partly sorted, so many lanes share a digit)
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = base + r * kThreads + threadIdx.x < n;
    const int digit = (kk[r] >> shift) & (radix - 1);
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (digit >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    if (valid && (peers & ((1ull << lane) - 1)) == 0) atomicAdd(&h[digit], __popcll(peers));
  }
  __syncthreads();
  for (int t = threadIdx.x; t < radix; t += kThreads) hist[t * nblocks + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(const int32_t* keys_in,
                                                             const int32_t* vals_in, int32_t n,
                                                             int shift, int bits, int nblocks,
                                                             const int32_t* hist_scan,
                                                             int32_t* keys_out, int32_t* vals_out) {
  __shared__ int32_t wcnt[kThreads / kWave][256];
  __shared__ int32_t gbase[256];
  __shared__ int32_t lbase[256];
  __shared__ int32_t scan_tmp[kThreads / kWave];
  // the block's keys and values in digit order (a local stable sort), written
  // out in runs of consecutive addresses per digit
  __shared__ int32_t skey[kSortTile];
  __shared__ int32_t sval[kSortTile];
  const int radix = 1 << bits;
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  for (int t = threadIdx.x; t < 256; t += kThreads) {
#pragma unroll
    for (int ww = 0; ww < kThreads / kWave; ++ww) wcnt[ww][t] = 0;
    gbase[t] = t < radix ? hist_scan[t * nblocks + blockIdx.x] : 0;
  }
  __syncthreads();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  const int64_t blk0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
  const int64_t base = blk0 + w * (kSortItems * kWave);
  int32_t kk[kSortItems], vv[kSortItems], loc[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t idx = base + r * kWave + lane;
    const bool valid = idx < n;
    const int32_t key = valid ? keys_in[idx] : 0;
    const int32_t val = valid ? (vals_in ? vals_in[idx] : static_cast<int32_t>(idx)) : 0;
    const int digit = (key >> shift) & (radix - 1);
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (digit >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    int before = 0;
    if (valid) before = wcnt[w][digit];
    const uint64_t lower = peers & lt_mask;
    if (valid && lower == 0) wcnt[w][digit] = before + __popcll(peers);
    kk[r] = key;
    vv[r] = val;
    loc[r] = valid ? before + __popcll(lower) : -1;
  }
  __syncthreads();
  // per digit: offsets of the waves inside the block, and the digit's start in
  // the block's digit order (an exclusive scan over the digits)
  {
    const int t = threadIdx.x;  // kThreads == 256 digits
    int sum = 0;
#pragma unroll
    for (int ww = 0; ww < kThreads / kWave; ++ww) {
      const int cnt = t < radix ? wcnt[ww][t] : 0;
      if (t < radix) wcnt[ww][t] = sum;
      sum += cnt;
    }
    int total;
    const int incl = block_incl_scan<kThreads>(sum, scan_tmp, total);
    lbase[t] = incl - sum;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (loc[r] < 0) continue;
    const int digit = (kk[r] >> shift) & (radix - 1);
    const int li = lbase[digit] + wcnt[w][digit] + loc[r];
    skey[li] = kk[r];
    sval[li] = vv[r];
  }
  __syncthreads();
  const int cnt = static_cast<int>(min(static_cast<int64_t>(kSortTile), static_cast<int64_t>(n) - blk0));
  for (int li = threadIdx.x; li < cnt; li += kThreads) {
    const int32_t key = skey[li];
    const int digit = (key >> shift) & (radix - 1);
    const int pos = gbase[digit] + (li - lbase[digit]);
    if (keys_out) keys_out[pos] = key;
    vals_out[pos] = sval[li];
  }
}

// Morton-ordered SoA copy (tile-major, Morton code inside the tile): the query
// order of both neighbour passes and the candidate layout of the nearest-higher
// search.  slot_of[p] is the tile-order slot of Morton position p (a permutation
// inside each tile, so every read stays within one tile's range); mid[p] its
// point id.
__global__ __launch_bounds__(kThreads) void k_gather_morton(const float* xs, const float* ws,
                                                             const int32_t* perm, const int32_t* slot_of,
                                                             int32_t n, int D, int64_t ld, float* xn,
                                                             float* wn, int32_t* mid) {
  const int p = logical_block(1) * kThreads + threadIdx.x;
  if (p >= n) return;
  const int32_t s = slot_of[p];
  for (int d = 0; d < D; ++d) xn[d * ld + p] = xs[d * ld + s];
  wn[p] = ws[s];
  mid[p] = perm[s];
}

// Morton order inside each tile: the tile-order slots [offsets[t], offsets[t+1])
// of tile t sorted by the Morton code in the low `bits` bits of their keys, one
// kernel in place of a radix sort of all slots by (tile id, code).  One wave per
// tile, tiles strided over the grid.  A tile of at most 64 points ranks each
// point against every other (lane j's code read as a wave-uniform value; equal
// codes in slot order).  A larger tile counts its codes in an LDS histogram of
// 2^KB entries (KB >= bits), scans it into bucket starts, and takes each slot's
// position from an LDS atomic add on its bucket's start, 64 slots at a time in
// slot order: equal codes follow the slot order between chunks, and inside a
// chunk the order in which the LDS serves the lanes.  Any order of the points of
// one Morton cell is a valid layout (no result depends on it).  slot_of[p] is
// the tile-order slot of Morton position p.  A workgroup is one wave, whose LDS
// operations execute in program order: a wavefront-scope fence keeps the
// compiler from moving them across each other, without a workgroup barrier's
// wait for the wave's outstanding global stores.
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

template <int KB>
__global__ __launch_bounds__(kWave) void k_tile_morton(const int32_t* keys, const int32_t* offsets, int32_t T,
                                                       int bits, int32_t* slot_of) {
  __shared__ int32_t hist[1 << KB];
  const int lane = __lane_id();
  const int32_t mask = (1 << bits) - 1;
  // the keys of kGroup chunks are loaded together, their loads in flight at once
  constexpr int kGroup = 8;
  for (int t = blockIdx.x; t < T; t += gridDim.x) {
    const int32_t o0 = uniform(offsets[t]), o1 = uniform(offsets[t + 1]);
    const int32_t c = o1 - o0;
    if (c <= 0) continue;
    if (c <= kWave) {
      const bool v = lane < c;
      const int32_t code = keys[v ? o0 + lane : o0] & mask;
      int rank = 0;
      for (int j = 0; j < c; ++j) {
        const int32_t cj = __builtin_amdgcn_readlane(code, j);
        rank += (cj < code || (cj == code && j < lane)) ? 1 : 0;
      }
      if (v && rank < c) slot_of[o0 + rank] = o0 + lane;
      continue;
    }
    auto load_group = [&](int32_t base, int32_t* kk) {
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int32_t s = base + u * kWave + lane;
        kk[u] = keys[s < o1 ? s : o1 - 1];  // clamped inside the tile
      }
    };
    for (int i = lane; i < (1 << KB); i += kWave) hist[i] = 0;
    wave_fence();
    for (int32_t base = o0; base < o1; base += kGroup * kWave) {
      int32_t kk[kGroup];
      load_group(base, kk);
#pragma unroll
      for (int u = 0; u < kGroup; ++u)
        if (base + u * kWave + lane < o1) atomicAdd(&hist[kk[u] & mask], 1);
    }
    wave_fence();
    // exclusive scan, one row of 64 entries at a time
    int32_t carry = 0;
    for (int i = 0; i < (1 << KB); i += kWave) {
      const int32_t h = hist[i + lane];
      const int32_t incl = wave_incl_scan(h);
      hist[i + lane] = carry + incl - h;
      carry += __builtin_amdgcn_readlane(incl, kWave - 1);
    }
    wave_fence();
    for (int32_t base = o0; base < o1; base += kGroup * kWave) {
      int32_t kk[kGroup];
      load_group(base, kk);
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const int32_t s = base + u * kWave + lane;
        if (s < o1) {
          const int32_t pos = atomicAdd(&hist[kk[u] & mask], 1);
          if (pos >= 0 && pos < c) slot_of[o0 + pos] = s;
        }
      }
    }
    wave_fence();
  }
}

// offsets[t] = number of points whose tile id is below t (lower bound of t in
// the sorted tile ids), t in [0, T]: the CSR offsets of the tiles.  Also clears
// the tiles' non-finite flags (set by k_gather) and the per-event counts of
// points without a nearest-higher.
__global__ void k_tile_offsets(const int32_t* skey, int32_t n, int32_t T, int32_t* offsets, int32_t* tnf,
                               int32_t* sc, int32_t n_events) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t < n_events) sc[t] = 0;
  if (t > T) return;
  if (t < T) tnf[t] = 0;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (skey[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  offsets[t] = lo;
}


__global__ void k_block_count(const int32_t* offsets, int32_t T, int32_t* cnt) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t < T) cnt[t] = (offsets[t + 1] - offsets[t] + kBlock - 1) / kBlock;
}

// Tile of every block of the Morton layout (the blocks are numbered tile by tile).
__global__ void k_block_tiles(const int32_t* blk_off, int32_t T, int32_t* btile) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= T) return;
  const int b1 = blk_off[t + 1];
  for (int b = blk_off[t]; b < b1; ++b) btile[b] = t;
}

// Bounding box of each block of 32 consecutive Morton-ordered points of a tile:
// kBoxLanes lanes per block, each reading every kBoxLanes-th point, then a
// reduction across the lane group.
__global__ __launch_bounds__(kThreads) void k_block_boxes(const float* xn, int64_t ld, int D,
                                                           const int32_t* offsets,
                                                           const int32_t* blk_off, const int32_t* btile,
                                                           int32_t T, float* bmin, float* bmax,
                                                           int64_t ldb) {
  const int64_t gid = static_cast<int64_t>(logical_block(1)) * kThreads + threadIdx.x;
  const int b = static_cast<int>(gid / kBoxLanes);
  const int sub = threadIdx.x % kBoxLanes;
  const int nblk = blk_off[T];
  if (b >= nblk) return;  // uniform over the lane group
  const int t = btile[b];
  const int start = offsets[t] + (b - blk_off[t]) * kBlock;
  const int end = min(start + kBlock, offsets[t + 1]);
  for (int d = 0; d < D; ++d) {
    float mn = __builtin_inff(), mx = -__builtin_inff();
#pragma unroll
    for (int k = 0; k < kBlock / kBoxLanes; ++k) {
      const int j = start + sub + k * kBoxLanes;
      if (j < end) {
        const float v = xn[d * ld + j];
        mn = v < mn ? v : mn;
        mx = mx < v ? v : mx;
      }
    }
#pragma unroll
    for (int off = kBoxLanes / 2; off > 0; off >>= 1) {
      const float omn = __shfl_xor(mn, off), omx = __shfl_xor(mx, off);
      mn = omn < mn ? omn : mn;
      mx = mx < omx ? omx : mx;
    }
    if (sub == 0) {
      bmin[d * ldb + b] = mn;
      bmax[d * ldb + b] = mx;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_tile_boxes(const float* bmin, const float* bmax,
                                                          int64_t ldb, int D,
                                                          const int32_t* blk_off, int32_t T,
                                                          float* tmin, float* tmax) {
  // kTileLanes lanes per tile split its blocks; a reduction across the group combines
  const int t = static_cast<int>((static_cast<int64_t>(logical_block(1)) * kThreads + threadIdx.x) / kTileLanes);
  if (t >= T) return;  // uniform over the lane group
  const int sub = threadIdx.x % kTileLanes;
  const int b0 = blk_off[t], b1 = blk_off[t + 1];
  for (int d = 0; d < D; ++d) {
    float mn = __builtin_inff(), mx = -__builtin_inff();
#pragma unroll 4
    for (int b = b0 + sub; b < b1; b += kTileLanes) {
      mn = bmin[d * ldb + b] < mn ? bmin[d * ldb + b] : mn;
      mx = mx < bmax[d * ldb + b] ? bmax[d * ldb + b] : mx;
    }
#pragma unroll
    for (int off = kTileLanes / 2; off > 0; off >>= 1) {
      const float omn = __shfl_xor(mn, off), omx = __shfl_xor(mx, off);
      mn = omn < mn ? omn : mn;
      mx = mx < omx ? omx : mx;
    }
    if (sub == 0) {
      tmin[d * T + t] = mn;
      tmax[d * T + t] = mx;
    }
  }
}

// Largest rank key of each block and of each tile (the density filter of the
// nearest-higher searches: a block whose largest key does not exceed a query's
// threshold holds no point higher than it).  kBoxLanes lanes per block read
// every kBoxLanes-th key of its Morton positions; then kTileLanes lanes per tile
// combine its blocks.  An empty tile gets 0, below every key.  Also clears the
// count of the first nearest-higher phase's unresolved points (*pending_count).
__global__ __launch_bounds__(kThreads) void k_block_maxkey(const uint64_t* hk, const int32_t* offsets,
                                                            const int32_t* blk_off, const int32_t* btile,
                                                            int32_t T, uint64_t* bmaxk, int32_t* pending_count) {
  const int64_t gid = static_cast<int64_t>(logical_block(1)) * kThreads + threadIdx.x;
  if (gid == 0) *pending_count = 0;
  const int b = static_cast<int>(gid / kBoxLanes);
  const int sub = threadIdx.x % kBoxLanes;
  if (b >= blk_off[T]) return;  // uniform over the lane group
  const int t = btile[b];
  const int start = offsets[t] + (b - blk_off[t]) * kBlock;
  const int end = min(start + kBlock, offsets[t + 1]);
  uint64_t m = 0;
#pragma unroll
  for (int k = 0; k < kBlock / kBoxLanes; ++k) {
    const int j = start + sub + k * kBoxLanes;
    if (j < end) {
      const uint64_t v = hk[j];
      m = v > m ? v : m;
    }
  }
#pragma unroll
  for (int off = kBoxLanes / 2; off > 0; off >>= 1) {
    const uint32_t lo = __shfl_xor(static_cast<uint32_t>(m), off);
    const uint32_t hi = __shfl_xor(static_cast<uint32_t>(m >> 32), off);
    const uint64_t o = (static_cast<uint64_t>(hi) << 32) | lo;
    m = o > m ? o : m;
  }
  if (sub == 0) bmaxk[b] = m;
}

__global__ __launch_bounds__(kThreads) void k_tile_maxkey(const uint64_t* bmaxk, const int32_t* blk_off,
                                                           int32_t T, uint64_t* tmaxk) {
  const int t = static_cast<int>((static_cast<int64_t>(logical_block(1)) * kThreads + threadIdx.x) / kTileLanes);
  if (t >= T) return;  // uniform over the lane group
  const int sub = threadIdx.x % kTileLanes;
  const int b1 = blk_off[t + 1];
  uint64_t m = 0;
#pragma unroll 4
  for (int b = blk_off[t] + sub; b < b1; b += kTileLanes) m = bmaxk[b] > m ? bmaxk[b] : m;
#pragma unroll
  for (int off = kTileLanes / 2; off > 0; off >>= 1) {
    const uint32_t lo = __shfl_xor(static_cast<uint32_t>(m), off);
    const uint32_t hi = __shfl_xor(static_cast<uint32_t>(m >> 32), off);
    const uint64_t o = (static_cast<uint64_t>(hi) << 32) | lo;
    m = o > m ? o : m;
  }
  if (sub == 0) tmaxk[t] = m;
}

__global__ __launch_bounds__(kThreads) void k_query_block_cost(const int32_t* kn, int mshift,
                                                                const int32_t* offsets, int32_t n,
                                                                int32_t* cost) {
  __shared__ int32_t part[kThreads / kWave];
  const int q = blockIdx.x * kThreads + threadIdx.x;
  int32_t c = 0;
  if (q < n) {
    const int t = kn[q] >> mshift;
    c = offsets[t + 1] - offsets[t];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (__lane_id() == 0) part[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int w = 0; w < kThreads / kWave; ++w) s += part[w];
    cost[blockIdx.x] = s;
  }
}


// Query blocks in decreasing cost, one workgroup: a counting sort on the cost
// quantised to 1/8 of an octave (heavy blocks start first, so the light ones
// fill in behind them and the pass does not end on a few late heavy blocks; the
// order inside a bucket is free -- the results do not depend on the launch
// order).  With xcd set, the blocks are first split into 8 contiguous ranges of
// the Morton order, one per XCD (the dispatch gives workgroup b to XCD b % 8),
// and ordered by cost inside each range: neighbouring queries -- which read the
// same tiles -- share one L2.
__device__ __forceinline__ int cost_bucket(int32_t cost) {
  const uint32_t c = min(static_cast<uint32_t>(cost > 0 ? cost : 0), 0xffffffu);
  if (c == 0) return 255;
  const int l = 31 - __builtin_clz(c);                              // floor(log2 c) <= 23
  const int frac = static_cast<int>((c << (31 - l)) >> 28) & 7;   // next three bits
  return 255 - (l * 8 + frac);                                      // 64..255, heaviest first
}

__global__ __launch_bounds__(1024) void k_query_block_order(const int32_t* cost, int32_t nb, int xcd,
                                                            int32_t* order) {
  __shared__ int32_t cnt[8 * 256];
  __shared__ int32_t sorted[kOrderMax];
  __shared__ int scan_part[1024 / kWave];
  const int tid = threadIdx.x;
  for (int i = tid; i < 8 * 256; i += 1024) cnt[i] = 0;
  __syncthreads();
  // XCD of a block: its chunk of the Morton order (32 chunks) modulo 8, so each
  // XCD holds four separate eighths-of-a-quarter of the domain (their costs even
  // out); a group's size may miss its XCD's share by a chunk, which only moves
  // those blocks to a neighbouring XCD
  const int chunk = (nb + 31) / 32;
  auto bucket_of = [&](int i) { return (xcd ? ((i / chunk) & 7) * 256 : 0) + cost_bucket(cost[i]); };
  for (int i = tid; i < nb; i += 1024) atomicAdd(&cnt[bucket_of(i)], 1);
  __syncthreads();
  // exclusive scan of the 2048 bucket counts, two per thread
  const int c0 = cnt[2 * tid], c1 = cnt[2 * tid + 1];
  int total;
  const int incl = block_incl_scan<1024>(c0 + c1, scan_part, total);
  cnt[2 * tid] = incl - c0 - c1;
  cnt[2 * tid + 1] = incl - c1;
  __syncthreads();
  for (int i = tid; i < nb; i += 1024) sorted[atomicAdd(&cnt[bucket_of(i)], 1)] = i;
  __syncthreads();
  const int q = nb >> 3, r = nb & 7;
  for (int b = tid; b < nb; b += 1024) {
    int src = b;
    if (xcd) {  // workgroup b runs on XCD b % 8: the (b / 8)-th block of that XCD's range
      const int x = b & 7, k = b >> 3;
      src = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
    }
    order[b] = sorted[src];
  }
}

// The event and chunk tables of a call over one event, written on the device
// from the descriptor passed by value (no host staging, no copy).
__global__ void k_single_event_tables(EventDesc ed, EventDesc* ev, Chunk* chunks) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) *ev = ed;
  if (c < ed.chunk_end) {
    const int32_t b = ed.begin + c * kChunk;
    chunks[c] = Chunk{0, b, min(b + kChunk, ed.begin + ed.n), 0};
  }
}

// Per event: seeds, capacity and the base of its seed ranks.  counters (totals
// over the events): [0] seed candidates, [1] grow-only capacity carried across
// single-event calls, [2] capacity used, [3] seeds, [4] stored seeds
// (clusters), [5] dropped seeds.
__global__ void k_capacity(const EventDesc* ev, int n_events, const int32_t* sc,
                           const int32_t* rank, int32_t* counters, int32_t* cap_used,
                           int32_t* rank_base, int unbounded, int persistent) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_events) return;
  const int b = ev[e].begin, n = ev[e].n;
  const int32_t cand = sc[e];
  int32_t cap = cand;  // SetupSeeds.hpp:14-18: a fresh seed array holds #(nh == -1)
  if (persistent) {  // a Clusterer keeps the largest capacity it has seen (one event)
    cap = counters[1] > cand ? counters[1] : cand;
    counters[1] = cap;
  }
  const int32_t ns = rank[b + n] - rank[b];
  const int32_t used = unbounded ? ns : cap;
  const int32_t stored = ns < used ? ns : used;
  cap_used[e] = used;
  rank_base[e] = rank[b];
  atomicAdd(&counters[0], cand);
  atomicAdd(&counters[2], used);
  atomicAdd(&counters[3], ns);
  atomicAdd(&counters[4], stored);
  atomicAdd(&counters[5], ns - stored);
}

// One synchronous pointer-jumping step, parent <- parent of parent: halves the
// distance of every node to its root (other threads' concurrent updates only
// replace a parent by one of its ancestors).  Each XCD takes a contiguous eighth
// of the points: parents lie in the point's event, so the random reads of one
// XCD stay within the events it walks, in its own L2.
__global__ __launch_bounds__(kThreads) void k_jump(int32_t* root, int32_t n) {
  const int i = logical_block(1) * kThreads + threadIdx.x;
  if (i < n) root[i] = root[root[i]];
}

// Pointer jumping: every hop follows the current parent array, which other
// threads only ever shorten; `hops` per round bounds the work of a round.  A
// thread that reaches its root writes the label -- the position of the root in
// its event's serial seed array, -1 if the root is not a stored seed
// (KernelAssignClusters, ClusteringKernels.hpp:252-279); one that runs out of
// hops raises the round's flag, and the next round (which returns at once when
// the previous flag is clear) goes on from where it stopped.  The first round
// also turns the nearest-higher into an index local to the event.
__global__ __launch_bounds__(kThreads) void k_labels(int32_t* root, int32_t n, int hops,
                                                      const int32_t* is_seed, const int32_t* rank,
                                                      const int32_t* pev, const EventDesc* ev,
                                                      const int32_t* cap_used,
                                                      const int32_t* rank_base, int32_t* cluster,
                                                      int32_t* nh_out, int32_t* flags, int round) {
  if (round > 0 && flags[round - 1] == 0) return;  // every thread finished before
  // later rounds run on a smaller grid, each thread striding over the points;
  // blocks mapped to XCDs as in k_jump
  for (int i = logical_block(1) * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const int e = pev[i];
    if (round == 0) {
      const int32_t nhv = nh_out[i];
      if (nhv >= 0) nh_out[i] = nhv - ev[e].begin;
    }
    int32_t r = root[i];
    bool at_root = false;
    for (int h = 0; h < hops; ++h) {
      const int32_t rr = root[r];
      if (rr == r) {
        at_root = true;
        break;
      }
      r = rr;
    }
    root[i] = r;
    if (at_root) {
      const int32_t local = rank[r] - rank_base[e];
      cluster[i] = (is_seed[r] && local < cap_used[e]) ? local : -1;
    } else {
      flags[round] = 1;
    }
  }
}

// ===========================================================================
// Post-processing: cluster -> points map and centroids.
// ===========================================================================
// Sort keys of the map: the label, outliers (negative labels) last.  Labels
// must lie below n; larger ones are treated as outliers.
__global__ void k_map_keys(const int32_t* cluster, int32_t n, int32_t* key) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    const int32_t c = cluster[i];
    key[i] = (c >= 0 && c < n) ? c : n;
  }
}

// offsets[c] = first sorted position whose key is >= c, for c in [0, n_clusters]
// (the sorted keys ascend, so every c is written by exactly one position).
// counts[0] = n_clusters = largest label + 1, counts[1] = clustered points.
__global__ void k_map_offsets(const int32_t* skey, int32_t n, int32_t* offsets, int32_t* counts) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  const int32_t cur = skey[k];
  const int32_t prev = k > 0 ? skey[k - 1] : -1;
  // labels in (prev, cur] start here; at the first outlier only n_clusters itself
  const int32_t hi = cur < n ? cur : (prev < n ? prev + 1 : prev);
  for (int32_t c = prev + 1; c <= hi; ++c) offsets[c] = k;
  const bool last = k == n - 1;
  if (cur < n && (last || skey[k + 1] >= n)) {
    counts[0] = cur + 1;
    counts[1] = k + 1;
    if (last) offsets[cur + 1] = n;
  }
}

// Centroid row c, coordinate d: the coordinates of the cluster's points summed in
// ascending point order from 0.f, divided by (float)size
// (utils/detail/cluster_centroid.hpp:38-65).
__global__ void k_centroids(Coords c, int ndim, const int32_t* offsets, const int32_t* indexes,
                            const int32_t* counts, int32_t max_clusters, float* out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const int32_t cl = static_cast<int32_t>(t / ndim);
  const int d = static_cast<int>(t % ndim);
  if (cl >= max_clusters) return;
  if (cl >= counts[0]) return;
  const int32_t b = offsets[cl], e = offsets[cl + 1];
  const float* x = c.p[d];
  float s = 0.f;
  for (int32_t k = b; k < e; ++k) s += x[indexes[k]];
  out[static_cast<int64_t>(cl) * ndim + d] = s / static_cast<float>(e - b);
}

// ===========================================================================
// Self-test arithmetic (parity of the device math with the host's).
// ===========================================================================
__global__ void k_selftest(int op, const float* a, const float* b, float* out, int64_t n) {
  __shared__ double s_wide[kExpWide];  // op 9: the density pass's wide table
  if (op == 9) {
    for (int k = threadIdx.x; k < kExpWide; k += blockDim.x) s_wide[k] = exp2_wide_entry(k - (kExpWide - 1));
    __syncthreads();
  }
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a[i];
  switch (op) {
    case 0: out[i] = engine_expf(x); break;
    case 1: out[i] = __builtin_sqrtf(x); break;
    case 2: out[i] = x / b[i]; break;
    case 3: out[i] = sqrt_lt_bound(x); break;
    case 4: out[i] = quotient_via_rcp(x, 1.0 / static_cast<double>(b[i])); break;
    case 5: out[i] = engine_expf_nonpos(x); break;
    case 8: out[i] = engine_expf_unit(x); break;
    // arguments clamped to [-14, 0], the range the density pass uses it on
    case 9: out[i] = engine_expf_unit_wide(__builtin_fminf(__builtin_fmaxf(x, -14.f), 0.f), s_wide + (kExpWide - 1)); break;
    case 10: out[i] = sqrt_rn_finite(x); break;  // finite x >= 0
    case 11: out[i] = strictly_nearer(x); break;  // finite x >= 0
    case 6:
    case 7: {
      float lo, hi;
      sqrt_class(__builtin_sqrtf(x), lo, hi);
      out[i] = op == 6 ? hi : lo;
      break;
    }
    default: out[i] = 0.f;
  }
}

// Counts the floats with bit patterns in [first, first + count) on which
// sqrt_rn_normal differs from the IEEE sqrtf.
__global__ void k_selftest_sqrt(uint32_t first, uint64_t count, unsigned long long* mismatches) {
  unsigned long long bad = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const float x = __uint_as_float(first + static_cast<uint32_t>(i));
    bad += __float_as_uint(sqrt_rn_normal(x)) != __float_as_uint(__builtin_sqrtf(x));
  }
  if (bad) atomicAdd(mismatches, bad);
}

// ===========================================================================
// Host side
// ===========================================================================
thread_local std::string g_error;

clue_status fail(clue_status st, const std::string& msg) {
  g_error = msg;
  return st;
}

void set_error(const std::string& msg) { g_error = msg; }

}  // namespace clue_hip

using namespace clue_hip;

namespace clue_hip {

// SetupTiles.hpp:23-26: ntiles = ceil(N / (float)ppb); npd = ceil(pow(ntiles, 1./D));
// ntiles = pow(npd, D) (double pow, as std::pow(int, double) resolves).
void grid_shape(int32_t D, int32_t n, int32_t ppb, int32_t& npd, int32_t& T) {
  int32_t nt = static_cast<int32_t>(std::ceil(static_cast<float>(n) / static_cast<float>(ppb)));
  npd = static_cast<int32_t>(std::ceil(std::pow(static_cast<double>(nt), 1. / D)));
  T = static_cast<int32_t>(std::pow(static_cast<double>(npd), static_cast<double>(D)));
}

// Largest float s with sqrtf(s) <= r (host sqrtf is correctly rounded, as the device's).
float sqrt_le_bound(float r) {
  if (!(r >= 0.f)) return -1.f;
  float t = r * r;
  if (!(t <= FLT_MAX)) t = FLT_MAX;
  while (t > 0.f && !(std::sqrt(t) <= r)) t = std::nextafter(t, -1.f);
  if (!(std::sqrt(t) <= r)) return -1.f;
  while (t < FLT_MAX && std::sqrt(std::nextafter(t, FLT_MAX)) <= r) t = std::nextafter(t, FLT_MAX);
  return t;
}

clue_status reserve(clue_plan_s* p, int32_t D, int32_t n, int32_t T, int32_t n_events,
                    int32_t n_chunks) {
  CLUE_HIP(hipSetDevice(p->device));
  // a plan over an arena lays every buffer out afresh, 256-byte aligned, and
  // allocates nothing; arena_used then holds the bytes this call needs (it may
  // exceed arena_bytes: the caller checks and refuses the call)
  if (p->arena_bytes) p->arena_used = 0;
  auto need = [p](DevBuf& b, size_t bytes) -> hipError_t {
    if (!p->arena_bytes) return b.ensure(bytes);
    const size_t a = std::max<size_t>(256, (bytes + 255) / 256 * 256);
    const bool fits = p->arena_used + a <= p->arena_bytes;
    b.p = fits ? p->arena + p->arena_used : nullptr;
    b.bytes = fits ? a : 0;
    p->arena_used += a;
    return hipSuccess;
  };
  const size_t N = static_cast<size_t>(n);
  const size_t nb = static_cast<size_t>(div_up(n, kSortTile));
  const size_t hist_entries = 256 * nb + 1;
  const size_t scan_max = std::max<size_t>({N + 1, hist_entries, static_cast<size_t>(T) + 1});
  const int64_t ld = div_up(n, 64) * 64;
  const size_t TT = static_cast<size_t>(T) + 1;
  const size_t E = static_cast<size_t>(n_events);
  CLUE_HIP(need(p->key, N * 4));
  CLUE_HIP(need(p->packed, N * static_cast<size_t>((D + 2 + 3) / 4 * 4) * 4));
  CLUE_HIP(need(p->key2, N * 4));
  CLUE_HIP(need(p->perm, N * 4));
  CLUE_HIP(need(p->perm2, N * 4));
  CLUE_HIP(need(p->offsets, TT * 4));
  CLUE_HIP(need(p->tnf, TT * 4));
  if (D > 3) {
    CLUE_HIP(need(p->qk, N * 4));
    CLUE_HIP(need(p->qk2, N * 4));
    CLUE_HIP(need(p->qv, N * 4));
    CLUE_HIP(need(p->qv2, N * 4));
  }
  CLUE_HIP(need(p->hist, hist_entries * 4));
  CLUE_HIP(need(p->part, (div_up(scan_max, kScanTile) + 1) * 4));
  CLUE_HIP(need(p->xs, static_cast<size_t>(ld) * D * 4));
  p->ld = ld;
  CLUE_HIP(need(p->ws, N * 4));
  CLUE_HIP(need(p->hw, N * 4));
  CLUE_HIP(need(p->rho_s, N * 4));
  CLUE_HIP(need(p->delta, N * 4));
  CLUE_HIP(need(p->flag, (N + 1) * 4));
  CLUE_HIP(need(p->root, N * 4));
  CLUE_HIP(need(p->nh_tmp, N * 4));
  CLUE_HIP(need(p->rho_tmp, N * 4));
  CLUE_HIP(need(p->seed_tmp, N * 4));
  CLUE_HIP(need(p->bbox_part, static_cast<size_t>(n_chunks) * D * 2 * 4));
  CLUE_HIP(need(p->events, E * sizeof(EventDesc)));
  CLUE_HIP(need(p->chunks, static_cast<size_t>(n_chunks) * sizeof(Chunk)));
  CLUE_HIP(need(p->pev, N * 4));
  CLUE_HIP(need(p->sc, E * 4));
  CLUE_HIP(need(p->cap_used, E * 4));
  CLUE_HIP(need(p->rank_base, E * 4));
  if (p->arena_bytes) {
    CLUE_HIP(need(p->counters, 16 * 4));  // cleared by a fill kernel on every arena call
  } else if (!p->counters.p) {
    CLUE_HIP(need(p->counters, 16 * 4));
    CLUE_HIP(hipMemset(p->counters.p, 0, 16 * 4));
  }
  CLUE_HIP(need(p->kn, N * 4));
  CLUE_HIP(need(p->kn2, N * 4));
  CLUE_HIP(need(p->pn, N * 4));
  CLUE_HIP(need(p->pn2, N * 4));
  CLUE_HIP(need(p->mid, N * 4));
  CLUE_HIP(need(p->xn, static_cast<size_t>(ld) * D * 4));
  CLUE_HIP(need(p->wn, N * 4));
  CLUE_HIP(need(p->hk, N * 8));
  CLUE_HIP(need(p->bcnt, TT * 4));
  CLUE_HIP(need(p->blk_off, TT * 4));
  const int64_t ldb = div_up(n, kBlock) + T + 1;
  CLUE_HIP(need(p->bbox_lo, static_cast<size_t>(ldb) * D * 4));
  CLUE_HIP(need(p->bbox_hi, static_cast<size_t>(ldb) * D * 4));
  CLUE_HIP(need(p->btile, static_cast<size_t>(ldb) * 4));
  p->ldb = ldb;
  CLUE_HIP(need(p->tbox_lo, static_cast<size_t>(T) * D * 4));
  CLUE_HIP(need(p->tbox_hi, static_cast<size_t>(T) * D * 4));
  CLUE_HIP(need(p->bmaxk, static_cast<size_t>(ldb) * 8));
  CLUE_HIP(need(p->tmaxk, static_cast<size_t>(T) * 8));
  CLUE_HIP(need(p->pending, (N + 64) * 4));
  CLUE_HIP(need(p->pending_count, 256));
  CLUE_HIP(need(p->qcost, static_cast<size_t>(div_up(n, kThreads)) * 4));
  CLUE_HIP(need(p->qorder, static_cast<size_t>(div_up(n, kThreads)) * 4));
  return CLUE_OK;
}

// Density launch order: estimated cost per query block, then the blocks in
// decreasing cost (xcd_remap bit 2: XCD groups of Morton chunks inside the
// heaviest-first order -- less than half the HBM traffic at the same speed).
clue_status launch_density_order(clue_plan_s* p, const PassArgs& a, int grid, hipStream_t st) {
  k_query_block_cost<<<grid, kThreads, 0, st>>>(a.kn, a.mshift, a.offsets, a.n, p->qcost.as<int32_t>());
  CLUE_LAUNCH_CHECK("query_block_cost");
  k_query_block_order<<<1, 1024, 0, st>>>(p->qcost.as<int32_t>(), grid, (p->tuning.xcd_remap >> 2) & 1,
                                          p->qorder.as<int32_t>());
  CLUE_LAUNCH_CHECK("query_block_order");
  return CLUE_OK;
}

// Largest rank key per Morton block and per tile (the nearest-higher filter).
clue_status launch_nh_maxkeys(clue_plan_s* p, const PassArgs& a, hipStream_t st) {
  k_block_maxkey<<<div_up((div_up(a.n, kBlock) + a.ntiles) * kBoxLanes, kThreads), kThreads, 0, st>>>(
      a.hk, a.offsets, a.blk_off, a.btile, a.ntiles, p->bmaxk.as<uint64_t>(), a.pending_count);
  CLUE_LAUNCH_CHECK("block_maxkey");
  k_tile_maxkey<<<div_up(static_cast<int64_t>(a.ntiles) * kTileLanes, kThreads), kThreads, 0, st>>>(
      p->bmaxk.as<uint64_t>(), a.blk_off, a.ntiles, p->tmaxk.as<uint64_t>());
  CLUE_LAUNCH_CHECK("tile_maxkey");
  return CLUE_OK;
}

clue_status scan_exclusive(const int32_t* in, int32_t* out, int64_t n, int32_t* part,
                           hipStream_t st) {
  const int nb = static_cast<int>(div_up(n, kScanTile));
  if (nb == 0) {
    k_fill_i32<<<1, 64, 0, st>>>(out, 0, 1);
    CLUE_LAUNCH_CHECK("fill");
    return CLUE_OK;
  }
  if (n <= kScanSmall) {
    k_scan_small<<<1, 1024, 0, st>>>(in, n, out);
    CLUE_LAUNCH_CHECK("scan_small");
    return CLUE_OK;
  }
  k_scan_reduce<<<nb, kThreads, 0, st>>>(in, n, part);
  CLUE_LAUNCH_CHECK("scan_reduce");
  // up to kScanRawBlocks blocks, each sums the partials before it (at most 4 per thread)
  const int raw = nb <= kScanRawBlocks;
  if (!raw) {
    k_scan_partials<<<1, 1024, 0, st>>>(part, nb, out + n);
    CLUE_LAUNCH_CHECK("scan_partials");
  }
  k_scan_down<<<nb, kThreads, 0, st>>>(in, n, part, out, raw);
  CLUE_LAUNCH_CHECK("scan_down");
  return CLUE_OK;
}


clue_status validate(const clue_params* prm, int32_t ndim, int32_t n, const float* const* coords,
                     const float* weights, const clue_kernel_desc* kernel,
                     const clue_metric_desc* metric, const int32_t* cluster_out) {
  if (!prm) return fail(CLUE_ERR_INVALID_ARGUMENT, "params is NULL");
  // detail/Clusterer.hpp:32-35
  if (prm->dc <= 0.f || prm->rhoc < 0.f || prm->dm <= 0.f || prm->seed_dc <= 0.f ||
      prm->points_per_tile <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Invalid clustering parameters. The parameters must be positive.");
  if (ndim < 1 || ndim > CLUE_MAX_DIM)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "number of dimensions must be in [1, 10]");
  // detail/PointsDevice.hpp:24-31
  if (n <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Number of points passed to PointsDevice constructor must be positive.");
  if (!coords || !weights || !cluster_out)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "coords, weights and cluster output are required");
  for (int d = 0; d < ndim; ++d)
    if (!coords[d]) return fail(CLUE_ERR_INVALID_ARGUMENT, "coordinate pointer is NULL");
  if (kernel && (kernel->type < 0 || kernel->type > 2))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown convolutional kernel type");
  if (kernel) {
    // The reference adds ((float)(d <= dc) * k) * w for every candidate of the
    // search box (ClusteringKernels.hpp:42); the engine skips the out-of-radius
    // terms, which are +-0 while k is finite for every distance.  Parameters
    // under which k can be infinite or NaN are refused.
    const float* kp = kernel->params;
    bool ok = true;
    if (kernel->type == CLUE_KERNEL_FLAT) {
      ok = std::isfinite(kp[0]);
    } else if (kernel->type == CLUE_KERNEL_GAUSSIAN) {
      const float den = (2.f * kp[1]) * kp[1];
      ok = std::isfinite(kp[0]) && std::isfinite(kp[2]) && std::isfinite(den) && den > 0.f;
    } else {
      ok = std::isfinite(kp[0]) && kp[0] > 0.f && std::isfinite(kp[1]);
    }
    if (!ok)
      return fail(CLUE_ERR_INVALID_ARGUMENT,
                  "kernel parameters must be finite, with (2*std)*std finite and positive (gaussian) "
                  "and a positive rate (exponential)");
  }
  if (metric && (metric->type < 0 || metric->type > 5))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown distance metric type");
  if (prm->seed_mode != CLUE_SEEDS_REFERENCE && prm->seed_mode != CLUE_SEEDS_UNBOUNDED)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "unknown seed mode");
  return CLUE_OK;
}

// Stable LSD counting sort of n int32 keys (< 2^total_bits) with their values
// (identity when vals is NULL), ping-ponging through (k0,k1) and (v0,v1).
clue_status radix_sort(clue_plan_s* p, hipStream_t st, int32_t n, int total_bits, int32_t* k0,
                       int32_t* k1, int32_t* v0, int32_t* v1, bool keep_keys,
                       int32_t** keys_sorted, int32_t** vals_sorted) {
  const int passes = std::max(1, (total_bits + 7) / 8);
  const int bpp = std::max(1, (total_bits + passes - 1) / passes);
  const int nblocks = static_cast<int>(div_up(n, kSortTile));
  int32_t* kin = k0;
  int32_t* vin = nullptr;
  int32_t* kout = k1;
  int32_t* vout = v0;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * bpp;
    const bool last = pass == passes - 1;
    k_radix_hist<<<nblocks, kThreads, 0, st>>>(kin, n, shift, bpp, nblocks, p->hist.as<int32_t>());
    CLUE_LAUNCH_CHECK("radix_hist");
    clue_status rc = scan_exclusive(p->hist.as<int32_t>(), p->hist.as<int32_t>(),
                                    static_cast<int64_t>(1 << bpp) * nblocks, p->part.as<int32_t>(), st);
    if (rc != CLUE_OK) return rc;
    k_radix_scatter<<<nblocks, kThreads, 0, st>>>(kin, vin, n, shift, bpp, nblocks,
                                                  p->hist.as<int32_t>(),
                                                  (last && !keep_keys) ? nullptr : kout, vout);
    CLUE_LAUNCH_CHECK("radix_scatter");
    int32_t* next_kout = (kout == k1) ? k0 : k1;
    int32_t* next_vout = (vout == v0) ? v1 : v0;
    kin = kout;
    vin = vout;
    kout = next_kout;
    vout = next_vout;
  }
  *keys_sorted = kin;
  *vals_sorted = vin;
  return CLUE_OK;
}

int bits_for(int64_t count) {
  int b = 0;
  while ((int64_t{1} << b) < count) ++b;
  return b;
}

clue_status mark(clue_plan_s* p, int k, hipStream_t st) {
  if (!p->timing) return CLUE_OK;
  if (!p->ev[k]) CLUE_HIP(hipEventCreate(&p->ev[k]));
  CLUE_HIP(hipEventRecord(p->ev[k], st));
  return CLUE_OK;
}

#define CLUE_MARK(k)                       \
  do {                                     \
    clue_status m_ = mark(p, (k), st);     \
    if (m_ != CLUE_OK) return m_;          \
  } while (0)

// A validated call over a batch of independent events (a plain call is a batch
// of one): event e is points [event_offsets[e], event_offsets[e+1]).
struct BatchCall {
  int32_t D = 0, n = 0, n_events = 0, n_chunks = 0, Ttot = 0;
  clue_kernel_desc kd{};
  clue_metric_desc md{};
  Wrap wrap{};
  bool wrapped_any = false;
  Coords c{};
  const float* weights = nullptr;
  int32_t* cluster_out = nullptr;
  int32_t* is_seed_out = nullptr;
  float* rho_out = nullptr;
  int32_t* nh_out = nullptr;
  bool persistent = false;  // the plan's grow-only seed capacity applies (plain calls)
  int32_t max_event_n = 0, max_event_tiles = 0;
};

clue_status describe_batch(const clue_params* prm, int32_t D, int32_t n_events, const int64_t* event_offsets,
                           const float* const* coords, const float* weights, const clue_kernel_desc* kernel,
                           const clue_metric_desc* metric, const uint8_t* wrapped, int32_t* cluster_out,
                           int32_t* is_seed_out, float* rho_out, int32_t* nh_out, bool persistent,
                           BatchCall& bc) {
  if (n_events <= 0 || !event_offsets)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "the batch needs at least one event");
  const int64_t n64 = event_offsets[n_events] - event_offsets[0];
  if (event_offsets[0] != 0 || n64 > INT32_MAX)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "event offsets must start at 0 and stay below 2^31 points");
  const int32_t n = static_cast<int32_t>(n64);
  clue_status rc = validate(prm, D, n, coords, weights, kernel, metric, cluster_out);
  if (rc != CLUE_OK) return rc;
  bc = BatchCall{};
  bc.D = D;
  bc.n = n;
  bc.n_events = n_events;
  bc.kd = kernel ? *kernel : clue_kernel_desc{CLUE_KERNEL_FLAT, {0.5f, 0.f, 0.f}};
  if (metric) bc.md = *metric;
  for (int d = 0; d < D; ++d) {
    bc.wrap.w[d] = wrapped ? (wrapped[d] ? 1 : 0) : 0;
    bc.wrapped_any = bc.wrapped_any || bc.wrap.w[d];
    bc.c.p[d] = coords[d];
  }
  bc.weights = weights;
  bc.cluster_out = cluster_out;
  bc.is_seed_out = is_seed_out;
  bc.rho_out = rho_out;
  bc.nh_out = nh_out;
  bc.persistent = persistent;
  // grid shapes with the reference's formulas
  int64_t total_tiles = 0, n_chunks64 = 0;
  for (int e = 0; e < n_events; ++e) {
    const int64_t ne = event_offsets[e + 1] - event_offsets[e];
    if (ne <= 0)
      return fail(CLUE_ERR_INVALID_ARGUMENT,
                  "Number of points passed to PointsDevice constructor must be positive.");
    int32_t npd, T;
    grid_shape(D, static_cast<int32_t>(ne), prm->points_per_tile, npd, T);
    if (npd <= 0 || T <= 0) return fail(CLUE_ERR_INVALID_ARGUMENT, "degenerate tile grid");
    total_tiles += T;
    n_chunks64 += div_up(ne, kChunk);
    bc.max_event_n = std::max(bc.max_event_n, static_cast<int32_t>(ne));
    bc.max_event_tiles = std::max(bc.max_event_tiles, T);
  }
  if (total_tiles >= (int64_t{1} << 30))
    return fail(CLUE_ERR_UNSUPPORTED, "more than 2^30 tiles in one call");
  bc.Ttot = static_cast<int32_t>(total_tiles);
  bc.n_chunks = static_cast<int32_t>(n_chunks64);
  return CLUE_OK;
}

size_t table_bytes(const BatchCall& bc) {
  return static_cast<size_t>(bc.n_events) * sizeof(EventDesc) + static_cast<size_t>(bc.n_chunks) * sizeof(Chunk);
}

// Descriptor of one event: its grid shape with the reference's formulas, tile
// base, integer strides npd^(D-1-d) and chunk range (extremes and tile sizes are
// filled on the device by the bounding-box kernels).
EventDesc event_desc(int32_t D, int32_t begin, int32_t n, int32_t ppb, int32_t tile_base, int32_t chunk_begin) {
  EventDesc ed;
  std::memset(&ed, 0, sizeof(ed));
  ed.begin = begin;
  ed.n = n;
  grid_shape(D, n, ppb, ed.npd, ed.ntiles);
  ed.tile_base = tile_base;
  int32_t s1 = 1;
  for (int d = D - 1; d >= 0; --d) {
    ed.stride[d] = s1;
    s1 *= ed.npd;
  }
  ed.chunk_begin = chunk_begin;
  ed.chunk_end = chunk_begin + static_cast<int32_t>(div_up(n, kChunk));
  return ed;
}

// Event descriptors (grid shape, tile base, integer strides, chunk range) and
// the chunk table, written to host memory laid out as [events | chunks].
void fill_tables(const BatchCall& bc, const clue_params* prm, const int64_t* event_offsets, void* host) {
  auto* hev = static_cast<EventDesc*>(host);
  auto* hch = reinterpret_cast<Chunk*>(hev + bc.n_events);
  int32_t tile_base = 0, chunk = 0;
  for (int e = 0; e < bc.n_events; ++e) {
    EventDesc& ed = hev[e];
    ed = event_desc(bc.D, static_cast<int32_t>(event_offsets[e]),
                    static_cast<int32_t>(event_offsets[e + 1] - event_offsets[e]), prm->points_per_tile,
                    tile_base, chunk);
    tile_base += ed.ntiles;
    for (int32_t b = ed.begin; b < ed.begin + ed.n; b += kChunk)
      hch[chunk++] = Chunk{e, b, std::min(b + kChunk, ed.begin + ed.n), 0};
  }
}

clue_status enqueue_batch(clue_plan_s* p, hipStream_t st, const clue_params* prm, const BatchCall& bc);

// Validate, upload the tables through the plan's pinned staging ring (a slot is
// reused once its previous copy has run) and enqueue the kernels.
clue_status cluster_batch(clue_plan_s* p, hipStream_t st, const clue_params* prm, int32_t D,
                          int32_t n_events, const int64_t* event_offsets,
                          const float* const* coords, const float* weights,
                          const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                          const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                          float* rho_out, int32_t* nh_out, bool persistent_capacity) {
  BatchCall bc;
  clue_status rc = describe_batch(prm, D, n_events, event_offsets, coords, weights, kernel, metric, wrapped,
                                  cluster_out, is_seed_out, rho_out, nh_out, persistent_capacity, bc);
  if (rc != CLUE_OK) return rc;
  const size_t tb = table_bytes(bc);
  clue_plan_s::StageSlot& slot = p->stage[p->stage_next];
  p->stage_next = (p->stage_next + 1) % clue_plan_s::kStageSlots;
  if (slot.done) CLUE_HIP(hipEventSynchronize(slot.done));
  if (slot.bytes < tb) {
    if (slot.host) CLUE_HIP(hipHostFree(slot.host));
    slot.host = nullptr;
    slot.bytes = 0;
    CLUE_HIP(hipHostMalloc(&slot.host, tb, hipHostMallocDefault));
    slot.bytes = tb;
  }
  if (!slot.done) CLUE_HIP(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming));
  fill_tables(bc, prm, event_offsets, slot.host);
  rc = reserve(p, D, bc.n, bc.Ttot, bc.n_events, bc.n_chunks);
  if (rc != CLUE_OK) return rc;
  CLUE_HIP(hipMemcpyAsync(p->events.p, slot.host, static_cast<size_t>(bc.n_events) * sizeof(EventDesc),
                          hipMemcpyHostToDevice, st));
  CLUE_HIP(hipMemcpyAsync(p->chunks.p, static_cast<EventDesc*>(slot.host) + bc.n_events,
                          static_cast<size_t>(bc.n_chunks) * sizeof(Chunk), hipMemcpyHostToDevice, st));
  CLUE_HIP(hipEventRecord(slot.done, st));
  rc = enqueue_batch(p, st, prm, bc);
  if (rc != CLUE_OK) return rc;
  // the introspection calls wait on this event, never on the caller's stream,
  // which may be gone by then
  if (!p->done) CLUE_HIP(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
  CLUE_HIP(hipEventRecord(p->done, st));
  return CLUE_OK;
}

// The kernels of one call; the event and chunk tables are already on the device.
clue_status enqueue_batch(clue_plan_s* p, hipStream_t st, const clue_params* prm, const BatchCall& bc) {
  const int32_t D = bc.D, n = bc.n, n_events = bc.n_events, n_chunks = bc.n_chunks, Ttot = bc.Ttot;
  const clue_kernel_desc& kd = bc.kd;
  const clue_metric_desc& md = bc.md;
  const Wrap& wrap = bc.wrap;
  const bool wrapped_any = bc.wrapped_any;
  const bool persistent_capacity = bc.persistent;
  const Coords& c = bc.c;
  const float* weights = bc.weights;
  int32_t* cluster_out = bc.cluster_out;
  clue_status rc = CLUE_OK;
  float* rho_buf = bc.rho_out ? bc.rho_out : p->rho_tmp.as<float>();
  int32_t* nh_buf = bc.nh_out ? bc.nh_out : p->nh_tmp.as<int32_t>();
  int32_t* seed_buf = bc.is_seed_out ? bc.is_seed_out : p->seed_tmp.as<int32_t>();
  int32_t* counters = p->counters.as<int32_t>();

  // The block search of the nearest-higher pass needs the Euclidean metric on a
  // grid without wrapping; other cases run the tile walk for both passes.
  const int tbits = bits_for(Ttot);
  int sub_bits = std::min(D == 1 ? 10 : (D == 2 ? 5 : 4), (31 - tbits) / D);
  if (p->tuning.morton_bits >= 0) sub_bits = std::min(p->tuning.morton_bits, (31 - tbits) / D);
  const bool small_tiles = n <= p->tuning.tile_call_points;
  const bool blocks = md.type == MEUCLID && !wrapped_any && D <= 3 && sub_bits >= 1 && !small_tiles;
  if (!blocks) sub_bits = -1;
  // small calls: each event's layout and labels in one workgroup (two launches
  // in place of about twenty)
  const bool fused = small_tiles && bc.max_event_n <= p->tuning.fused_event_points &&
                     bc.max_event_n <= kFusedMaxPoints && bc.max_event_tiles <= kFusedMaxTiles;
  const float h = kd.type == CLUE_KERNEL_FLAT ? kd.params[0] : 0.f;
  FusedArgs fa{};
  if (fused) {
    fa.c = c;
    fa.w = weights;
    fa.ev = p->events.as<EventDesc>();
    fa.wrap = wrap;
    fa.pev = p->pev.as<int32_t>();
    fa.perm = p->perm.as<int32_t>();
    fa.offsets = p->offsets.as<int32_t>();
    fa.xs = p->xs.as<float>();
    fa.ld = p->ld;
    fa.ws = p->ws.as<float>();
    fa.hw = p->hw.as<float>();
    fa.h = h;
    fa.tnf = p->tnf.as<int32_t>();
    fa.sc = p->sc.as<int32_t>();
    fa.counters = counters;
    fa.clear_cap = p->reset_capacity && persistent_capacity ? 1 : 0;
    fa.rho = rho_buf;
    fa.delta = p->delta.as<float>();
    fa.nh = nh_buf;
    fa.is_seed = seed_buf;
    fa.cluster = cluster_out;
    fa.seed_dc = prm->seed_dc;
    fa.rhoc = prm->rhoc;
    for (int d = 0; d < CLUE_MAX_DIM; ++d) fa.mp[d] = md.params[d];
    fa.unbounded = prm->seed_mode == CLUE_SEEDS_UNBOUNDED ? 1 : 0;
    fa.persistent = persistent_capacity ? 1 : 0;
  }
  int32_t* sorted_keys = nullptr;
  int32_t* perm = nullptr;     // slot -> point id
  int32_t* mkeys = nullptr;    // tile id << morton_shift of each Morton position
  int morton_shift = sub_bits * D;
  int32_t* slot_of = nullptr;  // Morton position -> tile-order slot
  if (fused) {
    CLUE_MARK(0);
    rc = dispatch_dim(D, [&](auto dc_) { return launch_fused_layout<decltype(dc_)::value>(fa, n_events, st); });
    if (rc != CLUE_OK) return rc;
    CLUE_MARK(1);
    CLUE_MARK(2);
    CLUE_MARK(3);
    perm = p->perm.as<int32_t>();
    p->last_perm = perm;
  } else {
  // ---- bounding boxes and tile sizes
  CLUE_MARK(0);
  k_bbox_partial<<<n_chunks, kThreads, 0, st>>>(c, p->chunks.as<Chunk>(), D, p->bbox_part.as<float>());
  CLUE_LAUNCH_CHECK("bbox_partial");
  k_bbox_final<<<n_events, kWave * D, 0, st>>>(p->bbox_part.as<float>(), D, n_events,
                                                p->events.as<EventDesc>());
  CLUE_LAUNCH_CHECK("bbox_final");

  // ---- tile ids + histogram (+ Morton keys, point -> event)
  CLUE_MARK(1);
  rc = dispatch_dim(D, [&](auto dc_) {
    return launch_bin<decltype(dc_)::value>(c, weights, n_chunks, p, wrap, sub_bits, st);
  });
  if (rc != CLUE_OK) return rc;

  // ---- stable counting sort by tile id: the reference's CSR order
  CLUE_MARK(2);
  rc = radix_sort(p, st, n, tbits, p->key.as<int32_t>(), p->key2.as<int32_t>(), p->perm.as<int32_t>(),
                  p->perm2.as<int32_t>(), true, &sorted_keys, &perm);
  if (rc != CLUE_OK) return rc;
  p->last_perm = perm;

  // ---- tile offsets
  CLUE_MARK(3);
  k_tile_offsets<<<div_up(std::max<int64_t>(Ttot + 1, n_events), kThreads), kThreads, 0, st>>>(
      sorted_keys, n, Ttot, p->offsets.as<int32_t>(), p->tnf.as<int32_t>(), p->sc.as<int32_t>(), n_events);
  CLUE_LAUNCH_CHECK("tile_offsets");

  // ---- tile-ordered SoA (+ Morton-ordered SoA and block boxes)
  rc = dispatch_dim(D, [&](auto dc_) {
    return launch_gather<decltype(dc_)::value>(p, perm, sorted_keys, n, h, blocks, st);
  });
  if (rc != CLUE_OK) return rc;
  if (blocks) {
    // Morton layout: the tile-ordered slots sorted by (tile id, Morton code inside
    // the tile) -- one kernel sorting each tile's slots when the code has at most
    // 12 bits (the tile of Morton position p is then the tile of slot p, whose
    // sorted keys the first sort left), else a radix sort of all of them
    if (sub_bits * D <= 12) {
      slot_of = p->pn.as<int32_t>();
      const int tgrid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(Ttot, 8192)));
      if (sub_bits * D <= 10)
        k_tile_morton<10><<<tgrid, kWave, 0, st>>>(p->kn2.as<int32_t>(), p->offsets.as<int32_t>(), Ttot,
                                                    sub_bits * D, slot_of);
      else
        k_tile_morton<12><<<tgrid, kWave, 0, st>>>(p->kn2.as<int32_t>(), p->offsets.as<int32_t>(), Ttot,
                                                    sub_bits * D, slot_of);
      CLUE_LAUNCH_CHECK("tile_morton");
      mkeys = sorted_keys;
      morton_shift = 0;
    } else {
      rc = radix_sort(p, st, n, tbits + sub_bits * D, p->kn2.as<int32_t>(), p->kn.as<int32_t>(),
                      p->pn.as<int32_t>(), p->pn2.as<int32_t>(), true, &mkeys, &slot_of);
      if (rc != CLUE_OK) return rc;
    }
    k_gather_morton<<<div_up(n, kThreads), kThreads, 0, st>>>(p->xs.as<float>(), p->ws.as<float>(), perm, slot_of,
                                                              n, D, p->ld, p->xn.as<float>(), p->wn.as<float>(),
                                                              p->mid.as<int32_t>());
    CLUE_LAUNCH_CHECK("gather_morton");
    k_block_count<<<div_up(Ttot, kThreads), kThreads, 0, st>>>(p->offsets.as<int32_t>(), Ttot,
                                                               p->bcnt.as<int32_t>());
    CLUE_LAUNCH_CHECK("block_count");
    rc = scan_exclusive(p->bcnt.as<int32_t>(), p->blk_off.as<int32_t>(), Ttot, p->part.as<int32_t>(), st);
    if (rc != CLUE_OK) return rc;
    k_block_tiles<<<div_up(Ttot, kThreads), kThreads, 0, st>>>(p->blk_off.as<int32_t>(), Ttot,
                                                                p->btile.as<int32_t>());
    CLUE_LAUNCH_CHECK("block_tiles");
    k_block_boxes<<<div_up((div_up(n, kBlock) + Ttot) * kBoxLanes, kThreads), kThreads, 0, st>>>(
        p->xn.as<float>(), p->ld, D, p->offsets.as<int32_t>(), p->blk_off.as<int32_t>(), p->btile.as<int32_t>(),
        Ttot, p->bbox_lo.as<float>(), p->bbox_hi.as<float>(), p->ldb);
    CLUE_LAUNCH_CHECK("block_boxes");
    k_tile_boxes<<<div_up(static_cast<int64_t>(Ttot) * kTileLanes, kThreads), kThreads, 0, st>>>(
        p->bbox_lo.as<float>(), p->bbox_hi.as<float>(), p->ldb, D, p->blk_off.as<int32_t>(), Ttot,
        p->tbox_lo.as<float>(), p->tbox_hi.as<float>());
    CLUE_LAUNCH_CHECK("tile_boxes");
  }
  }  // the layout of calls not fused

  PassArgs a{};
  for (int d = 0; d < D; ++d) a.xs[d] = p->xs.as<float>() + d * p->ld;
  a.ws = p->ws.as<float>();
  a.hw = p->hw.as<float>();
  a.rho_s = p->rho_s.as<float>();
  a.perm = perm;
  a.offsets = p->offsets.as<int32_t>();
  a.ev = p->events.as<EventDesc>();
  a.pev = p->pev.as<int32_t>();
  a.sc = p->sc.as<int32_t>();
  a.wrap = wrap;
  a.n = n;
  a.n_events = n_events;
  for (int k = 0; k < 3; ++k) a.kp[k] = kd.params[k];
  a.kden_rcp = 1.0 / static_cast<double>((2.f * kd.params[1]) * kd.params[1]);
  // Gaussian: -(d-mu)^2 / den <= 0 when 1/den is finite and positive; exponential:
  // -lambda * d <= 0 when lambda >= 0 (d is finite and >= 0 for in-radius candidates)
  a.exp_nonpos = (kd.type == CLUE_KERNEL_GAUSSIAN && std::isfinite(a.kden_rcp) && a.kden_rcp > 0.0 &&
                  std::isfinite(kd.params[0])) ||
                 (kd.type == CLUE_KERNEL_EXPONENTIAL && kd.params[0] >= 0.f && std::isfinite(kd.params[0]));
  // in-radius distances lie in [0, dc]: bound the exp argument from below (with
  // room for the float roundings of d - mu, its square, the quotient and -lambda*d)
  if (a.exp_nonpos) {
    const double dc = prm->dc;
    double lowest;
    if (kd.type == CLUE_KERNEL_GAUSSIAN) {
      const double mu = std::fabs(static_cast<double>(kd.params[0]));
      const double m = std::max(mu, dc + mu);
      lowest = m * m * a.kden_rcp;
    } else {
      lowest = static_cast<double>(kd.params[0]) * dc;
    }
    const float amp = kd.type == CLUE_KERNEL_GAUSSIAN ? kd.params[2] : kd.params[1];
    a.exp_unit = amp == 1.0f && lowest * 1.001 <= 14.0;
    // the exponential's unit evaluation also runs below a squared distance of
    // 2^-96 (where the fast square root is not exact); a rate of at most 2^22
    // keeps the argument there above -2^-25, whose exp is exactly 1
    if (kd.type == CLUE_KERNEL_EXPONENTIAL && !(kd.params[0] <= 0x1p22f)) a.exp_unit = 0;
    if (a.exp_unit && kd.type == CLUE_KERNEL_GAUSSIAN && kd.params[0] == 0.f) {
      // (2 sigma) sigma = 2^e with 2^-e a float: -(d*d) / den is d*d scaled exactly;
      // 2^-e <= 2^70 keeps the argument above -2^-25 at every squared distance
      // below 2^-96, where this evaluation runs too
      const float den = (2.f * kd.params[1]) * kd.params[1];
      int e = 0;
      const float mant = std::frexp(den, &e);  // den = mant * 2^e, mant in [0.5, 1)
      if (std::isfinite(den) && den > 0.f && mant == 0.5f && e - 1 <= 149 && e - 1 >= -70) a.exp_unit = 2;
    }
  }
  for (int d = 0; d < CLUE_MAX_DIM; ++d) a.mp[d] = md.params[d];
  if (blocks) {
    for (int d = 0; d < D; ++d) {
      a.xn[d] = p->xn.as<float>() + d * p->ld;
      a.bmin[d] = p->bbox_lo.as<float>() + d * p->ldb;
      a.bmax[d] = p->bbox_hi.as<float>() + d * p->ldb;
      a.tmin[d] = p->tbox_lo.as<float>() + static_cast<int64_t>(d) * Ttot;
      a.tmax[d] = p->tbox_hi.as<float>() + static_cast<int64_t>(d) * Ttot;
    }
    a.wn = p->wn.as<float>();
    a.sn = slot_of;
    a.pn = p->mid.as<int32_t>();
    a.kn = mkeys;
    a.mshift = morton_shift;
    a.blk_off = p->blk_off.as<int32_t>();
    a.btile = p->btile.as<int32_t>();
    a.bmaxk = p->bmaxk.as<uint64_t>();
    a.tmaxk = p->tmaxk.as<uint64_t>();
    a.pending = p->pending.as<int32_t>();
    a.pending_count = p->pending_count.as<int32_t>();
  }
  a.ntiles = Ttot;
  a.xcd = p->tuning.xcd_remap;
  a.dfilter = p->tuning.density_filter;
  a.nh_lone = p->tuning.nh_lone_lanes;
  a.tnf = p->tnf.as<int32_t>();
  // rank keys of the densities: Morton order for the block search, tile order for the tile walk
  a.hk_out = p->hk.as<uint64_t>();
  a.hk = p->hk.as<uint64_t>();
  a.rho_s_out = p->rho_s.as<float>();
  a.rho_out = rho_buf;
  a.nh_out = nh_buf;
  a.delta_out = p->delta.as<float>();
  rc = dispatch_dim(D, [&](auto dc_) {
    return dispatch_metric_passes<decltype(dc_)::value>(md.type, a, kd.type, prm->dc, prm->dm, st, p,
                                                        blocks);
  });
  if (rc != CLUE_OK) return rc;

  // ---- seeds, ranks, labels
  CLUE_MARK(6);
  if (fused) {
    rc = dispatch_dim(D, [&](auto dc_) {
      return dispatch_fused_labels<decltype(dc_)::value>(md.type, fa, n_events, st);
    });
    if (rc != CLUE_OK) return rc;
    if (persistent_capacity) p->reset_capacity = false;
    CLUE_MARK(7);
    CLUE_MARK(8);
    p->has_run = true;
    p->last_n = n;
    p->last_d = D;
    p->last_T = Ttot;
    p->last_events = n_events;
    return CLUE_OK;
  }
  ClassifyArgs ca{};
  ca.c = c;
  ca.rho = rho_buf;
  ca.delta = p->delta.as<float>();
  ca.nh = nh_buf;
  ca.is_seed = seed_buf;
  ca.flag = p->flag.as<int32_t>();
  ca.root = p->root.as<int32_t>();
  ca.n = n;
  ca.seed_dc = prm->seed_dc;
  ca.rhoc = prm->rhoc;
  for (int d = 0; d < CLUE_MAX_DIM; ++d) ca.mp[d] = md.params[d];
  // totals of this call ([2, 8) totals, [8, 11) label-round flags); [1], the
  // grow-only capacity, carries over unless reset
  ca.counters = counters;
  ca.clear_cap = p->reset_capacity && persistent_capacity ? 1 : 0;
  if (persistent_capacity) p->reset_capacity = false;
  rc = dispatch_dim(D, [&](auto dc_) {
    return dispatch_metric_classify<decltype(dc_)::value>(md.type, ca, st);
  });
  if (rc != CLUE_OK) return rc;
  rc = scan_exclusive(p->flag.as<int32_t>(), p->flag.as<int32_t>(), n, p->part.as<int32_t>(), st);
  if (rc != CLUE_OK) return rc;
  k_capacity<<<div_up(n_events, kThreads), kThreads, 0, st>>>(
      p->events.as<EventDesc>(), n_events, p->sc.as<int32_t>(), p->flag.as<int32_t>(), counters,
      p->cap_used.as<int32_t>(), p->rank_base.as<int32_t>(), prm->seed_mode == CLUE_SEEDS_UNBOUNDED ? 1 : 0,
      persistent_capacity ? 1 : 0);
  CLUE_LAUNCH_CHECK("capacity");
  CLUE_MARK(7);
  // Two synchronous halvings first shorten the chains each thread then follows;
  // rounds with 4096 hops each cover any forest depth below 4096^rounds.
  for (int r = 0; r < 2; ++r) {
    k_jump<<<div_up(n, kThreads), kThreads, 0, st>>>(p->root.as<int32_t>(), n);
    CLUE_LAUNCH_CHECK("jump");
  }
  const int hops = 4096;
  int rounds = 1;
  for (int64_t span = hops; span < n; span *= hops) ++rounds;
  for (int r = 0; r < rounds; ++r) {
    const int64_t lgrid = r == 0 ? div_up(n, kThreads) : std::min<int64_t>(div_up(n, kThreads), 1024);
    k_labels<<<lgrid, kThreads, 0, st>>>(
        p->root.as<int32_t>(), n, hops, seed_buf, p->flag.as<int32_t>(), p->pev.as<int32_t>(),
        p->events.as<EventDesc>(), p->cap_used.as<int32_t>(), p->rank_base.as<int32_t>(), cluster_out,
        nh_buf, counters + 8, r);
    CLUE_LAUNCH_CHECK("labels");
  }

  CLUE_MARK(8);
  p->has_run = true;
  p->last_n = n;
  p->last_d = D;
  p->last_T = Ttot;
  p->last_events = n_events;
  return CLUE_OK;
}

clue_status make_clusters(clue_plan_s* p, hipStream_t st, const clue_params* prm, int32_t D,
                          int32_t n, const float* const* coords, const float* weights,
                          const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                          const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                          float* rho_out, int32_t* nh_out) {
  if (n <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Number of points passed to PointsDevice constructor must be positive.");
  const int64_t offsets[2] = {0, n};
  return cluster_batch(p, st, prm, D, 1, offsets, coords, weights, kernel, metric, wrapped, cluster_out,
                       is_seed_out, rho_out, nh_out, true);
}

// A captured batched call: its own plan (scratch, tables) and the instantiated
// graph of the call's kernels.
struct BatchGraph {
  clue_plan_s* plan = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Scratch for batched calls of up to npoints points in up to max_events events
// of up to ndim dimensions at 128 points per tile or more, so that such calls
// allocate nothing (the served engine's slots).  An event of nt tiles has at
// most 2^D * nt tiles after the reference's rounding to npd^D; the bound is taken
// with D at most 3 (larger grids of tiny high-dimensional events still grow the
// buffers on their first call).  The pinned staging ring is sized for the tables.
// ---------------------------------------------------------------------------
// The no-allocation entry point's workspace: one array per GPU, in the code
// object's device memory, held from the library's load on that GPU until the
// process ends.  80 MiB hold a call of 262,144 points in any dimension 1..10 at
// 16 points per tile or more (72 MiB at ten dimensions); larger calls are
// refused.  Host code never uses the array's name as an address (on the host it
// is only a stand-in): its device address comes from hipGetSymbolAddress, once
// per device, and every buffer of the call is carved from that address.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t{80} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

struct StaticWorkspace {
  std::mutex mu;          // one host caller at a time per device
  unsigned char* base = nullptr;
  clue_plan_s plan;       // its buffers point into the device's scratch array
};
StaticWorkspace g_static_ws[kMaxDevices];

clue_status make_clusters_noalloc(int32_t device, hipStream_t st, const clue_params* prm, int32_t D, int32_t n,
                                  const float* const* coords, const float* weights, const clue_kernel_desc* kernel,
                                  const clue_metric_desc* metric, const uint8_t* wrapped, int32_t* cluster_out,
                                  int32_t* is_seed_out, float* rho_out, int32_t* nh_out) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(CLUE_ERR_RUNTIME, "No device available!");
  if (device < 0 || device >= count || device >= kMaxDevices)
    return fail(CLUE_ERR_OUT_OF_RANGE, "device index out of range");
  if (n <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Number of points passed to PointsDevice constructor must be positive.");
  StaticWorkspace& ws = g_static_ws[device];
  std::lock_guard<std::mutex> lock(ws.mu);
  CLUE_HIP(hipSetDevice(device));
  if (!ws.base) {
    void* ptr = nullptr;
    CLUE_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(CLUE_ERR_RUNTIME, "hipGetSymbolAddress returned no address for the workspace");
    ws.base = static_cast<unsigned char*>(ptr);
    ws.plan.device = device;
    ws.plan.arena = ws.base;
    ws.plan.arena_bytes = kScratchBytes;
  }
  const int64_t offsets[2] = {0, n};
  BatchCall bc;
  // a fresh Clusterer per call: the event's own seed capacity (persistent = false)
  clue_status rc = describe_batch(prm, D, 1, offsets, coords, weights, kernel, metric, wrapped, cluster_out,
                                  is_seed_out, rho_out, nh_out, false, bc);
  if (rc != CLUE_OK) return rc;
  clue_plan_s* p = &ws.plan;
  rc = reserve(p, D, n, bc.Ttot, 1, bc.n_chunks);
  if (rc != CLUE_OK) return rc;
  if (p->arena_used > p->arena_bytes)
    return fail(CLUE_ERR_OUT_OF_RANGE,
                "clue_make_clusters_noalloc: " + std::to_string(n) + " points in " + std::to_string(D) +
                    " dimensions (" + std::to_string(bc.Ttot) + " tiles) need " + std::to_string(p->arena_used) +
                    " bytes of workspace; the static workspace holds " + std::to_string(p->arena_bytes) +
                    " (262,144 points in any dimension at 16 points per tile or more)");
  const EventDesc ed = event_desc(D, 0, n, prm->points_per_tile, 0, 0);
  k_single_event_tables<<<div_up(bc.n_chunks, kThreads), kThreads, 0, st>>>(ed, p->events.as<EventDesc>(),
                                                                              p->chunks.as<Chunk>());
  CLUE_LAUNCH_CHECK("single_event_tables");
  // the workspace keeps the previous call's contents: clear the counters
  k_fill_i32<<<1, 64, 0, st>>>(p->counters.as<int32_t>(), 0, 16);
  CLUE_LAUNCH_CHECK("fill");
  return enqueue_batch(p, st, prm, bc);
}

clue_status reserve_batches(clue_plan_t p, int32_t ndim, int32_t npoints, int32_t max_events) {
  if (!p || ndim < 1 || ndim > CLUE_MAX_DIM || npoints <= 0 || max_events <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid reserve size");
  const int64_t e = std::min<int64_t>(max_events, npoints);
  const int64_t tiles = (int64_t{1} << std::min(ndim, 3)) * (div_up(npoints, 128) + e);
  const int64_t chunks = div_up(npoints, kChunk) + e;
  if (tiles > INT32_MAX / 2) return fail(CLUE_ERR_INVALID_ARGUMENT, "reserve size too large");
  CLUE_HIP(hipSetDevice(p->device));
  clue_status rc = reserve(p, ndim, npoints, static_cast<int32_t>(tiles), static_cast<int32_t>(e),
                           static_cast<int32_t>(chunks));
  if (rc != CLUE_OK) return rc;
  BatchCall bc;
  bc.n_events = static_cast<int32_t>(e);
  bc.n_chunks = static_cast<int32_t>(chunks);
  const size_t tb = table_bytes(bc);
  for (auto& slot : p->stage) {
    if (slot.done) CLUE_HIP(hipEventSynchronize(slot.done));
    if (slot.bytes < tb) {
      if (slot.host) CLUE_HIP(hipHostFree(slot.host));
      slot.host = nullptr;
      slot.bytes = 0;
      CLUE_HIP(hipHostMalloc(&slot.host, tb, hipHostMallocDefault));
      slot.bytes = tb;
    }
  }
  return CLUE_OK;
}

}  // namespace clue_hip

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* clue_last_error_message(void) { return g_error.c_str(); }
const char* clue_version(void) { return "clue-mi355x 0.1 (gfx950)"; }

clue_status clue_device_count(int32_t* count) {
  if (!count) return fail(CLUE_ERR_INVALID_ARGUMENT, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return fail(CLUE_ERR_RUNTIME, std::string("No device available! ") + hipGetErrorString(e));
  }
  *count = c;
  return CLUE_OK;
}

clue_status clue_set_device(int32_t device) {
  CLUE_HIP(hipSetDevice(device));
  return CLUE_OK;
}

clue_status clue_stream_create(int32_t device, void** stream) {
  if (!stream) return fail(CLUE_ERR_INVALID_ARGUMENT, "stream is NULL");
  CLUE_HIP(hipSetDevice(device));
  hipStream_t s;
  CLUE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return CLUE_OK;
}

clue_status clue_stream_destroy(void* stream) {
  if (stream) CLUE_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_stream_synchronize(void* stream) {
  CLUE_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_malloc_device(int32_t device, size_t bytes, void** ptr) {
  if (!ptr) return fail(CLUE_ERR_INVALID_ARGUMENT, "ptr is NULL");
  CLUE_HIP(hipSetDevice(device));
  CLUE_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return CLUE_OK;
}

clue_status clue_free_device(void* ptr) {
  if (ptr) CLUE_HIP(hipFree(ptr));
  return CLUE_OK;
}

clue_status clue_malloc_host(size_t bytes, void** ptr) {
  if (!ptr) return fail(CLUE_ERR_INVALID_ARGUMENT, "ptr is NULL");
  CLUE_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return CLUE_OK;
}

clue_status clue_free_host(void* ptr) {
  if (ptr) CLUE_HIP(hipHostFree(ptr));
  return CLUE_OK;
}

clue_status clue_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  CLUE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  CLUE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  CLUE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_memset_device(void* dst, int32_t value, size_t bytes, void* stream) {
  CLUE_HIP(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_plan_create(int32_t device, clue_plan_t* plan) {
  if (!plan) return fail(CLUE_ERR_INVALID_ARGUMENT, "plan is NULL");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(CLUE_ERR_RUNTIME, "No device available!");
  if (device < 0 || device >= count) return fail(CLUE_ERR_OUT_OF_RANGE, "device index out of range");
  clue_plan_s* p = new (std::nothrow) clue_plan_s();
  if (!p) return fail(CLUE_ERR_OUT_OF_MEMORY, "plan allocation failed");
  p->device = device;
  *plan = p;
  return CLUE_OK;
}

clue_status clue_plan_destroy(clue_plan_t plan) {
  if (!plan) return CLUE_OK;
  (void)hipSetDevice(plan->device);
  if (plan->done) {
    (void)hipEventSynchronize(plan->done);
    (void)hipEventDestroy(plan->done);
  }
  for (int k = 0; DevBuf* b = plan->all_buffers(k); ++k) b->release();
  for (auto& slot : plan->stage) {
    if (slot.done) (void)hipEventSynchronize(slot.done);
    if (slot.host) (void)hipHostFree(slot.host);
    if (slot.done) (void)hipEventDestroy(slot.done);
  }
  for (hipEvent_t& e : plan->ev)
    if (e) (void)hipEventDestroy(e);
  delete plan;
  return CLUE_OK;
}

clue_status clue_plan_reset_capacity(clue_plan_t plan) {
  if (!plan) return fail(CLUE_ERR_INVALID_ARGUMENT, "plan is NULL");
  plan->reset_capacity = true;  // applied in stream order by the next plain call
  return CLUE_OK;
}

clue_status clue_plan_reserve(clue_plan_t plan, int32_t ndim, int32_t npoints) {
  if (!plan) return fail(CLUE_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (ndim < 1 || ndim > CLUE_MAX_DIM || npoints <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid reserve size");
  int32_t npd, T;
  grid_shape(ndim, npoints, 128, npd, T);
  return reserve(plan, ndim, npoints, T, 1, static_cast<int32_t>(div_up(npoints, kChunk)));
}

clue_status clue_plan_device_bytes(clue_plan_t plan, size_t* bytes) {
  if (!plan || !bytes) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  size_t s = 0;
  for (int k = 0; DevBuf* b = plan->all_buffers(k); ++k) s += b->bytes;
  *bytes = s;
  return CLUE_OK;
}

clue_status clue_make_clusters(clue_plan_t plan, void* stream, const clue_params* params,
                               int32_t ndim, int32_t n, const float* const* coords_dev,
                               const float* weights_dev, const clue_kernel_desc* kernel,
                               const clue_metric_desc* metric, const uint8_t* wrapped,
                               int32_t* cluster_out, int32_t* is_seed_out, float* rho_out,
                               int32_t* nearest_higher_out) {
  if (!plan) return fail(CLUE_ERR_INVALID_ARGUMENT, "plan is NULL");
  CLUE_HIP(hipSetDevice(plan->device));
  return make_clusters(plan, static_cast<hipStream_t>(stream), params, ndim, n, coords_dev,
                       weights_dev, kernel, metric, wrapped, cluster_out, is_seed_out, rho_out,
                       nearest_higher_out);
}

clue_status clue_make_clusters_noalloc(int32_t device, void* stream, const clue_params* params, int32_t ndim,
                                       int32_t n, const float* const* coords_dev, const float* weights_dev,
                                       const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                       const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                                       float* rho_out, int32_t* nearest_higher_out) {
  return make_clusters_noalloc(device, static_cast<hipStream_t>(stream), params, ndim, n, coords_dev, weights_dev,
                               kernel, metric, wrapped, cluster_out, is_seed_out, rho_out, nearest_higher_out);
}

clue_status clue_make_clusters_batched(clue_plan_t plan, void* stream, const clue_params* params,
                                       int32_t ndim, int32_t n_events, const int64_t* event_offsets,
                                       const float* const* coords_dev, const float* weights_dev,
                                       const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                       const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                                       float* rho_out, int32_t* nearest_higher_out) {
  if (!plan) return fail(CLUE_ERR_INVALID_ARGUMENT, "plan is NULL");
  CLUE_HIP(hipSetDevice(plan->device));
  return cluster_batch(plan, static_cast<hipStream_t>(stream), params, ndim, n_events, event_offsets,
                       coords_dev, weights_dev, kernel, metric, wrapped, cluster_out, is_seed_out, rho_out,
                       nearest_higher_out, false);
}

clue_status clue_main_run(float dc, float rhoc, float dm, float seed_dc, int32_t points_per_tile,
                          const uint8_t* wrapped, const float* data, int32_t* results,
                          const clue_kernel_desc* kernel, int32_t ndim, int32_t n,
                          int32_t block_size, int32_t device) {
  (void)block_size;  // the engine picks its own launch shapes
  if (!data || !results) return fail(CLUE_ERR_INVALID_ARGUMENT, "data and results are required");
  if (ndim < 1 || ndim > CLUE_MAX_DIM) return fail(CLUE_ERR_INVALID_ARGUMENT, "Ndim must be in [1, 10]");
  if (n <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT,
                "Number of points passed to PointsDevice constructor must be positive.");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(CLUE_ERR_RUNTIME, "no HIP device available");
  if (device < 0 || device >= count) return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid device id");
  // One context per device for the whole process (like the reference's
  // function-static queue): a plan, a stream and grow-only device buffers,
  // used by one call at a time and kept until the process ends.
  struct MainRunCtx {
    std::mutex mu;
    clue_plan_t plan = nullptr;
    void* stream = nullptr;
    DevBuf in, out;
  };
  static std::mutex ctx_mu;
  static MainRunCtx* ctxs[64] = {};
  if (device >= 64) return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid device id");
  MainRunCtx* ctx;
  {
    std::lock_guard<std::mutex> g(ctx_mu);
    if (!ctxs[device]) ctxs[device] = new MainRunCtx();
    ctx = ctxs[device];
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  clue_status rc = CLUE_OK;
  if (!ctx->plan) rc = clue_plan_create(device, &ctx->plan);
  if (rc == CLUE_OK && !ctx->stream) rc = clue_stream_create(device, &ctx->stream);
  if (rc != CLUE_OK) return rc;
  CLUE_HIP(hipSetDevice(device));
  const size_t nn = static_cast<size_t>(n);
  CLUE_HIP(ctx->in.ensure((ndim + 1) * nn * 4));
  CLUE_HIP(ctx->out.ensure(nn * 4));
  float* dbuf = ctx->in.as<float>();
  int32_t* dout = ctx->out.as<int32_t>();
  clue_params prm{dc, rhoc, dm, seed_dc, points_per_tile, CLUE_SEEDS_REFERENCE};
  rc = clue_plan_reset_capacity(ctx->plan);  // a fresh Clusterer per call
  if (rc == CLUE_OK) rc = clue_memcpy_h2d(dbuf, data, (ndim + 1) * nn * 4, ctx->stream);
  if (rc == CLUE_OK) {
    const float* cp[CLUE_MAX_DIM];
    for (int d = 0; d < ndim; ++d) cp[d] = dbuf + d * nn;
    rc = clue_make_clusters(ctx->plan, ctx->stream, &prm, ndim, n, cp, dbuf + ndim * nn, kernel, nullptr,
                            wrapped, dout, nullptr, nullptr, nullptr);
  }
  if (rc == CLUE_OK) rc = clue_memcpy_d2h(results, dout, nn * 4, ctx->stream);
  const std::string keep = g_error;
  const clue_status rs = clue_stream_synchronize(ctx->stream);
  if (rc == CLUE_OK) return rs;
  g_error = keep;
  return rc;
}

clue_status clue_plan_last_stats(clue_plan_t plan, clue_stats* stats) {
  if (!plan || !stats) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!plan->has_run || !plan->done) return fail(CLUE_ERR_INVALID_ARGUMENT, "the plan has not run");
  CLUE_HIP(hipSetDevice(plan->device));
  CLUE_HIP(hipEventSynchronize(plan->done));
  int32_t c[16];
  CLUE_HIP(hipMemcpy(c, plan->counters.p, sizeof(c), hipMemcpyDeviceToHost));
  stats->seed_candidates = c[0];
  stats->seed_capacity = c[2];
  stats->n_seeds = c[3];
  stats->n_clusters = c[4];
  stats->n_dropped_seeds = c[5];
  return CLUE_OK;
}

clue_status clue_plan_last_grid(clue_plan_t plan, clue_grid_info* grid) {
  if (!plan || !grid) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!plan->has_run || !plan->done) return fail(CLUE_ERR_INVALID_ARGUMENT, "the plan has not run");
  CLUE_HIP(hipSetDevice(plan->device));
  CLUE_HIP(hipEventSynchronize(plan->done));
  EventDesc g;  // the first event of the last call (the only one of a plain call)
  CLUE_HIP(hipMemcpy(&g, plan->events.p, sizeof(g), hipMemcpyDeviceToHost));
  std::memset(grid, 0, sizeof(*grid));
  grid->ndim = plan->last_d;
  grid->npoints = g.n;
  grid->npd = g.npd;
  grid->ntiles = g.ntiles;
  for (int d = 0; d < plan->last_d; ++d) {
    grid->minmax[2 * d] = g.mn[d];
    grid->minmax[2 * d + 1] = g.mx[d];
    grid->tile_size[d] = g.ts[d];
  }
  return CLUE_OK;
}

clue_status clue_plan_copy_tiles(clue_plan_t plan, int32_t* offsets, int32_t* indexes) {
  if (!plan || !offsets || !indexes) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!plan->has_run || !plan->done) return fail(CLUE_ERR_INVALID_ARGUMENT, "the plan has not run");
  CLUE_HIP(hipSetDevice(plan->device));
  CLUE_HIP(hipEventSynchronize(plan->done));
  CLUE_HIP(hipMemcpy(offsets, plan->offsets.p, (static_cast<size_t>(plan->last_T) + 1) * 4,
                     hipMemcpyDeviceToHost));
  CLUE_HIP(hipMemcpy(indexes, plan->last_perm, static_cast<size_t>(plan->last_n) * 4,
                     hipMemcpyDeviceToHost));
  return CLUE_OK;
}

clue_status clue_plan_copy_delta(clue_plan_t plan, float* delta) {
  if (!plan || !delta) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!plan->has_run || !plan->done) return fail(CLUE_ERR_INVALID_ARGUMENT, "the plan has not run");
  CLUE_HIP(hipSetDevice(plan->device));
  CLUE_HIP(hipEventSynchronize(plan->done));
  CLUE_HIP(hipMemcpy(delta, plan->delta.p, static_cast<size_t>(plan->last_n) * 4,
                     hipMemcpyDeviceToHost));
  return CLUE_OK;
}

clue_status clue_plan_set_tuning(clue_plan_t plan, const clue_tuning* tuning) {
  if (!plan || !tuning) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!(tuning->nh_first_radius >= 0.f) || tuning->morton_bits > 10 ||
      tuning->xcd_remap < 0 || tuning->xcd_remap > 7 || tuning->pending_blocks < 0 ||
      (tuning->pending_group != 0 && tuning->pending_group != 8 && tuning->pending_group != 16 &&
       tuning->pending_group != 32 && tuning->pending_group != 64) ||
      tuning->density_order < -1 || tuning->density_order > 1 || tuning->small_call_points < 0 ||
      tuning->wave_dims < 0 || tuning->wave_dims > CLUE_MAX_DIM || tuning->tile_call_points < 0 ||
      tuning->density_filter < 0 || tuning->density_filter > 1 || tuning->fused_event_points < 0 ||
      tuning->fused_event_points > kFusedMaxPoints || tuning->split_call_points < 0 ||
      tuning->nh_lone_lanes < 0 || tuning->nh_lone_lanes > kWave)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid tuning values");
  plan->tuning = *tuning;
  return CLUE_OK;
}

clue_status clue_plan_get_tuning(clue_plan_t plan, clue_tuning* tuning) {
  if (!plan || !tuning) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  *tuning = plan->tuning;
  return CLUE_OK;
}

clue_status clue_plan_set_timing(clue_plan_t plan, int32_t enable) {
  if (!plan) return fail(CLUE_ERR_INVALID_ARGUMENT, "plan is NULL");
  plan->timing = enable != 0;
  return CLUE_OK;
}

clue_status clue_plan_stage_times(clue_plan_t plan, float* ms, int32_t capacity, int32_t* count) {
  if (!plan || !count) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  *count = kStages;
  if (!ms) return CLUE_OK;
  if (!plan->has_run || !plan->timing) return fail(CLUE_ERR_INVALID_ARGUMENT, "no timed call");
  CLUE_HIP(hipSetDevice(plan->device));
  CLUE_HIP(hipEventSynchronize(plan->ev[kStages]));
  for (int k = 0; k < kStages && k < capacity; ++k)
    CLUE_HIP(hipEventElapsedTime(&ms[k], plan->ev[k], plan->ev[k + 1]));
  return CLUE_OK;
}

const char* clue_stage_name(int32_t stage) {
  return (stage >= 0 && stage < kStages) ? kStageNames[stage] : "";
}

clue_status clue_cluster_map(clue_plan_t plan, void* stream, int32_t n, const int32_t* cluster_dev,
                             int32_t* offsets_dev, int32_t* indexes_dev, int32_t* counts_dev) {
  if (!plan || !cluster_dev || !offsets_dev || !indexes_dev || !counts_dev)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n <= 0 || n >= INT32_MAX) return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid number of labels");
  CLUE_HIP(hipSetDevice(plan->device));
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t N = static_cast<size_t>(n);
  const size_t nb = static_cast<size_t>(div_up(n, kSortTile));
  const int bits = bits_for(static_cast<int64_t>(n) + 1);
  const size_t hist_entries = static_cast<size_t>(1) << std::min(8, bits);
  CLUE_HIP(plan->mkey.ensure(N * 4));
  CLUE_HIP(plan->mkey2.ensure(N * 4));
  CLUE_HIP(plan->mval.ensure(N * 4));
  CLUE_HIP(plan->mval2.ensure(N * 4));
  CLUE_HIP(plan->hist.ensure((hist_entries * nb + 1) * 4));
  CLUE_HIP(plan->part.ensure((div_up(static_cast<int64_t>(hist_entries * nb + 1), kScanTile) + 1) * 4));
  k_fill_i32<<<1, 64, 0, st>>>(counts_dev, 0, 2);
  CLUE_LAUNCH_CHECK("fill");
  k_map_keys<<<div_up(n, kThreads), kThreads, 0, st>>>(cluster_dev, n, plan->mkey.as<int32_t>());
  CLUE_LAUNCH_CHECK("map_keys");
  int32_t* skeys = nullptr;
  int32_t* perm = nullptr;
  clue_status rc = radix_sort(plan, st, n, bits, plan->mkey.as<int32_t>(), plan->mkey2.as<int32_t>(),
                              plan->mval.as<int32_t>(), plan->mval2.as<int32_t>(), true, &skeys, &perm);
  if (rc != CLUE_OK) return rc;
  k_map_offsets<<<div_up(n, kThreads), kThreads, 0, st>>>(skeys, n, offsets_dev, counts_dev);
  CLUE_LAUNCH_CHECK("map_offsets");
  CLUE_HIP(hipMemcpyAsync(indexes_dev, perm, N * 4, hipMemcpyDeviceToDevice, st));
  return CLUE_OK;
}

clue_status clue_cluster_centroids(clue_plan_t plan, void* stream, int32_t ndim, const float* const* coords_dev,
                                   const int32_t* offsets_dev, const int32_t* indexes_dev, const int32_t* counts_dev,
                                   int32_t max_clusters, float* centroids_dev) {
  if (!plan || !coords_dev || !offsets_dev || !indexes_dev || !counts_dev || !centroids_dev)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (ndim < 1 || ndim > CLUE_MAX_DIM) return fail(CLUE_ERR_INVALID_ARGUMENT, "Ndim must be in [1, 10]");
  if (max_clusters <= 0) return CLUE_OK;
  for (int d = 0; d < ndim; ++d)
    if (!coords_dev[d]) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL coordinate pointer");
  CLUE_HIP(hipSetDevice(plan->device));
  const hipStream_t st = static_cast<hipStream_t>(stream);
  Coords c{};
  for (int d = 0; d < ndim; ++d) c.p[d] = coords_dev[d];
  const int64_t threads = static_cast<int64_t>(max_clusters) * ndim;
  k_centroids<<<div_up(threads, kThreads), kThreads, 0, st>>>(c, ndim, offsets_dev, indexes_dev, counts_dev,
                                                              max_clusters, centroids_dev);
  CLUE_LAUNCH_CHECK("centroids");
  return CLUE_OK;
}

clue_status clue_selftest_sqrt(uint32_t first_bits, uint64_t count, uint64_t* mismatches) {
  if (!mismatches) return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  if (count == 0 || first_bits + (count - 1) < first_bits) return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid range");
  unsigned long long* d = nullptr;
  CLUE_HIP(hipMalloc(&d, sizeof(*d)));
  hipError_t e = hipMemset(d, 0, sizeof(*d));
  if (e == hipSuccess) {
    k_selftest_sqrt<<<4096, kThreads>>>(first_bits, count, d);
    e = hipGetLastError();
  }
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(CLUE_ERR_RUNTIME, hipGetErrorString(e));
  *mismatches = h;
  return CLUE_OK;
}

clue_status clue_selftest_math(int32_t op, const float* a_dev, const float* b_dev, float* out_dev,
                               int64_t n, void* stream) {
  if (n <= 0) return CLUE_OK;
  if (!a_dev || !out_dev || ((op == 2 || op == 4) && !b_dev))
    return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  k_selftest<<<div_up(n, kThreads), kThreads, 0, static_cast<hipStream_t>(stream)>>>(op, a_dev, b_dev,
                                                                                   out_dev, n);
  CLUE_LAUNCH_CHECK("selftest");
  return CLUE_OK;
}

clue_status clue_batch_graph_create(int32_t device, const clue_params* params, int32_t ndim,
                                    int32_t n_events, const int64_t* event_offsets,
                                    const float* const* coords_dev, const float* weights_dev,
                                    const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                    const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                                    float* rho_out, int32_t* nearest_higher_out, const clue_tuning* tuning,
                                    clue_batch_graph_t* graph) {
  if (!graph) return fail(CLUE_ERR_INVALID_ARGUMENT, "graph is NULL");
  *graph = nullptr;
  BatchCall bc;
  clue_status rc = describe_batch(params, ndim, n_events, event_offsets, coords_dev, weights_dev, kernel, metric,
                                  wrapped, cluster_out, is_seed_out, rho_out, nearest_higher_out, false, bc);
  if (rc != CLUE_OK) return rc;
  clue_plan_t plan = nullptr;
  rc = clue_plan_create(device, &plan);
  if (rc != CLUE_OK) return rc;
  auto bg = std::make_unique<BatchGraph>();
  bg->plan = plan;
  auto cleanup = [&]() {
    const std::string keep = g_error;
    if (bg->exec) (void)hipGraphExecDestroy(bg->exec);
    if (bg->graph) (void)hipGraphDestroy(bg->graph);
    clue_plan_destroy(bg->plan);
    g_error = keep;
  };
  if (tuning) rc = clue_plan_set_tuning(plan, tuning);
  if (rc == CLUE_OK) rc = reserve(plan, ndim, bc.n, bc.Ttot, bc.n_events, bc.n_chunks);
  if (rc != CLUE_OK) {
    cleanup();
    return rc;
  }
  // The tables live in the graph's own plan: upload them once, before capture.
  std::vector<unsigned char> host(table_bytes(bc));
  fill_tables(bc, params, event_offsets, host.data());
  hipStream_t cs = nullptr;
  hipError_t e = hipMemcpy(plan->events.p, host.data(), static_cast<size_t>(bc.n_events) * sizeof(EventDesc),
                           hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(plan->chunks.p, host.data() + static_cast<size_t>(bc.n_events) * sizeof(EventDesc),
                  static_cast<size_t>(bc.n_chunks) * sizeof(Chunk), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    if (cs) (void)hipStreamDestroy(cs);
    cleanup();
    return fail(CLUE_ERR_RUNTIME, std::string("graph capture: ") + hipGetErrorString(e));
  }
  rc = enqueue_batch(plan, cs, params, bc);
  e = hipStreamEndCapture(cs, &bg->graph);
  (void)hipStreamDestroy(cs);
  if (rc == CLUE_OK && e != hipSuccess) rc = fail(CLUE_ERR_RUNTIME, std::string("graph capture: ") + hipGetErrorString(e));
  if (rc == CLUE_OK) {
    e = hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = fail(CLUE_ERR_RUNTIME, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  if (rc != CLUE_OK) {
    cleanup();
    return rc;
  }
  plan->has_run = true;
  plan->last_n = bc.n;
  plan->last_d = ndim;
  plan->last_T = bc.Ttot;
  plan->last_events = bc.n_events;
  *graph = reinterpret_cast<clue_batch_graph_t>(bg.release());
  return CLUE_OK;
}

clue_status clue_batch_graph_launch(clue_batch_graph_t graph, void* stream) {
  if (!graph) return fail(CLUE_ERR_INVALID_ARGUMENT, "graph is NULL");
  auto* bg = reinterpret_cast<BatchGraph*>(graph);
  CLUE_HIP(hipSetDevice(bg->plan->device));
  CLUE_HIP(hipGraphLaunch(bg->exec, static_cast<hipStream_t>(stream)));
  return CLUE_OK;
}

clue_status clue_batch_graph_destroy(clue_batch_graph_t graph) {
  if (!graph) return CLUE_OK;
  auto* bg = reinterpret_cast<BatchGraph*>(graph);
  (void)hipSetDevice(bg->plan->device);
  (void)hipDeviceSynchronize();
  if (bg->exec) (void)hipGraphExecDestroy(bg->exec);
  if (bg->graph) (void)hipGraphDestroy(bg->graph);
  bg->plan->has_run = false;
  clue_plan_destroy(bg->plan);
  delete bg;
  return CLUE_OK;
}

clue_status clue_autotune_candidates(int32_t* points_per_tile, int32_t capacity, int32_t* count) {
  if (!count) return fail(CLUE_ERR_INVALID_ARGUMENT, "count is NULL");
  constexpr int32_t k = static_cast<int32_t>(sizeof(kTunePointsPerTile) / sizeof(kTunePointsPerTile[0]));
  *count = k;
  if (points_per_tile)
    for (int32_t i = 0; i < k && i < capacity; ++i) points_per_tile[i] = kTunePointsPerTile[i];
  return CLUE_OK;
}

// Coordinate descent over the launch tuning (and, when asked, the points per
// tile) on the given batch: every candidate is run once untimed and five times
// timed with HIP events on `stream`; the fastest median wins.
clue_status clue_autotune_batched(clue_plan_t plan, void* stream, clue_params* params, int32_t ndim,
                                  int32_t n_events, const int64_t* event_offsets,
                                  const float* const* coords_dev, const float* weights_dev,
                                  const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                  const uint8_t* wrapped, int32_t flags, clue_tuning* best, float* best_ms) {
  if (!plan || !params || !best || !event_offsets || n_events <= 0)
    return fail(CLUE_ERR_INVALID_ARGUMENT, "NULL argument");
  CLUE_HIP(hipSetDevice(plan->device));
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = event_offsets[n_events];
  if (n <= 0 || n > INT32_MAX) return fail(CLUE_ERR_INVALID_ARGUMENT, "invalid event offsets");
  int32_t* labels = nullptr;
  CLUE_HIP(hipMalloc(&labels, static_cast<size_t>(n) * 4));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipError_t he = hipEventCreate(&t0);
  if (he == hipSuccess) he = hipEventCreate(&t1);
  const clue_tuning saved = plan->tuning;
  clue_params prm = *params;
  clue_status rc = he == hipSuccess ? CLUE_OK : fail(CLUE_ERR_RUNTIME, hipGetErrorString(he));
  auto run_once = [&](float* ms) -> clue_status {
    if (ms && hipEventRecord(t0, st) != hipSuccess) return fail(CLUE_ERR_RUNTIME, "event record");
    clue_status r = cluster_batch(plan, st, &prm, ndim, n_events, event_offsets, coords_dev, weights_dev, kernel,
                                  metric, wrapped, labels, nullptr, nullptr, nullptr, false);
    if (r != CLUE_OK) return r;
    if (ms) {
      if (hipEventRecord(t1, st) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
          hipEventElapsedTime(ms, t0, t1) != hipSuccess)
        return fail(CLUE_ERR_RUNTIME, "event timing");
    }
    return CLUE_OK;
  };
  auto measure = [&](float* ms) -> clue_status {
    clue_status r = run_once(nullptr);
    float v[5];
    for (int k = 0; r == CLUE_OK && k < 5; ++k) r = run_once(&v[k]);
    if (r != CLUE_OK) return r;
    std::sort(v, v + 5);
    *ms = v[2];
    return CLUE_OK;
  };
  clue_tuning cur = saved;
  float cur_ms = 0.f;
  if (rc == CLUE_OK) rc = measure(&cur_ms);
  // points per tile (changes the grid and hence the results' summation order)
  if (rc == CLUE_OK && (flags & CLUE_TUNE_POINTS_PER_TILE)) {
    const int32_t keep = prm.points_per_tile;
    int32_t best_ppb = keep;
    for (int32_t ppb : kTunePointsPerTile) {
      if (ppb == keep || rc != CLUE_OK) continue;
      prm.points_per_tile = ppb;
      float ms;
      rc = measure(&ms);
      if (rc == CLUE_OK && ms < cur_ms) {
        cur_ms = ms;
        best_ppb = ppb;
      }
    }
    prm.points_per_tile = best_ppb;
  }
  auto try_field = [&](auto setter, std::initializer_list<float> values) {
    for (float v : values) {
      if (rc != CLUE_OK) return;
      clue_tuning t = cur;
      setter(t, v);
      if (std::memcmp(&t, &cur, sizeof(t)) == 0) continue;
      plan->tuning = t;
      float ms;
      rc = measure(&ms);
      if (rc == CLUE_OK && ms < cur_ms) {
        cur_ms = ms;
        cur = t;
      }
    }
    plan->tuning = cur;
  };
  auto set_radius = [](clue_tuning& t, float v) { t.nh_first_radius = v; };
  auto set_xcd = [](clue_tuning& t, float v) { t.xcd_remap = static_cast<int32_t>(v); };
  auto set_group = [](clue_tuning& t, float v) { t.pending_group = static_cast<int32_t>(v); };
  try_field(set_radius, {0.08f, 0.12f, 0.18f, 0.25f, 0.35f, 0.5f, 0.75f, 1.0f});
  try_field([](clue_tuning& t, float v) { t.nh_lone_lanes = static_cast<int32_t>(v); }, {0.f, 2.f, 8.f});
  try_field(set_xcd, {0.f, 1.f, 2.f, 3.f, 4.f, 5.f, 6.f, 7.f});
  try_field(set_group, {8.f, 16.f, 32.f, 64.f});
  try_field([](clue_tuning& t, float v) { t.pending_blocks = static_cast<int32_t>(v); },
            {2048.f, 4096.f, 8192.f, 16384.f});
  try_field([](clue_tuning& t, float v) { t.density_order = static_cast<int32_t>(v); }, {0.f, 1.f});
  try_field([](clue_tuning& t, float v) { t.density_filter = static_cast<int32_t>(v); }, {0.f, 1.f});
  try_field([](clue_tuning& t, float v) { t.split_call_points = static_cast<int32_t>(v); }, {0.f, 1073741824.f});
  try_field(set_radius, {cur.nh_first_radius * 0.8f, cur.nh_first_radius * 1.25f});
  (void)hipStreamSynchronize(st);
  (void)hipFree(labels);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  plan->tuning = saved;
  if (rc != CLUE_OK) return rc;
  *best = cur;
  params->points_per_tile = prm.points_per_tile;
  if (best_ms) *best_ms = cur_ms;
  return CLUE_OK;
}

}  // extern "C"
