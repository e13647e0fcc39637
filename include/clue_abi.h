/*
 * clue_abi.h -- C ABI of the MI355X CLUE engine (libclue_hip.so).
 *
 * Plain pointers and sizes only.  All device pointers are HIP device memory on
 * the plan's GPU; streams are hipStream_t passed as void*.  Every call returns a
 * clue_status; on failure clue_last_error_message() describes the error of the
 * calling thread.  The C++20 drop-in headers (CLUEstering/CLUEstering.hpp)
 * translate the status codes back into the reference's exception types
 * (std::invalid_argument, std::out_of_range, std::runtime_error).
 *
 * Which reference interface each entry point replaces is cited per function,
 * as file:line into CLUEstering 2.9.0 (alpaka3 port).
 */
#ifndef CLUE_ABI_H
#define CLUE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLUE_MAX_DIM 10

typedef enum {
  CLUE_OK = 0,
  CLUE_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  CLUE_ERR_OUT_OF_RANGE = 2,     /* std::out_of_range */
  CLUE_ERR_RUNTIME = 3,          /* std::runtime_error (HIP failure, no device) */
  CLUE_ERR_OUT_OF_MEMORY = 4,
  CLUE_ERR_UNSUPPORTED = 5
} clue_status;

/* Convolutional kernels: core/ConvolutionalKernel.hpp:14-85 (FlatKernel,
 * GaussianKernel, ExponentialKernel).  params: flat {h} | gaussian {mean, std,
 * amplitude} | exponential {avg, amplitude}.  The engine evaluates exp() as the
 * correctly rounded float exponential (double-precision table method). */
enum { CLUE_KERNEL_FLAT = 0, CLUE_KERNEL_GAUSSIAN = 1, CLUE_KERNEL_EXPONENTIAL = 2 };
typedef struct {
  int32_t type;
  float params[3];
} clue_kernel_desc;

/* Distance metrics: core/DistanceMetrics.hpp:33-230.  params hold the per-
 * dimension weights (weighted metrics) or periods (periodic Euclidean). */
enum {
  CLUE_METRIC_EUCLIDEAN = 0,
  CLUE_METRIC_WEIGHTED_EUCLIDEAN = 1,
  CLUE_METRIC_PERIODIC_EUCLIDEAN = 2,
  CLUE_METRIC_MANHATTAN = 3,
  CLUE_METRIC_CHEBYSHEV = 4,
  CLUE_METRIC_WEIGHTED_CHEBYSHEV = 5
};
typedef struct {
  int32_t type;
  float params[CLUE_MAX_DIM];
} clue_metric_desc;

/* Seed storage.  REFERENCE reproduces SeedArrayView::push_back
 * (data_structures/internal/SeedArray.hpp:33-41): the seed array holds
 * max(previous capacity, #points with no nearest-higher) entries and seeds past
 * it, in ascending point order, are dropped (their clusters become -1).
 * UNBOUNDED keeps every seed. */
enum { CLUE_SEEDS_REFERENCE = 0, CLUE_SEEDS_UNBOUNDED = 1 };

/* Clusterer parameters: core/Clusterer.hpp:40-49 (dc, seed_dc, rhoc, dm, pPBin).
 * The reference validation applies (detail/Clusterer.hpp:32-35). */
typedef struct {
  float dc;
  float rhoc;
  float dm;
  float seed_dc;
  int32_t points_per_tile;
  int32_t seed_mode;
} clue_params;

/* Per-call counters, read back with clue_plan_last_stats (synchronises). */
typedef struct {
  int64_t seed_candidates; /* points with nh == -1 after the nearest-higher stage */
  int64_t seed_capacity;   /* capacity used for this call */
  int64_t n_seeds;         /* points flagged as seeds */
  int64_t n_clusters;      /* stored seeds = clusters */
  int64_t n_dropped_seeds; /* seeds lost to the capacity (reference mode) */
} clue_stats;

/* Grid of the last call: npd tiles per dimension, npd^ndim tiles; minmax in
 * CoordinateExtremes order (min0,max0,min1,max1,...). */
typedef struct {
  int32_t ndim;
  int32_t npoints;
  int32_t npd;
  int32_t ntiles;
  float minmax[2 * CLUE_MAX_DIM];
  float tile_size[CLUE_MAX_DIM];
} clue_grid_info;

/* Launch tuning of a plan.  Every value leaves the results unchanged (they only
 * steer work order and pruning); the autotuner picks them by timing.
 *  nh_first_radius     first nearest-higher phase radius, as a fraction of dm;
 *                      0 auto (0.3 in 2D and 3D, 0.15 otherwise)
 *  morton_bits         bits per dimension of the in-tile Morton order, -1 auto
 *  xcd_remap           bit 0: density, bit 1: nearest-higher -- give each
 *                      XCD a contiguous range of the pass's query blocks, so
 *                      neighbouring queries share one L2; bit 2: the same for
 *                      the density's heaviest-first order (each XCD four
 *                      separate chunks of the Morton order, heaviest first)
 *  pending_blocks      workgroups of the second nearest-higher phase, 0 auto (8192 for calls of 2^20 points or more, else 2048)
 *  pending_group       lanes per unresolved point in that phase: 8, 16, 32, 64;
 *                      0 auto (8 in 2D and 3D, 16 otherwise)
 *  density_order       1: launch the density pass's query blocks heaviest first
 *                      (estimated from their tiles' sizes), 0: in Morton order,
 *                      -1: heaviest first in 1D and 2D only
 *  small_call_points   calls of at most this many points (all events together)
 *                      run the density pass and the tile-walk nearest-higher
 *                      search with one wave per query, which shortens the
 *                      longest query's chain when few queries fill the GPU;
 *                      0: never by size
 *  wave_dims           calls in more dimensions than this run them with one
 *                      wave per query whatever their size (lanes walking their
 *                      own boxes diverge there); 10: never by dimension
 *  tile_call_points    calls of at most this many points in 1-3 dimensions skip
 *                      the Morton layout and the block search of the
 *                      nearest-higher pass and walk the tiles instead (one wave
 *                      per query when small_call_points holds them); 0: never
 *  density_filter      1: the density pass's union walk (1-3 dimensions,
 *                      Euclidean) drops the candidates beyond the radius from
 *                      the box of the wave's queries before evaluating any;
 *                      0: every candidate of the union is evaluated
 *  fused_event_points  calls that skip the Morton layout (tile_call_points)
 *                      and whose events all have at most this many points
 *                      (and 2,048 tiles) build each event's layout, and later
 *                      its seeds and labels, in one workgroup per event (two
 *                      launches in place of about twenty); at most 32,768;
 *                      0: never
 *  split_call_points   calls of at most this many points in 1-3 dimensions
 *                      (Euclidean, with the Morton layout) share each wave's 64
 *                      queries among four waves in the density pass (with
 *                      density_filter) and the first nearest-higher phase,
 *                      which shortens the longest query's chain when the call's
 *                      waves do not fill the GPU; it takes precedence over
 *                      small_call_points for the density pass; 0: never
 *  nh_lone_lanes       first nearest-higher phase: a Morton block of the
 *                      search that at most this many lanes of a wave still
 *                      need is scanned across the wave (each needing lane's
 *                      32 candidates on half the lanes, then a minimum over
 *                      them) instead of one candidate at a time for the whole
 *                      wave; 0: never, at most 64 */
typedef struct {
  float nh_first_radius;
  int32_t morton_bits;
  int32_t xcd_remap;
  int32_t pending_blocks;
  int32_t pending_group;
  int32_t density_order;
  int32_t small_call_points;
  int32_t wave_dims;
  int32_t tile_call_points;
  int32_t density_filter;
  int32_t fused_event_points;
  int32_t split_call_points;
  int32_t nh_lone_lanes;
} clue_tuning;

typedef struct clue_plan_s* clue_plan_t;
typedef struct clue_batch_graph_s* clue_batch_graph_t;

/* clue_autotune_batched flags */
#define CLUE_TUNE_POINTS_PER_TILE 1

/* ---- errors ------------------------------------------------------------- */
const char* clue_last_error_message(void);
const char* clue_version(void);

/* ---- devices and streams: internal/alpaka/devices.hpp:108-133 (DevicePool),
 *      utils/get_queue.hpp:14-27 (get_queue), alpaka::onHost::wait ---------- */
clue_status clue_device_count(int32_t* count);
clue_status clue_set_device(int32_t device);
clue_status clue_stream_create(int32_t device, void** stream);
clue_status clue_stream_destroy(void* stream);
clue_status clue_stream_synchronize(void* stream);

/* ---- buffers: internal/alpaka/memory.hpp:34-67 (make_device_buffer /
 *      make_host_buffer), PointsDevice.hpp:19-99 (copyToHost/copyToDevice) --- */
clue_status clue_malloc_device(int32_t device, size_t bytes, void** ptr);
clue_status clue_free_device(void* ptr);
clue_status clue_malloc_host(size_t bytes, void** ptr); /* pinned */
clue_status clue_free_host(void* ptr);
clue_status clue_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
clue_status clue_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
clue_status clue_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
clue_status clue_memset_device(void* dst, int32_t value, size_t bytes, void* stream);

/* ---- persistent clustering state: the Clusterer's grow-only scratch
 *      (optional<Tiles>, optional<SeedArray>, optional<Followers>,
 *      core/Clusterer.hpp:40-49; SetupTiles.hpp:28-37, SetupSeeds.hpp:14-18).
 *      A plan is bound to one device and is not reentrant. ------------------ */
clue_status clue_plan_create(int32_t device, clue_plan_t* plan);
clue_status clue_plan_destroy(clue_plan_t plan);
/* Forget the grow-only seed capacity (a fresh Clusterer). */
clue_status clue_plan_reset_capacity(clue_plan_t plan);
/* Pre-size the scratch for up to npoints points in ndim dimensions. */
clue_status clue_plan_reserve(clue_plan_t plan, int32_t ndim, int32_t npoints);
/* Device bytes currently held by the plan. */
clue_status clue_plan_device_bytes(clue_plan_t plan, size_t* bytes);
/* Launch tuning (no reference counterpart: the reference exposes only the
 * alpaka block size, which the engine does not need). */
clue_status clue_plan_set_tuning(clue_plan_t plan, const clue_tuning* tuning);
clue_status clue_plan_get_tuning(clue_plan_t plan, clue_tuning* tuning);

/* ---- the clustering path: Clusterer::make_clusters(queue, PointsDevice&,
 *      metric, kernel, block_size), core/detail/Clusterer.hpp:84-95 and
 *      make_clusters_impl :134-217 (tile fill, K4 density, K5 nearest-higher,
 *      K6 seeds, K7 labels).
 *
 * coords_dev: host array of ndim device pointers, coordinate d of point i at
 *             coords_dev[d][i] (PointsDevice SoA, detail/PointsDevice.hpp:24-52).
 * weights_dev: n device floats.
 * wrapped:    host array of ndim flags (Clusterer::setWrappedCoordinates), or NULL.
 * Outputs (device, n entries each, original point order): cluster is required;
 * rho, nearest_higher (after seed promotion: seeds hold -1), is_seed may be NULL.
 * Enqueued on `stream`; nothing is synchronised.  Results equal the reference's
 * serial backend (labels in ascending-seed order, i.e. the serial labels). ---- */
clue_status clue_make_clusters(clue_plan_t plan, void* stream, const clue_params* params,
                               int32_t ndim, int32_t n, const float* const* coords_dev,
                               const float* weights_dev, const clue_kernel_desc* kernel,
                               const clue_metric_desc* metric, const uint8_t* wrapped,
                               int32_t* cluster_out, int32_t* is_seed_out, float* rho_out,
                               int32_t* nearest_higher_out);

/* clue_make_clusters without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same inputs,
 * outputs and results as clue_make_clusters on a fresh plan (a fresh Clusterer:
 * the seed capacity is this call's own), but every intermediate buffer lives in
 * one static 80 MiB device workspace per GPU, held from the library's load on
 * that GPU until the process ends; no hipMalloc, hipFree, host-pinned
 * allocation or event creation happens on any call, and no host copy (the event
 * table is written by a kernel).  Calls needing more than the workspace
 * (beyond 262,144 points in any dimension 1..10 at 16 points per tile or more)
 * return CLUE_ERR_OUT_OF_RANGE.  The workspace keeps the last call's contents
 * and is shared by every call on that GPU: calls in one stream are ordered,
 * calls that overlap on one GPU (other streams, other threads' streams) race
 * and must not be made.  Enqueued on `stream` on `device`; nothing is
 * synchronised. */
clue_status clue_make_clusters_noalloc(int32_t device, void* stream, const clue_params* params, int32_t ndim,
                                       int32_t n, const float* const* coords_dev, const float* weights_dev,
                                       const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                       const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                                       float* rho_out, int32_t* nearest_higher_out);

/* Batched clustering of independent events in one set of launches.  No
 * reference counterpart: the reference clusters one event per make_clusters
 * call, reusing the Clusterer's scratch (SetupTiles.hpp:28-37,
 * SetupSeeds.hpp:14-18, SetupFollowers.hpp:14-22).
 * event_offsets: host array of n_events+1 ascending offsets, event_offsets[0]
 * = 0; event e is points [event_offsets[e], event_offsets[e+1]) of the arrays
 * (each event at least one point).  Every event is clustered exactly as a
 * make_clusters call of a fresh Clusterer on that event alone: its own grid,
 * its own seed capacity, labels 0..k-1 and nearest_higher indexes local to the
 * event.  The plan's grow-only capacity is neither used nor changed.  Same
 * output arrays as clue_make_clusters, in batch order. */
clue_status clue_make_clusters_batched(clue_plan_t plan, void* stream, const clue_params* params,
                                       int32_t ndim, int32_t n_events, const int64_t* event_offsets,
                                       const float* const* coords_dev, const float* weights_dev,
                                       const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                       const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                                       float* rho_out, int32_t* nearest_higher_out);

/* A batched call captured once as a HIP graph and replayed on any stream (no
 * reference counterpart; the reference launches every kernel of every call).
 * The graph owns its scratch.  Replays read the coordinate and weight buffers
 * given here and write the given outputs: their values may change between
 * replays, the buffers and the event offsets may not.  tuning may be NULL. */
clue_status clue_batch_graph_create(int32_t device, const clue_params* params, int32_t ndim,
                                    int32_t n_events, const int64_t* event_offsets,
                                    const float* const* coords_dev, const float* weights_dev,
                                    const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                    const uint8_t* wrapped, int32_t* cluster_out, int32_t* is_seed_out,
                                    float* rho_out, int32_t* nearest_higher_out, const clue_tuning* tuning,
                                    clue_batch_graph_t* graph);
clue_status clue_batch_graph_launch(clue_batch_graph_t graph, void* stream);
clue_status clue_batch_graph_destroy(clue_batch_graph_t graph);

/* Autotuner (no reference counterpart): times batched calls on the given
 * events and returns the fastest launch tuning in *best (results do not depend
 * on it) and its median time in *best_ms.  With CLUE_TUNE_POINTS_PER_TILE in
 * flags it also picks params->points_per_tile (the reference's pPBin, which
 * does change the tile grid and so the summation order of the densities:
 * results then equal the reference's at the chosen value).  The plan's own
 * tuning is left unchanged. */
clue_status clue_autotune_batched(clue_plan_t plan, void* stream, clue_params* params, int32_t ndim,
                                  int32_t n_events, const int64_t* event_offsets,
                                  const float* const* coords_dev, const float* weights_dev,
                                  const clue_kernel_desc* kernel, const clue_metric_desc* metric,
                                  const uint8_t* wrapped, int32_t flags, clue_tuning* best, float* best_ms);

/* The points-per-tile values clue_autotune_batched tries with
 * CLUE_TUNE_POINTS_PER_TILE: *count receives their number, the first
 * min(count, capacity) go to points_per_tile (which may be NULL). */
clue_status clue_autotune_candidates(int32_t* points_per_tile, int32_t capacity, int32_t* count);

/* ---- served clustering (no reference counterpart: the reference clusters
 *      one event per synchronous make_clusters call).  A server owns `slots`
 *      batch slots (stream, plan, device and pinned buffers), sized once so
 *      that all of them fit device_budget_bytes; a scheduler thread batches
 *      pending requests of one class (dimension, parameters, kernel, metric,
 *      wrapping) into a free slot in arrival order, and requests that arrive
 *      while the slots are busy join the next batch.  Every request is
 *      clustered exactly as a fresh Clusterer would cluster it alone. ------ */
typedef struct clue_server_s* clue_server_t;

typedef struct {
  size_t device_budget_bytes; /* 0: 32 GiB */
  int32_t slots;              /* batches in flight, 0: 4 */
  int32_t max_batch_points;   /* per slot before the budget cap, 0: 8 Mi */
  int32_t max_batch_events;   /* 0: 4096 */
  int32_t max_dim;            /* largest dimension served, 0: 3 */
  int32_t small_batch_points; /* > 0: requests of at most this many points are
                                 batched only with each other, up to this many
                                 points in all (unless many wait), so that a small
                                 request never waits for a large one of its batch,
                                 and one slot is kept for them; 0: one size class */
} clue_server_config;

/* One event.  coords: host (ndim, n) SoA, weights: host n floats; outputs:
 * host arrays of n (cluster required, the others may be NULL).  The host
 * buffers must stay valid until clue_server_wait returns for the ticket. */
typedef struct {
  int32_t ndim;
  int32_t n;
  const float* coords;
  const float* weights;
  clue_params params;
  clue_kernel_desc kernel;
  clue_metric_desc metric;
  uint8_t wrapped[CLUE_MAX_DIM];
  int32_t* cluster_out;
  int32_t* is_seed_out;
  float* rho_out;
  int32_t* nearest_higher_out;
} clue_request;

typedef struct {
  int64_t submit_ns; /* steady clock */
  int64_t launch_ns;
  int64_t done_ns;
  int32_t batch_events; /* events in the request's batch */
} clue_request_timing;

typedef struct {
  int64_t events_submitted;
  int64_t events_done;
  int64_t points_done;
  int64_t batches;
  int64_t max_batch_events;
  int64_t max_batch_points;
  int64_t slot_capacity_points;
  int64_t device_bytes;
} clue_server_stats;

clue_status clue_server_create(int32_t device, const clue_server_config* config, clue_server_t* server);
/* Queue one request (non-blocking); CLUE_ERR_OUT_OF_RANGE if it exceeds a slot. */
clue_status clue_server_submit(clue_server_t server, const clue_request* request, uint64_t* ticket);
/* Block until the request is done; its outputs are then in the caller's buffers. */
clue_status clue_server_wait(clue_server_t server, uint64_t ticket, clue_request_timing* timing);
clue_status clue_server_get_stats(clue_server_t server, clue_server_stats* stats);
/* Finish every submitted request, then free the server. */
clue_status clue_server_destroy(clue_server_t server);

/* Host-buffer convenience: the reference's Python FFI entry
 * mainRun(dc, rhoc, dm, seed_dc, pPBin, wrapped, data, results, kernel, Ndim,
 * n_points, block_size, device_id), CLUEstering/BindingModules/serial/
 * binding_cpu.cpp:35-52.  data: host float[(ndim+1)*n] SoA (coords then
 * weights); results: host int32[n].  Copies in, clusters as a fresh Clusterer
 * would on `device`, copies labels out, synchronises.  The device's plan, stream
 * and buffers are kept for the process's later calls (one call at a time per
 * device; concurrent callers wait). */
clue_status clue_main_run(float dc, float rhoc, float dm, float seed_dc, int32_t points_per_tile,
                          const uint8_t* wrapped, const float* data, int32_t* results,
                          const clue_kernel_desc* kernel, int32_t ndim, int32_t n,
                          int32_t block_size, int32_t device);

/* ---- post-processing on the device: the cluster -> points map of
 *      get_clusters / make_associator (utils/detail/get_clusters.hpp:21-46,
 *      internal/MakeAssociator.hpp:13-42) with ClusterProperties' n_clusters
 *      and cluster sizes (ClusterProperties.hpp:19-51), and cluster_centroids
 *      (utils/detail/cluster_centroid.hpp:38-65).  Stream-ordered, nothing is
 *      synchronised; they use the plan's scratch, not its clustering state. */
/* n labels (negative: outlier).  offsets_dev: n+1 entries, indexes_dev: n
 * entries.  On completion counts_dev[0] = n_clusters (largest label + 1),
 * counts_dev[1] = clustered points; offsets[0..n_clusters] delimit each
 * cluster's points, listed in ascending order in indexes[0..counts[1]) (the
 * outliers follow, also ascending); cluster sizes are offset differences. */
clue_status clue_cluster_map(clue_plan_t plan, void* stream, int32_t n, const int32_t* cluster_dev,
                             int32_t* offsets_dev, int32_t* indexes_dev, int32_t* counts_dev);
/* Row c of centroids_dev (ndim floats) for c < min(max_clusters, n_clusters):
 * each coordinate of the cluster's points summed from 0.f in ascending point
 * order, divided by (float)size -- the reference's float arithmetic. */
clue_status clue_cluster_centroids(clue_plan_t plan, void* stream, int32_t ndim, const float* const* coords_dev,
                                   const int32_t* offsets_dev, const int32_t* indexes_dev, const int32_t* counts_dev,
                                   int32_t max_clusters, float* centroids_dev);

/* ---- introspection of the last call (synchronise the plan's last stream) -- */
clue_status clue_plan_last_stats(clue_plan_t plan, clue_stats* stats);
clue_status clue_plan_last_grid(clue_plan_t plan, clue_grid_info* grid);
/* Tile CSR of the last call, host buffers: offsets[ntiles+1], indexes[npoints]
 * (Tiles / AssociationMap offsets+indexes, AssociationMap.hpp:122-191). */
clue_status clue_plan_copy_tiles(clue_plan_t plan, int32_t* offsets, int32_t* indexes);
/* Nearest-higher distance of the last call in original order (FLT_MAX when the
 * point has no nearest-higher), host buffer of npoints floats. */
clue_status clue_plan_copy_delta(clue_plan_t plan, float* delta);

/* ---- per-stage timing (HIP events recorded on the call's stream; no
 *      reference counterpart -- the reference has no built-in profiling). ---- */
clue_status clue_plan_set_timing(clue_plan_t plan, int32_t enable);
/* Milliseconds per stage of the last timed call; *count receives the number of
 * stages (ms may be NULL to query it).  Synchronises on the last event. */
clue_status clue_plan_stage_times(clue_plan_t plan, float* ms, int32_t capacity, int32_t* count);
const char* clue_stage_name(int32_t stage);

/* ---- diagnostics used by the parity tests: the engine's device arithmetic
 *      on device arrays (op 0 expf as the kernels use it, 1 sqrtf, 2 division,
 *      3 sqrt bound, 4 quotient through a double reciprocal, 5 expf of
 *      non-positive arguments, 6 / 7 the largest / smallest float whose sqrtf
 *      equals sqrtf(a), 8 expf of arguments in [-14, 0] as the density pass
 *      evaluates it for unit-amplitude kernels with the 64-entry table, 9 the
 *      same with the wide table the density pass reads, arguments clamped to
 *      [-14, 0], 10 the nearest-higher searches' sqrtf of finite a >= 0, 11
 *      their bound below which a squared distance has a smaller sqrtf than
 *      sqrtf(a), finite a >= 0). --------------------------------------------------- */
clue_status clue_selftest_math(int32_t op, const float* a_dev, const float* b_dev, float* out_dev,
                               int64_t n, void* stream);
/* Number of floats with bit patterns in [first_bits, first_bits + count) on
 * which the kernels' fast correctly rounded sqrt differs from IEEE sqrtf
 * (synchronous). */
clue_status clue_selftest_sqrt(uint32_t first_bits, uint64_t count, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif
#endif
