/*
 * clue_oracle.h -- CPU restatement of CLUEstering's serial backend.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libclue_hip.so and the headers under include/) never links it.
 *
 * Every function restates the reference algorithm of CLUEstering 2.9.0
 * (alpaka3 port); citations are file:line into that tree.  The arithmetic is
 * IEEE-strict: the library is compiled with -ffp-contract=off and without any
 * fast-math flag, so each float operation below is one correctly rounded
 * IEEE-754 operation, in the order the reference writes it.
 */
#ifndef CLUE_ORACLE_H
#define CLUE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_DIM 10

/* Convolutional kernels, core/ConvolutionalKernel.hpp:14-85 */
enum { ORACLE_KERNEL_FLAT = 0, ORACLE_KERNEL_GAUSSIAN = 1, ORACLE_KERNEL_EXPONENTIAL = 2 };
/* Distance metrics, core/DistanceMetrics.hpp:33-230 */
enum {
  ORACLE_METRIC_EUCLIDEAN = 0,
  ORACLE_METRIC_WEIGHTED_EUCLIDEAN = 1,
  ORACLE_METRIC_PERIODIC_EUCLIDEAN = 2,
  ORACLE_METRIC_MANHATTAN = 3,
  ORACLE_METRIC_CHEBYSHEV = 4,
  ORACLE_METRIC_WEIGHTED_CHEBYSHEV = 5
};
/* Which exp() the Gaussian/Exponential kernels use.
 *  ENGINE: the engine's documented expf (double-precision table method,
 *          correctly rounded for all but a vanishing set of inputs); the GPU
 *          engine computes bit-identical values.
 *  LIBM:   the host C library's expf, which is what the reference's serial
 *          backend calls through alpaka::math::exp on the CPU. */
enum { ORACLE_EXP_ENGINE = 0, ORACLE_EXP_LIBM = 1 };
/* Seed storage: REFERENCE truncates at the seed-candidate capacity exactly as
 * SeedArrayView::push_back does (SeedArray.hpp:33-41); UNBOUNDED stores all. */
enum { ORACLE_SEEDS_REFERENCE = 0, ORACLE_SEEDS_UNBOUNDED = 1 };

typedef struct {
  float dc, rhoc, dm, seed_dc;
  int32_t points_per_tile;
  int32_t kernel_type;
  float kernel_params[3];     /* flat: h | gauss: mean, std, amplitude | exp: avg, amplitude */
  int32_t metric_type;
  float metric_params[ORACLE_MAX_DIM]; /* weights or periods, per dimension */
  uint8_t wrapped[ORACLE_MAX_DIM];
  int32_t exp_mode;
  int32_t seed_mode;
  int64_t seed_capacity_floor; /* grow-only capacity carried over from earlier calls */
} oracle_params;

typedef struct {
  int32_t npd, ntiles;
  float minmax[2 * ORACLE_MAX_DIM]; /* CoordinateExtremes layout: min0,max0,min1,max1,... */
  float tile_size[ORACLE_MAX_DIM];
} oracle_grid;

typedef struct {
  int64_t seed_candidates; /* #points with nh == -1 after the nearest-higher stage */
  int64_t seed_capacity;   /* capacity actually used for the seed array */
  int64_t n_seeds;         /* points flagged is_seed */
  int64_t n_stored_seeds;  /* seeds that fit the capacity = number of clusters */
  int64_t n_dropped_seeds; /* seeds lost to the capacity */
  int64_t max_dfs_stack;   /* deepest DFS stack the reference's per-seed walk reaches */
  int64_t n_clustered;     /* points with cluster != -1 */
  int64_t candidate_pairs_density;
  int64_t candidate_pairs_nh;
  int64_t inradius_pairs_density; /* pairs with d <= dc (self included) */
  int64_t inradius_pairs_nh;      /* pairs with d <= dm (self included) */
} oracle_stats;

/* Grid sizing: SetupTiles.hpp:16-49 and ComputeTiles.hpp:13-27. */
int oracle_grid_setup(int32_t ndim, int32_t n, const float* coords, int32_t points_per_tile,
                      oracle_grid* grid);

/* Tile CSR (Tiles::fill -> HostAssociationMap-style counting sort, serial
 * order = ascending point id inside each tile).  offsets: ntiles+1, indexes: n. */
int oracle_tiles(int32_t ndim, int32_t n, const float* coords, const oracle_params* p,
                 const oracle_grid* grid, int32_t* offsets, int32_t* indexes);

/* Full serial pipeline (Clusterer::make_clusters_impl, core/detail/Clusterer.hpp:134-171).
 * coords: ndim x n SoA (coordinate d of point i at coords[d*n+i]).
 * Every output pointer may be NULL except cluster. delta (optional) receives the
 * nearest-higher distance (FLT_MAX when nh == -1). */
int oracle_make_clusters(int32_t ndim, int32_t n, const float* coords, const float* weights,
                         const oracle_params* p, float* rho, int32_t* nh, float* delta,
                         int32_t* is_seed, int32_t* cluster, oracle_stats* stats,
                         int32_t* tile_offsets, int32_t* tile_indexes, oracle_grid* grid_out);

/* Same semantics, density/nearest-higher/assignment parallelised with OpenMP
 * (the reference's CpuOmpBlocks executor, internal/alpaka/devices.hpp:17-19).
 * Results are identical to the serial call.  nthreads <= 0 uses the default. */
int oracle_make_clusters_omp(int32_t ndim, int32_t n, const float* coords,
                             const float* weights, const oracle_params* p, int32_t nthreads,
                             float* rho, int32_t* nh, int32_t* is_seed, int32_t* cluster,
                             oracle_stats* stats);

/* All outputs of oracle_make_clusters (delta, tile CSR and grid included), the
 * density / nearest-higher / assignment loops on nthreads OpenMP threads (<= 0:
 * the default).  Results are identical to the serial call. */
int oracle_make_clusters_par(int32_t ndim, int32_t n, const float* coords, const float* weights,
                             const oracle_params* p, int32_t nthreads, float* rho, int32_t* nh, float* delta,
                             int32_t* is_seed, int32_t* cluster, oracle_stats* stats, int32_t* tile_offsets,
                             int32_t* tile_indexes, oracle_grid* grid_out);

/* Margins of a clustering result: the points whose decisions lie within a
 * relative band (1e-6 in the tests) of a flip.  rho is the pipeline's float
 * density (the oracle's or the engine's, which are bit-identical).
 *  rhoc_band                 |rho_i - rhoc| <= band * rhoc (the seed test rho >= rhoc,
 *                            ClusteringKernels.hpp:233)
 *  nh_distance_ties          the two nearest higher candidates within dm at equal
 *                            float distance: the visit order decides (:139)
 *  nh_distance_near_ties     their distances within band (relative), ties included
 *  nh_density_near_ties      a candidate j != i at distance <= the chosen one's (dm
 *                            when none) with |rho_j - rho_i| <= band * max(|rho_i|,
 *                            |rho_j|): the "higher" predicate (:132-133) could flip
 *  density_max_rel_err_fp64  max_i |rho_i - rho64_i| / |rho64_i|, rho64 the same sum
 *                            (membership by the float distance <= dc) with distance,
 *                            kernel and accumulation in double
 *  density_above_band_fp64   points whose relative error exceeds band
 *  dc_boundary_pairs         candidate pairs whose float and double distances fall
 *                            on different sides of dc
 * Euclidean metric only (returns -2 otherwise). */
typedef struct {
  int64_t rhoc_band;
  int64_t nh_distance_ties;
  int64_t nh_distance_near_ties;
  int64_t nh_density_near_ties;
  double density_max_rel_err_fp64;
  int64_t density_above_band_fp64;
  int64_t dc_boundary_pairs;
} oracle_margins;
int oracle_parity_margins(int32_t ndim, int32_t n, const float* coords, const float* weights,
                          const oracle_params* p, const float* rho, int32_t nthreads, double band,
                          oracle_margins* out);

/* Brute-force checkers for tiny inputs (no tiles):
 * FP64 density rho_i = sum_{j: d_ij <= dc} w_j * k_ij (kernel and distance in double). */
int oracle_bruteforce_density_fp64(int32_t ndim, int32_t n, const float* coords,
                                   const float* weights, const oracle_params* p, double* rho);
/* O(N^2) nearest-higher with the serial predicate; candidates restricted to the
 * same search box the tiles would give, ties broken by (distance, tile-visit slot).
 * Needs the FP32 rho of the pipeline as input. */
int oracle_bruteforce_nh(int32_t ndim, int32_t n, const float* coords, const float* rho,
                         const oracle_params* p, int32_t* nh, float* delta);

/* The engine exponential (same definition the GPU engine computes). */
float oracle_expf_engine(float x);
/* The same value evaluated in long double and rounded once to float; used to
 * check that the engine exponential is correctly rounded. */
float oracle_expf_longdouble(float x);
void oracle_expf_engine_batch(const float* x, float* y, int64_t n);
void oracle_expf_libm_batch(const float* x, float* y, int64_t n);
/* Distance between two points with the reference operation order (right fold). */
float oracle_distance(int32_t ndim, const float* a, const float* b, const oracle_params* p);

/* CSV reader with std::stof semantics (utils/detail/read_csv.hpp:17-45):
 * header line skipped; per line ndim coordinates then the weight.
 * Returns the number of points, or -1 on error.  If coords is NULL only counts. */
int64_t oracle_read_csv(const char* path, int32_t ndim, float* coords, float* weights,
                        int64_t capacity);

/* Post-processing: cluster -> points map (returns n_clusters; offsets n+1,
 * indexes n) and centroids (n_clusters rows of ndim) with the reference's
 * serial arithmetic. */
int32_t oracle_cluster_map(int32_t n, const int32_t* cluster, int32_t* offsets, int32_t* indexes,
                           int32_t* clustered);
void oracle_cluster_centroids(int32_t ndim, int32_t n, const float* coords, const int32_t* cluster,
                              int32_t n_clusters, const int32_t* offsets, float* centroids);

/* Plant a known mistake (tests only; 0 = none).  See clue_oracle.cpp. */
void oracle_set_mutation(int32_t mutation);

#ifdef __cplusplus
}
#endif
#endif
