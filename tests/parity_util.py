"""Shared helpers of the GPU parity tests: run the engine through the C ABI and
the oracle on the same inputs, and compare every output."""
import os

import numpy as np

import cluestering_alpaka3_amd as clue
import oracle

# host threads for the oracle's parallel loops: the process's CPU share, at most 16
ORACLE_THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def run_engine(coords, weights, dc, rhoc, dm=None, seed_dc=None, ppb=128, kernel=("flat", 0.5),
               metric="euclidean", metric_params=None, wrapped=None, seed_mode=0, plan=None, tuning=None):
    """Engine outputs plus the plan's introspection (tile CSR, grid, counters, delta)."""
    print(f"case: ndim={np.atleast_2d(coords).shape[0]} n={np.atleast_2d(coords).shape[1]} dc={dc} "
          f"rhoc={rhoc} dm={dm} ppb={ppb} kernel={kernel} metric={metric} wrapped={wrapped} tuning={tuning}",
          flush=True)
    own = plan is None
    plan = plan or clue.Plan(0)
    if tuning:
        plan.set_tuning(**tuning)
    try:
        got = clue.cluster_points(coords, weights, dc, rhoc, dm, seed_dc, ppb, kernel, metric, metric_params,
                                  wrapped, seed_mode, plan=plan)
        tiles = plan.tiles()
        grid = plan.grid()
        stats = plan.stats()
        delta = plan.delta()
    finally:
        if own:
            plan.close()
    return got, tiles, grid, stats, delta


def run_both(coords, weights, dc, rhoc, dm=None, seed_dc=None, ppb=128, kernel=("flat", 0.5),
             metric="euclidean", metric_params=None, wrapped=None, seed_mode=0, plan=None, tuning=None,
             parallel=False):
    got, tiles, grid, stats, delta = run_engine(coords, weights, dc, rhoc, dm, seed_dc, ppb, kernel, metric,
                                                metric_params, wrapped, seed_mode, plan, tuning)
    p = oracle.make_params(dc, rhoc, dm, seed_dc, ppb, kernel, metric, metric_params, wrapped,
                           seed_mode=seed_mode)
    if parallel:
        ref = oracle.make_clusters_par(coords, weights, p, ORACLE_THREADS)
    else:
        ref = oracle.make_clusters(coords, weights, p)
    return got, ref, tiles, grid, stats, delta


def assert_same(got, ref, tiles, grid, stats, delta):
    """Bit-exact equality of every output: grid, tile CSR, densities, nearest-higher
    distances and ids, seed flags, labels and the seed counters."""
    assert grid["npd"] == ref.grid["npd"] and grid["ntiles"] == ref.grid["ntiles"]
    np.testing.assert_array_equal(grid["minmax"], ref.grid["minmax"])
    np.testing.assert_array_equal(grid["tile_size"], ref.grid["tile_size"])
    np.testing.assert_array_equal(tiles[0], ref.tile_offsets)
    np.testing.assert_array_equal(tiles[1], ref.tile_indexes)
    np.testing.assert_array_equal(got.rho.view(np.int32), ref.rho.view(np.int32))
    np.testing.assert_array_equal(delta.view(np.int32), ref.delta.view(np.int32))
    np.testing.assert_array_equal(got.nearest_higher, ref.nh)
    np.testing.assert_array_equal(got.is_seed, ref.is_seed)
    np.testing.assert_array_equal(got.cluster, ref.cluster)
    assert stats["seed_candidates"] == ref.stats["seed_candidates"]
    assert stats["n_seeds"] == ref.stats["n_seeds"]
    assert stats["n_clusters"] == ref.stats["n_stored_seeds"]
    assert stats["n_dropped_seeds"] == ref.stats["n_dropped_seeds"]
