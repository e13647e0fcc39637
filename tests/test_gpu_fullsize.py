"""Parity at the workloads' full sizes: the 2D 1,048,576-point Gaussian, the 3D
4,194,304-point exponential and the 10D 262,144-point flat workload, every
output of the engine (grid, tile CSR, densities, nearest-higher distances and
ids, seed flags, labels, seed counters) against the oracle bit for bit; the
same results against the oracle's libm-exp mode (the host expf the reference's
serial backend calls, core/detail/ConvolutionalKernel.hpp:43,63) with labels,
nearest-highers and seeds bit-exact and densities within 1e-6; and the margins
(points within 1e-6 of the rhoc test or of a nearest-higher tie, the FP64
density error) counted and reported.  Then the 4096-event calorimeter batch
(autotuned, as the benchmark runs it) event by event against the oracle.

The oracle runs its loops on the process's CPU share (at most 16 threads); each
workload's oracle results are computed once per session."""
import concurrent.futures
import json
import os

import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
import oracle
from cluestering_alpaka3_amd import datasets
from parity_util import ORACLE_THREADS, assert_same, run_engine

pytestmark = pytest.mark.gpu

FULL = ["2d_1m_gauss", "3d_4m_exp", "10d_256k_flat"]
BAND = 1e-6  # relative band of the densities
_cache = {}


def workload_results(name):
    """(points, engine outputs, oracle ENGINE-mode result, margins), computed once."""
    if name not in _cache:
        ps = datasets.workload(name)
        prm = datasets.WORKLOADS[name]
        print(f"full-size: {name} n={ps.n} ndim={ps.ndim} sha256={ps.sha256()[:16]}", flush=True)
        eng = run_engine(ps.coords, ps.weights, prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"])
        p = oracle.make_params(prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"])
        print(f"full-size: {name} engine done; oracle on {ORACLE_THREADS} threads", flush=True)
        ref = oracle.make_clusters_par(ps.coords, ps.weights, p, ORACLE_THREADS)
        print(f"full-size: {name} oracle done; margins", flush=True)
        margins = oracle.parity_margins(ps.coords, ps.weights, p, ref.rho, ORACLE_THREADS, BAND)
        _cache[name] = (ps, prm, eng, ref, margins)
    return _cache[name]


@pytest.mark.parametrize("name", FULL)
def test_fullsize_parity(name):
    ps, prm, eng, ref, margins = workload_results(name)
    got, tiles, grid, stats, delta = eng
    assert_same(got, ref, tiles, grid, stats, delta)
    # the DFS depth the reference's per-seed walk reaches (its stack holds 256)
    assert ref.stats["max_dfs_stack"] < 256
    print("margins " + json.dumps({"workload": name, "n": ps.n, "max_dfs_stack": ref.stats["max_dfs_stack"],
                                   **margins}), flush=True)


@pytest.mark.parametrize("name", [n for n in FULL if datasets.WORKLOADS[n]["kernel"][0] != "flat"])
def test_fullsize_against_libm_exp(name):
    ps, prm, eng, ref, margins = workload_results(name)
    got = eng[0]
    p = oracle.make_params(prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"], exp_mode=oracle.EXP_LIBM)
    print(f"full-size: {name} libm-exp oracle", flush=True)
    lib = oracle.make_clusters_par(ps.coords, ps.weights, p, ORACLE_THREADS, with_tiles=False)
    rel = np.abs(got.rho.astype(np.float64) - lib.rho) / np.abs(lib.rho.astype(np.float64))
    differing = int(np.count_nonzero(got.rho != lib.rho))
    print("libm " + json.dumps({"workload": name, "densities_differing": differing,
                                "max_rel_diff": float(rel.max()),
                                "rhoc_band": margins["rhoc_band"],
                                "nh_density_near_ties": margins["nh_density_near_ties"]}), flush=True)
    assert rel.max() <= BAND
    np.testing.assert_array_equal(got.nearest_higher, lib.nh)
    np.testing.assert_array_equal(got.is_seed, lib.is_seed)
    np.testing.assert_array_equal(got.cluster, lib.cluster)


def calorimeter_batch(n_events, seed):
    """The benchmark's batch: sizes lognormal (median 5000) clipped to [500, 50000]."""
    sizes = datasets.event_sizes(n_events, seed)
    evs = [datasets.calorimeter_event(int(n), 7_000_000 + k) for k, n in enumerate(sizes)]
    coords = np.ascontiguousarray(np.concatenate([e.coords for e in evs], axis=1))
    weights = np.ascontiguousarray(np.concatenate([e.weights for e in evs]))
    off = np.zeros(sizes.size + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return coords, weights, off


def check_events(coords, weights, off, got, ppb):
    """Every event against a fresh oracle run, on the oracle thread pool (ctypes
    releases the GIL during each call)."""
    p = oracle.make_params(1.3, 2.0, 2.6, ppb=ppb)

    def one(e):
        a, b = int(off[e]), int(off[e + 1])
        ref = oracle.make_clusters(coords[:, a:b], weights[a:b], p, with_tiles=False)
        ok = (np.array_equal(got.rho[a:b].view(np.int32), ref.rho.view(np.int32)) and
              np.array_equal(got.nearest_higher[a:b], ref.nh) and
              np.array_equal(got.is_seed[a:b], ref.is_seed) and
              np.array_equal(got.cluster[a:b], ref.cluster))
        return e, ok

    with concurrent.futures.ThreadPoolExecutor(ORACLE_THREADS) as ex:
        bad = [e for e, ok in ex.map(one, range(off.size - 1)) if not ok]
    assert not bad, f"{len(bad)} events differ from the oracle, first {bad[:10]}"


def test_batched_4096_events_autotuned():
    coords, weights, off = calorimeter_batch(4096, 2024)
    n, D = coords.shape[1], coords.shape[0]
    print(f"batch: 4096 events n={n}", flush=True)
    st = clue.Stream(0)
    dcoords = clue.DeviceArray.from_host(coords)
    dw = clue.DeviceArray.from_host(weights)
    cp = [dcoords.offset(d * n) for d in range(D)]
    kd, md = clue.kernel_desc(("flat", 0.5)), clue.metric_desc("euclidean")
    params = clue.make_params(1.3, 2.0, 2.6)
    plan = clue.Plan(0)
    try:
        k = max(1, min(int(np.searchsorted(off, 4_000_000)), off.size - 1))
        tuning, ppb, ms = plan.autotune_batched(st, params, D, off[: k + 1], cp, dw.ptr, kd, md,
                                                tune_points_per_tile=True)
        print(f"autotuned: ppb={ppb} tuning={tuning} {ms:.3f} ms", flush=True)
        assert ppb in clue.autotune_candidates()
        plan.set_tuning(**tuning)
        got = clue.cluster_events(coords, weights, off, 1.3, 2.0, 2.6, ppb=ppb, plan=plan)
    finally:
        plan.close()
        dcoords.free()
        dw.free()
    check_events(coords, weights, off, got, ppb)


@pytest.mark.parametrize("ppb", [16, 128])
def test_batched_calorimeter_fixed_points_per_tile(ppb):
    # the benchmark's event generator and parameters at fixed points per tile
    # (16 is what the autotuner picks on this batch), untuned and tuned launches
    coords, weights, off = calorimeter_batch(256, 99)
    print(f"batch: 256 events n={coords.shape[1]} ppb={ppb}", flush=True)
    for tuning in (None, dict(nh_first_radius=0.5, xcd_remap=3, pending_group=8, density_order=0)):
        plan = clue.Plan(0)
        try:
            if tuning:
                plan.set_tuning(**tuning)
            got = clue.cluster_events(coords, weights, off, 1.3, 2.0, 2.6, ppb=ppb, plan=plan)
        finally:
            plan.close()
        check_events(coords, weights, off, got, ppb)
