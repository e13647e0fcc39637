"""Batched clustering (clue_make_clusters_batched): every event of a batch gets
exactly the result a fresh Clusterer gives on that event alone -- checked
against the oracle event by event, bit for bit, and against plain calls."""
import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
import oracle
from cluestering_alpaka3_amd import datasets

pytestmark = pytest.mark.gpu


def make_batch(sizes, seed):
    evs = [datasets.calorimeter_event(int(n), seed + k) for k, n in enumerate(sizes)]
    coords = np.concatenate([e.coords for e in evs], axis=1)
    weights = np.concatenate([e.weights for e in evs])
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return coords, weights, off


def check_against_oracle(coords, weights, off, dc, rhoc, dm=None, seed_dc=None, ppb=128, kernel=("flat", 0.5),
                         metric="euclidean", metric_params=None, wrapped=None, seed_mode=0):
    print(f"batch: events={len(off) - 1} n={coords.shape[1]} ndim={coords.shape[0]} dc={dc} rhoc={rhoc} "
          f"dm={dm} kernel={kernel} metric={metric} wrapped={wrapped}", flush=True)
    got = clue.cluster_events(coords, weights, off, dc, rhoc, dm, seed_dc, ppb, kernel, metric, metric_params,
                              wrapped, seed_mode)
    p = oracle.make_params(dc, rhoc, dm, seed_dc, ppb, kernel, metric, metric_params, wrapped,
                           seed_mode=seed_mode)
    for e in range(len(off) - 1):
        a, b = int(off[e]), int(off[e + 1])
        ref = oracle.make_clusters(coords[:, a:b], weights[a:b], p, with_tiles=False)
        ctx = f"event {e} ({b - a} points)"
        np.testing.assert_array_equal(got.rho[a:b].view(np.int32), ref.rho.view(np.int32), err_msg=ctx)
        np.testing.assert_array_equal(got.nearest_higher[a:b], ref.nh, err_msg=ctx)
        np.testing.assert_array_equal(got.is_seed[a:b], ref.is_seed, err_msg=ctx)
        np.testing.assert_array_equal(got.cluster[a:b], ref.cluster, err_msg=ctx)
    return got


def test_batched_calorimeter_events():
    sizes = datasets.event_sizes(40, 7, median=1500.0, lo=1, hi=12000)
    coords, weights, off = make_batch(sizes, 100)
    check_against_oracle(coords, weights, off, 4.0, 2.5, 4.0)


def test_batched_tiny_and_degenerate_events():
    # one-point events, identical points (zero tile size), two points, between regular ones
    parts = [np.zeros((2, 1), np.float32), np.full((2, 300), 7.5, np.float32),
             np.array([[0.0, 0.5], [0.0, 0.0]], np.float32)]
    ev = datasets.calorimeter_event(3000, 5)
    parts = [ev.coords, parts[0], parts[1], ev.coords[:, :777], parts[2], parts[0]]
    coords = np.ascontiguousarray(np.concatenate(parts, axis=1))
    weights = np.ones(coords.shape[1], np.float32)
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([q.shape[1] for q in parts])
    check_against_oracle(coords, weights, off, 4.0, 0.5, 4.0)


@pytest.mark.parametrize("kernel", [("gaussian", 0.0, 1.0, 1.0), ("exponential", 1.0, 1.0)])
@pytest.mark.parametrize("metric,mp", [("euclidean", None), ("manhattan", None),
                                       ("weighted_euclidean", [0.5, 2.0])])
def test_batched_kernels_metrics(kernel, metric, mp):
    sizes = datasets.event_sizes(12, 11, median=1200.0, lo=1, hi=6000)
    coords, weights, off = make_batch(sizes, 300)
    check_against_oracle(coords, weights, off, 4.0, 2.5, 4.0, kernel=kernel, metric=metric, metric_params=mp)


def test_batched_wrapped_and_3d():
    sizes = datasets.event_sizes(10, 13, median=1000.0, lo=1, hi=5000)
    coords, weights, off = make_batch(sizes, 500)
    check_against_oracle(coords, weights, off, 4.0, 2.5, 4.0, wrapped=[1, 0])
    parts = [datasets.blobs(int(n), 3, 6, 0.0, 60.0, 1.5, 0.1, ("uniform", 0.5, 1.5), 40 + k)
             for k, n in enumerate([2000, 1, 4500, 50, 3000])]
    coords3 = np.ascontiguousarray(np.concatenate([q.coords for q in parts], axis=1))
    w3 = np.concatenate([q.weights for q in parts])
    off3 = np.zeros(len(parts) + 1, np.int64)
    off3[1:] = np.cumsum([q.n for q in parts])
    check_against_oracle(coords3, w3, off3, 1.5, 3.0, 2.5)


def test_batched_seed_capacity_per_event():
    # dm > seed_dc: seeds may outnumber the points without a nearest-higher, and
    # each event truncates against its own fresh capacity
    parts = [datasets.blobs(n, 2, 30, 0.0, 150.0, 1.0, 0.0, ("const", 1.0), 90 + k)
             for k, n in enumerate([8000, 3000, 12000])]
    coords = np.ascontiguousarray(np.concatenate([q.coords for q in parts], axis=1))
    weights = np.concatenate([q.weights for q in parts])
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([q.n for q in parts])
    check_against_oracle(coords, weights, off, 0.6, 1.0, 5.0, seed_dc=0.6)
    check_against_oracle(coords, weights, off, 0.6, 1.0, 5.0, seed_dc=0.6, seed_mode=1)


def test_batched_equals_plain_calls_and_leaves_capacity():
    sizes = datasets.event_sizes(16, 21, median=2000.0, lo=1, hi=9000)
    coords, weights, off = make_batch(sizes, 700)
    plan = clue.Plan(0)
    try:
        # the plan's grow-only capacity from a plain call survives a batched call
        first = clue.cluster_points(coords[:, : int(off[1])], weights[: int(off[1])], 4.0, 2.5, 4.0, plan=plan)
        cap = plan.stats()["seed_capacity"]
        got = clue.cluster_events(coords, weights, off, 4.0, 2.5, 4.0, plan=plan)
        again = clue.cluster_points(coords[:, : int(off[1])], weights[: int(off[1])], 4.0, 2.5, 4.0, plan=plan)
        assert plan.stats()["seed_capacity"] == cap
        np.testing.assert_array_equal(again.cluster, first.cluster)
    finally:
        plan.close()
    for e in range(len(sizes)):
        a, b = int(off[e]), int(off[e + 1])
        alone = clue.cluster_points(coords[:, a:b], weights[a:b], 4.0, 2.5, 4.0)
        np.testing.assert_array_equal(got.cluster[a:b], alone.cluster, err_msg=f"event {e}")
        np.testing.assert_array_equal(got.nearest_higher[a:b], alone.nearest_higher, err_msg=f"event {e}")
        np.testing.assert_array_equal(got.rho[a:b].view(np.int32), alone.rho.view(np.int32), err_msg=f"event {e}")


def test_batched_invalid_arguments():
    c = np.zeros((2, 10), np.float32)
    w = np.ones(10, np.float32)
    with pytest.raises(clue.InvalidArgument):
        clue.cluster_events(c, w, [0, 5, 5, 10], 1.0, 1.0)   # empty event
    with pytest.raises(clue.InvalidArgument):
        clue.cluster_events(c, w, [1, 10], 1.0, 1.0)         # does not start at 0
