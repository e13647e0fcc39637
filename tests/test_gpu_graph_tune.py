"""Batched calls captured as HIP graphs, the autotuner, and the stream-ordered
capacity reset: each checked against direct calls and the oracle."""
import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
import oracle
from cluestering_alpaka3_amd import datasets

pytestmark = pytest.mark.gpu


def calorimeter_batch(n_events, seed, median=1500.0):
    sizes = datasets.event_sizes(n_events, seed, median=median, lo=1, hi=10000)
    evs = [datasets.calorimeter_event(int(n), seed * 1000 + k) for k, n in enumerate(sizes)]
    coords = np.ascontiguousarray(np.concatenate([e.coords for e in evs], axis=1))
    weights = np.concatenate([e.weights for e in evs])
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return coords, weights, off


class Outputs:
    def __init__(self, n):
        self.arrays = [clue.DeviceArray((n,), t) for t in (np.int32, np.int32, np.float32, np.int32)]

    @property
    def ptrs(self):
        return [a.ptr for a in self.arrays]

    def download(self, stream):
        return [a.download(stream) for a in self.arrays]


def test_graph_replays_equal_direct_calls():
    coords, weights, off = calorimeter_batch(48, 3)
    coords2, weights2, off2 = calorimeter_batch(48, 3)  # same sizes, same values
    coords2 = coords2 + np.float32(17.25)                # shifted values, same shape
    n, D = coords.shape[1], coords.shape[0]
    print(f"graph: events={off.size - 1} n={n}", flush=True)
    st = clue.Stream(0)
    dcoords = clue.DeviceArray.from_host(coords)
    dw = clue.DeviceArray.from_host(weights)
    params = clue.make_params(1.3, 2.0, 2.6)
    kd, md = clue.kernel_desc(("flat", 0.5)), clue.metric_desc("euclidean")
    cp = [dcoords.offset(d * n) for d in range(D)]
    direct, graph_out = Outputs(n), Outputs(n)
    plan = clue.Plan(0)
    g = clue.BatchGraph(params, D, off, cp, dw.ptr, kd, md, None, *graph_out.ptrs)
    try:
        plan.make_clusters_batched_device(st, params, D, off, cp, dw.ptr, kd, md, None, *direct.ptrs)
        want = direct.download(st)
        for rep in range(3):
            g.replay(st)
            got = graph_out.download(st)
            for a, b in zip(got, want):
                np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32), err_msg=f"replay {rep}")
            # eager work on the same stream between replays
            plan.make_clusters_batched_device(st, params, D, off, cp, dw.ptr, kd, md, None, *direct.ptrs)
        # new values in the same buffers
        dcoords.upload(coords2, st)
        dw.upload(weights2, st)
        g.replay(st)
        got = graph_out.download(st)
        plan.make_clusters_batched_device(st, params, D, off, cp, dw.ptr, kd, md, None, *direct.ptrs)
        want2 = direct.download(st)
        for a, b in zip(got, want2):
            np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))
    finally:
        g.close()
        plan.close()


def test_autotuner_result_and_parity():
    coords, weights, off = calorimeter_batch(24, 5)
    n, D = coords.shape[1], coords.shape[0]
    print(f"autotune: events={off.size - 1} n={n}", flush=True)
    st = clue.Stream(0)
    dcoords = clue.DeviceArray.from_host(coords)
    dw = clue.DeviceArray.from_host(weights)
    params = clue.make_params(1.3, 2.0, 2.6)
    kd, md = clue.kernel_desc(("flat", 0.5)), clue.metric_desc("euclidean")
    cp = [dcoords.offset(d * n) for d in range(D)]
    plan = clue.Plan(0)
    try:
        before = plan.get_tuning()
        tuning, ppb, ms = plan.autotune_batched(st, params, D, off, cp, dw.ptr, kd, md, tune_points_per_tile=True)
        assert plan.get_tuning() == before
        assert ppb in clue.autotune_candidates() and ms > 0
        plan.set_tuning(**tuning)
        got = clue.cluster_events(coords, weights, off, 1.3, 2.0, 2.6, ppb=ppb, plan=plan)
    finally:
        plan.close()
    p = oracle.make_params(1.3, 2.0, 2.6, ppb=ppb)
    for e in range(off.size - 1):
        a, b = int(off[e]), int(off[e + 1])
        ref = oracle.make_clusters(coords[:, a:b], weights[a:b], p, with_tiles=False)
        np.testing.assert_array_equal(got.rho[a:b].view(np.int32), ref.rho.view(np.int32))
        np.testing.assert_array_equal(got.cluster[a:b], ref.cluster)


def test_reset_capacity_is_a_fresh_clusterer():
    big = datasets.blobs(20000, 2, 60, 0.0, 200.0, 1.0, 0.0, ("const", 1.0), 99)
    small = datasets.blobs(20000, 2, 60, 0.0, 200.0, 1.0, 0.0, ("const", 1.0), 98)
    plan = clue.Plan(0)
    try:
        clue.cluster_points(big.coords, big.weights, 0.6, 1.0, 5.0, 0.6, plan=plan)
        plan.reset_capacity()
        got = clue.cluster_points(small.coords, small.weights, 0.6, 1.0, 5.0, 0.6, plan=plan)
        st = plan.stats()
    finally:
        plan.close()
    ref = oracle.make_clusters(small.coords, small.weights, oracle.make_params(0.6, 1.0, 5.0, 0.6))
    np.testing.assert_array_equal(got.cluster, ref.cluster)
    assert st["seed_capacity"] == ref.stats["seed_candidates"]
