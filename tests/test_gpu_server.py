"""Served clustering: every request of a mixed 2D/3D stream, batched with the
others, gets exactly the result it gets alone (and the oracle's)."""
import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
import oracle
from cluestering_alpaka3_amd import datasets

pytestmark = pytest.mark.gpu


def alone(ps, prm):
    return clue.cluster_points(ps.coords, ps.weights, prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"])


def test_mixed_requests_equal_alone_and_oracle():
    events = [datasets.served_event(k, 11, lo=500, hi=30000) for k in range(30)]
    print(f"server: {len(events)} events, {sum(e[0].n for e in events)} points", flush=True)
    srv = clue.Server(0, slots=2)
    try:
        tickets = [srv.submit(ps.coords, ps.weights, p["dc"], p["rhoc"], p["dm"], kernel=p["kernel"])
                   for ps, p in events]
        results = [t.result() for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["events_done"] == len(events)
    assert st["batches"] < len(events)  # the burst was batched
    assert max(t.timing["batch_events"] for t in tickets) > 1
    for k, ((ps, p), got) in enumerate(zip(events, results)):
        ref = alone(ps, p)
        np.testing.assert_array_equal(got.cluster, ref.cluster, err_msg=f"event {k}")
        np.testing.assert_array_equal(got.rho.view(np.int32), ref.rho.view(np.int32), err_msg=f"event {k}")
        np.testing.assert_array_equal(got.nearest_higher, ref.nearest_higher, err_msg=f"event {k}")
        np.testing.assert_array_equal(got.is_seed, ref.is_seed, err_msg=f"event {k}")
    for k in np.argsort([ps.n for ps, _ in events])[:4]:
        ps, p = events[k]
        o = oracle.make_clusters(ps.coords, ps.weights, oracle.make_params(p["dc"], p["rhoc"], p["dm"],
                                                                           kernel=p["kernel"]), with_tiles=False)
        np.testing.assert_array_equal(results[k].cluster, o.cluster)
        np.testing.assert_array_equal(results[k].rho.view(np.int32), o.rho.view(np.int32))


def test_requests_arriving_while_busy_join_later_batches():
    events = [datasets.served_event(k, 12, lo=2000, hi=60000) for k in range(16)]
    srv = clue.Server(0, slots=1)
    try:
        tickets = []
        outs = [np.full(ps.n, -7, np.int32) if k % 2 == 0 else None for k, (ps, _) in enumerate(events)]
        for (ps, p), out in zip(events, outs):
            tickets.append(srv.submit(ps.coords, ps.weights, p["dc"], p["rhoc"], p["dm"], kernel=p["kernel"],
                                      outputs="labels", cluster_out=out))
        res = [t.result() for t in tickets]
        with pytest.raises(clue.InvalidArgument):
            srv.submit(events[0][0].coords, events[0][0].weights, 1.0, 1.0, cluster_out=np.empty(3, np.int32))
    finally:
        srv.close()
    for (ps, p), got, out in zip(events, res, outs):
        np.testing.assert_array_equal(got.cluster, alone(ps, p).cluster)
        if out is not None:  # the caller's array received the labels
            np.testing.assert_array_equal(out, got.cluster)


def test_budget_caps_slot_capacity():
    srv = clue.Server(0, budget_bytes=96 << 20, slots=2, max_dim=3)
    try:
        cap = srv.stats()["slot_capacity_points"]
        assert 1024 <= cap < (1 << 20)
        assert srv.stats()["device_bytes"] <= 96 << 20
        big = datasets.blobs(cap + 1, 2, 4, 0.0, 100.0, 1.0, 0.1, ("const", 1.0), 5)
        with pytest.raises(clue.OutOfRange):
            srv.submit(big.coords, big.weights, 1.0, 2.0)
        small = [datasets.blobs(int(cap // 3), 2, 4, 0.0, 100.0, 1.0, 0.1, ("const", 1.0), 6 + k) for k in range(5)]
        ts = [srv.submit(q.coords, q.weights, 1.0, 2.0, 2.0) for q in small]
        for q, t in zip(small, ts):
            np.testing.assert_array_equal(t.result().cluster,
                                          clue.cluster_points(q.coords, q.weights, 1.0, 2.0, 2.0).cluster)
    finally:
        srv.close()


@pytest.mark.parametrize("slots", [1, 3])
def test_small_request_class(slots):
    # small_batch_points: small requests batch only with each other (one slot kept
    # for them); results still equal the requests alone
    events = [datasets.served_event(k, 13, lo=500, hi=80000) for k in range(24)]
    srv = clue.Server(0, slots=slots, small_batch_points=16384)
    try:
        tickets = [srv.submit(ps.coords, ps.weights, p["dc"], p["rhoc"], p["dm"], kernel=p["kernel"])
                   for ps, p in events]
        results = [t.result() for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["events_done"] == len(events)
    for k, ((ps, p), got) in enumerate(zip(events, results)):
        ref = alone(ps, p)
        np.testing.assert_array_equal(got.cluster, ref.cluster, err_msg=f"event {k}")
        np.testing.assert_array_equal(got.rho.view(np.int32), ref.rho.view(np.int32), err_msg=f"event {k}")
        np.testing.assert_array_equal(got.nearest_higher, ref.nearest_higher, err_msg=f"event {k}")
