"""Post-processing (cluster -> points map, cluster sizes, centroids): the oracle's
restatement pinned by independent numpy computations (CPU), and the device
implementation against the oracle (GPU)."""
import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
import oracle
from cluestering_alpaka3_amd import datasets


def numpy_map(labels):
    lab = np.asarray(labels)
    ok = lab >= 0
    nc = int(lab.max()) + 1 if ok.any() else 0
    sizes = np.bincount(lab[ok], minlength=nc) if nc else np.zeros(0, np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    key = np.where(ok, lab, np.iinfo(np.int32).max)
    indexes = np.argsort(key, kind="stable")[: int(ok.sum())].astype(np.int32)
    return offsets, indexes


def numpy_centroids_sequential(coords, labels):
    """Sequential float32 sums (np.cumsum accumulates left to right) / float32(size)."""
    off, _ = numpy_map(labels)
    nc = off.size - 1
    out = np.zeros((nc, coords.shape[0]), np.float32)
    for c in range(nc):
        sel = labels == c
        for d in range(coords.shape[0]):
            x = coords[d][sel].astype(np.float32)
            s = np.cumsum(x, dtype=np.float32)[-1] if x.size else np.float32(0)
            out[c, d] = np.float32(s) / np.float32(x.size)
    return out


def labelled_points(n=20000, seed=3):
    ps = datasets.blobs(n, 2, 12, 0.0, 300.0, ("uniform", 1.0, 6.0), 0.1, ("uniform", 0.5, 1.5), seed)
    rng = np.random.default_rng(seed)
    labels = rng.integers(-1, 12, n).astype(np.int32)
    return ps.coords, labels


def test_oracle_map_matches_numpy():
    rng = np.random.default_rng(0)
    for trial in range(50):
        n = int(rng.integers(1, 400))
        # CLUE labels are ranks of seeds, so they lie below the number of points
        labels = rng.integers(-1, min(n, max(1, int(rng.integers(1, 40)))), n).astype(np.int32)
        if trial % 7 == 0:
            labels[:] = -1
        off, idx = oracle.cluster_map(labels)
        want_off, want_idx = numpy_map(labels)
        np.testing.assert_array_equal(off, want_off)
        np.testing.assert_array_equal(idx, want_idx)


def test_oracle_centroids_sequential_order():
    coords, labels = labelled_points()
    got = oracle.cluster_centroids(coords, labels)
    np.testing.assert_array_equal(got.view(np.int32), numpy_centroids_sequential(coords, labels).view(np.int32))
    # close to the float64 means, and sensitive to the summation order: a pairwise
    # float32 sum gives different bits for some cluster
    off, _ = numpy_map(labels)
    f64 = np.array([[coords[d][labels == c].astype(np.float64).mean() for d in range(2)]
                    for c in range(off.size - 1)])
    np.testing.assert_allclose(got, f64, rtol=1e-5)
    pairwise = np.array([[np.float32(coords[d][labels == c].sum(dtype=np.float32)) / np.float32((labels == c).sum())
                          for d in range(2)] for c in range(off.size - 1)], np.float32)
    assert np.any(pairwise.view(np.int32) != got.view(np.int32))


@pytest.mark.gpu
def test_device_map_and_centroids_match_oracle(ref_data):
    cases = []
    ps = datasets.workload("2d_1m_gauss", n=200000, seed=5)
    prm = datasets.WORKLOADS["2d_1m_gauss"]
    res = clue.cluster_points(ps.coords, ps.weights, prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"])
    cases.append((ps.coords, res.cluster))
    c, w = ref_data("data_32768")
    cases.append((c, clue.cluster_points(c, w, 1.3, 10.0, 1.3).cluster))
    cases.append(labelled_points())
    cases.append((np.zeros((3, 50), np.float32), -np.ones(50, np.int32)))           # no cluster
    cases.append((np.arange(64, dtype=np.float32)[None, :], np.zeros(64, np.int32)))  # one cluster
    gaps = np.full(100, -1, np.int32)
    gaps[::7] = 3
    gaps[1::11] = 9                                                                  # empty labels between
    cases.append((np.linspace(0, 1, 300, dtype=np.float32).reshape(3, 100), gaps))
    for k, (coords, labels) in enumerate(cases):
        print(f"post-processing case {k}: ndim={coords.shape[0]} n={labels.size}", flush=True)
        off, idx = clue.cluster_map(labels)
        want_off, want_idx = oracle.cluster_map(labels)
        np.testing.assert_array_equal(off, want_off)
        np.testing.assert_array_equal(idx, want_idx)
        cen = clue.cluster_centroids(coords, labels)
        want = oracle.cluster_centroids(coords, labels)
        assert cen.shape == want.shape
        np.testing.assert_array_equal(cen.view(np.int32), want.view(np.int32))
