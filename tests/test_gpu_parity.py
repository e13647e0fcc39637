"""Parity of the HIP engine (through the C ABI) with the oracle, the CPU
restatement of the reference's serial backend.  Integer outputs (tile CSR,
nearest-higher, seed flags, labels) must be bit-exact; densities are compared
bit-exactly too, because the engine accumulates in the reference's visit order
and evaluates exp() with the same correctly rounded definition as the oracle's
ENGINE mode."""
import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
import oracle
from cluestering_alpaka3_amd import datasets
from parity_util import assert_same, run_both, run_engine

pytestmark = pytest.mark.gpu


def test_fast_sqrt_exhaustive():
    # every float in [2^-96, FLT_MAX]: the kernels' fast sqrt equals IEEE sqrtf
    first = int(np.float32(2.0 ** -96).view(np.uint32))
    last = int(np.finfo(np.float32).max.view(np.uint32))
    assert clue.selftest_sqrt(first, last - first + 1) == 0
    # the nearest-higher searches' square root of any finite distance: below 2^-96
    # (every 11th float there, subnormals included, and both ends) on an argument
    # scaled by 2^64 into the range above, and the range above itself
    tiny = np.concatenate([np.arange(0, first, 11, dtype=np.uint64),
                           np.array([0, 1, first - 1, first], np.uint64)]).astype(np.uint32)
    for k in range(0, tiny.size, 1 << 23):
        xs = tiny[k:k + (1 << 23)].view(np.float32)
        np.testing.assert_array_equal(clue.selftest_math(10, xs).view(np.int32), np.sqrt(xs).view(np.int32))
    big = np.arange(first, last + 1, 997, dtype=np.uint64).astype(np.uint32).view(np.float32)
    np.testing.assert_array_equal(clue.selftest_math(10, big).view(np.int32), np.sqrt(big).view(np.int32))


def test_device_math_matches_host():
    # expf: dense sweep of the whole range the kernels can meet, plus specials
    x = np.concatenate([np.linspace(-111.0, 90.0, 400001, dtype=np.float32),
                        np.array([0.0, -0.0, 1e-30, -1e-30, 88.72, 88.73, -103.9, -104.0, np.inf, -np.inf],
                                 np.float32)])
    f32 = np.float32
    specials = np.array([np.nan, 89.0, np.nextafter(f32(89), f32(np.inf)), -110.0,
                         np.nextafter(f32(-110), f32(-np.inf)), np.finfo(f32).max, -np.finfo(f32).max,
                         -87.34, -103.97, -103.98, 88.7228, 88.7229], f32)
    x = np.concatenate([x, specials])
    got = clue.selftest_math(0, x)
    want = oracle.expf_engine(x)
    np.testing.assert_array_equal(got.view(np.int32), want.view(np.int32))
    # every 61st float of [-110, 89] (about 37 million arguments)
    lo_neg = np.float32(-110.0).view(np.uint32)  # negative floats: bit patterns 0x80000000..lo_neg
    hi_pos = np.float32(89.0).view(np.uint32)
    bits = np.concatenate([np.arange(0x80000000, int(lo_neg) + 1, 61, dtype=np.uint64),
                           np.arange(0, int(hi_pos) + 1, 61, dtype=np.uint64)]).astype(np.uint32)
    for k in range(0, bits.size, 1 << 23):
        xs = bits[k:k + (1 << 23)].view(np.float32)
        np.testing.assert_array_equal(clue.selftest_math(0, xs).view(np.int32),
                                      oracle.expf_engine(xs).view(np.int32))
    # the non-positive-argument variant on every 61st non-positive float (-0 to -inf)
    neg = np.arange(0x80000000, 0xFF800001, 61, dtype=np.uint64).astype(np.uint32)
    neg = np.concatenate([neg, np.array([0x80000000, 0xFF800000], np.uint32)])
    for k in range(0, neg.size, 1 << 23):
        xs = neg[k:k + (1 << 23)].view(np.float32)
        np.testing.assert_array_equal(clue.selftest_math(5, xs).view(np.int32),
                                      oracle.expf_engine(xs).view(np.int32))
    # the density pass's evaluation for unit-amplitude kernels, arguments in [-14, 0]
    # (every 7th float there, about 157 million, plus both ends), with the 64-entry
    # table and with the wide table the density pass reads
    lo14 = int(np.float32(-14.0).view(np.uint32))
    unit = np.concatenate([np.arange(0x80000000, lo14 + 1, 7, dtype=np.uint64),
                           np.array([0x80000000, lo14], np.uint64)]).astype(np.uint32)
    for k in range(0, unit.size, 1 << 23):
        xs = unit[k:k + (1 << 23)].view(np.float32)
        want = oracle.expf_engine(xs).view(np.int32)
        np.testing.assert_array_equal(clue.selftest_math(8, xs).view(np.int32), want)
        np.testing.assert_array_equal(clue.selftest_math(9, xs).view(np.int32), want)
    # sqrtf and division are correctly rounded on both sides
    rng = np.random.default_rng(1)
    a = np.abs(rng.standard_normal(200000).astype(np.float32)) * np.float32(1e3)
    a[:1000] = rng.random(1000).astype(np.float32) * np.float32(1e-38)  # subnormal range
    np.testing.assert_array_equal(clue.selftest_math(1, a).view(np.int32), np.sqrt(a).view(np.int32))
    b = (rng.random(200000).astype(np.float32) + np.float32(0.01)) * np.float32(7.3)
    np.testing.assert_array_equal(clue.selftest_math(2, a, b).view(np.int32), (a / b).view(np.int32))
    # the Gaussian argument's quotient through a double reciprocal equals the IEEE float division
    num = -(rng.standard_normal(300000).astype(np.float32) ** 2) * np.float32(3.0)
    num[:2000] = -(rng.random(2000).astype(np.float32) * np.float32(1e-30))
    num[2000:4000] = -np.ldexp(rng.random(2000), -130).astype(np.float32)
    den = (rng.random(300000).astype(np.float32) * np.float32(20) + np.float32(1e-3))
    den[4000:6000] = np.float32(2.0)
    den[6000:8000] = (rng.random(2000).astype(np.float32) * np.float32(1e-38))
    np.testing.assert_array_equal(clue.selftest_math(4, num, den).view(np.int32), (num / den).view(np.int32))
    # sqrt_lt_bound: largest s with sqrtf(s) < delta
    delta = np.abs(rng.standard_normal(20000).astype(np.float32)) * np.float32(10) + np.float32(1e-3)
    bnd = clue.selftest_math(3, delta)
    assert np.all(np.sqrt(bnd) < delta)
    nxt = np.nextafter(bnd, np.float32(np.inf))
    assert np.all(np.sqrt(nxt) >= delta)


def test_sqrt_class_bounds():
    # ops 6 / 7: the interval of squared distances whose sqrtf equals sqrtf(s), used
    # by the nearest-higher searches to compare distances in the squared domain
    f32 = np.float32
    rng = np.random.default_rng(7)
    bits = np.arange(0, int(f32(1e4).view(np.uint32)) + 1, 97, dtype=np.uint64).astype(np.uint32)
    pw = np.array([f32(2.0) ** k for k in range(-140, 64)], f32).view(np.uint32)
    near = (pw[:, None].astype(np.int64) + np.arange(-3, 4)[None, :]).ravel()
    near = near[(near >= 0)].astype(np.uint32)
    rnd = (rng.random(200000) * 64.0).astype(f32)
    s = np.concatenate([bits.view(f32), near.view(f32), rnd, np.array([0.0, 1e-45, 3e-45, 1e-40], f32)])
    d = np.sqrt(s)
    for k in range(0, s.size, 1 << 23):
        ss, dd = s[k:k + (1 << 23)], d[k:k + (1 << 23)]
        hi = clue.selftest_math(6, ss)
        lo = clue.selftest_math(7, ss)
        assert np.all(lo <= ss) and np.all(ss <= hi)
        assert np.array_equal(np.sqrt(hi), dd) and np.array_equal(np.sqrt(lo), dd)
        assert np.all(np.sqrt(np.nextafter(hi, f32(np.inf))) > dd)
        below = np.nextafter(lo, f32(-np.inf))
        assert np.all((lo == 0) | (np.sqrt(np.maximum(below, f32(0))) < dd))
        # op 11: every square below strictly_nearer(s) has a smaller sqrtf, and the
        # bound lies within 2^-19 of s (it only spares square roots)
        nn = clue.selftest_math(11, ss)
        pos = nn > 0
        assert np.all(np.sqrt(np.nextafter(nn[pos], f32(-np.inf))) < dd[pos])
        big = ss >= f32(2.0 ** -100)
        assert np.all(nn[big] >= ss[big] * f32(1 - 2.0 ** -19))


# reduced sizes with the local density of the full workloads (box and blob count
# scaled with n); the full sizes are in test_gpu_fullsize.py
WORKLOAD_CASES = [("2d_16k_flat", 16384), ("2d_1m_gauss", 65536), ("3d_4m_exp", 65536),
                  ("10d_256k_flat", 16384)]


@pytest.mark.parametrize("name,n", WORKLOAD_CASES)
def test_workload_parity(name, n):
    ps = datasets.workload_scaled(name, n)
    prm = datasets.WORKLOADS[name]
    assert_same(*run_both(ps.coords, ps.weights, prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"]))


TUNINGS = [dict(nh_first_radius=0.05, morton_bits=5),
           dict(nh_first_radius=1.0, morton_bits=1),
           dict(nh_first_radius=0.5, morton_bits=-1),
           dict(xcd_remap=0, pending_blocks=1, pending_group=8),
           dict(xcd_remap=3, pending_blocks=7, pending_group=64),
           dict(xcd_remap=2, pending_blocks=0, nh_first_radius=0.01, pending_group=32, density_order=0),
           dict(density_order=1, xcd_remap=0, pending_group=16),
           dict(density_order=1, xcd_remap=7),
           dict(small_call_points=0, wave_dims=10, tile_call_points=0),
           dict(small_call_points=1 << 30, nh_first_radius=0.3, tile_call_points=0),
           dict(small_call_points=1 << 30, tile_call_points=1 << 30),
           dict(small_call_points=0, tile_call_points=1 << 30),
           dict(density_filter=0, density_order=1),
           dict(fused_event_points=0),
           dict(split_call_points=1 << 30, xcd_remap=0),
           dict(split_call_points=1 << 30, small_call_points=0, xcd_remap=7),
           dict(split_call_points=0, small_call_points=0),
           dict(nh_lone_lanes=0),
           dict(nh_lone_lanes=64, nh_first_radius=1.0, split_call_points=0),
           dict(nh_lone_lanes=64, split_call_points=1 << 30)]


@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("name,n", [("2d_1m_gauss", 50000), ("3d_4m_exp", 50000)])
def test_tunings_do_not_change_results(name, n, tuning):
    ps = datasets.workload_scaled(name, n, seed=11)
    prm = datasets.WORKLOADS[name]
    assert_same(*run_both(ps.coords, ps.weights, prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"],
                          tuning=tuning))


def test_config1_golden(config1_golden):
    g = config1_golden
    res = clue.cluster_points(g["coords"], g["weights"], 1.5, 5.0, 3.0, kernel=("flat", 0.5))
    np.testing.assert_array_equal(res.rho.view(np.int32), g["rho"].view(np.int32))
    np.testing.assert_array_equal(res.nearest_higher, g["nh"])
    np.testing.assert_array_equal(res.is_seed, g["is_seed"])
    np.testing.assert_array_equal(res.cluster, g["cluster"])


# Parameters the reference's own tests use on its datasets
# (tests/test_clustering.cpp:11-131, tests/test_*_dataset.py).
REF_CASES = [
    ("data_1024", 1.5, 10.0, 1.5), ("data_4096", 1.5, 10.0, 1.5), ("data_16384", 1.5, 10.0, 1.5),
    ("data_32768", 1.5, 10.0, 1.5), ("data_32768", 1.3, 10.0, 1.3),
    ("toyDetector_1000", 4.0, 2.5, 4.0), ("toyDetector_10000", 4.0, 2.5, 4.0),
    ("sissa_1000", 20.0, 10.0, 20.0), ("sissa_4000", 20.0, 10.0, 20.0),
    ("aniso_1000", 25.0, 5.0, 23.0), ("blob", 1.0, 5.0, 2.0), ("moons", 21.0, 10.0, 21.0),
    ("square", 1.0, 2.0, 1.6), ("box", 1.0, 2.0, 1.6), ("data_dim_3", 1.5, 10.0, 1.5),
    ("data_dim_10", 3.0, 10.0, 6.0),
]


@pytest.mark.parametrize("name,dc,rhoc,dm", REF_CASES)
def test_reference_dataset_parity(ref_data, name, dc, rhoc, dm):
    c, w = ref_data(name)
    assert_same(*run_both(c, w, dc, rhoc, dm))


def test_known_answer_twenty_clusters(ref_data):
    # tests/test_host_points.cpp:437-450 and tests/test_device_points.cpp:252-268
    c, w = ref_data("data_32768")
    res = clue.cluster_points(c, w, 1.3, 10.0, 1.3)
    assert int(res.cluster.max()) + 1 == 20


@pytest.mark.parametrize("metric,mp", [("weighted_euclidean", [0.5, 2.0]), ("periodic_euclidean", [200.0, 0.0]),
                                       ("manhattan", None), ("chebyshev", None),
                                       ("weighted_chebyshev", [1.5, 0.75])])
@pytest.mark.parametrize("kernel", [("flat", 0.5), ("gaussian", 0.5, 1.2, 2.0), ("exponential", 0.8, 1.5)])
def test_metrics_and_kernels(ref_data, metric, mp, kernel):
    c, w = ref_data("toyDetector_1000")
    assert_same(*run_both(c, w, 4.0, 2.5, 4.0, kernel=kernel, metric=metric, metric_params=mp))
    assert_same(*run_both(c, w, 4.0, 2.5, 4.0, kernel=kernel, metric=metric, metric_params=mp,
                          tuning=dict(small_call_points=0, wave_dims=10, tile_call_points=0)))


@pytest.mark.parametrize("kernel", [("gaussian", 0.0, 1.0, 1.0), ("exponential", 1.0, 1.0)])
def test_kernels_euclidean(kernel):
    ps = datasets.workload("2d_1m_gauss", n=32768, seed=77)
    assert_same(*run_both(ps.coords, ps.weights, 2.0, 8.0, 4.0, kernel=kernel))


# small calls run the neighbour passes with one wave per query (small_call_points),
# larger ones with one lane per query: both on every dimension
@pytest.mark.parametrize("lanes", [False, True])
@pytest.mark.parametrize("ndim", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_every_dimension(ndim, lanes):
    ps = datasets.blobs(3000, ndim, 6, 0.0, 40.0, 1.0, 0.1, ("uniform", 0.5, 1.5), 50 + ndim)
    tuning = dict(small_call_points=0, wave_dims=10, tile_call_points=0) if lanes else None
    assert_same(*run_both(ps.coords, ps.weights, 1.5, 3.0, 2.5, kernel=("gaussian", 0.3, 1.0, 1.0), tuning=tuning))


def test_seed_capacity_truncation():
    # dm > seed_dc: seeds can outnumber the points without a nearest-higher; the
    # reference drops the highest-index seeds (SeedArray.hpp:33-41)
    ps = datasets.blobs(20000, 2, 60, 0.0, 200.0, 1.0, 0.0, ("const", 1.0), 99)
    got, ref, *rest = run_both(ps.coords, ps.weights, 0.6, 1.0, 5.0, seed_dc=0.6)
    assert ref.stats["n_dropped_seeds"] > 0
    assert_same(got, ref, *rest)
    got2, ref2, *rest2 = run_both(ps.coords, ps.weights, 0.6, 1.0, 5.0, seed_dc=0.6, seed_mode=1)
    assert ref2.stats["n_dropped_seeds"] == 0
    assert_same(got2, ref2, *rest2)


def test_grow_only_capacity_across_calls():
    # SetupSeeds.hpp:14-18: a Clusterer keeps the largest capacity it has seen
    big = datasets.blobs(20000, 2, 60, 0.0, 200.0, 1.0, 0.0, ("const", 1.0), 99)
    small = datasets.blobs(20000, 2, 60, 0.0, 200.0, 1.0, 0.0, ("const", 1.0), 98)
    plan = clue.Plan(0)
    try:
        clue.cluster_points(small.coords, small.weights, 0.6, 1.0, 5.0, 0.6, plan=plan)
        cap_small = plan.stats()["seed_capacity"]
        clue.cluster_points(big.coords, big.weights, 0.6, 1.0, 5.0, 0.6, plan=plan)
        st = plan.stats()
        p1 = oracle.make_params(0.6, 1.0, 5.0, 0.6)
        r_big_fresh = oracle.make_clusters(big.coords, big.weights, p1)
        p2 = oracle.make_params(0.6, 1.0, 5.0, 0.6, seed_capacity_floor=cap_small)
        r_big = oracle.make_clusters(big.coords, big.weights, p2)
        assert st["seed_capacity"] == max(cap_small, r_big_fresh.stats["seed_candidates"])
        assert st["n_clusters"] == r_big.stats["n_stored_seeds"]
    finally:
        plan.close()


def test_wrapped_coordinates(ref_data):
    c, w = ref_data("opposite_angles")
    # the reference's known answer (tests/test_clustering.cpp:133-150): wrapped
    # coordinate 1, PeriodicEuclidean with periods {0, 2*pi_v<float>}, one cluster
    two_pi = float(np.float32(2.0) * np.float32(np.pi))
    got, ref, *rest = run_both(c, w, 0.2, 5.0, 0.2, wrapped=[0, 1], metric="periodic_euclidean",
                               metric_params=[0.0, two_pi])
    assert_same(got, ref, *rest)
    assert int(got.cluster.max()) + 1 == 1
    assert_same(*run_both(c, w, 0.2, 5.0, 0.2, wrapped=[0, 1]))
    ps = datasets.blobs(5000, 2, 8, 0.0, 50.0, 1.0, 0.1, ("const", 1.0), 4242)
    assert_same(*run_both(ps.coords, ps.weights, 1.0, 3.0, 2.0, wrapped=[1, 1]))


def test_in_radius_outside_box():
    # q = 0.75 and c = -0.25 - 2^-25 are exactly dc = 1 apart after rounding
    # (fl(q - c) = 1), but c's tile lies below q's search box: its lower edge
    # fl(q - dc) = -0.25 is a tile boundary (x tiles of 0.5 from -0.75).  The
    # reference never visits c for q; the density's filtered walk must not add it.
    f32 = np.float32
    q, c = f32(0.75), f32(-0.25) - f32(2.0 ** -25)
    x = np.array([q, c, -0.75, 0.0, 0.5, -0.5, 0.25, 0.7, -0.7], f32)
    y = np.array([0.0, 0.0, 1.0, 2.0, 0.5, 1.5, 1.0, 2.0, 0.2], f32)
    assert np.sqrt(f32(q - c) * f32(q - c)) <= f32(1.0)
    ts = (x.max() - x.min()) / f32(3)
    assert int((f32(q - f32(1.0)) - x.min()) / ts) == 1 and int((c - x.min()) / ts) == 0
    coords = np.stack([x, y])
    w = np.ones(9, f32)
    for tuning in (None, dict(small_call_points=0, wave_dims=10, tile_call_points=0),
                   dict(small_call_points=0, wave_dims=10, tile_call_points=0, density_filter=0)):
        got, ref, *rest = run_both(coords, w, 1.0, 0.5, 1.0, ppb=1, tuning=tuning)
        assert ref.grid["npd"] == 3 and ref.rho[0] == f32(1.0) + f32(0.5)  # itself and (0.5, 0.5)
        assert_same(got, ref, *rest)


@pytest.mark.parametrize("fused", [32768, 0])
def test_small_calls_fused_and_not(ref_data, fused):
    # calls that skip the Morton layout: each event's layout and labels in one
    # workgroup (up to fused_event_points points per event) or through the
    # call-wide kernels
    t = dict(fused_event_points=fused)
    c, w = ref_data("toyDetector_1000")
    for kernel in (("flat", 0.5), ("gaussian", 0.0, 1.0, 1.0), ("exponential", 1.0, 1.0)):
        assert_same(*run_both(c, w, 4.0, 2.5, 4.0, kernel=kernel, tuning=t))
    for ndim in (1, 3, 7, 10):
        ps = datasets.blobs(6000, ndim, 5, 0.0, 30.0, 1.0, 0.1, ("uniform", 0.5, 1.5), 90 + ndim)
        assert_same(*run_both(ps.coords, ps.weights, 1.5, 3.0, 2.5, kernel=("exponential", 0.7, 1.0), tuning=t))
    # largest fused event (32,768 points) and a grid near the 2,048-tile limit
    ps = datasets.workload_scaled("2d_1m_gauss", 32768, seed=5)
    assert_same(*run_both(ps.coords, ps.weights, 2.0, 8.0, 4.0, kernel=("gaussian", 0.0, 1.0, 1.0), tuning=t))
    ps = datasets.blobs(30000, 2, 40, 0.0, 300.0, 1.5, 0.2, ("const", 1.0), 17)
    got, ref, *rest = run_both(ps.coords, ps.weights, 1.3, 2.0, 2.6, ppb=16, tuning=t)
    assert ref.grid["ntiles"] <= 2048
    assert_same(got, ref, *rest)
    # wrapped coordinates and seeds beyond the capacity
    ps = datasets.blobs(5000, 2, 8, 0.0, 50.0, 1.0, 0.1, ("const", 1.0), 4242)
    assert_same(*run_both(ps.coords, ps.weights, 1.0, 3.0, 2.0, wrapped=[1, 1], tuning=t))
    ps = datasets.blobs(20000, 2, 60, 0.0, 200.0, 1.0, 0.0, ("const", 1.0), 99)
    got, ref, *rest = run_both(ps.coords, ps.weights, 0.6, 1.0, 5.0, seed_dc=0.6, tuning=t)
    assert ref.stats["n_dropped_seeds"] > 0
    assert_same(got, ref, *rest)


def ring_ties(ndim, centres, per_ring, seed):
    """Light centres, each inside a shell of heavier points at radius 2.5-3.5
    rounded to float: their distances to the centre fall within a few ulps, so a
    centre's nearest-higher search meets equal sqrtf distances (ties broken by
    the visit order) and distances one ulp apart whose squares differ by ulps."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(centres ** (1.0 / ndim)))
    grid = np.stack(np.meshgrid(*[np.arange(side)] * ndim, indexing="ij")).reshape(ndim, -1)[:, :centres]
    cen = (grid * 9.0 - 9.0 * (side - 1) / 2).astype(np.float64)
    cen += rng.uniform(-0.3, 0.3, cen.shape)
    pts, w = [cen.astype(np.float32)], [np.full(cen.shape[1], 0.25, np.float32)]
    for k in range(cen.shape[1]):
        u = rng.normal(size=(ndim, per_ring))
        u /= np.linalg.norm(u, axis=0)
        r = rng.uniform(2.5, 3.5)
        pts.append((cen[:, k:k + 1] + r * u).astype(np.float32))
        w.append(rng.choice(np.float32([1.0, 1.5, 2.0]), per_ring))
    order = rng.permutation(sum(p.shape[1] for p in pts))
    return np.concatenate(pts, axis=1)[:, order], np.concatenate(w)[order]


NH_PATHS = [dict(small_call_points=0, tile_call_points=0),                       # block search, one phase
            dict(small_call_points=0, tile_call_points=0, nh_first_radius=0.3),  # centres left to the second
            dict(small_call_points=0, tile_call_points=0, nh_first_radius=0.3, pending_group=64),
            dict(small_call_points=0, tile_call_points=1 << 30),                 # tile walk, a lane per query
            dict(small_call_points=1 << 30, tile_call_points=1 << 30),           # tile walk, a wave per query
            dict(small_call_points=1 << 30, tile_call_points=0),
            dict(small_call_points=0, tile_call_points=0, split_call_points=1 << 30),
            dict(small_call_points=0, tile_call_points=0, nh_lone_lanes=0),  # point by point only
            dict(small_call_points=0, tile_call_points=0, nh_lone_lanes=64, nh_first_radius=1.0),  # across only
            dict(small_call_points=0, tile_call_points=0, nh_lone_lanes=64, split_call_points=1 << 30)]


_RING_CACHE = {}


def ring_case(ndim):
    """The ring set in ndim dimensions and the oracle's result (one per process)."""
    if ndim not in _RING_CACHE:
        c, w = ring_ties(ndim, 64, 1000 if ndim == 2 else 1500, 40 + ndim)
        args = (0.05, 0.2, 4.0, None, 32, ("flat", 0.5), "euclidean", None, None)
        _RING_CACHE[ndim] = (c, w, oracle.make_clusters(c, w, oracle.make_params(*args, seed_mode=0)))
    return _RING_CACHE[ndim]


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("tuning", NH_PATHS)
def test_nearest_higher_sqrt_ties(ndim, tuning):
    c, w, ref = ring_case(ndim)
    got, *rest = run_engine(c, w, 0.05, 0.2, 4.0, ppb=32, tuning=tuning)
    # the centres' nearest-highers are shell points at a distance shared by others
    d = ref.delta[w == np.float32(0.25)]
    assert np.all((d >= 2.5 * 0.99) & (d <= 3.5 * 1.01))
    assert_same(got, ref, *rest)


def test_edge_cases():
    # one point; identical points (zero tile size); two points; tiny clusters
    assert_same(*run_both(np.zeros((2, 1), np.float32), np.ones(1, np.float32), 1.0, 0.5))
    assert_same(*run_both(np.full((3, 257), 2.5, np.float32), np.ones(257, np.float32), 1.0, 0.5))
    assert_same(*run_both(np.array([[0.0, 0.5]], np.float32), np.array([1.0, 2.0], np.float32), 1.0, 0.5))
    rng = np.random.default_rng(3)
    x = np.round(rng.random((2, 3000)) * 20).astype(np.float32)  # many exact ties in rho and distance
    assert_same(*run_both(x, np.ones(3000, np.float32), 1.0, 2.0, 2.0))
    assert_same(*run_both(x, np.ones(3000, np.float32), 1.0, 2.0, 2.0, ppb=7))
    assert_same(*run_both(x, np.ones(3000, np.float32), 1.0, 2.0, 2.0, ppb=100000))



@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_dense_tile_and_flat_dimension(ndim):
    # three 8k-point blobs inside one tile (two far points spread the grid:
    # hundreds of 32-point blocks in one tile, most tiles empty), then the same
    # points with one coordinate constant (zero extent, zero tile size there)
    rng = np.random.default_rng(11 + ndim)
    x = rng.normal(size=(ndim, 8000)) * 0.3
    x[0] += 1.2 * (np.arange(8000) % 3)
    x = x.astype(np.float32)
    x[:, 0] = -500.0
    x[:, 1] = 500.0
    w = rng.uniform(0.5, 1.5, 8000).astype(np.float32)
    assert_same(*run_both(x, w, 0.2, 3.0, 0.4))
    if ndim > 1:
        x[ndim - 1, :] = 7.0
        assert_same(*run_both(x, w, 0.2, 3.0, 0.4, kernel=("gaussian", 0.0, 0.5, 1.0)))

def assert_same_special(got, ref, tiles, grid, stats, delta):
    # densities may be NaN here: equal as values (NaN matches NaN), not as bits,
    # since the payload of a NaN made by inf - inf differs between host and GPU
    np.testing.assert_array_equal(tiles[1], ref.tile_indexes)
    np.testing.assert_array_equal(got.rho, ref.rho)
    np.testing.assert_array_equal(delta, ref.delta)
    np.testing.assert_array_equal(got.nearest_higher, ref.nh)
    np.testing.assert_array_equal(got.is_seed, ref.is_seed)
    np.testing.assert_array_equal(got.cluster, ref.cluster)
    assert stats["seed_candidates"] == ref.stats["seed_candidates"]


@pytest.mark.parametrize("case", ["signed", "zero", "negzero", "nan", "inf", "signed_gauss"])
def test_nonpositive_and_special_densities(case):
    # the "higher" relation at rho <= 0 (no tie-break), at rho == +inf (tie-break
    # by id) and with NaN densities (never higher, nothing higher than them)
    rng = np.random.default_rng(11)
    n = 3000
    x = np.round(rng.random((2, n)) * 30).astype(np.float32) * np.float32(0.5)
    kernel = ("flat", 0.5)
    if case == "signed" or case == "signed_gauss":
        w = rng.integers(-1, 2, n).astype(np.float32)
        kernel = ("gaussian", 0.0, 1.0, 1.0) if case == "signed_gauss" else kernel
    elif case == "zero":
        w = np.zeros(n, np.float32)
    elif case == "negzero":
        w = np.where(rng.random(n) < 0.5, np.float32(-0.0), np.float32(0.0)).astype(np.float32)
        w[::97] = 1.0
    elif case == "nan":
        w = np.ones(n, np.float32)
        w[rng.choice(n, 20, replace=False)] = np.nan
    else:
        w = np.ones(n, np.float32)
        w[rng.choice(n, 20, replace=False)] = np.inf
        w[rng.choice(n, 5, replace=False)] = -np.inf
    assert_same_special(*run_both(x, w, 1.0, 1.0, 2.0, kernel=kernel))
    assert_same_special(*run_both(x, w, 1.0, 1.0, 2.0, kernel=kernel, tuning=dict(nh_first_radius=1.0)))
    # the tile walk (other metrics, wrapped grids, more than three dimensions)
    assert_same_special(*run_both(x, w, 1.0, 1.0, 2.0, kernel=kernel, metric="manhattan"))
    assert_same_special(*run_both(x, w, 1.0, 1.0, 2.0, kernel=kernel, wrapped=[1, 0]))
    x4 = np.concatenate([x, x[::-1]], axis=0)
    assert_same_special(*run_both(x4, w, 1.5, 1.0, 2.5, kernel=kernel))


@pytest.mark.parametrize("offset,scale", [(1.0e6, 1.0), (-3.0e5, 1.0), (0.0, 1.0e-3), (0.0, 250.0)])
def test_far_and_scaled_coordinates(offset, scale):
    # large coordinates (the first phase's box margin of 2^-20 |q|), tiny and large
    # scales (squared distances far from 1), dm below dc, seed_dc above dm, one
    # point per tile
    rng = np.random.default_rng(5)
    ps = datasets.workload("2d_16k_flat", n=6000)
    x = (ps.coords.astype(np.float64) * scale + offset).astype(np.float32)
    w = (rng.random(6000) + 0.5).astype(np.float32)
    dc, dm = np.float32(1.5 * scale), np.float32(3.0 * scale)
    assert_same(*run_both(x, w, float(dc), 3.0, float(dm)))
    assert_same(*run_both(x, w, float(dc), 3.0, float(dm) * 0.5, seed_dc=float(dm) * 1.5))
    assert_same(*run_both(x, w, float(dc), 3.0, float(dm), ppb=1, kernel=("gaussian", 0.0, float(scale), 1.0)))


def tiny_clusters(scale, tiny, n_wide=500, n_tiny=300, seed=9):
    """n_wide points uniform in [0, scale]^2 and two groups of n_tiny points at the
    origin, within `tiny` and 2^-10 `tiny` of it: squared distances below 2^-96
    (normal and subnormal) besides ordinary ones."""
    rng = np.random.default_rng(seed)
    wide = rng.random((2, n_wide)) * scale
    t1 = rng.random((2, n_tiny)) * tiny
    t2 = rng.random((2, n_tiny)) * tiny * 2.0 ** -10
    x = np.concatenate([wide, t1, t2], axis=1).astype(np.float32)
    w = (rng.random(x.shape[1]) + 0.5).astype(np.float32)
    return np.ascontiguousarray(x), w


@pytest.mark.parametrize("case", ["gauss_unit", "exp_unit", "gauss_steep", "exp_steep"])
def test_tiny_distances(case):
    # squared distances below 2^-96, where the kernels' fast square root is not
    # exact: the unit evaluations run there and must give exactly 1 (Gaussian of
    # mean 0 with 1/((2 sigma) sigma) <= 2^70, exponential with a rate <= 2^22);
    # steeper kernels, whose value there is not 1, keep the exact square root
    if case == "gauss_unit":
        x, w = tiny_clusters(5.0, 1e-16)
        kernel, dc = ("gaussian", 0.0, 1.0, 1.0), 2.0
    elif case == "exp_unit":
        x, w = tiny_clusters(5.0, 1e-16)
        kernel, dc = ("exponential", 4.0, 1.0), 2.0
    elif case == "gauss_steep":  # (2 sigma) sigma = 2^-81: 1/den = 2^81
        x, w = tiny_clusters(2e-12, 1e-15)
        kernel, dc = ("gaussian", 0.0, 2.0 ** -41, 1.0), 2e-12
    else:  # rate 2^23
        x, w = tiny_clusters(1.5e-6, 1e-15)
        kernel, dc = ("exponential", 2.0 ** 23, 1.0), 1.5e-6
    # one wave per query (small calls), one lane per query with the block search
    # and with the tile walk, and the split walk
    for tuning in (None, dict(small_call_points=0, tile_call_points=0, split_call_points=0),
                   dict(small_call_points=0, tile_call_points=1 << 30),
                   dict(small_call_points=0, tile_call_points=0, split_call_points=1 << 30)):
        got, ref, *rest = run_both(x, w, dc, 3.0, 2 * dc, kernel=kernel, tuning=tuning)
        assert_same(got, ref, *rest)
    d2 = (x[0, -1] - x[0, -2]) ** 2 + (x[1, -1] - x[1, -2]) ** 2
    assert d2 < 2.0 ** -96


def test_invalid_arguments():
    with pytest.raises(clue.InvalidArgument):
        clue.cluster_points(np.zeros((2, 4), np.float32), np.ones(4, np.float32), -1.0, 1.0)
    with pytest.raises(clue.InvalidArgument):
        clue.cluster_points(np.zeros((2, 4), np.float32), np.ones(4, np.float32), 1.0, -1.0)
    with pytest.raises(clue.InvalidArgument):
        clue.cluster_points(np.zeros((11, 4), np.float32), np.ones(4, np.float32), 1.0, 1.0)


def test_main_run_matches(ref_data):
    c, w = ref_data("toyDetector_1000")
    n = c.shape[1]
    data = np.concatenate([c.reshape(-1), w]).astype(np.float32)
    results = np.empty(n, np.int32)
    clue.main_run(4.0, 2.5, 4.0, 4.0, 128, [0, 0], data, results, ("flat", 0.5), 2, n, 256, 0)
    ref = oracle.make_clusters(c, w, oracle.make_params(4.0, 2.5, 4.0))
    np.testing.assert_array_equal(results, ref.cluster)


def test_main_run_repeated_sizes_and_dims():
    # the device context of mainRun is kept between calls: grow, shrink, change the
    # dimension, and a fresh seed capacity every call
    cases = [("2d_16k_flat", 4000, 2), ("3d_4m_exp", 30000, 3), ("2d_16k_flat", 1500, 2), ("10d_256k_flat", 3000, 10)]
    for name, n, ndim in cases:
        ps = datasets.workload(name, n=n)
        prm = datasets.WORKLOADS[name]
        data = np.concatenate([ps.coords.reshape(-1), ps.weights]).astype(np.float32)
        results = np.full(n, -7, np.int32)
        seed_dc = prm["dc"]
        print(f"case: mainRun {name} n={n}", flush=True)
        clue.main_run(prm["dc"], prm["rhoc"], prm["dm"], seed_dc, 128, [0] * ndim, data, results, prm["kernel"],
                      ndim, n, 256, 0)
        ref = oracle.make_clusters(ps.coords, ps.weights, oracle.make_params(prm["dc"], prm["rhoc"], prm["dm"],
                                                                            kernel=prm["kernel"]))
        np.testing.assert_array_equal(results, ref.cluster, err_msg=name)


def test_deterministic_repeat():
    ps = datasets.workload("2d_1m_gauss", n=100000, seed=5)
    a = clue.cluster_points(ps.coords, ps.weights, 2.0, 8.0, 4.0, kernel=("gaussian", 0.0, 1.0, 1.0))
    b = clue.cluster_points(ps.coords, ps.weights, 2.0, 8.0, 4.0, kernel=("gaussian", 0.0, 1.0, 1.0))
    for x, y in zip((a.rho, a.nearest_higher, a.is_seed, a.cluster), (b.rho, b.nearest_higher, b.is_seed, b.cluster)):
        np.testing.assert_array_equal(x, y)
