"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg compare the HIP engine against.  The product never
imports this module.  The oracle restates CLUEstering 2.9.0's serial backend;
see clue_oracle.cpp for the file:line of every stage it follows.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
# the same sources under the reference's Release-style flags (fast math): for
# timing the CPU baseline only, never as the parity reference
_RELEASE_PATH = os.path.join(_HERE, "liboracle_release.so")

MAX_DIM = 10
KERNELS = {"flat": 0, "gaussian": 1, "exponential": 2}
METRICS = {
    "euclidean": 0,
    "weighted_euclidean": 1,
    "periodic_euclidean": 2,
    "manhattan": 3,
    "chebyshev": 4,
    "weighted_chebyshev": 5,
}
EXP_ENGINE, EXP_LIBM = 0, 1
SEEDS_REFERENCE, SEEDS_UNBOUNDED = 0, 1


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("dc", ctypes.c_float),
        ("rhoc", ctypes.c_float),
        ("dm", ctypes.c_float),
        ("seed_dc", ctypes.c_float),
        ("points_per_tile", ctypes.c_int32),
        ("kernel_type", ctypes.c_int32),
        ("kernel_params", ctypes.c_float * 3),
        ("metric_type", ctypes.c_int32),
        ("metric_params", ctypes.c_float * MAX_DIM),
        ("wrapped", ctypes.c_uint8 * MAX_DIM),
        ("exp_mode", ctypes.c_int32),
        ("seed_mode", ctypes.c_int32),
        ("seed_capacity_floor", ctypes.c_int64),
    ]


class OracleGrid(ctypes.Structure):
    _fields_ = [
        ("npd", ctypes.c_int32),
        ("ntiles", ctypes.c_int32),
        ("minmax", ctypes.c_float * (2 * MAX_DIM)),
        ("tile_size", ctypes.c_float * MAX_DIM),
    ]


class OracleStats(ctypes.Structure):
    _fields_ = [
        (name, ctypes.c_int64)
        for name in (
            "seed_candidates",
            "seed_capacity",
            "n_seeds",
            "n_stored_seeds",
            "n_dropped_seeds",
            "max_dfs_stack",
            "n_clustered",
            "candidate_pairs_density",
            "candidate_pairs_nh",
            "inradius_pairs_density",
            "inradius_pairs_nh",
        )
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class OracleMargins(ctypes.Structure):
    _fields_ = [
        ("rhoc_band", ctypes.c_int64),
        ("nh_distance_ties", ctypes.c_int64),
        ("nh_distance_near_ties", ctypes.c_int64),
        ("nh_density_near_ties", ctypes.c_int64),
        ("density_max_rel_err_fp64", ctypes.c_double),
        ("density_above_band_fp64", ctypes.c_int64),
        ("dc_boundary_pairs", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: (float(getattr(self, name)) if t is ctypes.c_double else int(getattr(self, name)))
                for name, t in self._fields_}


_libs = {}


def lib(release=False):
    path = _RELEASE_PATH if release else _LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        f32p, i32p, f64p = P(ctypes.c_float), P(ctypes.c_int32), P(ctypes.c_double)
        L.oracle_grid_setup.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p, ctypes.c_int32, P(OracleGrid)]
        L.oracle_tiles.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p, P(OracleParams), P(OracleGrid), i32p, i32p]
        L.oracle_make_clusters.argtypes = [
            ctypes.c_int32, ctypes.c_int32, f32p, f32p, P(OracleParams),
            f32p, i32p, f32p, i32p, i32p, P(OracleStats), i32p, i32p, P(OracleGrid)]
        L.oracle_make_clusters_omp.argtypes = [
            ctypes.c_int32, ctypes.c_int32, f32p, f32p, P(OracleParams), ctypes.c_int32,
            f32p, i32p, i32p, i32p, P(OracleStats)]
        L.oracle_make_clusters_par.argtypes = [
            ctypes.c_int32, ctypes.c_int32, f32p, f32p, P(OracleParams), ctypes.c_int32,
            f32p, i32p, f32p, i32p, i32p, P(OracleStats), i32p, i32p, P(OracleGrid)]
        L.oracle_parity_margins.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p, f32p, P(OracleParams), f32p,
                                            ctypes.c_int32, ctypes.c_double, P(OracleMargins)]
        L.oracle_bruteforce_density_fp64.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p, f32p, P(OracleParams), f64p]
        L.oracle_bruteforce_nh.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p, f32p, P(OracleParams), i32p, f32p]
        L.oracle_expf_engine.argtypes = [ctypes.c_float]
        L.oracle_expf_engine.restype = ctypes.c_float
        L.oracle_expf_longdouble.argtypes = [ctypes.c_float]
        L.oracle_expf_longdouble.restype = ctypes.c_float
        L.oracle_expf_engine_batch.argtypes = [f32p, f32p, ctypes.c_int64]
        L.oracle_expf_libm_batch.argtypes = [f32p, f32p, ctypes.c_int64]
        L.oracle_distance.argtypes = [ctypes.c_int32, f32p, f32p, P(OracleParams)]
        L.oracle_distance.restype = ctypes.c_float
        L.oracle_read_csv.argtypes = [ctypes.c_char_p, ctypes.c_int32, f32p, f32p, ctypes.c_int64]
        L.oracle_read_csv.restype = ctypes.c_int64
        L.oracle_set_mutation.argtypes = [ctypes.c_int32]
        L.oracle_cluster_map.argtypes = [ctypes.c_int32, i32p, i32p, i32p, i32p]
        L.oracle_cluster_map.restype = ctypes.c_int32
        L.oracle_cluster_centroids.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p, i32p, ctypes.c_int32, i32p, f32p]
        _libs[path] = L
    return _libs[path]


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(dc, rhoc, dm=None, seed_dc=None, ppb=128, kernel=("flat", 0.5), metric="euclidean",
                metric_params=None, wrapped=None, exp_mode=EXP_ENGINE, seed_mode=SEEDS_REFERENCE,
                seed_capacity_floor=0):
    """Parameters with the reference defaults: dm = seed_dc = dc, pPBin = 128
    (core/Clusterer.hpp:84-90 via Clusterer ctor, core/detail/Clusterer.hpp:24-36)."""
    p = OracleParams()
    p.dc = dc
    p.rhoc = rhoc
    p.dm = dc if dm is None else dm
    p.seed_dc = dc if seed_dc is None else seed_dc
    p.points_per_tile = ppb
    name = kernel[0]
    p.kernel_type = KERNELS[name]
    for i, v in enumerate(kernel[1:]):
        p.kernel_params[i] = v
    p.metric_type = METRICS[metric]
    if metric_params is not None:
        for i, v in enumerate(metric_params):
            p.metric_params[i] = v
    if wrapped is not None:
        for i, v in enumerate(wrapped):
            p.wrapped[i] = int(v)
    p.exp_mode = exp_mode
    p.seed_mode = seed_mode
    p.seed_capacity_floor = seed_capacity_floor
    return p


@dataclass
class OracleResult:
    rho: np.ndarray
    nh: np.ndarray
    delta: np.ndarray
    is_seed: np.ndarray
    cluster: np.ndarray
    stats: dict
    tile_offsets: np.ndarray
    tile_indexes: np.ndarray
    grid: dict = field(default_factory=dict)


def _soa(coords):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if c.ndim == 1:
        c = c[None, :]
    return c


def grid(coords, ppb=128):
    c = _soa(coords)
    g = OracleGrid()
    rc = lib().oracle_grid_setup(c.shape[0], c.shape[1], _fp(c), ppb, ctypes.byref(g))
    if rc != 0:
        raise ValueError("oracle_grid_setup failed")
    d = c.shape[0]
    return {"npd": g.npd, "ntiles": g.ntiles, "minmax": np.array(g.minmax[: 2 * d], np.float32),
            "tile_size": np.array(g.tile_size[:d], np.float32)}


def make_clusters(coords, weights, params, with_tiles=True, release=False) -> OracleResult:
    """Serial pipeline: coords is (ndim, n) float32 SoA, weights (n,).  release=True runs
    the fast-math build (timing only)."""
    c = _soa(coords)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    ndim, n = c.shape
    rho = np.empty(n, np.float32)
    nh = np.empty(n, np.int32)
    delta = np.empty(n, np.float32)
    is_seed = np.empty(n, np.int32)
    cluster = np.empty(n, np.int32)
    st = OracleStats()
    g = OracleGrid()
    gi = grid(c, params.points_per_tile)
    off = np.empty(gi["ntiles"] + 1, np.int32) if with_tiles else None
    idx = np.empty(n, np.int32) if with_tiles else None
    rc = lib(release).oracle_make_clusters(ndim, n, _fp(c), _fp(w), ctypes.byref(params), _fp(rho), _ip(nh),
                                    _fp(delta), _ip(is_seed), _ip(cluster), ctypes.byref(st),
                                    _ip(off), _ip(idx), ctypes.byref(g))
    if rc != 0:
        raise ValueError("oracle_make_clusters: invalid parameters")
    return OracleResult(rho, nh, delta, is_seed, cluster, st.as_dict(), off, idx, gi)


def make_clusters_par(coords, weights, params, nthreads=0, with_tiles=True) -> OracleResult:
    """make_clusters with every output, its loops on nthreads OpenMP threads (same results)."""
    c = _soa(coords)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    ndim, n = c.shape
    rho = np.empty(n, np.float32)
    nh = np.empty(n, np.int32)
    delta = np.empty(n, np.float32)
    is_seed = np.empty(n, np.int32)
    cluster = np.empty(n, np.int32)
    st = OracleStats()
    g = OracleGrid()
    gi = grid(c, params.points_per_tile)
    off = np.empty(gi["ntiles"] + 1, np.int32) if with_tiles else None
    idx = np.empty(n, np.int32) if with_tiles else None
    rc = lib().oracle_make_clusters_par(ndim, n, _fp(c), _fp(w), ctypes.byref(params), nthreads, _fp(rho),
                                        _ip(nh), _fp(delta), _ip(is_seed), _ip(cluster), ctypes.byref(st),
                                        _ip(off), _ip(idx), ctypes.byref(g))
    if rc != 0:
        raise ValueError("oracle_make_clusters_par: invalid parameters")
    return OracleResult(rho, nh, delta, is_seed, cluster, st.as_dict(), off, idx, gi)


def parity_margins(coords, weights, params, rho, nthreads=0, band=1e-6):
    """Points within `band` (relative) of a decision flip; see oracle_margins in clue_oracle.h."""
    c = _soa(coords)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    r = np.ascontiguousarray(rho, dtype=np.float32)
    m = OracleMargins()
    rc = lib().oracle_parity_margins(c.shape[0], c.shape[1], _fp(c), _fp(w), ctypes.byref(params), _fp(r),
                                     nthreads, band, ctypes.byref(m))
    if rc != 0:
        raise ValueError("oracle_parity_margins: invalid arguments (Euclidean metric only)")
    return m.as_dict()


def make_clusters_omp(coords, weights, params, nthreads=0, release=False):
    c = _soa(coords)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    ndim, n = c.shape
    rho = np.empty(n, np.float32)
    nh = np.empty(n, np.int32)
    is_seed = np.empty(n, np.int32)
    cluster = np.empty(n, np.int32)
    st = OracleStats()
    rc = lib(release).oracle_make_clusters_omp(ndim, n, _fp(c), _fp(w), ctypes.byref(params), nthreads,
                                        _fp(rho), _ip(nh), _ip(is_seed), _ip(cluster), ctypes.byref(st))
    if rc != 0:
        raise ValueError("oracle_make_clusters_omp: invalid parameters")
    return OracleResult(rho, nh, None, is_seed, cluster, st.as_dict(), None, None)


def bruteforce_density_fp64(coords, weights, params):
    c = _soa(coords)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = np.empty(c.shape[1], np.float64)
    rc = lib().oracle_bruteforce_density_fp64(c.shape[0], c.shape[1], _fp(c), _fp(w), ctypes.byref(params), _dp(out))
    if rc != 0:
        raise ValueError("bruteforce density failed")
    return out


def bruteforce_nh(coords, rho, params):
    c = _soa(coords)
    r = np.ascontiguousarray(rho, dtype=np.float32)
    nh = np.empty(c.shape[1], np.int32)
    delta = np.empty(c.shape[1], np.float32)
    rc = lib().oracle_bruteforce_nh(c.shape[0], c.shape[1], _fp(c), _fp(r), ctypes.byref(params), _ip(nh), _fp(delta))
    if rc != 0:
        raise ValueError("bruteforce nh failed")
    return nh, delta


def expf_engine(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().oracle_expf_engine_batch(_fp(x), _fp(y), x.size)
    return y


def expf_libm(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().oracle_expf_libm_batch(_fp(x), _fp(y), x.size)
    return y


def expf_longdouble(x):
    return np.array([lib().oracle_expf_longdouble(float(v)) for v in np.asarray(x, np.float32)], np.float32)


def distance(a, b, params):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib().oracle_distance(a.size, _fp(a), _fp(b), ctypes.byref(params)))


def read_csv(path, ndim):
    n = lib().oracle_read_csv(str(path).encode(), ndim, None, None, 0)
    if n < 0:
        raise RuntimeError(f"could not read {path}")
    coords = np.empty((ndim, n), np.float32)
    w = np.empty(n, np.float32)
    lib().oracle_read_csv(str(path).encode(), ndim, _fp(coords), _fp(w), n)
    return coords, w


def cluster_map(cluster):
    """(offsets[n_clusters+1], indexes[clustered]) of the host association map."""
    c = np.ascontiguousarray(cluster, dtype=np.int32)
    n = c.size
    off = np.zeros(max(n, int(c.max(initial=-1)) + 1) + 1, np.int32)
    idx = np.zeros(max(n, 1), np.int32)
    cnt = ctypes.c_int32(0)
    nc = lib().oracle_cluster_map(n, _ip(c), _ip(off), _ip(idx), ctypes.byref(cnt))
    return off[: nc + 1].copy(), idx[: cnt.value].copy()


def cluster_centroids(coords, cluster):
    """(n_clusters, ndim) centroids with the reference's serial float arithmetic."""
    x = _soa(coords)
    c = np.ascontiguousarray(cluster, dtype=np.int32)
    off, _ = cluster_map(c)
    nc = off.size - 1
    out = np.zeros((max(nc, 1), x.shape[0]), np.float32)
    lib().oracle_cluster_centroids(x.shape[0], x.shape[1], _fp(x), _ip(c), nc, _ip(off), _fp(out))
    return out[:nc]


class mutation:
    """Context manager planting a known mistake in the oracle (pinning tests)."""

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        lib().oracle_set_mutation(self.code)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_mutation(0)
        return False
