"""Deterministic synthetic point sets for the clustering workloads.

The generator is counter-based (splitmix64) with Box-Muller normals evaluated in
double and rounded once to float32, so a (config, seed) pair gives the same
arrays on every machine with IEEE doubles.  The blob/noise split follows the
reference's benchmark generator (benchmark/utils/generation.hpp:81-87):
per_blob = floor(N * (1 - noise) / n_blobs), the remainder is uniform noise over
the bounding box.  Points are returned in a random order (shuffled), coords as
an (ndim, n) float32 SoA array and weights as (n,) float32.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    """Counter-based splitmix64 stream."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def next_u64(self, n):
        idx = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            z = self.state + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        u1 = 1.0 - self.uniform(n)  # (0, 1]
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    def exponential(self, n, mean):
        return -mean * np.log(1.0 - self.uniform(n))

    def permutation(self, n):
        return np.argsort(self.next_u64(n), kind="stable")


@dataclass
class PointSet:
    coords: np.ndarray  # (ndim, n) float32
    weights: np.ndarray  # (n,) float32
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def ndim(self):
        return self.coords.shape[0]

    @property
    def n(self):
        return self.coords.shape[1]

    def sha256(self):
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.coords).tobytes())
        h.update(np.ascontiguousarray(self.weights).tobytes())
        return h.hexdigest()


def blobs(n, ndim, n_blobs, lo, hi, sigma, noise, weights, seed, name=""):
    """Gaussian blobs plus uniform noise.

    sigma: float (isotropic), ("uniform", a, b) per blob, or ("uniform_axis", a, b)
    per blob and axis.  weights: ("const", v), ("uniform", a, b),
    ("exponential", mean) or ("exponential_plus", mean, offset).
    """
    rng = SplitMix64(seed)
    per_blob = int(np.floor(n * (1.0 - noise) / n_blobs))
    n_noise = n - per_blob * n_blobs
    centres = rng.uniform(n_blobs * ndim, lo, hi).reshape(n_blobs, ndim)
    if isinstance(sigma, tuple) and sigma[0] == "uniform":
        sig = np.repeat(rng.uniform(n_blobs, sigma[1], sigma[2])[:, None], ndim, axis=1)
    elif isinstance(sigma, tuple) and sigma[0] == "uniform_axis":
        sig = rng.uniform(n_blobs * ndim, sigma[1], sigma[2]).reshape(n_blobs, ndim)
    else:
        sig = np.full((n_blobs, ndim), float(sigma))
    pts = np.empty((n, ndim), np.float64)
    z = rng.normal(per_blob * n_blobs * ndim).reshape(n_blobs, per_blob, ndim)
    pts[: per_blob * n_blobs] = (centres[:, None, :] + z * sig[:, None, :]).reshape(-1, ndim)
    pts[per_blob * n_blobs:] = rng.uniform(n_noise * ndim, lo, hi).reshape(n_noise, ndim)
    kind = weights[0]
    if kind == "const":
        w = np.full(n, weights[1], np.float64)
    elif kind == "uniform":
        w = rng.uniform(n, weights[1], weights[2])
    elif kind == "exponential":
        w = rng.exponential(n, weights[1])
    elif kind == "exponential_plus":
        w = rng.exponential(n, weights[1]) + weights[2]
    else:
        raise ValueError(kind)
    order = rng.permutation(n)
    coords = np.ascontiguousarray(pts[order].T.astype(np.float32))
    return PointSet(coords, w[order].astype(np.float32), name,
                    {"n_blobs": n_blobs, "noise": noise, "seed": seed, "per_blob": per_blob})


# ---------------------------------------------------------------------------
# The benchmark workloads
# ---------------------------------------------------------------------------
# Clustering parameters per workload: (dc, rhoc, dm, kernel)
WORKLOADS = {
    "2d_16k_flat": dict(dc=1.5, rhoc=5.0, dm=3.0, kernel=("flat", 0.5)),
    "2d_1m_gauss": dict(dc=2.0, rhoc=8.0, dm=4.0, kernel=("gaussian", 0.0, 1.0, 1.0)),
    "3d_4m_exp": dict(dc=1.5, rhoc=6.0, dm=3.0, kernel=("exponential", 1.0, 1.0)),
    "10d_256k_flat": dict(dc=3.0, rhoc=10.0, dm=6.0, kernel=("flat", 0.5)),
}


def workload(name, seed=None, n=None):
    """Points of a named workload (optionally with a different size for tests)."""
    if name == "2d_16k_flat":
        N = n or 16384
        return blobs(N, 2, 10, 0.0, 100.0, 1.5, 0.10, ("const", 1.0), 1001 if seed is None else seed, name)
    if name == "2d_1m_gauss":
        N = n or 1048576
        return blobs(N, 2, 200, 0.0, 1000.0, ("uniform", 1.0, 4.0), 0.20, ("uniform", 0.5, 1.5),
                     2002 if seed is None else seed, name)
    if name == "3d_4m_exp":
        N = n or 4194304
        return blobs(N, 3, 1000, 0.0, 500.0, ("uniform_axis", 0.5, 3.0), 0.15, ("exponential", 1.0),
                     3003 if seed is None else seed, name)
    if name == "10d_256k_flat":
        N = n or 262144
        return blobs(N, 10, 50, 0.0, 100.0, 1.0, 0.05, ("const", 1.0), 3004 if seed is None else seed, name)
    raise KeyError(name)


def workload_scaled(name, n, seed=None):
    """A named workload at n points with its local density kept: the box side
    scales by (n / N)^(1/D) and the number of blobs by n / N (N the workload's
    full size), as in the served mix -- so a reduced test sees the full
    workload's tile occupancy and neighbour counts, not a sparser regime."""
    full = {"2d_16k_flat": (16384, 2, 10, 100.0), "2d_1m_gauss": (1048576, 2, 200, 1000.0),
            "3d_4m_exp": (4194304, 3, 1000, 500.0), "10d_256k_flat": (262144, 10, 50, 100.0)}
    N, D, nb_full, side_full = full[name]
    if n is None or n == N:
        return workload(name, seed=seed)
    f = n / N
    side = side_full * f ** (1.0 / D)
    nb = max(1, int(round(nb_full * f)))
    if name == "2d_16k_flat":
        return blobs(n, 2, nb, 0.0, side, 1.5, 0.10, ("const", 1.0), 1001 if seed is None else seed, name)
    if name == "2d_1m_gauss":
        return blobs(n, 2, nb, 0.0, side, ("uniform", 1.0, 4.0), 0.20, ("uniform", 0.5, 1.5),
                     2002 if seed is None else seed, name)
    if name == "3d_4m_exp":
        return blobs(n, 3, nb, 0.0, side, ("uniform_axis", 0.5, 3.0), 0.15, ("exponential", 1.0),
                     3003 if seed is None else seed, name)
    return blobs(n, 10, nb, 0.0, side, 1.0, 0.05, ("const", 1.0), 3004 if seed is None else seed, name)


def calorimeter_event(n, seed):
    """A toy calorimeter layer: 5-40 showers in [-150,150]^2 with radial Gaussian
    spread 2, plus uniform noise hits; energies exponential (mean 0.5) + 0.1."""
    rng = SplitMix64(seed)
    n_showers = int(5 + np.floor(rng.uniform(1)[0] * 36))
    n_noise = max(1, int(np.floor(0.1 * n)))
    per = (n - n_noise) // n_showers
    n_noise = n - per * n_showers
    centres = rng.uniform(n_showers * 2, -150.0, 150.0).reshape(n_showers, 2)
    z = rng.normal(per * n_showers * 2).reshape(n_showers, per, 2) * 2.0
    pts = np.empty((n, 2), np.float64)
    pts[: per * n_showers] = (centres[:, None, :] + z).reshape(-1, 2)
    pts[per * n_showers:] = rng.uniform(n_noise * 2, -150.0, 150.0).reshape(n_noise, 2)
    w = rng.exponential(n, 0.5) + 0.1
    order = rng.permutation(n)
    return PointSet(np.ascontiguousarray(pts[order].T.astype(np.float32)), w[order].astype(np.float32),
                    "calorimeter", {"showers": n_showers})


def event_sizes(n_events, seed, median=5000.0, sigma_log=1.0, lo=500, hi=50000):
    rng = SplitMix64(seed)
    z = rng.normal(n_events)
    return np.clip(np.round(median * np.exp(sigma_log * z)), lo, hi).astype(np.int64)


def served_event(k, seed, lo=1000, hi=200000, frac_2d=0.7):
    """Event k of the served mix: 70% 2D, 30% 3D, sizes log-uniform in [lo, hi],
    drawn from the 2D 1M-point and 3D 4M-point generators with the number of
    blobs and the box scaled with the size (so local densities stay those of the
    full workloads), exponential weights (mean 1).  Returns (PointSet, params)."""
    rng = SplitMix64(seed * 1_000_003 + k)
    u = rng.uniform(2)
    n = int(np.floor(np.exp(np.log(lo) + u[0] * (np.log(hi) - np.log(lo)))))
    if u[1] < frac_2d:
        f = n / 1048576.0
        side = 1000.0 * np.sqrt(f)
        nb = max(1, int(round(200 * f)))
        ps = blobs(n, 2, nb, 0.0, side, ("uniform", 1.0, 4.0), 0.20, ("exponential", 1.0),
                   seed * 7 + k, "served_2d")
        return ps, WORKLOADS["2d_1m_gauss"]
    f = n / 4194304.0
    side = 500.0 * f ** (1.0 / 3.0)
    nb = max(1, int(round(1000 * f)))
    ps = blobs(n, 3, nb, 0.0, side, ("uniform_axis", 0.5, 3.0), 0.15, ("exponential", 1.0), seed * 7 + k,
               "served_3d")
    return ps, WORKLOADS["3d_4m_exp"]
