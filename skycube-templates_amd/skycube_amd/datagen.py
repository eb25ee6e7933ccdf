"""Seeded synthetic skycube inputs (float32, row-major n x d).

The reference reads its inputs from CSV files made by an external generator
that is not part of it (README.md:63-64); the workloads below are defined
here, by distribution, size and seed, so every machine regenerates the same
bytes (numpy's PCG64 stream is platform independent).

Workloads:
  independent_distinct  n=100,000, d=8: U[0,1) i.i.d., every value distinct
                        within its column (seed 1)
  anticorrelated_plane  n=500,000, d=12: per point c ~ N(0.5, 0.05); offsets
                        u_j ~ U(-h, h), h = 0.5, re-centred to sum 0; x_j =
                        clip(c + u_j, 0, nextafter(1, 0)) (seed 2)
  integer_ties          n=1,000,000, d=16: x_0..x_7 uniform integers 0..255,
                        x_8..x_15 = clip(round(x_0 + N(0, 20)), 0, 255),
                        stored as float32 (seed 3); subspaces capped at 8 dims
  batched instances     4,096 instances, n log-uniform in [1000, 65536], d
                        uniform in {4..10}, distribution independent /
                        correlated / anticorrelated with equal probability
                        (seed 4)
  serving stream        Poisson arrivals; n log-normal (median 20,000,
                        sigma 1.0) clipped to [500, 500,000], d uniform in
                        {6..14}, mixed distributions (seed 5)
"""
from __future__ import annotations

import numpy as np

ONE_BELOW = np.nextafter(np.float32(1.0), np.float32(0.0))
ANTI_SPREAD_H = 0.5
CORR_SIGMA = 0.05


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def independent(n: int, d: int, seed: int, distinct: bool = False) -> np.ndarray:
    """U[0,1) i.i.d. float32; with distinct=True every column has no repeats
    (repeats are redrawn until none is left)."""
    rng = _rng(seed)
    x = rng.random((n, d), dtype=np.float32)
    if distinct:
        for j in range(d):
            col = x[:, j]
            while True:
                _, first = np.unique(col, return_index=True)
                if first.size == n:
                    break
                dup = np.ones(n, dtype=bool)
                dup[first] = False
                col[dup] = rng.random(int(dup.sum()), dtype=np.float32)
            x[:, j] = col
    return np.ascontiguousarray(x)


def anticorrelated(n: int, d: int, seed: int, h: float = ANTI_SPREAD_H,
                   centre_sd: float = 0.05) -> np.ndarray:
    """Points near the plane sum(x) = d/2: centre c ~ N(0.5, centre_sd) per
    point, uniform offsets in (-h, h) re-centred to sum 0, clipped to [0,1)."""
    rng = _rng(seed)
    c = rng.normal(0.5, centre_sd, size=(n, 1))
    u = rng.uniform(-h, h, size=(n, d))
    u -= u.mean(axis=1, keepdims=True)
    x = np.clip(c + u, 0.0, 1.0).astype(np.float32)
    return np.ascontiguousarray(np.minimum(x, ONE_BELOW))


def correlated(n: int, d: int, seed: int, sigma: float = CORR_SIGMA) -> np.ndarray:
    """Points near the diagonal: c ~ U[0,1) per point, x_j = clip(c + N(0,
    sigma), 0, 1) with 1 mapped to the largest float32 below 1."""
    rng = _rng(seed)
    c = rng.random((n, 1))
    x = np.clip(c + rng.normal(0.0, sigma, size=(n, d)), 0.0, 1.0).astype(np.float32)
    return np.ascontiguousarray(np.minimum(x, ONE_BELOW))


def integer_ties(n: int, d: int, seed: int, sd: float = 20.0) -> np.ndarray:
    """Half the dimensions uniform integers 0..255, half correlated with
    dimension 0 (clip(round(x_0 + N(0, sd)), 0, 255)); float32."""
    rng = _rng(seed)
    h = d // 2 if d > 1 else 1
    x = np.empty((n, d), dtype=np.float32)
    x[:, :h] = rng.integers(0, 256, size=(n, h)).astype(np.float32)
    if d > h:
        noise = rng.normal(0.0, sd, size=(n, d - h))
        x[:, h:] = np.clip(np.rint(x[:, :1] + noise), 0, 255).astype(np.float32)
    return np.ascontiguousarray(x)


def by_name(dist: str, n: int, d: int, seed: int) -> np.ndarray:
    if dist == "independent":
        return independent(n, d, seed)
    if dist == "correlated":
        return correlated(n, d, seed)
    if dist == "anticorrelated":
        return anticorrelated(n, d, seed, centre_sd=0.05)
    if dist == "integer":
        return integer_ties(n, d, seed)
    raise ValueError(f"unknown distribution {dist!r}")


# The three plain-skycube workloads: name -> (n, d, max_d, seed, maker)
WORKLOADS = {
    "independent_distinct": dict(n=100_000, d=8, max_d=8, seed=1),
    "anticorrelated_plane": dict(n=500_000, d=12, max_d=12, seed=2),
    "integer_ties": dict(n=1_000_000, d=16, max_d=8, seed=3),
    # n = 1M, d = 16 independent data (distinct values): 635,442 points in
    # the extended skyline
    "independent_16d": dict(n=1_000_000, d=16, max_d=8, seed=6),
}


def workload(name: str, n: int | None = None) -> np.ndarray:
    """Rows of a named workload; n < the workload's size gives a prefix-sized
    instance drawn from the same generator and seed (not a row prefix)."""
    w = WORKLOADS[name]
    size = w["n"] if n is None else int(n)
    if name == "independent_distinct":
        return independent(size, w["d"], w["seed"], distinct=True)
    if name == "anticorrelated_plane":
        return anticorrelated(size, w["d"], w["seed"])
    if name == "integer_ties":
        return integer_ties(size, w["d"], w["seed"])
    if name == "independent_16d":
        return independent(size, w["d"], w["seed"], distinct=True)
    raise KeyError(name)


DISTS = ("independent", "correlated", "anticorrelated")


def batch_specs(count: int = 4096, seed: int = 4, n_lo: int = 1000,
                n_hi: int = 65536, d_lo: int = 4, d_hi: int = 10):
    """Instance list of the batched workload: (n, d, dist, seed) tuples."""
    rng = _rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), size=count))).astype(np.int64)
    ds = rng.integers(d_lo, d_hi + 1, size=count)
    kinds = rng.integers(0, 3, size=count)
    seeds = rng.integers(1, 2**31 - 1, size=count)
    return [(int(ns[i]), int(ds[i]), DISTS[int(kinds[i])], int(seeds[i])) for i in range(count)]


def batch_instance(spec) -> np.ndarray:
    n, d, dist, s = spec
    return by_name(dist, n, d, s)


def serving_stream(count: int, qps: float, seed: int = 5, median_n: int = 20000,
                   sigma: float = 1.0, n_lo: int = 500, n_hi: int = 500_000,
                   d_lo: int = 6, d_hi: int = 14):
    """Request list of the serving workload: (arrival_s, n, d, dist, seed)."""
    rng = _rng(seed)
    gaps = rng.exponential(1.0 / qps, size=count)
    t = np.cumsum(gaps) - gaps[0]
    ns = np.clip(np.rint(np.exp(rng.normal(np.log(median_n), sigma, size=count))), n_lo, n_hi).astype(np.int64)
    ds = rng.integers(d_lo, d_hi + 1, size=count)
    kinds = rng.integers(0, 3, size=count)
    seeds = rng.integers(1, 2**31 - 1, size=count)
    return [(float(t[i]), int(ns[i]), int(ds[i]), DISTS[int(kinds[i])], int(seeds[i]))
            for i in range(count)]
