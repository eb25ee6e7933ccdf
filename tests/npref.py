"""Independent numpy statements of the skycube, used to check the C oracle.

Written from the definition, not from the oracle's code:
  q dominates p in s  iff  s is a subset of {i: q_i <= p_i} and s meets
  {i: q_i < p_i};  strictly iff s is a subset of {i: q_i < p_i}.
Bitmaps use the reference's layout: subspace s is bit s%32 of word s//32
(src/skycube/NodeAuxilliary.h:58-60).
"""
from __future__ import annotations

import numpy as np


def words_per_point(d: int) -> int:
    return 1 << (d - 5) if d >= 5 else 1


def pack_bits(bits: np.ndarray) -> np.ndarray:
    """bool [n, S] -> uint32 [n, W] with bit s of row i = bits[i, s]."""
    n, S = bits.shape
    W = max(1, S // 32)
    full = np.zeros((n, W * 32), dtype=np.uint64)
    full[:, :S] = bits
    full = full.reshape(n, W, 32) << np.arange(32, dtype=np.uint64)
    return full.sum(axis=2).astype(np.uint32)


def unpack_bits(words: np.ndarray, S: int) -> np.ndarray:
    words = np.asarray(words, dtype=np.uint64)
    n, W = words.shape
    b = (words[:, :, None] >> np.arange(32, dtype=np.uint64)) & 1
    return b.reshape(n, W * 32)[:, :S].astype(bool)


def brute(x: np.ndarray):
    """(dominated [n, 2^d], strictly dominated [n, 2^d]) as bool arrays."""
    x = np.asarray(x, dtype=np.float32)
    n, d = x.shape
    S = 1 << d
    q = x[None, :, :]
    p = x[:, None, :]
    w = (1 << np.arange(d)).astype(np.int64)
    lt = ((q < p) * w).sum(axis=2)          # [p, q]
    le = ((q <= p) * w).sum(axis=2)
    np.fill_diagonal(lt, 0)
    np.fill_diagonal(le, 0)
    s = np.arange(S, dtype=np.int64)[None, None, :]
    le3 = le[:, :, None]
    lt3 = lt[:, :, None]
    dom = (((s & ~le3) == 0) & ((s & lt3) != 0)).any(axis=1)
    sdom = (((s & ~lt3) == 0) & (s != 0)).any(axis=1)
    return dom, sdom


def brute_rows(x: np.ndarray):
    """The same as brute(), one point at a time (memory O(n * 2^d)), for
    d up to 16."""
    x = np.asarray(x, dtype=np.float32)
    n, d = x.shape
    S = 1 << d
    w = (1 << np.arange(d)).astype(np.int64)
    s = np.arange(S, dtype=np.int64)[None, :]
    dom = np.zeros((n, S), dtype=bool)
    sdom = np.zeros((n, S), dtype=bool)
    for i in range(n):
        others = np.delete(x, i, axis=0)
        lt = ((others < x[i]) * w).sum(axis=1)[:, None]
        le = ((others <= x[i]) * w).sum(axis=1)[:, None]
        dom[i] = (((s & ~le) == 0) & ((s & lt) != 0)).any(axis=0)
        sdom[i] = (((s & ~lt) == 0) & (s != 0)).any(axis=0)
    return dom, sdom


def scope(d: int, max_d: int) -> np.ndarray:
    s = np.arange(1 << d)
    pc = np.array([bin(v).count("1") for v in s])
    return (s >= 1) & (pc <= max_d)


def skyline_1d(col: np.ndarray) -> np.ndarray:
    """Points of a 1-D cuboid skyline: every point holding the minimum."""
    return np.flatnonzero(col == col.min())


def skyline_2d(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """2-D skyline by a sweep in ascending a: p survives iff its b is below
    every b of a strictly smaller a, and minimal among equal a."""
    order = np.lexsort((b, a))
    keep = []
    best_before = np.inf
    i = 0
    n = a.size
    while i < n:
        j = i
        while j < n and a[order[j]] == a[order[i]]:
            j += 1
        grp = order[i:j]
        gmin = b[grp].min()
        for k in grp:
            if b[k] == gmin and b[k] < best_before:
                keep.append(k)
        best_before = min(best_before, gmin)
        i = j
    return np.sort(np.array(keep, dtype=np.int64))


def dense_rank(x: np.ndarray) -> np.ndarray:
    """Column-wise dense ranks as float32: order and ties preserved."""
    out = np.empty_like(x, dtype=np.float32)
    for j in range(x.shape[1]):
        _, inv = np.unique(x[:, j], return_inverse=True)
        out[:, j] = inv.astype(np.float32)
    return out


def tie_heavy(rng: np.random.Generator, n: int, d: int, levels: int = 4) -> np.ndarray:
    return rng.integers(0, levels, size=(n, d)).astype(np.float32)
