"""Pins the CPU oracle (oracle/skyorc.c) before anything is checked against it.

The reference ships no golden outputs and its STSC oracle cannot be built
here (it needs Boost), so the oracle is pinned by:
  * an independent numpy brute force over all pairs and subspaces,
  * agreement of its three algorithms (brute force, the STSC lattice walk,
    the per-point closure) with each other,
  * closed forms: 1-D skylines are argmin sets, 2-D skylines a sweep,
  * invariants: duplicates share bitmaps; permutations and order-preserving
    transforms of the columns leave the skycube unchanged; sky(s) is inside
    ext(s + j); every cuboid skyline is non-empty,
  * planted mistakes, which these checks must catch.
"""
import numpy as np
import pytest

import npref
from oracle import skyorc


def _in_scope_equal(a, b, d, max_d):
    S = 1 << d
    m = npref.scope(d, max_d)
    A = npref.unpack_bits(a, S)[:, m]
    B = npref.unpack_bits(b, S)[:, m]
    return np.array_equal(A, B)


CASES = [(n, d, lv, seed) for seed, (n, d, lv) in enumerate(
    [(1, 3, 4), (2, 2, 2), (7, 1, 3), (12, 2, 3), (25, 3, 3), (30, 4, 4), (40, 5, 3),
     (33, 6, 5), (24, 7, 3), (20, 8, 2), (60, 3, 1000), (50, 6, 1000)])]


# d = 9..16: fewer points, every pair still decides every subspace
WIDE = [(n, d, lv, 40 + seed) for seed, (n, d, lv) in enumerate(
    [(24, 9, 3), (22, 10, 2), (20, 11, 3), (18, 12, 2), (18, 13, 3), (16, 14, 2),
     (16, 15, 3), (16, 16, 2), (12, 16, 1000), (20, 16, 3)])]


@pytest.mark.parametrize("n,d,levels,seed", WIDE)
@pytest.mark.parametrize("cap", [None, 3, 8])
def test_wide_d_match_numpy(n, d, levels, seed, cap):
    """Brute force, the STSC walk and the per-point closure against the
    numpy statement at d = 9..16 (the word-level closures of the oracle act
    from d = 6 on, its 2^8-entry tables from d = 9 on), with and without -m."""
    x = npref.tie_heavy(np.random.default_rng(seed), n, d, levels)
    max_d = d if cap is None else min(cap, d)
    rdom, rsdom = npref.brute_rows(x)
    S = 1 << d
    m = npref.scope(d, max_d)
    dom, sdom = skyorc.brute(x)
    assert np.array_equal(npref.unpack_bits(dom, S)[:, 1:], rdom[:, 1:])
    assert np.array_equal(npref.unpack_bits(sdom, S)[:, 1:], rsdom[:, 1:])
    st_dom, st_sdom, _, _ = skyorc.stsc(x, max_d)
    assert np.array_equal(npref.unpack_bits(st_dom, S)[:, m], rdom[:, m])
    assert np.array_equal(npref.unpack_bits(st_sdom, S)[:, m], rsdom[:, m])
    ext, mdom = skyorc.mdmc(x, max_d)
    assert np.array_equal(ext, np.flatnonzero(~rsdom[:, S - 1]))
    assert np.array_equal(npref.unpack_bits(mdom, S)[:, m], rdom[ext][:, m])
    assert npref.unpack_bits(mdom, S)[:, ~m].all()


@pytest.mark.parametrize("n,d,levels,seed", CASES)
def test_brute_matches_numpy(n, d, levels, seed):
    x = npref.tie_heavy(np.random.default_rng(seed), n, d, levels)
    dom, sdom = skyorc.brute(x)
    rdom, rsdom = npref.brute(x)
    S = 1 << d
    assert np.array_equal(npref.unpack_bits(dom, S)[:, 1:], rdom[:, 1:])
    assert np.array_equal(npref.unpack_bits(sdom, S)[:, 1:], rsdom[:, 1:])


@pytest.mark.parametrize("n,d,levels,seed", CASES)
@pytest.mark.parametrize("cap", [None, 2])
def test_stsc_and_mdmc_match_brute(n, d, levels, seed, cap):
    x = npref.tie_heavy(np.random.default_rng(100 + seed), n, d, levels)
    max_d = d if cap is None else min(cap, d)
    dom, sdom = skyorc.brute(x)
    sdom_stsc = skyorc.stsc(x, max_d)
    assert _in_scope_equal(sdom_stsc[0], dom, d, max_d)
    assert _in_scope_equal(sdom_stsc[1], sdom, d, max_d)
    ext, mdom = skyorc.mdmc(x, max_d)
    full = (1 << d) - 1
    # the extended skyline is exactly the points not strictly dominated
    assert np.array_equal(ext, np.flatnonzero(npref.unpack_bits(sdom, 1 << d)[:, full] == 0))
    assert _in_scope_equal(mdom, dom[ext], d, max_d)
    # bits outside the scope are reported as "not a member"
    out = ~npref.scope(d, max_d)
    assert npref.unpack_bits(mdom, 1 << d)[:, out].all()


@pytest.mark.parametrize("seed", range(6))
def test_skyalign_matches_brute(seed):
    rng = np.random.default_rng(200 + seed)
    n, d = 40, 5
    x = npref.tie_heavy(rng, n, d, 3)
    dom, sdom = npref.brute(x)
    for sub in (1, 3, 0b10110, 31):
        st = skyorc.skyalign(x, sub)
        want = np.where(sdom[:, sub], 2, np.where(dom[:, sub], 1, 0))
        assert np.array_equal(st, want)
    # candidate subsets: only candidates compete
    cand = np.sort(rng.choice(n, size=25, replace=False)).astype(np.uint32)
    st = skyorc.skyalign(x, 0b1011, cand)
    cdom, csdom = npref.brute(x[cand])
    assert np.array_equal(st, np.where(csdom[:, 0b1011], 2, np.where(cdom[:, 0b1011], 1, 0)))


def test_closed_forms_1d_2d():
    rng = np.random.default_rng(7)
    x = rng.integers(0, 40, size=(3000, 6)).astype(np.float32)
    ext, dom = skyorc.mdmc(x)
    import skycube_amd  # members(): the getsubspace query
    for j in range(6):
        got = skycube_amd.members(dom, ext, 1 << j)
        assert np.array_equal(np.sort(got), npref.skyline_1d(x[:, j]))
    for a in range(6):
        for b in range(a + 1, 6):
            got = skycube_amd.members(dom, ext, (1 << a) | (1 << b))
            assert np.array_equal(np.sort(got), npref.skyline_2d(x[:, a], x[:, b]))


def test_invariants():
    rng = np.random.default_rng(11)
    base = rng.integers(0, 12, size=(600, 7)).astype(np.float32)
    x = np.concatenate([base, base[:50]])  # 50 exact duplicates
    ext, dom = skyorc.mdmc(x)
    pos = {int(p): i for i, p in enumerate(ext)}
    for i in range(50):
        a, b = i, 600 + i
        assert (a in pos) == (b in pos)
        if a in pos:
            assert np.array_equal(dom[pos[a]], dom[pos[b]])
    # permutation of the points
    perm = rng.permutation(x.shape[0])
    ext2, dom2 = skyorc.mdmc(x[perm])
    back = perm[ext2]
    order = np.argsort(back)
    assert np.array_equal(back[order], ext)
    assert np.array_equal(dom2[order], dom)
    # order-preserving transform of every column (dense ranks, then x*4-7)
    ext3, dom3 = skyorc.mdmc(npref.dense_rank(x) * np.float32(4) - np.float32(7))
    assert np.array_equal(ext3, ext) and np.array_equal(dom3, dom)
    # sky(s) inside ext(s + j) and every cuboid skyline non-empty
    d = 7
    sd, ssd, skyn, extn = skyorc.stsc(x, d)
    S = 1 << d
    D = npref.unpack_bits(sd, S)
    SD = npref.unpack_bits(ssd, S)
    for s in range(1, S):
        assert skyn[s] >= 1
        for j in range(d):
            if not s >> j & 1:
                t = s | 1 << j
                assert not np.any(~D[:, s] & SD[:, t])


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_planted_mistakes_are_caught(mutation):
    x = npref.tie_heavy(np.random.default_rng(5), 30, 4, 3)
    rdom, rsdom = npref.brute(x)
    skyorc.set_mutation(mutation)
    try:
        dom, sdom = skyorc.brute(x)
        ext, mdom = skyorc.mdmc(x)
    finally:
        skyorc.set_mutation(0)
    bad_brute = not (np.array_equal(npref.unpack_bits(dom, 16)[:, 1:], rdom[:, 1:])
                     and np.array_equal(npref.unpack_bits(sdom, 16)[:, 1:], rsdom[:, 1:]))
    bad_mdmc = not (np.array_equal(ext, np.flatnonzero(~rsdom[:, 15]))
                    and _in_scope_equal(mdom, npref.pack_bits(rdom)[ext], 4, 4))
    assert bad_brute and bad_mdmc


def test_score_tie_detector():
    # q dominates p in s={0,1} with equal FP32 sums: 1.0 + 2^-24 rounds to 1.0
    x = np.array([[1.0, 0.0], [1.0, np.float32(2.0 ** -25)], [3.0, 3.0]], dtype=np.float32)
    ext = skyorc.ext_skyline(x)
    assert skyorc.score_tie_count(x, ext) >= 1
    y = np.array([[1.0, 0.0], [1.0, 0.5], [3.0, 3.0]], dtype=np.float32)
    assert skyorc.score_tie_count(y, skyorc.ext_skyline(y)) == 0


def test_score_tie_resolution():
    # q = (0.5, 2^24, 0, 1) dominates p = (1, 2^24, 1, 0) in s = {0, 1} and
    # in s = {1, 2}, and both norms round to 2^24 in each; r = (0.25, 2^24 -
    # 2, 0.5, 1) dominates p in both with a smaller norm, so STSC's early
    # exit would not matter (column 3 keeps p, column 2 keeps q in the
    # extended skyline)
    big = np.float32(2.0 ** 24)
    p, q = [1.0, big, 1.0, 0.0], [0.5, big, 0.0, 1.0]
    r = [0.25, big - 2, 0.5, 1.0]
    x = np.array([p, q], dtype=np.float32)
    ext = skyorc.ext_skyline(x)
    assert skyorc.score_tie_count(x, ext, 2) >= 1
    assert skyorc.score_tie_unresolved(x, ext, 2) >= 1
    y = np.array([p, q, r], dtype=np.float32)
    ey = skyorc.ext_skyline(y)
    assert ey.tolist() == [0, 1, 2]
    assert skyorc.score_tie_count(y, ey, 2) >= 1
    assert skyorc.score_tie_unresolved(y, ey, 2) == 0


def test_nan_rejected():
    x = np.zeros((4, 3), dtype=np.float32)
    x[2, 1] = np.nan
    with pytest.raises(RuntimeError):
        skyorc.mdmc(x)


@pytest.mark.parametrize("d", [2, 5, 7, 11])
def test_infinities_match_numpy(d):
    """Rows with +inf and -inf, some holding both (a NaN Manhattan norm,
    which the oracle orders as -inf): every algorithm of the oracle against
    the numpy statement."""
    rng = np.random.default_rng(60 + d)
    n = 60 if d <= 7 else 30
    x = rng.integers(0, 4, size=(n, d)).astype(np.float32)
    x[rng.random(x.shape) < 0.08] = np.inf
    x[rng.random(x.shape) < 0.08] = -np.inf
    x[0, :2] = [-np.inf, np.inf]
    S = 1 << d
    rdom, rsdom = npref.brute_rows(x)
    dom, sdom = skyorc.brute(x)
    assert np.array_equal(npref.unpack_bits(dom, S)[:, 1:], rdom[:, 1:])
    assert np.array_equal(npref.unpack_bits(sdom, S)[:, 1:], rsdom[:, 1:])
    st_dom, st_sdom, _, _ = skyorc.stsc(x, d)
    assert np.array_equal(npref.unpack_bits(st_dom, S)[:, 1:], rdom[:, 1:])
    ext, mdom = skyorc.mdmc(x, d)
    assert np.array_equal(ext, np.flatnonzero(~rsdom[:, S - 1]))
    assert np.array_equal(npref.unpack_bits(mdom, S)[:, 1:], rdom[ext][:, 1:])
    for sub in (1, S - 1, (S - 1) & 0b10110):
        st = skyorc.skyalign(x, sub)
        assert np.array_equal(st, np.where(rsdom[:, sub], 2, np.where(rdom[:, sub], 1, 0)))
