"""Parity of the HIP path with the CPU oracle, through the C ABI.

Bit-exact: the ext pids and every membership word must be identical (bits of
out-of-scope subspaces are 1 on both sides).
"""
import json
import os
import sys

import numpy as np
import pytest

import npref
import skycube_amd
from oracle import skyorc
from skycube_amd import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_mdmc(ctx, x, max_d=None, tag=""):
    n, d = x.shape
    _say(f"mdmc parity {tag} n={n} d={d} max_d={max_d}")
    ext, dom = ctx.mdmc(x, max_d)
    oext, odom = skyorc.mdmc(x, d if max_d is None else max_d)
    assert np.array_equal(ext, oext), f"ext differs: {ext.size} vs {oext.size}"
    if not np.array_equal(dom, odom):
        bad = np.flatnonzero((dom != odom).any(axis=1))
        raise AssertionError(f"{bad.size} of {dom.shape[0]} bitmaps differ (first pid {ext[bad[0]]})")
    return ext, dom


@pytest.mark.parametrize("d", list(range(1, 17)))
def test_mdmc_every_d_ties(ctx, d):
    rng = np.random.default_rng(1000 + d)
    n = 400 if d <= 12 else 150
    check_mdmc(ctx, npref.tie_heavy(rng, n, d, 3), tag="ties")


@pytest.mark.parametrize("order,max_d", [(-1, None), (2, 9), (3, None)])
def test_mdmc_d16_rare_ties(order, max_d):
    """d = 16 data that is tie-free but for a few planted pairs: the
    membership runs without its tie lattices, and the points whose exact
    tests meet a residual tie (q equal to p in one column, below it in the
    rest) go to the full kernel."""
    rng = np.random.default_rng(2600 + order)
    x = rng.random((2500, 16), dtype=np.float32) * np.float32(0.5) + np.float32(0.25)
    p = rng.choice(2500, size=80, replace=False)
    q = x[p] - np.float32(1e-3)
    cols = rng.integers(1, 16, size=80)
    q[np.arange(80), cols] = x[p, cols]  # tied in one column ...
    q[:, 0] = x[p, 0]  # ... and in column 0
    x = np.concatenate([x, q]).astype(np.float32)
    with skycube_amd.Context(0) as c:
        c.set_option(skycube_amd.SKY_OPT_MEMBER_INDEX, order)
        check_mdmc(c, x, max_d, tag=f"rare ties order {order}")


@pytest.mark.parametrize("d", [2, 5, 8, 11, 14, 16])
def test_mdmc_every_d_continuous(ctx, d):
    rng = np.random.default_rng(2000 + d)
    check_mdmc(ctx, rng.random((700, d), dtype=np.float32), tag="uniform")


@pytest.mark.parametrize("d", [3, 4, 6, 8, 10, 12, 13, 14, 15, 16])
def test_mdmc_large_ext_with_ties(ctx, d):
    """More than 2048 extended-skyline points, so the code-pruned scan runs,
    with ties (anticorrelated values rounded to a coarse grid)."""
    x = datagen.anticorrelated(7000, d, 6000 + d)
    x = np.floor(x * 40).astype(np.float32)
    ext, _ = check_mdmc(ctx, x, tag="anti-ties")
    assert ext.size > 2048 or d <= 4


@pytest.mark.parametrize("max_d", [16, 8])
def test_mdmc_d16_many_minimum_columns(ctx, max_d):
    """d = 16 with most points at the column minimum in more than 8 columns
    (values 0, 1, 2 with 0 likely): the membership launch defers them to its
    second pass with the full boundary table; both passes must agree with
    the oracle, as must the points computed in the first."""
    rng = np.random.default_rng(16016)
    x = rng.choice(np.array([0, 1, 2], dtype=np.float32), size=(900, 16), p=[0.72, 0.2, 0.08])
    x[:300] += rng.random((300, 16), dtype=np.float32) * 0.5  # fewer ties in a third
    at_min = (x == x.min(axis=0)).sum(axis=1)
    assert (at_min > 8).sum() > 100 and (at_min <= 8).sum() > 100
    check_mdmc(ctx, x, max_d, tag="d16-minima")


@pytest.mark.parametrize("max_d", [16, 8])
def test_mdmc_d16_tie_free_with_minima(ctx, max_d):
    """d = 16, continuous values clipped at 0: the only ties are at the
    column minima (boundary columns), so the launch leaves out its tie
    lattices, while the many points with more than 8 columns at 0 are listed
    before it and computed beside it with the full boundary table."""
    rng = np.random.default_rng(16017)
    x = np.maximum(rng.standard_normal((1500, 16)).astype(np.float32), np.float32(0.0))
    at_min = (x == 0).sum(axis=1)
    assert (at_min > 8).sum() > 100 and (at_min <= 8).sum() > 100
    check_mdmc(ctx, x, max_d, tag="d16-tie-free-minima")


@pytest.mark.parametrize("d,max_d", [(6, 1), (9, 3), (12, 5), (16, 4)])
def test_mdmc_max_d(ctx, d, max_d):
    rng = np.random.default_rng(3000 + d)
    check_mdmc(ctx, rng.integers(0, 6, size=(500, d)).astype(np.float32), max_d, tag="cap")


def test_mdmc_edge_cases(ctx):
    # empty input
    ext, dom = ctx.mdmc(np.zeros((0, 4), dtype=np.float32))
    assert ext.size == 0 and dom.shape == (0, 1)
    # a single point is in every cuboid skyline
    ext, dom = ctx.mdmc(np.array([[0.5, 0.25, 1.0]], dtype=np.float32))
    assert ext.tolist() == [0]
    assert np.array_equal(dom[0], ~skycube_amd.in_scope_mask(3))
    # all points equal: all are members everywhere
    x = np.full((300, 7), 3.0, dtype=np.float32)
    ext, dom = ctx.mdmc(x)
    assert ext.size == 300
    assert (dom == ~skycube_amd.in_scope_mask(7)[None, :]).all()
    # -0.0 and +0.0 are equal values
    x = np.array([[0.0, 1.0], [-0.0, 1.0], [1.0, 0.0]], dtype=np.float32)
    check_mdmc(ctx, x, tag="signed zero")
    # negative and large magnitudes
    rng = np.random.default_rng(9)
    x = (rng.standard_normal((500, 6)) * 1e6).astype(np.float32)
    check_mdmc(ctx, x, tag="wide range")


def test_mdmc_rejects_nan(ctx):
    x = np.zeros((10, 3), dtype=np.float32)
    x[3, 2] = np.nan
    with pytest.raises(skycube_amd.SkyError) as e:
        ctx.mdmc(x)
    assert e.value.code == -2


def test_mdmc_rejects_bad_d(ctx):
    with pytest.raises(skycube_amd.SkyError):
        ctx.mdmc(np.zeros((4, 17), dtype=np.float32))


def test_mdmc_workload_independent_full(ctx):
    """The full independent workload: n=100,000, d=8, all 255 cuboids."""
    x = datagen.workload("independent_distinct")
    ext, _ = check_mdmc(ctx, x, tag="independent_distinct full")
    assert ext.size == 9984


def test_mdmc_workload_anticorrelated_sample(ctx):
    check_mdmc(ctx, datagen.workload("anticorrelated_plane", 20000), tag="anticorrelated 20k")


def test_mdmc_workload_integer_ties_sample(ctx):
    check_mdmc(ctx, datagen.workload("integer_ties", 60000), max_d=8, tag="integer_ties 60k")


def test_mdmc_batch_mix(ctx):
    """A few instances of the batched workload, one call each."""
    for spec in datagen.batch_specs(24, seed=4)[:24:3]:
        n, d, dist, s = spec
        check_mdmc(ctx, datagen.batch_instance((min(n, 6000), d, dist, s)), tag=f"{dist}")


@pytest.mark.parametrize("seed", range(4))
def test_skyalign_matches_oracle(ctx, seed):
    rng = np.random.default_rng(4000 + seed)
    n, d = 3000, 7
    x = rng.integers(0, 9, size=(n, d)).astype(np.float32)
    for sub in (1, 0b11, 0b1010101, 0b1111111, int(rng.integers(1, 128))):
        for cand in (None, np.sort(rng.choice(n, size=1200, replace=False)).astype(np.uint32)):
            _say(f"skyalign seed={seed} sub={sub:#x} cand={'all' if cand is None else cand.size}")
            sky, ext = ctx.skyalign(x, sub, cand)
            st = skyorc.skyalign(x, sub, cand)
            ids = np.arange(n, dtype=np.uint32) if cand is None else cand
            assert np.array_equal(sky, np.sort(ids[st == 0]))
            assert np.array_equal(ext, np.sort(ids[st == 1]))


@pytest.mark.parametrize("n,d,max_d", [(300, 5, 5), (400, 6, 3), (250, 8, 8)])
def test_sdsc_matches_stsc(ctx, n, d, max_d):
    rng = np.random.default_rng(5000 + d)
    x = rng.integers(0, 7, size=(n, d)).astype(np.float32)
    _say(f"sdsc n={n} d={d} max_d={max_d}")
    so, sky, eo, ext = ctx.sdsc(x, max_d)
    dom, sdom, skyn, extn = skyorc.stsc(x, max_d)
    S = 1 << d
    D = npref.unpack_bits(dom, S)
    SD = npref.unpack_bits(sdom, S)
    for s in range(1, S):
        if bin(s).count("1") > max_d and s != S - 1:
            assert so[s + 1] == so[s] and eo[s + 1] == eo[s]
            continue
        got_sky = sky[so[s]:so[s + 1]]
        got_ext = ext[eo[s]:eo[s + 1]]
        assert np.array_equal(got_sky, np.flatnonzero(~D[:, s])), s
        assert np.array_equal(got_ext, np.flatnonzero(D[:, s] & ~SD[:, s])), s


@pytest.mark.parametrize("d,max_d,n", [(3, 3, 500), (6, 6, 2000), (9, 4, 3000), (12, 12, 1500)])
def test_hashcube_export(ctx, d, max_d, n):
    """Device-grouped HashCube equals the host grouping of the bitmaps, and
    its query gives every cuboid skyline."""
    x = datagen.integer_ties(n, d, 7000 + d)
    _say(f"hashcube d={d} max_d={max_d} n={n}")
    ext, dom = ctx.mdmc(x, max_d)
    gw, gv, go, ids = ctx.mdmc_hashcube()
    hw, hv, ho, hids = skycube_amd.hashcube_from_bitmaps(ext, dom, d)
    assert np.array_equal(gw, hw) and np.array_equal(gv, hv)
    assert np.array_equal(go, ho) and np.array_equal(ids, hids)
    scope = npref.scope(d, max_d)
    for s in np.flatnonzero(scope)[:200]:
        assert np.array_equal(skycube_amd.hashcube_query(gw, gv, go, ids, s),
                              np.sort(skycube_amd.members(dom, ext, s)))


@pytest.mark.parametrize("d,max_d,n", [(4, 4, 700), (10, 6, 2500), (16, 8, 1500)])
def test_hashcube_export_single_call(ctx, d, max_d, n):
    """The export into caller buffers sized in advance (one call: the ids
    start copying as soon as they are sorted, and the per-row entry counts
    come from the membership kernel) equals the host grouping."""
    x = datagen.integer_ties(n, d, 7100 + d)
    _say(f"hashcube single call d={d} max_d={max_d} n={n}")
    ext, dom = ctx.mdmc(x, max_d)
    W = dom.shape[1]
    cap_e = len(ext) * (W + 1) + 1
    gw = np.zeros(cap_e + 1, dtype=np.uint32)
    gv = np.zeros(cap_e + 1, dtype=np.uint32)
    go = np.zeros(cap_e + 2, dtype=np.uint64)
    ids = np.zeros(cap_e, dtype=np.uint32)
    g, t = ctx.hashcube_into(gw.ctypes.data, gv.ctypes.data, go.ctypes.data, ids.ctypes.data,
                             cap_e + 1, cap_e)
    hw, hv, ho, hids = skycube_amd.hashcube_from_bitmaps(ext, dom, d)
    assert np.array_equal(gw[:g], hw) and np.array_equal(gv[:g], hv)
    assert np.array_equal(go[:g + 1], ho) and np.array_equal(ids[:t], hids)
    # the copies again, from the groups already built
    ids2 = np.zeros(cap_e, dtype=np.uint32)
    assert ctx.hashcube_into(gw.ctypes.data, gv.ctypes.data, go.ctypes.data, ids2.ctypes.data,
                             cap_e + 1, cap_e) == (g, t)
    assert np.array_equal(ids2[:t], hids)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("d", [3, 6, 9, 12, 14, 16])
def test_mdmc_member_candidate_order(d, mode):
    """Every membership candidate order (score-order scan, median-mask
    buckets, code-order groups, groups narrowed to subgroups) gives the
    oracle's bitmaps, with ties and without."""
    rng = np.random.default_rng(2000 + 17 * d + mode)
    n = 3000 if d <= 12 else 1200
    with skycube_amd.Context(0) as c:
        c.set_option(skycube_amd.SKY_OPT_MEMBER_INDEX, mode)
        check_mdmc(c, npref.tie_heavy(rng, n, d, 4), tag=f"order {mode} ties")
        x = rng.random((n, d), dtype=np.float32)
        check_mdmc(c, x, max(1, d - 3), tag=f"order {mode} continuous")
        check_mdmc(c, datagen.anticorrelated(n, d, 50 + d), tag=f"order {mode} anticorrelated")


def test_mdmc_independent_16d_sampled(ctx):
    """n = 1M, d = 16 independent data: the extended skyline's size is the
    oracle's (its full run is in the golden digests, checked word for word
    by test_golden.py), and sampled points are checked here directly:
    sampled points of the extended skyline have no strict dominator and
    their bitmaps equal the oracle's; sampled points outside it have one."""
    w = datagen.WORKLOADS["independent_16d"]
    x = datagen.workload("independent_16d")
    _say(f"independent_16d n={x.shape[0]} d={x.shape[1]} max_d={w['max_d']}")
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "golden.json")))["full"]["independent_16d"]
    ext, dom = ctx.mdmc(x, w["max_d"])
    assert ext.size == gold["n_ext"]
    rng = np.random.default_rng(7)
    smp = np.sort(rng.choice(ext.size, size=48, replace=False))
    odom = skyorc.mdmc_sample(x, ext, smp, w["max_d"])
    assert np.array_equal(dom[smp], odom)
    inside = np.zeros(x.shape[0], dtype=bool)
    inside[ext] = True
    for i in ext[smp[:16]]:
        assert not np.all(x < x[i], axis=1).any(), i
    for i in rng.choice(np.flatnonzero(~inside), size=16, replace=False):
        assert np.all(x < x[i], axis=1).any(), i


def test_infinities(ctx):
    """Rows with +inf and -inf, some holding both (a NaN Manhattan norm): a
    weak dominator such as (-inf, +inf) of p = (0, +inf) must be found in
    the skyalign and SDSC paths too, whose survivor pass cuts large buckets
    at the point's norm."""
    rng = np.random.default_rng(77)
    x = rng.integers(0, 6, size=(2000, 2)).astype(np.float32)
    x[:40, 1] = np.inf
    x[40] = [-np.inf, np.inf]
    x[41] = [0.0, np.inf]
    x[42] = [np.inf, -np.inf]
    sky, ext = ctx.skyalign(x, 0b11)
    st = skyorc.skyalign(x, 0b11)
    assert np.array_equal(sky, np.flatnonzero(st == 0))
    assert np.array_equal(ext, np.flatnonzero(st == 1))
    assert 41 not in sky.tolist()
    check_mdmc(ctx, x, tag="inf 2d")
    y = rng.integers(0, 4, size=(900, 7)).astype(np.float32)
    y[rng.random(y.shape) < 0.04] = np.inf
    y[rng.random(y.shape) < 0.04] = -np.inf
    check_mdmc(ctx, y, tag="inf 7d")
    check_mdmc(ctx, y, 3, tag="inf 7d cap")
    so, sk, eo, ex = ctx.sdsc(y[:300], 4)
    dom, sdom, _, _ = skyorc.stsc(y[:300], 4)
    D = npref.unpack_bits(dom, 128)
    SD = npref.unpack_bits(sdom, 128)
    for s in range(1, 128):
        if bin(s).count("1") > 4 and s != 127:
            continue
        assert np.array_equal(sk[so[s]:so[s + 1]], np.flatnonzero(~D[:, s])), s
        assert np.array_equal(ex[eo[s]:eo[s + 1]], np.flatnonzero(D[:, s] & ~SD[:, s])), s


@pytest.mark.parametrize("n,d,max_d", [(3000, 12, 12), (2500, 13, 6), (2000, 14, 5),
                                       (1500, 16, 4), (1200, 15, 15)])
def test_sdsc_wide_d_matches_stsc(ctx, n, d, max_d):
    """The device lattice walk at d = 12..16, with and without -m caps,
    against the oracle's STSC walk: every cuboid's sky and ext lists."""
    rng = np.random.default_rng(7000 + d + max_d)
    x = datagen.anticorrelated(n, d, 70 + d) if d == 12 else \
        rng.integers(0, 5, size=(n, d)).astype(np.float32)
    _say(f"sdsc n={n} d={d} max_d={max_d}")
    so, sky, eo, ext = ctx.sdsc(x, max_d)
    dom, sdom, _, _ = skyorc.stsc(x, max_d)
    S = 1 << d
    D = npref.unpack_bits(dom, S)
    SD = npref.unpack_bits(sdom, S)
    for s in range(1, S):
        if bin(s).count("1") > max_d and s != S - 1:
            assert so[s + 1] == so[s] and eo[s + 1] == eo[s], s
            continue
        assert np.array_equal(sky[so[s]:so[s + 1]], np.flatnonzero(~D[:, s])), s
        assert np.array_equal(ext[eo[s]:eo[s + 1]], np.flatnonzero(D[:, s] & ~SD[:, s])), s


def test_sdsc_anticorrelated_full_is_golden(ctx):
    """The full anticorrelated workload (n = 500,000, d = 12, 4,095 cuboids)
    through the device SDSC walk: the skyline sizes hash to the golden value
    and the skylines, turned into per-point bitmaps, hash to the golden
    digest of the oracle's MDMC result."""
    import hashlib
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "golden.json")))["full"]["anticorrelated_plane"]
    x = datagen.workload("anticorrelated_plane")
    _say("sdsc anticorrelated_plane n=500000 d=12")
    so, sky, eo, ext = ctx.sdsc(x, 12)
    S = 1 << 12
    sizes = np.zeros(S, dtype=np.int64)
    sizes[1:] = np.diff(so.astype(np.int64))[1:S]
    assert int(sizes[1:].sum()) == gold["sky_size_total"]
    assert hashlib.sha256(sizes.astype("<i8").tobytes()).hexdigest() == gold["sky_sizes_sha256"]
    full = S - 1
    pts = np.sort(np.concatenate([sky[so[full]:so[full + 1]], ext[eo[full]:eo[full + 1]]]))
    assert pts.size == gold["n_ext"]
    pos = np.full(x.shape[0], -1, dtype=np.int64)
    pos[pts] = np.arange(pts.size)
    dom = np.full((pts.size, S // 32), 0xFFFFFFFF, dtype=np.uint32)
    for s in range(1, S):
        ids = sky[so[s]:so[s + 1]]
        dom[pos[ids], s >> 5] &= np.uint32(~(1 << (s & 31)) & 0xFFFFFFFF)
    h = hashlib.sha256(pts.astype("<u4").tobytes() + dom.astype("<u4").tobytes()).hexdigest()
    assert h == gold["sha256"]


def test_pinned_host_buffers(ctx):
    """Rows read from, and ids written to, page-locked memory from
    sky_host_alloc give the oracle's result, and the buffers free cleanly."""
    import ctypes
    L = skycube_amd.load()
    L.sky_host_alloc.restype = ctypes.c_void_p
    L.sky_host_alloc.argtypes = [ctypes.c_size_t]
    L.sky_host_free.argtypes = [ctypes.c_void_p]
    L.sky_host_free.restype = None
    rng = np.random.default_rng(77)
    x = npref.tie_heavy(rng, 4000, 10, 6)
    n, d = x.shape
    prow = L.sky_host_alloc(x.nbytes)
    pext = L.sky_host_alloc(4 * n)
    assert prow and pext
    try:
        rows = np.ctypeslib.as_array((ctypes.c_float * x.size).from_address(prow)).reshape(n, d)
        rows[:] = x
        ext_buf = np.ctypeslib.as_array((ctypes.c_uint32 * n).from_address(pext))
        _say("pinned host rows and ids")
        ne = ctx.mdmc_raw(prow, n, d, 7, pext)
        ext = ext_buf[:ne].copy()
        _, dom = ctx.mdmc_fetch(ne, d)
        oext, odom = skyorc.mdmc(x, 7)
        assert np.array_equal(ext, oext) and np.array_equal(dom, odom)
    finally:
        L.sky_host_free(prow)
        L.sky_host_free(pext)
    L.sky_host_free(None)  # a null pointer is ignored
