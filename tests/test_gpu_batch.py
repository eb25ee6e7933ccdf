"""Batched MDMC (many instances, one pass of launches per dimensionality):
every instance bit-exact against the single-instance path and the oracle;
HIP-graph replays identical to the eager pass."""
import numpy as np
import pytest

from oracle import skyorc
from skycube_amd import datagen

pytestmark = pytest.mark.gpu


def _instances(count, seed, n_hi=3000):
    specs = datagen.batch_specs(count, seed=seed, n_lo=200, n_hi=n_hi, d_lo=2, d_hi=12)
    mats = [datagen.batch_instance(s) for s in specs]
    # a few tie-heavy instances
    rng = np.random.default_rng(seed)
    for k in range(0, count, 7):
        mats[k] = np.floor(mats[k] * 16).astype(np.float32)
    return mats


def test_batch_matches_single_and_oracle(ctx):
    mats = _instances(40, 11)
    caps = [max(1, m.shape[1] - (k % 3)) for k, m in enumerate(mats)]
    print(f"batch of {len(mats)} instances", flush=True)
    res = ctx.batch_mdmc(mats, caps)
    for k, (x, (ext, dom)) in enumerate(zip(mats, res)):
        sext, sdom = ctx.mdmc(x, caps[k])
        assert np.array_equal(ext, sext), k
        assert np.array_equal(dom, sdom), k
        if k % 5 == 0:
            oext, odom = skyorc.mdmc(x, caps[k])
            assert np.array_equal(ext, oext) and np.array_equal(dom, odom), k


def test_batch_graph_replays(ctx):
    mats = _instances(24, 12, n_hi=2000)
    eager = ctx.batch_mdmc(mats)
    for rep in range(3):
        print(f"graph replay {rep}", flush=True)
        got = ctx.batch_mdmc(mats, use_graph=True)
        for (e1, d1), (e2, d2) in zip(eager, got):
            assert np.array_equal(e1, e2) and np.array_equal(d1, d2)
        # eager work on the same context between replays
        ctx.mdmc(mats[rep], None)
        ctx.batch_mdmc(mats[:5])


def test_batch_edge_cases(ctx):
    one = np.array([[1.0, 2.0, 3.0]], dtype=np.float32)
    dup = np.full((50, 4), 2.0, dtype=np.float32)
    res = ctx.batch_mdmc([one, dup, one.copy()])
    assert res[0][0].tolist() == [0] and res[2][0].tolist() == [0]
    assert res[1][0].size == 50


def test_batch_options_do_not_change_results(ctx):
    """Extreme settings of every tuning option give the same bitmaps."""
    import skycube_amd
    mats = _instances(20, 13, n_hi=2500)
    ref = ctx.batch_mdmc(mats)
    settings = [(skycube_amd.SKY_OPT_COVER, 64), (skycube_amd.SKY_OPT_PRUNERS, 64),
                (skycube_amd.SKY_OPT_MIN_BUCKET, 1), (skycube_amd.SKY_OPT_MEMBER_GRID, 64),
                (skycube_amd.SKY_OPT_TILE, 1)]
    for order in (0, 2, 3):  # buckets (min_bucket 1), code-order groups, subgroups
        with skycube_amd.Context(0) as c:
            c.set_option(skycube_amd.SKY_OPT_BATCH_ORDER, order)
            for opt, val in settings:
                c.set_option(opt, val)
            got = c.batch_mdmc(mats)
            for k, ((e1, d1), (e2, d2)) in enumerate(zip(ref, got)):
                assert np.array_equal(e1, e2) and np.array_equal(d1, d2), (order, k)
    with skycube_amd.Context(0) as c:
        for opt, val in settings:
            c.set_option(opt, val)
            assert c.get_option(opt) == val
        for graph in (False, True):
            got = c.batch_mdmc(mats, use_graph=graph)
            for k, ((e1, d1), (e2, d2)) in enumerate(zip(ref, got)):
                assert np.array_equal(e1, e2) and np.array_equal(d1, d2), k
        # the single path under the same settings
        for k in (0, 7, 13):
            e, dm = c.mdmc(mats[k])
            assert np.array_equal(e, ref[k][0]) and np.array_equal(dm, ref[k][1]), k
        with pytest.raises(skycube_amd.SkyError):
            c.set_option(skycube_amd.SKY_OPT_COVER, 1)
        c.set_option(skycube_amd.SKY_OPT_COVER, 0)  # back to the size rule
        e, dm = c.mdmc(mats[3])
        assert np.array_equal(e, ref[3][0]) and np.array_equal(dm, ref[3][1])
        with pytest.raises(skycube_amd.SkyError):
            c.set_option(99, 1)


def test_batch_autotune(ctx):
    import skycube_amd
    mats = _instances(48, 14, n_hi=3000)
    ref = ctx.batch_mdmc(mats)
    with skycube_amd.Context(0) as c:
        chosen = c.batch_autotune(mats)
        print(f"autotune: {chosen}", flush=True)
        dims = sorted({m.shape[1] for m in mats})
        assert sorted(int(k) for k in chosen) == dims
        for d, t in chosen.items():
            assert c.get_option(skycube_amd.SKY_OPT_COVER | (int(d) << 8)) == t["cover"]
            assert t["tile"] in (0, 1) and t["grid"] in (256, 512, 1024, 2048)
            assert t["order"] in (0, 2)
        got = c.batch_mdmc(mats, use_graph=True)
        for k, ((e1, d1), (e2, d2)) in enumerate(zip(ref, got)):
            assert np.array_equal(e1, e2) and np.array_equal(d1, d2), k


def test_batch_wide_and_narrow_d(ctx):
    """d = 1 and d = 13..16 through the batched kernels (512-thread
    workgroups, the full boundary table), with a d = 16 tie-heavy instance
    whose points mostly hold more than 8 column minima: every instance
    equals the single path, a third equal the oracle, under both tile widths
    and both candidate orders."""
    import skycube_amd
    rng = np.random.default_rng(31)
    mats = []
    for d in (1, 13, 14, 15, 16, 1, 16):
        n = int(rng.integers(200, 1500))
        mats.append(rng.integers(0, 5, size=(n, d)).astype(np.float32))
        mats.append(rng.random((n, d), dtype=np.float32))
    x = rng.choice(np.array([0, 1, 2], dtype=np.float32), size=(700, 16), p=[0.72, 0.2, 0.08])
    assert ((x == x.min(axis=0)).sum(axis=1) > 8).mean() > 0.5
    mats.append(x)
    mats.append(datagen.anticorrelated(2500, 14, 77))
    caps = [None if k % 3 else max(1, m.shape[1] - 5) for k, m in enumerate(mats)]
    alone = [ctx.mdmc(m, caps[k]) for k, m in enumerate(mats)]
    for k in range(0, len(mats), 3):
        print(f"oracle instance {k}: {mats[k].shape}", flush=True)
        oe, od = skyorc.mdmc(mats[k], caps[k] or mats[k].shape[1])
        assert np.array_equal(alone[k][0], oe) and np.array_equal(alone[k][1], od), k
    for tile in (0, 1):
        for order in (0, 2, 3):
            with skycube_amd.Context(0) as c:
                c.set_option(skycube_amd.SKY_OPT_TILE, tile)
                c.set_option(skycube_amd.SKY_OPT_BATCH_ORDER, order)
                print(f"batch tile={tile} order={order}", flush=True)
                got = c.batch_mdmc(mats, [m or 0 for m in caps])
                for k, ((e1, d1), (e2, d2)) in enumerate(zip(alone, got)):
                    assert np.array_equal(e1, e2) and np.array_equal(d1, d2), (tile, order, k)


def test_batch_dev_separate_allocations(ctx):
    """Instances in separate device tensors (sky_batch_mdmc_dev) give the
    concatenated batch's results; two calls in a row, the second smaller and
    with other pointers, so the asynchronously staged descriptors of the
    first are not reused by mistake."""
    import torch
    mats = _instances(30, 21, n_hi=2500)
    caps = [max(1, m.shape[1] - (k % 4)) for k, m in enumerate(mats)]
    want = ctx.batch_mdmc(mats, caps)
    for rnd, sel in enumerate([list(range(30)), list(range(3, 30, 3))]):
        print(f"dev batch round {rnd}: {len(sel)} instances", flush=True)
        tens = [torch.from_numpy(mats[k]).to("cuda:0") for k in sel]
        torch.cuda.synchronize()
        tot = ctx.batch_mdmc_dev([t.data_ptr() for t in tens], [mats[k].shape[0] for k in sel],
                                 [mats[k].shape[1] for k in sel], [caps[k] for k in sel])
        got = ctx.batch_fetch(len(sel))
        assert tot == sum(int(want[k][0].size) for k in sel)
        for j, k in enumerate(sel):
            assert np.array_equal(got[j][0], want[k][0]), (rnd, k)
            assert np.array_equal(got[j][1], want[k][1]), (rnd, k)
        del tens
