"""Skycube server (continuous batching over the batched MDMC pass): every
request's result is identical to the same rows computed alone, whatever
batch it joined; HBM-budget admission and input rejection."""
import threading

import numpy as np
import pytest

from skycube_amd import Server, SkyError, datagen

pytestmark = pytest.mark.gpu


def _requests(count, seed, n_hi=6000):
    specs = datagen.batch_specs(count, seed=seed, n_lo=50, n_hi=n_hi, d_lo=2, d_hi=12)
    mats = [datagen.batch_instance(s) for s in specs]
    for k in range(0, count, 5):  # tie-heavy requests
        mats[k] = np.floor(mats[k] * 12).astype(np.float32)
    caps = [None if k % 4 else max(1, m.shape[1] - 2) for k, m in enumerate(mats)]
    return mats, caps


def _check_alone(ctx, mats, caps, got):
    for k, (x, (ext, dom)) in enumerate(zip(mats, got)):
        sext, sdom = ctx.mdmc(x, caps[k])
        assert np.array_equal(ext, sext), k
        assert np.array_equal(dom, sdom), k


def test_server_results_equal_alone(ctx):
    mats, caps = _requests(48, 21)
    with Server(0, workers=2) as srv:
        ids = [srv.submit(x, caps[k]) for k, x in enumerate(mats)]
        print(f"submitted {len(ids)} requests", flush=True)
        out = [srv.result(i, timeout_ms=60000) for i in ids]
        st = srv.stats()
    print(f"batches {st['batches']}, sizes {sorted({o[3] for o in out})}", flush=True)
    assert max(o[3] for o in out) > 1  # requests were served in shared batches
    assert st["batches"] < len(mats)
    assert all(o[2] > 0 for o in out)
    _check_alone(ctx, mats, caps, [(o[0], o[1]) for o in out])


def test_server_concurrent_clients(ctx):
    mats, caps = _requests(32, 22, n_hi=4000)
    got = [None] * len(mats)
    errors = []

    def client(ks):
        try:
            with_ids = [(k, srv.submit(mats[k], caps[k])) for k in ks]
            for k, i in with_ids:
                e, dm, _, _ = srv.result(i, timeout_ms=60000)
                got[k] = (e, dm)
        except Exception as exc:  # reported below
            errors.append(exc)

    with Server(0, workers=2) as srv:
        th = [threading.Thread(target=client, args=(list(range(j, len(mats), 4)),))
              for j in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join(120)
    assert not errors, errors
    _check_alone(ctx, mats, caps, got)


def test_server_budget_admission(ctx):
    mats, caps = _requests(24, 23, n_hi=8000)
    need = max(Server.footprint(x.shape[0], x.shape[1]) for x in mats)
    # a share that holds about two of the largest requests at once
    budget = 2 * (2 * need + (8 << 20))
    with Server(0, hbm_budget=budget, workers=2) as srv:
        ids = [srv.submit(x, caps[k]) for k, x in enumerate(mats)]
        out = [srv.result(i, timeout_ms=60000) for i in ids]
        big = np.random.default_rng(0).random((200000, 16), dtype=np.float32)
        with pytest.raises(SkyError) as ei:
            srv.submit(big)
        assert ei.value.code == -3
        st = srv.stats()
    print(f"budget {budget} B: {st['batches']} batches", flush=True)
    assert st["batches"] >= len(mats) // 8
    _check_alone(ctx, mats, caps, [(o[0], o[1]) for o in out])


def test_server_rejects_bad_input():
    with Server(0, workers=1) as srv:
        x = np.ones((10, 3), dtype=np.float32)
        x[4, 1] = np.nan
        with pytest.raises(SkyError) as ei:
            srv.submit(x)
        assert ei.value.code == -2
        with pytest.raises(SkyError) as ei:
            srv.submit(np.ones((5, 17), dtype=np.float32))
        assert ei.value.code == -1
        # empty and single-point requests
        e0, d0, _, _ = srv.result(srv.submit(np.zeros((0, 4), dtype=np.float32)), 30000)
        e1, d1, _, _ = srv.result(srv.submit(np.ones((1, 4), dtype=np.float32)), 30000)
        assert e0.size == 0 and e1.tolist() == [0]
        with pytest.raises(SkyError):
            srv.result(987654321, timeout_ms=10)


def test_server_drains_on_close(ctx):
    mats, caps = _requests(12, 24, n_hi=3000)
    srv = Server(0, workers=2)
    ids = [srv.submit(x) for x in mats]
    first = srv.result(ids[0], 60000)
    srv.close()  # waits for the queued requests
    sext, sdom = ctx.mdmc(mats[0])
    assert np.array_equal(first[0], sext) and np.array_equal(first[1], sdom)
