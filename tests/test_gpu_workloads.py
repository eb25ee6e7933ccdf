"""The batched and serving configurations at full size, through the C ABI.

Batched: the 4,096-instance workload (n log-uniform in [1000, 65536], d in
4..10, independent / correlated / anticorrelated, seed 4), tuned by the
autotuner on a sample and replayed from its HIP graph: every instance equals
the single-instance path (sky_mdmc), and a 5% sample stratified by
dimensionality and distribution equals the CPU oracle.

Serving: the 600-request Poisson stream (n log-normal, median 20,000, clipped
to [500, 500,000], d in 6..14, seed 5) submitted at its arrival times to the
continuous-batching server: every result equals the same rows computed alone
(sky_mdmc), and a sample stratified by dimensionality equals the CPU oracle.
"""
import collections
import sys
import threading
import time

import numpy as np
import pytest

import skycube_amd
from oracle import skyorc
from skycube_amd import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg4():
    specs = datagen.batch_specs(4096, seed=4)
    _say("cfg4: generating 4096 instances")
    mats = [datagen.batch_instance(s) for s in specs]
    n = np.array([m.shape[0] for m in mats], dtype=np.uint32)
    d = np.array([m.shape[1] for m in mats], dtype=np.uint32)
    flat = np.ascontiguousarray(np.concatenate([m.ravel() for m in mats]))
    with skycube_amd.Context(0) as c:
        sample = np.arange(0, len(mats), 8)
        sflat = np.ascontiguousarray(np.concatenate([mats[k].ravel() for k in sample]))
        _say("cfg4: autotune on every 8th instance")
        tuned = c.batch_autotune_raw(sflat.ctypes.data, False, n[sample], d[sample])
        _say(f"cfg4: tuned {tuned}")
        _say("cfg4: graph capture + replay")
        c.batch_mdmc_raw(flat.ctypes.data, False, n, d, None, True)
        tot = c.batch_mdmc_raw(flat.ctypes.data, False, n, d, None, True)
        res = c.batch_fetch(len(mats))
    _say(f"cfg4: {tot} extended-skyline points")
    return specs, mats, res


def test_cfg4_batch_equals_single(ctx, cfg4):
    specs, mats, res = cfg4
    for k, x in enumerate(mats):
        if k % 512 == 0:
            _say(f"cfg4 single path {k}/{len(mats)}")
        e, dm = ctx.mdmc(x)
        assert np.array_equal(res[k][0], e), k
        assert np.array_equal(res[k][1], dm), k


def _strata4():
    specs = datagen.batch_specs(4096, seed=4)
    by = collections.defaultdict(list)
    for k, (n, d, dist, _) in enumerate(specs):
        by[(d, dist)].append(k)
    rng = np.random.default_rng(44)
    out = []
    for key in sorted(by):
        ks = by[key]
        pick = sorted(rng.choice(ks, size=int(np.ceil(0.05 * len(ks))), replace=False).tolist())
        out.append((key[0], key[1], pick))
    return out


@pytest.mark.parametrize("d,dist,pick", _strata4(), ids=lambda v: str(v) if not isinstance(v, list) else f"{len(v)}")
def test_cfg4_oracle_sample(cfg4, d, dist, pick):
    specs, mats, res = cfg4
    _say(f"cfg4 oracle d={d} {dist}: {len(pick)} instances")
    for k in pick:
        x = mats[k]
        oe, od = skyorc.mdmc(x, x.shape[1])
        assert np.array_equal(res[k][0], oe), k
        assert np.array_equal(res[k][1], od), k


@pytest.fixture(scope="module")
def cfg5():
    qps = 500.0
    stream = datagen.serving_stream(600, qps, seed=5)
    _say("cfg5: generating 600 requests")
    mats = [datagen.by_name(dist, n, d, sd) for (_, n, d, dist, sd) in stream]
    got = [None] * len(mats)
    ids = [None] * len(mats)
    bsz = [0] * len(mats)
    errors = []
    with skycube_amd.Server(0, workers=4) as srv:
        def collect():
            try:
                for k in range(len(mats)):
                    while ids[k] is None:
                        time.sleep(0.0002)
                    e, dm, _, b = srv.result(ids[k], timeout_ms=120000)
                    got[k] = (e, dm)
                    bsz[k] = b
            except Exception as exc:  # reported by the test
                errors.append(repr(exc))

        th = threading.Thread(target=collect, daemon=True)
        th.start()
        t0 = time.perf_counter()
        _say(f"cfg5: submitting at Poisson arrivals, {qps}/s")
        for k, (arr, *_r) in enumerate(stream):
            now = time.perf_counter() - t0
            if arr > now:
                time.sleep(arr - now)
            ids[k] = srv.submit(mats[k])
        th.join(300)
        stats = srv.stats()
    _say(f"cfg5: {stats['batches']} batches, mean batch {np.mean(bsz):.2f}, max {max(bsz)}")
    return stream, mats, got, bsz, errors


def test_cfg5_every_result_equals_alone(ctx, cfg5):
    stream, mats, got, bsz, errors = cfg5
    assert not errors, errors
    assert max(bsz) > 1  # requests shared batches
    for k, x in enumerate(mats):
        if k % 100 == 0:
            _say(f"cfg5 alone {k}/{len(mats)}")
        e, dm = ctx.mdmc(x)
        assert np.array_equal(got[k][0], e), k
        assert np.array_equal(got[k][1], dm), k


def _strata5():
    stream = datagen.serving_stream(600, 500.0, seed=5)
    by = collections.defaultdict(list)
    for k, (_, n, d, dist, _) in enumerate(stream):
        # the oracle's cost grows with the extended skyline: requests it
        # finishes in seconds (anticorrelated ones are nearly all skyline)
        lim = 30000 if dist != "anticorrelated" else (12000 if d <= 10 else 5000)
        if n <= lim:
            by[d].append(k)
    rng = np.random.default_rng(55)
    return [(d, sorted(rng.choice(by[d], size=min(8, len(by[d])), replace=False).tolist()))
            for d in sorted(by)]


@pytest.mark.parametrize("d,pick", _strata5(), ids=lambda v: str(v) if not isinstance(v, list) else f"{len(v)}")
def test_cfg5_oracle_sample(cfg5, d, pick):
    stream, mats, got, bsz, errors = cfg5
    assert not errors, errors
    for k in pick:
        x = mats[k]
        _say(f"cfg5 oracle request {k}: n={x.shape[0]} d={d} {stream[k][3]}")
        oe, od = skyorc.mdmc(x, d)
        assert np.array_equal(got[k][0], oe), k
        assert np.array_equal(got[k][1], od), k
