"""Benchmark of the skycube engine on one MI355X: one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                  [--no-batch] [--no-serve] [--no-sdsc] [--no-cpu-baseline]

The headline (`value`) is ms per skycube of integer_ties, n = 1,000,000,
d = 16, subspaces of at most 8 dimensions (39,202 cuboids; integers 0..255
stored as FP32, seed 3), timed over the reference's own window
(src/main.h:324-336: the rows already in host memory after Init, then
Execute -- rows up to the device, the skycube, the HashCube materialised in
host memory).  Like the reference's Init, which copies the rows into its own
flat buffer, the rows are copied once, untimed, into pinned host memory; the
HashCube lands in pinned host buffers.  A step is one such window; W steps
warm up, K are timed.  The same line carries:

  stages_ms       device-only time of the skycube (rows already in HBM,
                  bitmaps left there) and its stages, from HIP events
  roofline        the membership kernel k_member against HBM bandwidth
  batch           4,096 instances (n log-uniform 1k..65k, d 4..10, mixed
                  distributions, seed 4): instances/s of the batched pass
                  replayed from its HIP graph, against one plain MDMC call
                  per instance on the same GPU
  serve           600 requests (Poisson arrivals, n log-normal median 20k,
                  d 6..14, seed 5) through the continuous-batching server
                  at about 70% busy: p50 / p99 latency, requests/s
  sdsc            the device SDSC lattice walk, ms, on the independent
                  (n = 100,000, d = 8) and anticorrelated (n = 500,000,
                  d = 12) workloads
  cpu_baseline    the reference's own CPU path (STSC: MommyQ per cuboid,
                  built from the reference sources into oracle/_ref) timed on
                  this host on a bounded sample of the same workload and
                  scaled up (see ref_cpu_baseline)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "skycube-templates_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "ms/skycube (n=1M,d=16) vs CPU; batched inst/s vs plain MDMC; p99 latency"
HBM_PEAK_GBS = 8000.0


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def cuboids_in_scope(d: int, max_d: int) -> int:
    return sum(math.comb(d, k) for k in range(1, min(d, max_d) + 1))


def member_bytes(n_ext: int, d: int, max_d: int) -> int:
    """Algorithmic bytes of one membership launch: each extended-skyline
    point's row read once (4d B), its id and its C-bit membership mask
    written once (4 + ceil(C/8) B)."""
    C = cuboids_in_scope(d, max_d)
    return n_ext * (4 * d + 4 + (C + 7) // 8)


# --------------------------------------------------------------------------
# headline: the reference's Execute() window

def bench_window(args, torch, skycube_amd, datagen, w, x):
    n, d, max_d = w["n"], w["d"], w["max_d"]
    rows_pin = torch.from_numpy(x).pin_memory()  # the rows after Init
    ext_pin = torch.empty(n, dtype=torch.int32).pin_memory()
    ctx = skycube_amd.Context(0)
    say("window: warm-up")
    ctx.mdmc_raw(rows_pin.data_ptr(), n, d, max_d, ext_pin.data_ptr())
    ng, nt = ctx.hashcube_sizes()
    cap_g, cap_e = ng + ng // 8 + 16, nt + nt // 8 + 16
    gw = torch.empty(cap_g, dtype=torch.int32).pin_memory()
    gv = torch.empty(cap_g, dtype=torch.int32).pin_memory()
    go = torch.empty(cap_g + 1, dtype=torch.int64).pin_memory()
    ids = torch.empty(cap_e, dtype=torch.int32).pin_memory()

    def step():
        ne = ctx.mdmc_raw(rows_pin.data_ptr(), n, d, max_d, ext_pin.data_ptr())
        g, e = ctx.hashcube_into(gw.data_ptr(), gv.data_ptr(), go.data_ptr(), ids.data_ptr(),
                                 cap_g, cap_e)
        return ne, g, e

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ne, g, e = step()
        times.append(time.perf_counter() - t0)
    torch.cuda.synchronize()
    # the parts of the window, one more time with the HashCube timed alone
    t0 = time.perf_counter()
    ctx.mdmc_raw(rows_pin.data_ptr(), n, d, max_d, ext_pin.data_ptr())
    t1 = time.perf_counter()
    ctx.hashcube_into(gw.data_ptr(), gv.data_ptr(), go.data_ptr(), ids.data_ptr(), cap_g, cap_e)
    t2 = time.perf_counter()
    ctx.close()
    return {"ms": float(np.mean(times)) * 1e3, "ms_min": float(np.min(times)) * 1e3,
            "n_ext": ne, "groups": g, "entries": e,
            "parts_ms": {"rows_up_skycube_ids_down": round((t1 - t0) * 1e3, 3),
                         "hashcube_to_host": round((t2 - t1) * 1e3, 3)}}


def bench_device(args, torch, skycube_amd, w, x):
    """The skycube alone: rows resident in HBM, bitmaps left there."""
    n, d, max_d = w["n"], w["d"], w["max_d"]
    xd = torch.from_numpy(x).to("cuda")
    torch.cuda.synchronize()
    ctx = skycube_amd.Context(0)
    ctx.set_timing(True)
    for _ in range(max(1, args.warmup)):
        ctx.mdmc_device(xd.data_ptr(), n, d, max_d)
    torch.cuda.synchronize()
    stages = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n_ext = ctx.mdmc_device(xd.data_ptr(), n, d, max_d)
        stages.append(ctx.stage_times())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ext, _ = ctx.mdmc_fetch(n_ext, d)
    ctx.close()
    st = {k: round(float(np.mean([s[k] for s in stages])), 3) for k in stages[0]}
    return ms, n_ext, st, ext


# --------------------------------------------------------------------------
# batched and serving configurations

def bench_batch(args, torch, skycube_amd, datagen):
    """4,096 instances, rows resident in HBM: (a) one plain MDMC call per
    instance, (b) the batched pass replayed from its HIP graph after the
    autotuner (on every 8th instance) picked the settings."""
    specs = datagen.batch_specs(args.instances, seed=4)
    say(f"batch: generating {len(specs)} instances")
    mats = [datagen.batch_instance(s) for s in specs]
    n = np.array([m.shape[0] for m in mats], dtype=np.uint32)
    d = np.array([m.shape[1] for m in mats], dtype=np.uint32)
    flat = np.concatenate([m.ravel() for m in mats])
    xd = torch.from_numpy(flat).to("cuda")
    torch.cuda.synchronize()
    base = xd.data_ptr()
    foff = np.concatenate([[0], np.cumsum(n.astype(np.int64) * d)])
    ctx = skycube_amd.Context(0)
    sample = np.arange(0, len(mats), 8)
    sflat = np.concatenate([mats[k].ravel() for k in sample])
    t0 = time.perf_counter()
    tuned = ctx.batch_autotune_raw(sflat.ctypes.data, False, n[sample], d[sample])
    t_tune = time.perf_counter() - t0
    plain_ctx = skycube_amd.Context(0)
    for t in range(min(8, len(mats))):
        plain_ctx.mdmc_device(base + 4 * int(foff[t]), int(n[t]), int(d[t]))
    torch.cuda.synchronize()
    say("batch: plain calls")
    t0 = time.perf_counter()
    for t in range(len(mats)):
        plain_ctx.mdmc_device(base + 4 * int(foff[t]), int(n[t]), int(d[t]))
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain_ctx.close()
    res = {}
    for label, graph in (("eager", False), ("graph", True)):
        say(f"batch: {label}")
        ctx.batch_mdmc_raw(base, True, n, d, None, graph)
        torch.cuda.synchronize()
        times = []
        for _ in range(args.batch_steps):
            t0 = time.perf_counter()
            tot = ctx.batch_mdmc_raw(base, True, n, d, None, graph)
            times.append(time.perf_counter() - t0)
        res[label] = (float(np.median(times)), tot)
    ctx.close()
    ips = {k: len(mats) / v[0] for k, v in res.items()}
    return {
        "value": round(ips["graph"], 1), "unit": "instances/s",
        "workload": (f"{len(mats)} instances, n log-uniform [1000, 65536], d in 4..10, "
                     f"independent / correlated / anticorrelated, seed 4, rows resident in HBM"),
        "points": int(n.sum()), "ext_points": res["graph"][1],
        "ms_per_batch": round(res["graph"][0] * 1e3, 2),
        "plain_mdmc": {"instances_per_s": round(len(mats) / t_plain, 1),
                       "ms_total": round(t_plain * 1e3, 1)},
        "batched_eager": {"instances_per_s": round(ips["eager"], 1),
                          "ms_total": round(res["eager"][0] * 1e3, 1)},
        "speedup_vs_plain": round(ips["graph"] * t_plain / len(mats), 2),
        "autotune": {"sample": f"every 8th instance ({sample.size})",
                     "seconds": round(t_tune, 2), "per_d": tuned},
    }


def _pct(a, q):
    return round(float(np.percentile(np.asarray(a), q)), 3) if len(a) else None


def _serve_stream(srv, mats, stream, keep_every):
    """Submits mats[k] at stream[k]'s arrival time while a collector thread
    takes the results in order: per-request latency (arrival -> result in
    host memory, ms), batch sizes, kept results, wall seconds, errors."""
    import threading

    count = len(mats)
    lat, lag, bsz = [None] * count, [0.0] * count, [0] * count
    ids, kept, errors = [None] * count, {}, []
    done = threading.Event()

    def collector():
        try:
            for k in range(count):
                while ids[k] is None:
                    time.sleep(0.0002)
                e, dm, ms, b = srv.result(ids[k])
                lat[k] = ms + lag[k]
                bsz[k] = b
                if keep_every and k % keep_every == 0:
                    kept[k] = (e, dm)
        except Exception as exc:  # reported in the JSON line
            errors.append(repr(exc))
        done.set()

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for k, (arr, *_rest) in enumerate(stream):
        now = time.perf_counter() - t0
        if arr > now:
            time.sleep(arr - now)
        lag[k] = max(0.0, (time.perf_counter() - t0 - arr) * 1e3)
        ids[k] = srv.submit(mats[k])
    done.wait()
    return lat, bsz, kept, time.perf_counter() - t0, errors


def bench_serve(args, torch, skycube_amd, datagen):
    """600 requests submitted from the host at Poisson arrival times to the
    continuous-batching server; the rate is searched (at most three runs) so
    that the server is busy about 70% of the time.  Latency = arrival ->
    result in host memory.  Every 10th result of the last run is checked
    against the same rows computed alone."""
    count = args.requests
    sizes = datagen.serving_stream(count, 1.0, seed=5)
    say(f"serve: generating {count} requests")
    mats = [datagen.by_name(dist, n, d, sd) for (_, n, d, dist, sd) in sizes]
    srv = skycube_amd.Server(0, hbm_budget=0, workers=args.workers)
    for i in [srv.submit(m) for m in mats]:  # warm-up pass
        srv.result(i)
    t0 = time.perf_counter()
    ids = [srv.submit(m) for m in mats]
    for i in ids:
        srv.result(i)
    t_closed = time.perf_counter() - t0
    capacity = count / t_closed
    qps = args.qps or 0.7 * capacity
    runs = []
    for attempt in range(1 if args.qps else 3):
        stream = datagen.serving_stream(count, qps, seed=5)
        st0 = srv.stats()
        lat, bsz, kept, t_open, errors = _serve_stream(srv, mats, stream, 10)
        st1 = srv.stats()
        busy = (st1["busy_ms"] - st0["busy_ms"]) / (t_open * 1e3)
        runs.append({"qps": round(qps, 1), "busy": round(busy, 3)})
        say(f"serve: run {attempt} at {qps:.0f}/s, busy {busy:.2f}")
        if args.qps or 0.6 <= busy <= 0.8 or errors:
            break
        qps = qps * min(4.0, 0.7 / max(busy, 0.05))
    srv.close()
    # the same stream at the same offered rate on a server with another
    # worker count (fewer workers: larger batches, longer waits)
    compare = None
    if args.compare_workers and args.compare_workers != args.workers and not errors:
        alt = skycube_amd.Server(0, hbm_budget=0, workers=args.compare_workers)
        for i in [alt.submit(m) for m in mats]:  # warm-up pass
            alt.result(i)
        st0 = alt.stats()
        lat2, bsz2, _, t2, err2 = _serve_stream(alt, mats, stream, 0)
        st1 = alt.stats()
        alt.close()
        good2 = [v for v in lat2 if v is not None]
        say(f"serve: {args.compare_workers} workers at {qps:.0f}/s")
        compare = {"workers": args.compare_workers, "offered_qps": round(qps, 1),
                   "latency_ms": {"p50": _pct(good2, 50), "p99": _pct(good2, 99),
                                  "max": _pct(good2, 100)},
                   "batch_size": {"mean": round(float(np.mean(bsz2)), 2), "max": int(max(bsz2))},
                   "server_busy_frac": round((st1["busy_ms"] - st0["busy_ms"]) / (t2 * 1e3), 3),
                   "errors": err2 or None}
    ok = 0
    with skycube_amd.Context(0) as ctx:
        for k, (e0, d0) in kept.items():
            e, dm = ctx.mdmc(mats[k])
            ok += int(np.array_equal(e, e0) and np.array_equal(dm, d0))
    good = [v for v in lat if v is not None]
    return {
        "value": _pct(good, 99), "unit": "ms (p99 latency)",
        "workload": (f"{count} requests, Poisson arrivals, n log-normal (median 20000, sigma 1.0, "
                     f"[500, 500000]), d in 6..14, mixed distributions, seed 5"),
        "latency_ms": {"p50": _pct(good, 50), "p90": _pct(good, 90), "p99": _pct(good, 99),
                       "max": _pct(good, 100), "mean": round(float(np.mean(good)), 3)},
        "requests_per_s": round(count / t_open, 1), "offered_qps": round(qps, 1),
        "closed_loop_capacity_per_s": round(capacity, 1), "server_busy_frac": runs[-1]["busy"],
        "rate_search": runs, "workers": args.workers,
        "batch_size": {"mean": round(float(np.mean(bsz)), 2), "max": int(max(bsz))},
        "verified_alone": f"{ok}/{len(kept)} sampled requests identical to the same rows alone",
        "errors": errors or None,
        "same_load_other_workers": compare,
    }


def bench_sdsc(args, torch, skycube_amd, datagen):
    """The device SDSC walk (rows resident in HBM, lattice left there)."""
    out = {}
    with skycube_amd.Context(0) as ctx:
        for wl in ("independent_distinct", "anticorrelated_plane"):
            w = datagen.WORKLOADS[wl]
            x = datagen.workload(wl)
            xd = torch.from_numpy(x).to("cuda")
            torch.cuda.synchronize()
            say(f"sdsc: {wl}")
            ctx.sdsc_device(xd.data_ptr(), w["n"], w["d"], w["max_d"])  # warm-up
            t0 = time.perf_counter()
            ctx.sdsc_device(xd.data_ptr(), w["n"], w["d"], w["max_d"])
            torch.cuda.synchronize()
            out[wl] = {"ms": round((time.perf_counter() - t0) * 1e3, 2),
                       "cuboids": cuboids_in_scope(w["d"], w["max_d"]),
                       "n": w["n"], "d": w["d"]}
            del xd
    return out


# --------------------------------------------------------------------------
# CPU baseline: the reference's STSC (MommyQ per cuboid), sampled

def ref_cpu_baseline(skycube_amd, x, max_d, threads, per_level=None):
    """The reference's CPU path (its STSC walk: MommyQ per cuboid, from the
    reference sources in oracle/_ref) timed on this host with `threads`
    threads, on a sample, scaled up:

      * the full space over all points, in full (MommyQ with every thread);
      * for each level L <= max_d, `threads` cuboids drawn at random, run
        together as the reference runs a level (one thread per cuboid), each
        over the candidate list the reference gives it: the full space's
        sky + ext where its parents were skipped (L = max_d < d), else the
        sky + ext of its parent with the fewest (the lists of those parents
        come from the GPU, one SkyAlign over the full space's sky + ext per
        parent);
      * level time = sum of the sampled cuboids' seconds x (cuboids of the
        level / sampled) / threads.
    """
    from oracle import skyref

    n, d = x.shape
    S = 1 << d
    full = S - 1
    rng = np.random.default_rng(0)
    sky, ext, t_full = skyref.cuboid(x, full, None, threads=threads)
    E = np.sort(np.concatenate([sky, ext])).astype(np.uint32)
    xe = np.ascontiguousarray(x[E])
    parts = {"full_space_s": round(t_full, 3)}
    total = t_full
    sample_wall = t_full
    ctx = skycube_amd.Context(0)
    memo = {}

    def lists(p):  # sky + ext of cuboid p over all points, ascending ids
        if p not in memo:
            s_, e_ = ctx.skyalign(xe, p)
            memo[p] = np.sort(E[np.concatenate([s_, e_])])
        return memo[p]

    for L in range(min(d - 1, max_d), 0, -1):
        subs = [s for s in range(1, S) if bin(s).count("1") == L]
        k = min(len(subs), per_level or threads)
        pick = [subs[i] for i in sorted(rng.choice(len(subs), size=k, replace=False))]
        cands = []
        for s in pick:
            if L + 1 > max_d and L + 1 < d:  # the parents' level was skipped
                cands.append(E)
                continue
            parents = [s | (1 << j) for j in range(d - 1, -1, -1) if not (s >> j) & 1]
            if len(parents) == 1 and parents[0] == full:
                cands.append(E)
                continue
            best = min(parents, key=lambda p: (lists(p).size, -p))
            cands.append(lists(best))
        wall, each = skyref.cuboids_timed(x, pick, cands, threads)
        sample_wall += wall
        est = float(each.sum()) * len(subs) / k / threads
        parts[f"level_{L}"] = {"cuboids": len(subs), "sampled": k,
                               "mean_candidates": int(np.mean([c.size for c in cands])),
                               "est_s": round(est, 2)}
        total += est
    ctx.close()
    return total, sample_wall, parts


# --------------------------------------------------------------------------

def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="integer_ties")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-sdsc", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-device", action="store_true",
                    help="only the device-resident skycube loop (for profiling)")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--requests", type=int, default=600)
    ap.add_argument("--qps", type=float, default=0.0)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--compare-workers", type=int, default=2,
                    help="also run the serving stream at the same rate with this many workers "
                         "(0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: every core of this host)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: every run here uses a single MI355X "
              f"(got --gpus {args.gpus}); the multi-GPU sharding is a design note, not a run.",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(max(1, args.steps)):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": max(1, args.steps)}))
        return 0

    import skycube_amd
    from skycube_amd import datagen

    w = datagen.WORKLOADS[args.workload]
    n, d, max_d = w["n"], w["d"], w["max_d"]
    x = datagen.workload(args.workload)
    dev_ms, n_ext, st, ext = bench_device(args, torch, skycube_amd, w, x)
    if args.only_device:
        print(json.dumps({"device_ms": round(dev_ms, 3), "stages_ms": st, "n_ext": n_ext}))
        return 0
    win = bench_window(args, torch, skycube_amd, datagen, w, x)
    member_ms = st["member"]
    B = member_bytes(n_ext, d, max_d)
    achieved = B / (member_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_member_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(args.workload)
    out = {
        "metric": METRIC, "value": round(win["ms"], 3), "unit": "ms/skycube", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(win["ms"], 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} n={n} d={d} max_d={max_d} seed={w['seed']}",
                   "cuboids": cuboids_in_scope(d, max_d), "n_ext": n_ext,
                   "window": ("the reference's Execute() window: rows in pinned host memory "
                              "(copied there untimed, as Init copies them) -> device -> skycube "
                              "-> HashCube (word, value) -> ids grouped on the device and "
                              "copied to pinned host memory"),
                   "hashcube": {"groups": win["groups"], "entries": win["entries"]}},
        "window_parts_ms": win["parts_ms"],
        "stages_ms": dict(st, device_total=round(dev_ms, 3)),
        "roofline": {"bound": "hbm",
                     "kernel": (f"k_member<d={d}>: the membership stage (its main launch and, "
                                "at d = 16, the short launch of its deferred points)"),
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "algorithmic_bytes": B},
    }
    if not args.no_batch:
        out["batch"] = bench_batch(args, torch, skycube_amd, datagen)
    if not args.no_serve:
        out["serve"] = bench_serve(args, torch, skycube_amd, datagen)
    if not args.no_sdsc:
        out["sdsc"] = bench_sdsc(args, torch, skycube_amd, datagen)
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        from oracle import skyref
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        if skyref.available():
            say(f"cpu baseline: reference STSC sample, {threads} threads")
            t0 = time.perf_counter()
            total, wall, parts = ref_cpu_baseline(skycube_amd, x, max_d, threads)
            out["cpu_baseline"] = {
                "value": round(total * 1e3, 1), "unit": "ms/skycube", "cores": threads,
                "kind": "reference",
                "sample": (f"the reference's STSC (MommyQ per cuboid, built from its sources) on "
                           f"the same input with {threads} threads: the full space in full, "
                           f"{threads} random cuboids per level with the reference's candidate "
                           f"lists, each level scaled by its cuboid count; "
                           f"{wall:.1f} s of sampled work"),
                "parts": parts, "sample_seconds": round(time.perf_counter() - t0, 1),
            }
            out["speedup_vs_cpu_baseline"] = round(total * 1e3 / win["ms"], 1)
            # the sampling estimator against a full run, on an input small
            # enough to run in full in seconds: the first 100,000 points of
            # the d = 16 independent workload, subspaces of at most 8
            # dimensions (39,202 cuboids)
            w1 = datagen.WORKLOADS["independent_16d"]
            x1 = datagen.workload("independent_16d", 100_000)
            say("cpu baseline: full reference run on 100,000 points of independent_16d")
            full_s = float(skyref.stsc(x1, w1["max_d"], threads=threads)[4].sum())  # the walk's levels
            est1, _, _ = ref_cpu_baseline(skycube_amd, x1, w1["max_d"], threads)
            out["cpu_baseline"]["sampling_check"] = {
                "workload": "independent_16d, first 100000 points, d=16 max_d=8 (39202 cuboids)",
                "full_run_ms": round(full_s * 1e3, 1), "sampled_estimate_ms": round(est1 * 1e3, 1),
                "estimate_over_full": round(est1 / full_s, 3)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
