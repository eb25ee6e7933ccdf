"""Benchmark: one skycube per step on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                  [--no-cpu-baseline] [--verify]

A step is one full skycube (MDMC: extended skyline + per-point cuboid
membership bitmaps) of the named workload, with the input already resident
in HBM and the result left in HBM.  Default workload: integer_ties, n =
1,000,000 points, d = 16, subspaces of at most 8 dimensions (39,202 cuboids),
integer values 0..255 stored as FP32 (seed 3).

Prints one JSON line: the headline metric (ms per skycube, lower is better),
the roofline of the dominant kernel (membership bitmaps) from HIP events on
the library's own stream, and the CPU baseline: the oracle port timed on
this host's cores on a bounded sample of the same workload and scaled up.
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
# wave64 VALU instructions per second the chip can issue: 256 CUs x 4 SIMDs,
# one instruction per SIMD every 4 clocks, at 2.4 GHz
VALU_PEAK_IPS = 256 * 4 * 2.4e9 / 4


def cuboids_in_scope(d: int, max_d: int) -> int:
    return sum(math.comb(d, k) for k in range(1, min(d, max_d) + 1))


def member_bytes(n_ext: int, d: int, max_d: int) -> int:
    """Algorithmic bytes of one membership launch: each extended-skyline
    point's row read once (4d B), its id and its C-bit membership mask
    written once (4 + ceil(C/8) B)."""
    C = cuboids_in_scope(d, max_d)
    return n_ext * (4 * d + 4 + (C + 7) // 8)


def cpu_baseline(x: np.ndarray, max_d: int, ext: np.ndarray, target_s: float = 8.0) -> dict:
    """The oracle port on this host: the extended skyline's prefilter in full
    and its survivor pass on every k-th survivor, then membership bitmaps of
    a random sample of extended-skyline points; each sampled part is scaled
    by its sampling factor."""
    from oracle import skyorc

    n, d = x.shape
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    # membership: calibrate the sample size on a small probe first
    probe = np.sort(rng.choice(ext.size, size=min(ext.size, 64), replace=False))
    t = time.perf_counter()
    skyorc.mdmc_sample(x, ext, probe, max_d)
    per_pt = (time.perf_counter() - t) / max(probe.size, 1)
    k = int(min(ext.size, max(64, target_s / max(per_pt, 1e-9))))
    smp = np.sort(rng.choice(ext.size, size=k, replace=False))
    t = time.perf_counter()
    skyorc.mdmc_sample(x, ext, smp, max_d)
    t_member = (time.perf_counter() - t) * ext.size / k
    # extended skyline: prefilter in full, survivor pass sampled
    stride = max(1, int(ext.size / max(1, k)))
    pr, se, ns = skyorc.ext_skyline_timed(x, stride)
    t_ext = pr + se * stride
    total_ms = (t_ext + t_member) * 1e3
    return {
        "value": round(total_ms, 1), "unit": "ms/skycube", "cores": threads, "kind": "port",
        "sample": (f"CPU oracle (C, OpenMP) on the same input: extended-skyline prefilter over "
                   f"all {n} points, survivor pass on every {stride}th of {ns} survivors "
                   f"(x{stride}), membership of {k} of {ext.size} extended-skyline points "
                   f"(x{ext.size / k:.1f}); {threads} threads"),
        "parts_ms": {"ext_skyline": round(t_ext * 1e3, 1), "membership": round(t_member * 1e3, 1)},
    }


def bench_batch(args, torch, skycube_amd, datagen) -> int:
    """Batched workload: instances of log-uniform n in [1000, 65536], d in
    {4..10}, independent / correlated / anticorrelated (seed 4).  Timed with
    the rows resident in HBM: (a) one plain MDMC call per instance, (b) the
    batched pass, (c) the batched pass replayed from its HIP graph."""
    specs = datagen.batch_specs(args.instances, seed=4)
    mats = [datagen.batch_instance(s) for s in specs]
    n = np.array([m.shape[0] for m in mats], dtype=np.uint32)
    d = np.array([m.shape[1] for m in mats], dtype=np.uint32)
    flat = np.concatenate([m.ravel() for m in mats])
    xd = torch.from_numpy(flat).to("cuda")
    torch.cuda.synchronize()
    base = xd.data_ptr()
    foff = np.concatenate([[0], np.cumsum(n.astype(np.int64) * d)])
    ctx = skycube_amd.Context(0)
    # autotuner on a sample of the batch (every 8th instance); its settings
    # (per dimensionality) are then used for the batched runs below
    sample = np.arange(0, len(mats), 8)
    sflat = np.concatenate([mats[k].ravel() for k in sample])
    t0 = time.perf_counter()
    tuned = ctx.batch_autotune_raw(sflat.ctypes.data, False, n[sample], d[sample])
    t_tune = time.perf_counter() - t0
    plain_ctx = skycube_amd.Context(0)  # default settings for the plain path
    # (a) plain: one MDMC per instance, rows already on the device
    for t in range(min(8, len(mats))):
        plain_ctx.mdmc_device(base + 4 * int(foff[t]), int(n[t]), int(d[t]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_ext = 0
    for t in range(len(mats)):
        plain_ext += plain_ctx.mdmc_device(base + 4 * int(foff[t]), int(n[t]), int(d[t]))
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain_ctx.close()
    # (b) batched, eager; (c) batched, graph replay
    res = {}
    for label, graph in (("eager", False), ("graph", True)):
        for _ in range(max(1, args.warmup)):
            ctx.batch_mdmc_raw(base, True, n, d, None, graph)
        torch.cuda.synchronize()
        times = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            tot = ctx.batch_mdmc_raw(base, True, n, d, None, graph)
            times.append(time.perf_counter() - t0)
        res[label] = (float(np.median(times)), tot)
    ctx.close()
    ips = {k: len(mats) / v[0] for k, v in res.items()}
    out = {
        "metric": METRIC, "value": round(ips["graph"], 1), "unit": "instances/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(res["graph"][0] * 1e3, 3), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched: {len(mats)} instances, n log-uniform [1000, 65536], "
                               f"d in 4..10, mixed distributions, seed 4",
                   "points": int(n.sum()), "ext_points": res["graph"][1]},
        "plain_mdmc": {"instances_per_s": round(len(mats) / t_plain, 1),
                       "ms_total": round(t_plain * 1e3, 1), "ext_points": plain_ext},
        "batched_eager": {"instances_per_s": round(ips["eager"], 1),
                          "ms_total": round(res["eager"][0] * 1e3, 1)},
        "batched_graph": {"instances_per_s": round(ips["graph"], 1),
                          "ms_total": round(res["graph"][0] * 1e3, 1)},
        "speedup_vs_plain": round(ips["graph"] * t_plain / len(mats), 2),
        "autotune": {"sample": f"every 8th instance ({sample.size})",
                     "seconds": round(t_tune, 2), "per_d": tuned},
    }
    print(json.dumps(out), flush=True)
    return 0


def _pct(a, q):
    """q-th percentile of a, rounded (None if empty)."""
    return round(float(np.percentile(np.asarray(a), q)), 3) if len(a) else None


def _serve_stream(srv, mats, stream, keep_every):
    """Submits mats[k] at stream[k]'s arrival time from this thread while a
    collector thread takes the results in order.  Returns per-request
    latency (arrival -> result in host memory, ms), batch sizes, the kept
    results, wall seconds and errors."""
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


def bench_serve(args, torch, skycube_amd, datagen) -> int:
    """Serving workload: a Poisson stream of skycube requests, n log-normal
    (median 20,000, sigma 1.0, clipped to [500, 500,000]), d in 6..14, mixed
    distributions (seed 5), submitted from the host at their arrival times
    to the continuous-batching server.  The arrival rate is searched (at most
    three runs) so that the server is busy about 70% of the time, unless
    --qps is given.  Latency = arrival -> result in host memory."""
    count = args.requests
    sizes = datagen.serving_stream(count, 1.0, seed=5)
    mats = [datagen.by_name(dist, n, d, sd) for (_, n, d, dist, sd) in sizes]
    pts = int(sum(m.shape[0] for m in mats))
    srv = skycube_amd.Server(0, hbm_budget=0, workers=args.workers)
    # warm-up: one closed-loop pass (kernels, device buffers, pinned slabs)
    for i in [srv.submit(m) for m in mats]:
        srv.result(i)
    # closed loop: every request submitted at once
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
        lat, bsz, kept, t_open, errors = _serve_stream(srv, mats, stream, args.verify_every)
        st1 = srv.stats()
        busy = (st1["busy_ms"] - st0["busy_ms"]) / (t_open * 1e3)
        runs.append({"qps": round(qps, 1), "busy": round(busy, 3)})
        print(json.dumps({"serve_run": attempt, "offered_qps": round(qps, 1),
                          "server_busy_frac": round(busy, 3)}), file=sys.stderr, flush=True)
        if args.qps or 0.6 <= busy <= 0.8 or errors:
            break
        qps = qps * min(4.0, 0.7 / max(busy, 0.05))
    srv.close()
    # the sampled results of the last run, each recomputed alone on a plain
    # context, must match
    verified = None
    if kept:
        ctx = skycube_amd.Context(0)
        ok = 0
        for k, (e0, d0) in kept.items():
            e, dm = ctx.mdmc(mats[k])
            ok += int(np.array_equal(e, e0) and np.array_equal(dm, d0))
        ctx.close()
        verified = f"{ok}/{len(kept)} sampled requests identical to the same rows computed alone"
    good = [v for v in lat if v is not None]
    out = {
        "metric": METRIC, "value": _pct(good, 99), "unit": "ms (p99 latency)", "n_gpus": 1,
        "steps": count, "warmup": count, "ms_per_step": round(t_open * 1e3 / count, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"serving: {count} requests, Poisson arrivals at {qps:.1f}/s, "
                               f"n log-normal (median 20000, sigma 1.0, [500, 500000]), "
                               f"d in 6..14, mixed distributions, seed 5",
                   "points": pts, "workers": args.workers},
        "latency_ms": {"p50": _pct(good, 50), "p90": _pct(good, 90), "p99": _pct(good, 99),
                       "max": _pct(good, 100), "mean": round(float(np.mean(good)), 3)},
        "requests_per_s": round(count / t_open, 1),
        "closed_loop_capacity_per_s": round(capacity, 1),
        "offered_qps": round(qps, 1),
        "server_busy_frac": runs[-1]["busy"],
        "rate_search": runs,
        "batch_size": {"mean": round(float(np.mean(bsz)), 2), "max": int(max(bsz))},
        "verified": verified, "errors": errors or None,
    }
    print(json.dumps(out), flush=True)
    return 0 if not errors else 1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="integer_ties")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hashcube", action="store_true",
                    help="single mode: skip the HashCube export timing")
    ap.add_argument("--verify", action="store_true", help="sha256 of the result vs the golden digest")
    ap.add_argument("--mode", default="single", choices=["single", "batch", "serve"],
                    help="single: one skycube per step (the headline); batch: the batched "
                         "workload; serve: the Poisson request stream")
    ap.add_argument("--instances", type=int, default=4096, help="batch mode: instance count")
    ap.add_argument("--requests", type=int, default=600, help="serve mode: request count")
    ap.add_argument("--qps", type=float, default=0.0, help="serve mode: arrival rate "
                    "(default 0.7 x the closed-loop capacity)")
    ap.add_argument("--workers", type=int, default=2, help="serve mode: server workers")
    ap.add_argument("--verify-every", type=int, default=10,
                    help="serve mode: check every k-th request against a plain call (0: off)")
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

    if args.mode == "batch":
        return bench_batch(args, torch, skycube_amd, datagen)
    if args.mode == "serve":
        return bench_serve(args, torch, skycube_amd, datagen)

    w = datagen.WORKLOADS[args.workload]
    n, d, max_d = w["n"], w["d"], w["max_d"]
    x = datagen.workload(args.workload)
    xd = torch.from_numpy(x).to("cuda")
    torch.cuda.synchronize()
    ctx = skycube_amd.Context(0)
    ctx.set_timing(True)
    ptr = xd.data_ptr()
    for _ in range(args.warmup):
        ctx.mdmc_device(ptr, n, d, max_d)
    torch.cuda.synchronize()
    stages = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n_ext = ctx.mdmc_device(ptr, n, d, max_d)
        stages.append(ctx.stage_times())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    st = {k: round(float(np.mean([s[k] for s in stages])), 3) for k in stages[0]}
    member_ms = st["member"]
    B = member_bytes(n_ext, d, max_d)
    achieved = B / (member_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_member_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(args.workload)
    # the kernel is bound by instruction issue, not bytes: its VALU
    # instruction count per launch (rocprofv3, profiles/) over the live time
    valu = None
    cpath = os.path.join(ROOT, "profiles", "k_member_counters.json")
    if os.path.exists(cpath):
        cnt = json.load(open(cpath)).get(args.workload)
        if cnt:
            ips = cnt["valu_insts"] / (member_ms * 1e-3)
            valu = {"bound": "valu", "achieved": round(ips / 1e9, 1), "peak": VALU_PEAK_IPS / 1e9,
                    "unit": "G wave64 VALU instructions/s", "frac": round(ips / VALU_PEAK_IPS, 3),
                    "insts_per_launch": cnt["valu_insts"], "source": "profiles/k_member_counters.json"}
    out = {
        "metric": METRIC, "value": round(ms_step, 3), "unit": "ms/skycube", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} n={n} d={d} max_d={max_d} seed={w['seed']}",
                   "cuboids": cuboids_in_scope(d, max_d), "n_ext": n_ext,
                   "result": "device-resident membership bitmaps (2^d bits per ext point)"},
        "stages_ms": st,
        "roofline": {"bound": "hbm", "kernel": f"k_member<d={d}> (code-order groups)",
                     "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic, "algorithmic_bytes": B},
        "valu_roofline": valu,
        # implementation-independent: ext-skyline pairs per second (pruning
        # can take it past any peak; for comparing implementations only)
        "ext_pair_rate_per_s": float("%.4g" % (n_ext * n_ext / (ms_step * 1e-3))),
    }
    # the same skycube through host buffers (sky_mdmc: rows up, ids and
    # bitmaps down over PCIe into pageable memory), for reference only
    t0 = time.perf_counter()
    h_ext, h_dom = ctx.mdmc(x, max_d)
    out["host_buffers_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
    out["host_buffers_bytes"] = int(x.nbytes + h_ext.nbytes + h_dom.nbytes)
    del h_ext, h_dom
    if not args.no_hashcube:
        # the reference's timed window (src/main.h:324-336: rows already in
        # host memory, result materialised as a HashCube): rows up, the
        # skycube, the HashCube grouped on the device and copied to host
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xd2 = torch.from_numpy(x).to("cuda")
        torch.cuda.synchronize()
        ctx.mdmc_device(xd2.data_ptr(), n, d, max_d)
        t1 = time.perf_counter()
        gw, gv, go, ids = ctx.mdmc_hashcube()
        t2 = time.perf_counter()
        out["reference_window_ms"] = round((t2 - t0) * 1e3, 1)
        out["hashcube"] = {"ms": round((t2 - t1) * 1e3, 1), "groups": int(gw.size),
                           "entries": int(ids.size),
                           "what": "(word, value) -> ids grouped on the device, copied to host"}
        del xd2, gw, gv, go, ids
    if args.verify:
        import hashlib
        gold = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
        ext, dom = ctx.mdmc_fetch(n_ext, d)
        h = hashlib.sha256(ext.astype("<u4").tobytes() + dom.astype("<u4").tobytes()).hexdigest()
        out["verified"] = h == gold.get("full", {}).get(args.workload, {}).get("sha256")
    if not args.no_cpu_baseline:
        ext, _ = ctx.mdmc_fetch(n_ext, d) if "ext" not in locals() else (ext, None)
        cb = cpu_baseline(x, max_d, ext)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_baseline"] = round(cb["value"] / ms_step, 1)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
