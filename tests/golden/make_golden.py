"""Writes the golden fixtures in tests/golden/ from the CPU oracle.

  python tests/golden/make_golden.py            NBA + seeded samples
  python tests/golden/make_golden.py --full     + full-size workload digests
  python tests/golden/make_golden.py --full-only NAME [NAME ...]
                                                only the full-size digests of
                                                the named workloads (kept
                                                entries of the others stay)
  python tests/golden/make_golden.py --copy-nba copies the NBA input from the
                                                reference checkout (once)

Inputs:
  nba-U-8-17264.csv.gz  the reference's only dataset (workloads/nba-U-8-17264
                        .csv, 17,264 points x 8 columns, many per-column ties),
                        gzipped; parsed with C strtof like the reference's
                        istringstream >> float (src/utilities/utilities.h:81-102)
  seeded samples        regenerated from skycube_amd.datagen at test time

Every expected output is computed by the oracle's per-point closure and, where
the size allows, checked here against its STSC lattice walk; the script stops
if the two disagree.  It also records how many (q, p, s) triples could
trigger STSC's FP32 score early exit (src/hybrid/mommyq.cpp:104) -- zero means
STSC itself computes exactly these sets.
"""
from __future__ import annotations

import argparse
import ctypes
import gzip
import hashlib
import json
import os
import shutil
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "skycube-templates_amd"))

from oracle import skyorc  # noqa: E402
from skycube_amd import datagen  # noqa: E402

NBA_GZ = os.path.join(HERE, "nba-U-8-17264.csv.gz")
REF_NBA = "/root/reference/workloads/nba-U-8-17264.csv"

# (name, workload, n, max_d) of the seeded samples kept in full
SAMPLES = [
    ("independent_5k", "independent_distinct", 5000, 8),
    ("anticorrelated_3k", "anticorrelated_plane", 3000, 12),
    ("integer_ties_8k", "integer_ties", 8000, 8),
    ("integer_ties_8k_m3", "integer_ties", 8000, 3),
]


def read_csv_strtof(path: str) -> np.ndarray:
    """Rows of a reference-format CSV (comma separated, optional trailing
    comma), each field converted with C strtof (correctly rounded)."""
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
    opener = gzip.open if path.endswith(".gz") else open
    rows = []
    with opener(path, "rb") as f:
        for line in f:
            fields = [t for t in line.strip().split(b",") if t != b""]
            if not fields:
                continue
            rows.append([libc.strtof(t, None) for t in fields])
    return np.array(rows, dtype=np.float32)


def sky_sizes(dom: np.ndarray, d: int) -> np.ndarray:
    """Per-cuboid skyline sizes: clear bits per subspace."""
    W = dom.shape[1] if dom.ndim == 2 else 1
    sizes = np.zeros(W * 32, dtype=np.int64)
    for b in range(32):
        sizes[b::32] = ((dom >> np.uint32(b)) & 1 == 0).sum(axis=0)
    return sizes[: 1 << d]


def digest(ext: np.ndarray, dom: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(ext, dtype="<u4").tobytes())
    h.update(np.ascontiguousarray(dom, dtype="<u4").tobytes())
    return h.hexdigest()


def cross_check_stsc(x: np.ndarray, max_d: int, ext: np.ndarray, dom: np.ndarray) -> None:
    n, d = x.shape
    sdom, _, _, _ = skyorc.stsc(x, max_d)
    S = 1 << d
    scope = np.array([1 <= s and bin(s).count("1") <= max_d for s in range(S)])
    from npref import unpack_bits  # tests/npref.py

    a = unpack_bits(sdom, S)[:, scope]
    b = np.ones((n, int(scope.sum())), dtype=bool)
    b[ext] = unpack_bits(dom, S)[:, scope]
    if not np.array_equal(a, b):
        raise SystemExit("oracle algorithms disagree: STSC walk vs per-point closure")


def make_nba() -> dict:
    x = read_csv_strtof(NBA_GZ)
    t = time.time()
    ext, dom = skyorc.mdmc(x, 8)
    cross_check_stsc(x, 8, ext, dom)
    ties = skyorc.score_tie_count(x, ext, 8)
    np.savez_compressed(os.path.join(HERE, "nba_mdmc.npz"), ext=ext, dom=dom,
                        sky_sizes=sky_sizes(dom, 8))
    print(f"nba: n={x.shape[0]} n_ext={ext.size} score_ties={ties} ({time.time() - t:.1f}s)")
    return {"n": int(x.shape[0]), "d": 8, "max_d": 8, "n_ext": int(ext.size),
            "sha256": digest(ext, dom), "score_tie_triples": ties}


def make_samples() -> dict:
    out, meta = {}, {}
    for name, wl, n, max_d in SAMPLES:
        x = datagen.workload(wl, n)
        ext, dom = skyorc.mdmc(x, max_d)
        cross_check_stsc(x, max_d, ext, dom)
        ties = skyorc.score_tie_count(x, ext, max_d)
        out[f"{name}_ext"] = ext
        out[f"{name}_dom"] = dom
        meta[name] = {"workload": wl, "n": n, "max_d": max_d, "n_ext": int(ext.size),
                      "sha256": digest(ext, dom), "score_tie_triples": ties}
        print(f"{name}: n_ext={ext.size} score_ties={ties}", flush=True)
    np.savez_compressed(os.path.join(HERE, "samples_mdmc.npz"), **out)
    return meta


def make_full(names=None) -> dict:
    meta = {}
    for wl, w in datagen.WORKLOADS.items():
        if names and wl not in names:
            continue
        t = time.time()
        x = datagen.workload(wl)
        ext, dom = skyorc.mdmc(x, w["max_d"])
        t1 = time.time() - t
        ties = skyorc.score_tie_count(x, ext, w["max_d"])
        # triples without a smaller-norm dominator (about an hour on 8 cores
        # for the 379 triples of the anticorrelated workload)
        unres = skyorc.score_tie_unresolved(x, ext, w["max_d"]) if ties else 0
        sizes = sky_sizes(dom, w["d"])
        meta[wl] = {"n": w["n"], "d": w["d"], "max_d": w["max_d"], "n_ext": int(ext.size),
                    "sha256": digest(ext, dom), "score_tie_triples": ties,
                    "score_tie_unresolved": unres,
                    "sky_sizes_sha256": hashlib.sha256(sizes.astype("<i8").tobytes()).hexdigest(),
                    "sky_size_total": int(sizes[1:].sum()), "oracle_seconds": round(t1, 1)}
        print(f"{wl}: {meta[wl]}", flush=True)
    return meta


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--copy-nba", action="store_true")
    ap.add_argument("--full-only", nargs="+", metavar="NAME")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    if args.copy_nba:
        with open(REF_NBA, "rb") as src, gzip.open(NBA_GZ, "wb", compresslevel=9) as dst:
            shutil.copyfileobj(src, dst)
    path = os.path.join(HERE, "golden.json")
    meta = json.load(open(path)) if os.path.exists(path) else {}
    if args.full_only:
        meta.setdefault("full", {}).update(make_full(args.full_only))
        with open(path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        return
    meta["nba"] = make_nba()
    meta["samples"] = make_samples()
    if args.full:
        meta["full"] = make_full()
    with open(path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
