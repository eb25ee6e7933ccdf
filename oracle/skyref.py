"""ctypes binding of the reference's own CPU skyline code (oracle/_ref/
libskyref.so, built by `make -C oracle ref` from the reference sources; see
oracle/ref_harness.cpp).

Test infrastructure, like the rest of oracle/: imported by tests/ (to pin
the oracle) and by bench.py's cpu_baseline leg (to time the reference's CPU
path).  The library exists only where it was built; callers check
available() first.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "libskyref.so")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)

# the reference driver's defaults: alpha block 1024, PQ filter of 128 points
# (src/main.cu:74-90, src/utilities/config.h:13-27)
ACCUM = 1024
PQ_SIZE = 128


def available() -> bool:
    return os.path.exists(_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_LIB_PATH)
        L.ref_stsc.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p, _u64p,
                               _f64p]
        L.ref_stsc_fetch.argtypes = [_u32p, _u32p]
        L.ref_cuboids_timed.argtypes = [_f32p, ctypes.c_uint32, _u32p, _u64p, _u32p,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, _f64p, _f64p]
        L.ref_cuboid.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p,
                                 _u32p, _u32p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def stsc(x: np.ndarray, max_d: int | None = None, threads: int = 8):
    """The reference's STSC lattice: (sky_off, sky_ids, ext_off, ext_ids,
    level_seconds); cuboid s's lists are ids[off[s]:off[s+1]], each sorted
    ascending here (MommyQ leaves them in its own order)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    S = 1 << d
    so = np.zeros(S + 1, dtype=np.uint64)
    eo = np.zeros(S + 1, dtype=np.uint64)
    lv = np.zeros(d + 1, dtype=np.float64)
    rc = lib().ref_stsc(_p(x, _f32p), n, d, max_d or d, threads, ACCUM, PQ_SIZE, _p(so, _u64p),
                        _p(eo, _u64p), _p(lv, _f64p))
    if rc != 0:
        raise RuntimeError(f"ref_stsc failed ({rc})")
    sky = np.zeros(int(so[S]) + 1, dtype=np.uint32)
    ext = np.zeros(int(eo[S]) + 1, dtype=np.uint32)
    lib().ref_stsc_fetch(_p(sky, _u32p), _p(ext, _u32p))
    for s in range(1, S):
        sky[so[s]:so[s + 1]].sort()
        ext[eo[s]:eo[s + 1]].sort()
    return so, sky[: int(so[S])], eo, ext[: int(eo[S])], lv


def cuboid(x: np.ndarray, sub: int, cand=None, threads: int = 1):
    """One cuboid by the reference's per-cuboid code: (sky ids, ext ids, seconds)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    c = np.arange(n, dtype=np.uint32) if cand is None else np.ascontiguousarray(cand, np.uint32)
    sky = np.zeros(max(c.size, 1), dtype=np.uint32)
    ext = np.zeros(max(c.size, 1), dtype=np.uint32)
    ns = ctypes.c_uint32()
    ne = ctypes.c_uint32()
    sec = ctypes.c_double()
    rc = lib().ref_cuboid(_p(x, _f32p), d, sub, _p(c, _u32p), c.size, threads, ACCUM, PQ_SIZE,
                          _p(sky, _u32p), ctypes.byref(ns), _p(ext, _u32p), ctypes.byref(ne),
                          ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError(f"ref_cuboid failed ({rc})")
    return np.sort(sky[: ns.value]), np.sort(ext[: ne.value]), sec.value


def cuboids_timed(x: np.ndarray, subs, cands, threads: int):
    """(wall seconds, seconds of each cuboid) of the cuboids subs[i] over
    candidates cands[i], run as one parallel level of the reference's walk."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    d = x.shape[1]
    subs = np.ascontiguousarray(subs, dtype=np.uint32)
    off = np.zeros(len(cands) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(c) for c in cands])
    ids = (np.concatenate([np.asarray(c, np.uint32) for c in cands]) if len(cands)
           else np.zeros(1, np.uint32))
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    wall = ctypes.c_double()
    each = np.zeros(max(subs.size, 1), dtype=np.float64)
    rc = lib().ref_cuboids_timed(_p(x, _f32p), d, _p(subs, _u32p), _p(off, _u64p), _p(ids, _u32p),
                                 subs.size, threads, ACCUM, PQ_SIZE, ctypes.byref(wall),
                                 _p(each, _f64p))
    if rc != 0:
        raise RuntimeError(f"ref_cuboids_timed failed ({rc})")
    return wall.value, each[: subs.size]
