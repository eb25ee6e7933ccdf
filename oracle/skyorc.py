"""ctypes binding of the CPU oracle (libskyorc.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libskyorc.so")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f32p = ctypes.POINTER(ctypes.c_float)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_words_per_point.argtypes = [ctypes.c_int]
        L.orc_skyalign.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                   _u32p, ctypes.c_uint32, _u8p]
        L.orc_ext_skyline.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int, _u32p, _u32p]
        L.orc_mdmc.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _u32p,
                               _u32p, _u32p]
        L.orc_brute.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int, _u32p, _u32p]
        L.orc_stsc.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _u32p,
                               _u32p, _u32p, _u32p]
        L.orc_score_tie_count.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _u32p,
                                          ctypes.c_uint32]
        L.orc_score_tie_count.restype = ctypes.c_uint64
        L.orc_score_tie_unresolved.argtypes = L.orc_score_tie_count.argtypes
        L.orc_score_tie_unresolved.restype = ctypes.c_uint64
        L.orc_set_mutation.argtypes = [ctypes.c_int]
        L.orc_ext_skyline_timed.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int,
                                            ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), _u32p]
        L.orc_mdmc_sample.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_uint32,
                                      _u32p, ctypes.c_uint32, _u32p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def words_per_point(d: int) -> int:
    return 1 << (d - 5) if d >= 5 else 1


def _rows(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    assert x.ndim == 2
    return x


def set_mutation(m: int) -> None:
    lib().orc_set_mutation(int(m))


def skyalign(x, sub: int, cand=None) -> np.ndarray:
    """Status per candidate: 0 sky, 1 ext, 2 strictly dominated."""
    x = _rows(x)
    n, d = x.shape
    if cand is None:
        c = n
        cp = None
    else:
        cand = np.ascontiguousarray(cand, dtype=np.uint32)
        c = cand.size
        cp = _ptr(cand, _u32p)
    st = np.zeros(max(c, 1), dtype=np.uint8)
    rc = lib().orc_skyalign(_ptr(x, _f32p), n, d, sub, cp, c, _ptr(st, _u8p))
    if rc != 0:
        raise RuntimeError(f"orc_skyalign rc={rc}")
    return st[:c]


def ext_skyline(x) -> np.ndarray:
    x = _rows(x)
    n, d = x.shape
    out = np.zeros(max(n, 1), dtype=np.uint32)
    m = np.zeros(1, dtype=np.uint32)
    rc = lib().orc_ext_skyline(_ptr(x, _f32p), n, d, _ptr(out, _u32p), _ptr(m, _u32p))
    if rc != 0:
        raise RuntimeError(f"orc_ext_skyline rc={rc}")
    return out[: int(m[0])].copy()


def mdmc(x, max_d: int | None = None):
    """(ext_idx ascending, dom words [n_ext, W]); bit s = dominated in s."""
    x = _rows(x)
    n, d = x.shape
    max_d = d if max_d is None else max_d
    W = words_per_point(d)
    ext = np.zeros(max(n, 1), dtype=np.uint32)
    m = np.zeros(1, dtype=np.uint32)
    dom = np.zeros(max(n, 1) * W, dtype=np.uint32)
    rc = lib().orc_mdmc(_ptr(x, _f32p), n, d, max_d, _ptr(ext, _u32p), _ptr(m, _u32p),
                        _ptr(dom, _u32p))
    if rc != 0:
        raise RuntimeError(f"orc_mdmc rc={rc}")
    k = int(m[0])
    return ext[:k].copy(), dom[: k * W].reshape(k, W).copy()


def brute(x):
    """(dom [n, W], sdom [n, W]) for every point and every subspace."""
    x = _rows(x)
    n, d = x.shape
    W = words_per_point(d)
    dom = np.zeros(max(n, 1) * W, dtype=np.uint32)
    sdom = np.zeros(max(n, 1) * W, dtype=np.uint32)
    rc = lib().orc_brute(_ptr(x, _f32p), n, d, _ptr(dom, _u32p), _ptr(sdom, _u32p))
    if rc != 0:
        raise RuntimeError(f"orc_brute rc={rc}")
    return dom[: n * W].reshape(n, W), sdom[: n * W].reshape(n, W)


def stsc(x, max_d: int | None = None):
    """(dom [n, W], sdom [n, W], sky sizes [2^d], ext sizes [2^d])."""
    x = _rows(x)
    n, d = x.shape
    max_d = d if max_d is None else max_d
    W = words_per_point(d)
    dom = np.zeros(max(n, 1) * W, dtype=np.uint32)
    sdom = np.zeros(max(n, 1) * W, dtype=np.uint32)
    skyn = np.zeros(1 << d, dtype=np.uint32)
    extn = np.zeros(1 << d, dtype=np.uint32)
    rc = lib().orc_stsc(_ptr(x, _f32p), n, d, max_d, _ptr(dom, _u32p), _ptr(sdom, _u32p),
                        _ptr(skyn, _u32p), _ptr(extn, _u32p))
    if rc != 0:
        raise RuntimeError(f"orc_stsc rc={rc}")
    return dom[: n * W].reshape(n, W), sdom[: n * W].reshape(n, W), skyn, extn


def score_tie_count(x, ext, max_d: int | None = None) -> int:
    x = _rows(x)
    n, d = x.shape
    max_d = d if max_d is None else max_d
    ext = np.ascontiguousarray(ext, dtype=np.uint32)
    count = int(lib().orc_score_tie_count(_ptr(x, _f32p), d, max_d, _ptr(ext, _u32p), ext.size))
    if count == 2**64 - 1:
        raise MemoryError("score_tie_count: no memory for the candidate rows")
    return count


def score_tie_unresolved(x, ext, max_d: int | None = None) -> int:
    """Score-tie triples with no smaller-norm dominator of p in s (the ones
    where STSC's early exit would change its answer)."""
    x = _rows(x)
    n, d = x.shape
    max_d = d if max_d is None else max_d
    ext = np.ascontiguousarray(ext, dtype=np.uint32)
    return int(lib().orc_score_tie_unresolved(_ptr(x, _f32p), d, max_d, _ptr(ext, _u32p),
                                               ext.size))


def mdmc_sample(x, ext, sample, max_d: int | None = None) -> np.ndarray:
    """Bitmaps [len(sample), W] of the ext points ext[sample] only."""
    x = _rows(x)
    n, d = x.shape
    max_d = d if max_d is None else max_d
    W = words_per_point(d)
    ext = np.ascontiguousarray(ext, dtype=np.uint32)
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    out = np.zeros(max(sample.size, 1) * W, dtype=np.uint32)
    rc = lib().orc_mdmc_sample(_ptr(x, _f32p), d, max_d, _ptr(ext, _u32p), ext.size,
                               _ptr(sample, _u32p), sample.size, _ptr(out, _u32p))
    if rc != 0:
        raise RuntimeError(f"orc_mdmc_sample rc={rc}")
    return out[: sample.size * W].reshape(sample.size, W)


def ext_skyline_timed(x, stride: int):
    """(prefilter seconds, sampled survivor-pass seconds, survivors)."""
    x = _rows(x)
    n, d = x.shape
    tp = ctypes.c_double()
    ts = ctypes.c_double()
    ns = np.zeros(1, dtype=np.uint32)
    rc = lib().orc_ext_skyline_timed(_ptr(x, _f32p), n, d, int(stride), ctypes.byref(tp),
                                     ctypes.byref(ts), _ptr(ns, _u32p))
    if rc != 0:
        raise RuntimeError(f"orc_ext_skyline_timed rc={rc}")
    return tp.value, ts.value, int(ns[0])
