"""Python host binding of the MI355X skycube engine.

A thin ctypes layer over the C ABI of ``include/skycube_amd.h``
(libskycube_amd.so, HIP kernels for gfx950).  It mirrors the reference's
algorithm objects (src/mdmc/MDMC.h:34-67, src/sdsc/SDSC.h:18-54): ``mdmc``
returns per-point cuboid membership bitmaps, ``skyalign`` one cuboid's
sky/ext split, ``sdsc`` the lattice of cuboid skylines.

There is no fallback: if the library is missing or the device call fails,
the call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "SkyError", "Context", "library_path", "load", "words_per_point",
    "in_scope_mask", "members", "SKY_OK", "hashcube_query", "hashcube_from_bitmaps", "Server",
]

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_PKG_ROOT, "libskycube_amd.so")

SKY_OK = 0
SKY_OPT_MEMBER_INDEX = 1
SKY_OPT_COVER = 2
SKY_OPT_PRUNERS = 3
SKY_OPT_MIN_BUCKET = 4
SKY_OPT_MEMBER_GRID = 5
SKY_OPT_TILE = 6
SKY_OPT_BATCH_ORDER = 7
SKY_OPT_COVER_HEAD = 8
_ERRORS = {
    -1: "SKY_ERR_ARG", -2: "SKY_ERR_NAN", -3: "SKY_ERR_OOM_BUDGET",
    -4: "SKY_ERR_DEVICE", -5: "SKY_ERR_CAPACITY", -6: "SKY_ERR_STATE",
}

# Exported symbols of libskycube_amd.so (the declarations of skycube_amd.h).
EXPORTS = (
    "sky_ctx_create", "sky_ctx_destroy", "sky_last_error", "sky_words_per_point",
    "sky_ctx_bytes_in_use", "sky_mdmc", "sky_mdmc_device", "sky_mdmc_fetch",
    "sky_mdmc_device_results", "sky_skyalign", "sky_sdsc", "sky_ctx_set_timing",
    "sky_ctx_stage_times", "sky_mdmc_hashcube", "sky_batch_mdmc", "sky_batch_fetch",
    "sky_batch_fetch_into", "sky_ctx_get_option", "sky_batch_autotune",
    "sky_ctx_release", "sky_ctx_set_option", "sky_server_create", "sky_server_destroy", "sky_server_footprint",
    "sky_server_submit", "sky_server_wait", "sky_server_result", "sky_server_stats",
    "sky_mdmc_device_noalloc", "sky_sdsc_device", "sky_sdsc_fetch", "sky_batch_mdmc_dev",
    "sky_ctx_stream", "sky_host_alloc", "sky_host_free",
)
STAGES = ("score_sort", "prune", "self", "compact", "member")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_lib = None


class SkyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


def library_path() -> str:
    return _LIB_PATH


def load():
    """Loads libskycube_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise FileNotFoundError(
            f"{_LIB_PATH} is missing: build it with `make -C {_PKG_ROOT}` "
            "(or __graft_entry__.build())")
    # torch, where installed, ships its own HIP runtime; loaded first, it is
    # the runtime the library binds too, so torch and the library share one
    # runtime in the process whichever the caller imports first
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(_LIB_PATH)
    L.sky_ctx_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
    L.sky_ctx_destroy.argtypes = [ctypes.c_void_p]
    L.sky_ctx_destroy.restype = None
    L.sky_last_error.argtypes = [ctypes.c_void_p]
    L.sky_last_error.restype = ctypes.c_char_p
    L.sky_words_per_point.argtypes = [ctypes.c_uint32]
    L.sky_words_per_point.restype = ctypes.c_uint32
    L.sky_ctx_bytes_in_use.argtypes = [ctypes.c_void_p]
    L.sky_ctx_bytes_in_use.restype = ctypes.c_size_t
    L.sky_mdmc.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                           ctypes.c_uint32, _u32p, _u32p, _u32p, ctypes.c_uint64]
    L.sky_mdmc_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, _u32p]
    L.sky_mdmc_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_uint64, _u32p]
    L.sky_mdmc_fetch.argtypes = [ctypes.c_void_p, _u32p, _u32p, ctypes.c_uint64]
    L.sky_mdmc_device_results.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(ctypes.c_void_p)]
    L.sky_skyalign.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p, _u32p, _u32p,
                               _u32p]
    L.sky_sdsc.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                           ctypes.c_uint32, _u64p, _u32p, ctypes.c_uint64, _u64p, _u32p,
                           ctypes.c_uint64]
    L.sky_sdsc_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32]
    L.sky_sdsc_fetch.argtypes = [ctypes.c_void_p, _u64p, _u32p, ctypes.c_uint64, _u64p, _u32p,
                                 ctypes.c_uint64]
    L.sky_mdmc_hashcube.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u32p, _u32p, _u64p, _u32p,
                                     ctypes.c_uint64, ctypes.c_uint64]
    L.sky_batch_mdmc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                  _u32p, _u32p, _u32p, ctypes.c_int, _u64p]
    L.sky_batch_mdmc_dev.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.c_uint32, _u32p, _u32p, _u32p, _u64p]
    L.sky_ctx_stream.argtypes = [ctypes.c_void_p]
    L.sky_ctx_stream.restype = ctypes.c_void_p
    L.sky_batch_fetch.argtypes = [ctypes.c_void_p, _u64p, _u32p, _u64p, _u32p, ctypes.c_uint64,
                                  ctypes.c_uint64]
    L.sky_batch_fetch_into.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_void_p)]
    L.sky_ctx_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.sky_ctx_stage_times.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_int]
    L.sky_ctx_release.argtypes = [ctypes.c_void_p]
    L.sky_ctx_set_option.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
    L.sky_ctx_get_option.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.sky_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_uint32, _u32p, _u32p, _u32p, ctypes.c_char_p,
                                     ctypes.c_size_t]
    L.sky_server_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
    L.sky_server_destroy.argtypes = [ctypes.c_void_p]
    L.sky_server_destroy.restype = None
    L.sky_server_footprint.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.sky_server_footprint.restype = ctypes.c_uint64
    L.sky_server_submit.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, _u64p]
    L.sky_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
    L.sky_server_result.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _u32p, _u32p, _u32p,
                                    ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), _u32p]
    L.sky_server_stats.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p,
                                   ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def words_per_point(d: int) -> int:
    """max(1, 2^d / 32): words of one membership bitmap."""
    return 1 << (d - 5) if d >= 5 else 1


def in_scope_mask(d: int, max_d: int | None = None) -> np.ndarray:
    """[W] uint32 words with bit s set iff 1 <= s < 2^d and popcount(s) <= max_d."""
    max_d = d if max_d is None else max_d
    W = words_per_point(d)
    s = np.arange(W * 32, dtype=np.uint64)
    pc = np.zeros_like(s)
    t = s.copy()
    while t.any():
        pc += t & 1
        t >>= 1
    ok = (s >= 1) & (s < (1 << d)) & (pc <= max_d)
    bits = ok.reshape(W, 32).astype(np.uint64) << np.arange(32, dtype=np.uint64)
    return bits.sum(axis=1).astype(np.uint32)


def members(dom: np.ndarray, ext_pids: np.ndarray, s: int) -> np.ndarray:
    """Skyline of cuboid s from membership bitmaps: the ext pids whose bit s
    is clear (the query of getsubspace, src/utilities/utilities.h:253-267)."""
    w, b = divmod(int(s), 32)
    return np.asarray(ext_pids)[((dom[:, w] >> np.uint32(b)) & 1) == 0]


def hashcube_query(group_word, group_value, group_off, pids, s: int) -> np.ndarray:
    """Skyline of cuboid s from a HashCube: the ids of every group of word
    s/32 whose value has bit s%32 clear (getsubspace,
    src/utilities/utilities.h:253-267), ascending."""
    w, b = divmod(int(s), 32)
    sel = np.flatnonzero((group_word == w) & (((group_value >> np.uint32(b)) & 1) == 0))
    parts = [pids[group_off[g]:group_off[g + 1]] for g in sel]
    return np.sort(np.concatenate(parts)) if parts else np.zeros(0, dtype=np.uint32)


def hashcube_from_bitmaps(ext_pids, dom, d: int):
    """Host grouping of membership bitmaps into HashCube groups (the same
    layout as Context.mdmc_hashcube), for checks."""
    W = dom.shape[1] if dom.ndim == 2 else 1
    words = [dom[:, j] for j in range(W)]
    if d >= 5:
        words.append(np.zeros(len(ext_pids), dtype=np.uint32))
    keys, ids = [], []
    for j, col in enumerate(words):
        keep = col != 0xFFFFFFFF
        keys.append((np.uint64(j) << np.uint64(32)) | col[keep].astype(np.uint64))
        ids.append(np.asarray(ext_pids)[keep])
    keys = np.concatenate(keys) if keys else np.zeros(0, np.uint64)
    ids = np.concatenate(ids) if ids else np.zeros(0, np.uint32)
    order = np.lexsort((ids, keys))
    keys, ids = keys[order], ids[order]
    uk, start = np.unique(keys, return_index=True)
    off = np.append(start, keys.size).astype(np.uint64)
    return (uk >> np.uint64(32)).astype(np.uint32), (uk & np.uint64(0xFFFFFFFF)).astype(np.uint32), off, ids.astype(np.uint32)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Context:
    """One device context: owns device memory within an HBM budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.sky_ctx_create(int(device), int(hbm_budget), ctypes.byref(h))
        if rc != SKY_OK:
            raise SkyError(rc, f"sky_ctx_create(device={device}) failed")
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None):
            self._L.sky_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def _check(self, rc: int):
        if rc != SKY_OK:
            raise SkyError(rc, self._L.sky_last_error(self._h).decode(errors="replace"))

    def set_timing(self, enable: bool = True) -> None:
        self._check(self._L.sky_ctx_set_timing(self._h, int(bool(enable))))

    def stage_times(self) -> dict:
        """Device ms per stage of the last MDMC call (HIP events on the
        context's stream; needs set_timing(True))."""
        ms = np.zeros(len(STAGES), dtype=np.float32)
        self._check(self._L.sky_ctx_stage_times(self._h, _ptr(ms, _f32p), ms.size))
        return {k: float(v) for k, v in zip(STAGES, ms)}

    def set_option(self, option: int, value: int) -> None:
        """sky_ctx_set_option; SKY_OPT_MEMBER_INDEX = 1."""
        self._check(self._L.sky_ctx_set_option(self._h, option, value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        self._check(self._L.sky_ctx_get_option(self._h, option, ctypes.byref(v)))
        return v.value

    def batch_autotune_raw(self, rows_ptr: int, on_device: bool, n, d, max_d=None) -> dict:
        """sky_batch_autotune over a batch given by a pointer to its rows;
        returns the chosen settings per dimensionality."""
        import json as _json
        n = np.ascontiguousarray(n, dtype=np.uint32)
        d = np.ascontiguousarray(d, dtype=np.uint32)
        m = None if max_d is None else np.ascontiguousarray(max_d, dtype=np.uint32)
        buf = ctypes.create_string_buffer(8192)
        self._check(self._L.sky_batch_autotune(self._h, ctypes.c_void_p(rows_ptr), int(on_device),
                                               n.size, _ptr(n, _u32p), _ptr(d, _u32p),
                                               None if m is None else _ptr(m, _u32p), buf,
                                               len(buf)))
        return _json.loads(buf.value.decode())

    def batch_autotune(self, instances, max_d=None) -> dict:
        mats = [self._rows(x) for x in instances]
        flat = np.ascontiguousarray(np.concatenate([m.ravel() for m in mats]) if mats
                                    else np.zeros(1, np.float32))
        return self.batch_autotune_raw(flat.ctypes.data, False, [m.shape[0] for m in mats],
                                       [m.shape[1] for m in mats], max_d)

    def release(self) -> None:
        """Frees the device buffers the context holds."""
        self._check(self._L.sky_ctx_release(self._h))

    def bytes_in_use(self) -> int:
        return int(self._L.sky_ctx_bytes_in_use(self._h))

    @staticmethod
    def _rows(rows) -> np.ndarray:
        x = np.ascontiguousarray(rows, dtype=np.float32)
        if x.ndim != 2:
            raise ValueError("rows must be an n x d array")
        return x

    def mdmc(self, rows, max_d: int | None = None):
        """(ext_pids ascending [n_ext], membership bitmaps [n_ext, W] uint32)."""
        x = self._rows(rows)
        n, d = x.shape
        W = words_per_point(d)
        ext = np.zeros(max(n, 1), dtype=np.uint32)
        ne = np.zeros(1, dtype=np.uint32)
        self._check(self._L.sky_mdmc(self._h, _ptr(x, _f32p), n, d, 0 if max_d is None else max_d,
                                     _ptr(ext, _u32p), _ptr(ne, _u32p), None, 0))
        k = int(ne[0])
        dom = np.zeros(max(k * W, 1), dtype=np.uint32)
        self._check(self._L.sky_mdmc_fetch(self._h, None, _ptr(dom, _u32p), dom.size))
        return ext[:k].copy(), dom[: k * W].reshape(k, W)

    def mdmc_raw(self, rows_ptr: int, n: int, d: int, max_d: int | None, ext_ptr: int) -> int:
        """sky_mdmc on raw host pointers (rows, and room for n ext ids): the
        bitmaps stay on the device.  Returns n_ext."""
        ne = np.zeros(1, dtype=np.uint32)
        self._check(self._L.sky_mdmc(self._h, ctypes.cast(ctypes.c_void_p(rows_ptr), _f32p), n, d,
                                     0 if max_d is None else max_d, ctypes.cast(ctypes.c_void_p(ext_ptr), _u32p),
                                     _ptr(ne, _u32p), None, 0))
        return int(ne[0])

    def hashcube_sizes(self):
        """(groups, entries) of the HashCube of the last MDMC result (built
        on the device if not built yet)."""
        ng = np.zeros(1, dtype=np.uint64)
        nt = np.zeros(1, dtype=np.uint64)
        self._check(self._L.sky_mdmc_hashcube(self._h, _ptr(ng, _u64p), _ptr(nt, _u64p), None,
                                              None, None, None, 0, 0))
        return int(ng[0]), int(nt[0])

    def hashcube_into(self, gw_ptr: int, gv_ptr: int, go_ptr: int, ids_ptr: int,
                      cap_groups: int, cap_entries: int):
        """The HashCube of the last MDMC result into caller buffers given by
        raw host pointers (uint32 words, uint32 values, uint64 offsets of
        cap_groups + 1, uint32 ids); returns (groups, entries)."""
        ng = np.zeros(1, dtype=np.uint64)
        nt = np.zeros(1, dtype=np.uint64)
        self._check(self._L.sky_mdmc_hashcube(self._h, _ptr(ng, _u64p), _ptr(nt, _u64p),
                                              ctypes.cast(ctypes.c_void_p(gw_ptr), _u32p),
                                              ctypes.cast(ctypes.c_void_p(gv_ptr), _u32p),
                                              ctypes.cast(ctypes.c_void_p(go_ptr), _u64p),
                                              ctypes.cast(ctypes.c_void_p(ids_ptr), _u32p), cap_groups,
                                              cap_entries))
        return int(ng[0]), int(nt[0])

    def mdmc_device(self, rows_dev_ptr: int, n: int, d: int, max_d: int | None = None) -> int:
        """Runs on rows already in device memory; results stay on the device."""
        ne = np.zeros(1, dtype=np.uint32)
        self._check(self._L.sky_mdmc_device(self._h, ctypes.c_void_p(rows_dev_ptr), n, d,
                                            0 if max_d is None else max_d, _ptr(ne, _u32p)))
        return int(ne[0])

    def mdmc_device_noalloc(self, rows_dev_ptr: int, n: int, d: int, max_d: int | None,
                            ext_dev_ptr: int, dom_dev_ptr: int, dom_cap_words: int) -> int:
        """sky_mdmc_device_noalloc: rows, ids and bitmaps all in caller-owned
        device memory; nothing is allocated.  Returns n_ext."""
        ne = np.zeros(1, dtype=np.uint32)
        self._check(self._L.sky_mdmc_device_noalloc(
            self._h, ctypes.c_void_p(rows_dev_ptr), n, d, 0 if max_d is None else max_d,
            ctypes.c_void_p(ext_dev_ptr), ctypes.c_void_p(dom_dev_ptr), dom_cap_words,
            _ptr(ne, _u32p)))
        return int(ne[0])

    def mdmc_fetch(self, n_ext: int, d: int):
        W = words_per_point(d)
        ext = np.zeros(max(n_ext, 1), dtype=np.uint32)
        dom = np.zeros(max(n_ext * W, 1), dtype=np.uint32)
        self._check(self._L.sky_mdmc_fetch(self._h, _ptr(ext, _u32p), _ptr(dom, _u32p), dom.size))
        return ext[:n_ext].copy(), dom[: n_ext * W].reshape(n_ext, W)

    def mdmc_device_results(self):
        e = ctypes.c_void_p()
        w = ctypes.c_void_p()
        self._check(self._L.sky_mdmc_device_results(self._h, ctypes.byref(e), ctypes.byref(w)))
        return e.value, w.value

    def mdmc_hashcube(self):
        """HashCube of the last MDMC result, grouped on the device:
        (group_word, group_value, group_off [n_groups+1], pids)."""
        ng = np.zeros(1, dtype=np.uint64)
        nt = np.zeros(1, dtype=np.uint64)
        self._check(self._L.sky_mdmc_hashcube(self._h, _ptr(ng, _u64p), _ptr(nt, _u64p), None,
                                              None, None, None, 0, 0))
        g, t = int(ng[0]), int(nt[0])
        gw = np.zeros(max(g, 1), dtype=np.uint32)
        gv = np.zeros(max(g, 1), dtype=np.uint32)
        go = np.zeros(g + 1, dtype=np.uint64)
        ids = np.zeros(max(t, 1), dtype=np.uint32)
        self._check(self._L.sky_mdmc_hashcube(self._h, _ptr(ng, _u64p), _ptr(nt, _u64p),
                                              _ptr(gw, _u32p), _ptr(gv, _u32p), _ptr(go, _u64p),
                                              _ptr(ids, _u32p), max(g, 1), max(t, 1)))
        return gw[:g], gv[:g], go, ids[:t]

    def batch_mdmc_raw(self, rows_ptr: int, on_device: bool, n, d, max_d=None,
                       use_graph: bool = False) -> int:
        """Runs a batch from a pointer to the concatenated rows; returns the
        total number of extended-skyline points (results stay on device)."""
        n = np.ascontiguousarray(n, dtype=np.uint32)
        d = np.ascontiguousarray(d, dtype=np.uint32)
        m = None if max_d is None else np.ascontiguousarray(max_d, dtype=np.uint32)
        tot = np.zeros(1, dtype=np.uint64)
        self._check(self._L.sky_batch_mdmc(self._h, ctypes.c_void_p(rows_ptr), int(on_device),
                                           n.size, _ptr(n, _u32p), _ptr(d, _u32p),
                                           None if m is None else _ptr(m, _u32p),
                                           int(bool(use_graph)), _ptr(tot, _u64p)))
        return int(tot[0])

    def batch_mdmc_dev(self, rows_dev_ptrs, n, d, max_d=None) -> int:
        """Runs a batch whose instances live in separate device allocations
        (one device pointer per instance); returns the total number of
        extended-skyline points (results stay on device)."""
        n = np.ascontiguousarray(n, dtype=np.uint32)
        d = np.ascontiguousarray(d, dtype=np.uint32)
        m = None if max_d is None else np.ascontiguousarray(max_d, dtype=np.uint32)
        ptrs = (ctypes.c_void_p * max(len(rows_dev_ptrs), 1))(*[int(p) for p in rows_dev_ptrs])
        tot = np.zeros(1, dtype=np.uint64)
        self._check(self._L.sky_batch_mdmc_dev(self._h, ptrs, n.size, _ptr(n, _u32p),
                                               _ptr(d, _u32p),
                                               None if m is None else _ptr(m, _u32p),
                                               _ptr(tot, _u64p)))
        return int(tot[0])

    def stream(self) -> int:
        """The context's HIP stream, as an integer handle."""
        return int(self._L.sky_ctx_stream(self._h) or 0)

    def batch_fetch(self, n_inst: int):
        """[(ext_pids, bitmaps)] of the last batch, in instance order."""
        eo = np.zeros(n_inst + 1, dtype=np.uint64)
        do = np.zeros(n_inst + 1, dtype=np.uint64)
        self._check(self._L.sky_batch_fetch(self._h, _ptr(eo, _u64p), None, _ptr(do, _u64p), None,
                                            0, 0))
        ids = np.zeros(max(int(eo[-1]), 1), dtype=np.uint32)
        words = np.zeros(max(int(do[-1]), 1), dtype=np.uint32)
        self._check(self._L.sky_batch_fetch(self._h, _ptr(eo, _u64p), _ptr(ids, _u32p),
                                            _ptr(do, _u64p), _ptr(words, _u32p), ids.size,
                                            words.size))
        out = []
        for t in range(n_inst):
            e = ids[int(eo[t]):int(eo[t + 1])].copy()
            w = words[int(do[t]):int(do[t + 1])]
            W = w.size // max(e.size, 1) if e.size else 0
            out.append((e, w.reshape(e.size, W) if e.size else w.reshape(0, 0)))
        return out

    def batch_mdmc(self, instances, max_d=None, use_graph: bool = False):
        """Skycubes of a list of n_t x d_t arrays in one batch:
        [(ext_pids, bitmaps)] per instance."""
        mats = [self._rows(x) for x in instances]
        n = [m.shape[0] for m in mats]
        d = [m.shape[1] for m in mats]
        flat = np.ascontiguousarray(np.concatenate([m.ravel() for m in mats]) if mats
                                    else np.zeros(1, np.float32))
        self.batch_mdmc_raw(flat.ctypes.data, False, n, d, max_d, use_graph)
        return self.batch_fetch(len(mats))

    def skyalign(self, rows, subspace: int, cand=None):
        """(sky ids ascending, ext-but-not-sky ids ascending) of cuboid `subspace`."""
        x = self._rows(rows)
        n, d = x.shape
        if cand is None:
            c = n
            cp = None
        else:
            cand = np.ascontiguousarray(cand, dtype=np.uint32)
            c = cand.size
            cp = _ptr(cand, _u32p)
        sky = np.zeros(max(c, 1), dtype=np.uint32)
        ext = np.zeros(max(c, 1), dtype=np.uint32)
        ns = np.zeros(1, dtype=np.uint32)
        ne = np.zeros(1, dtype=np.uint32)
        self._check(self._L.sky_skyalign(self._h, _ptr(x, _f32p), n, d, int(subspace), cp, c,
                                         _ptr(sky, _u32p), _ptr(ns, _u32p), _ptr(ext, _u32p),
                                         _ptr(ne, _u32p)))
        return sky[: int(ns[0])].copy(), ext[: int(ne[0])].copy()

    def sdsc(self, rows, max_d: int | None = None):
        """CSR lattice: (sky_off [2^d+1], sky_ids, ext_off [2^d+1], ext_ids)."""
        x = self._rows(rows)
        n, d = x.shape
        S = 1 << d
        so = np.zeros(S + 1, dtype=np.uint64)
        eo = np.zeros(S + 1, dtype=np.uint64)
        rc = self._L.sky_sdsc(self._h, _ptr(x, _f32p), n, d, 0 if max_d is None else max_d,
                              _ptr(so, _u64p), None, 0, _ptr(eo, _u64p), None, 0)
        if rc != -5:
            self._check(rc)
        return self.sdsc_fetch(d)

    def sdsc_device(self, rows_dev_ptr: int, n: int, d: int, max_d: int | None = None) -> None:
        """The SDSC lattice of rows already in device memory, kept on the
        device (sdsc_fetch copies it out)."""
        self._check(self._L.sky_sdsc_device(self._h, ctypes.c_void_p(rows_dev_ptr), n, d,
                                            0 if max_d is None else max_d))

    def sdsc_fetch(self, d: int):
        """CSR lattice of the last SDSC call: (sky_off, sky_ids, ext_off, ext_ids)."""
        S = 1 << d
        so = np.zeros(S + 1, dtype=np.uint64)
        eo = np.zeros(S + 1, dtype=np.uint64)
        rc = self._L.sky_sdsc_fetch(self._h, _ptr(so, _u64p), None, 0, _ptr(eo, _u64p), None, 0)
        if rc != -5:
            self._check(rc)
        sky = np.zeros(max(int(so[S]), 1), dtype=np.uint32)
        ext = np.zeros(max(int(eo[S]), 1), dtype=np.uint32)
        self._check(self._L.sky_sdsc_fetch(self._h, _ptr(so, _u64p), _ptr(sky, _u32p), sky.size,
                                           _ptr(eo, _u64p), _ptr(ext, _u32p), ext.size))
        return so, sky[: int(so[S])].copy(), eo, ext[: int(eo[S])].copy()


class Server:
    """Continuous-batching skycube server (sky_server_* of the C ABI).

    ``submit`` queues one skycube request and returns its id at once;
    ``result`` blocks until it is done and returns (ext_pids, dom, latency_ms,
    batch_size), the same arrays ``Context.mdmc`` returns for those rows."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, workers: int = 4,
                 max_batch: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.sky_server_create(device, hbm_budget, workers, max_batch, ctypes.byref(h))
        if rc != SKY_OK:
            raise SkyError(rc, "server creation failed")
        self._h = h
        self._d = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sky_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def footprint(n: int, d: int) -> int:
        return int(load().sky_server_footprint(n, d))

    def submit(self, rows, max_d: int | None = None) -> int:
        x = Context._rows(rows)
        n, d = x.shape
        rid = ctypes.c_uint64()
        rc = self._lib.sky_server_submit(self._h, _ptr(x, _f32p), n, d, max_d or 0,
                                         ctypes.byref(rid))
        if rc != SKY_OK:
            raise SkyError(rc, f"request of {n} x {d} rejected")
        self._d[rid.value] = d
        return rid.value

    def result(self, rid: int, timeout_ms: int = -1):
        rc = self._lib.sky_server_wait(self._h, rid, timeout_ms)
        if rc != SKY_OK:
            self._d.pop(rid, None)
            raise SkyError(rc, f"request {rid} failed or timed out")
        ne = ctypes.c_uint32()
        lat = ctypes.c_double()
        bs = ctypes.c_uint32()
        rc = self._lib.sky_server_result(self._h, rid, None, ctypes.byref(ne), None, 0,
                                         ctypes.byref(lat), ctypes.byref(bs))
        if rc != SKY_OK:
            raise SkyError(rc, f"request {rid}: no result")
        W = words_per_point(self._d.pop(rid))
        ext = np.zeros(max(ne.value, 1), dtype=np.uint32)
        dom = np.zeros((max(ne.value, 1), W), dtype=np.uint32)
        rc = self._lib.sky_server_result(self._h, rid, _ptr(ext, _u32p), ctypes.byref(ne),
                                         _ptr(dom, _u32p), dom.size, None, None)
        if rc != SKY_OK:
            raise SkyError(rc, f"request {rid}: result copy failed")
        return ext[:ne.value], dom[:ne.value], lat.value, bs.value

    def stats(self) -> dict:
        b, m, q = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        busy = ctypes.c_double()
        self._lib.sky_server_stats(self._h, ctypes.byref(b), ctypes.byref(m), ctypes.byref(q),
                                   ctypes.byref(busy))
        return {"batches": b.value, "max_concurrent": m.value, "queued": q.value,
                "busy_ms": busy.value}
