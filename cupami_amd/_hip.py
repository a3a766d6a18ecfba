"""ctypes boundary to ``libcupami_hip.so`` (declared in ``include/cupami_hip.h``).

Status codes map to the reference's exception types (``cuApriori.py:458-471``,
``_opencl_bitsets.py:92-135``): invalid arguments raise ``ValueError``, HIP
failures and missing devices ``RuntimeError``, device allocation failures
``MemoryError``.  A missing or unloadable library raises ``ImportError`` --
there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading
from collections.abc import Mapping
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("CUPAMI_HIP_LIB", _HERE / "libcupami_hip.so"))

OK, ERR_INVALID, ERR_RUNTIME, ERR_NOMEM, ERR_NODEVICE, ERR_LAYOUT = 0, -1, -2, -3, -4, -5
NO_PARENT = 0xFFFFFFFF

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol include/cupami_hip.h declares
SIGNATURES = {
    "cpm_version": (ctypes.c_int, []),
    "cpm_last_error": (ctypes.c_char_p, []),
    "cpm_device_count": (ctypes.c_int, []),
    "cpm_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "cpm_ctx_destroy": (ctypes.c_int, [_vp]),
    "cpm_ctx_synchronize": (ctypes.c_int, [_vp]),
    "cpm_ctx_stream": (_vp, [_vp]),
    "cpm_bits_build": (ctypes.c_int, [_vp, _u64p, _u32p, ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "cpm_bits_build_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint64,
                                             ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "cpm_bits_build_device_batch": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(_vp),
                                                   ctypes.POINTER(_vp), _u64p, _u64p, _u32p,
                                                   ctypes.POINTER(_vp)]),
    "cpm_bits_build_batch": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            _u64p, _u64p, _u32p, ctypes.POINTER(_vp)]),
    "cpm_bits_build_csr": (ctypes.c_int, [_vp, _u64p, ctypes.c_uint64, _u32p, ctypes.c_uint64,
                                          ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(_vp)]),
    "cpm_bits_build_csr_device": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64,
                                                 ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(_vp)]),
    "cpm_bits_labels": (ctypes.c_int, [_vp, _vp, _u32p]),
    "cpm_baskets_parse_text": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char,
                                              ctypes.POINTER(_vp)]),
    "cpm_baskets_info": (ctypes.c_int, [_vp, _u64p, _u64p, _u64p, _u32p]),
    "cpm_baskets_download": (ctypes.c_int, [_vp, _vp, _u64p, _u32p]),
    "cpm_bits_build_baskets": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "cpm_baskets_destroy": (ctypes.c_int, [_vp]),
    "cpm_ingest_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "cpm_patterns_keys": (ctypes.c_int, [_vp, ctypes.c_char_p, _u64p, ctypes.c_uint32, ctypes.c_char, _vp,
                                         ctypes.c_uint64, _u64p]),
    "cpm_patterns_save": (ctypes.c_int, [_vp, ctypes.c_char_p, _u64p, ctypes.c_uint32, ctypes.c_char_p]),
    "cpm_bits_destroy": (ctypes.c_int, [_vp]),
    "cpm_bits_shape": (ctypes.c_int, [_vp, _u64p, _u32p, _u64p, _u64p]),
    "cpm_bits_download": (ctypes.c_int, [_vp, _vp, _u64p]),
    "cpm_count_supports": (ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint64, ctypes.c_uint32,
                                          _u64p]),
    "cpm_count_supports_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32, _vp]),
    "cpm_intersect_extensions": (ctypes.c_int, [_vp, _u64p, _u64p, ctypes.c_uint32,
                                                ctypes.c_uint64, ctypes.c_uint64, _u64p, _u64p,
                                                _u32p, _u32p]),
    "cpm_pair_gram": (ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint32, _u32p]),
    "cpm_periodic_eval": (ctypes.c_int, [_vp, _vp, _i64p, ctypes.c_int64, _u32p, ctypes.c_uint64,
                                         ctypes.c_uint32, _u64p, _i64p]),
    "cpm_eclat_mine": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.POINTER(_vp)]),
    "cpm_pfeclat_mine": (ctypes.c_int, [_vp, _vp, _i64p, ctypes.c_int64, ctypes.c_uint64,
                                        ctypes.c_int64, ctypes.POINTER(_vp)]),
    "cpm_pfeclat_mine_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_uint64,
                                        ctypes.c_int64, ctypes.POINTER(_vp)]),
    "cpm_unique_pairs": (ctypes.c_int, [_vp, _u64p, _u32p, ctypes.c_uint64, _u64p, _u32p, _u64p]),
    "cpm_fptree_build": (ctypes.c_int, [_vp, _u64p, _u32p, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_uint64, _u64p,
                                        ctypes.POINTER(_vp)]),
    "cpm_fptree_conditional": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64,
                                              ctypes.POINTER(_vp)]),
    "cpm_fptree_info": (ctypes.c_int, [_vp, _u64p, _u32p, ctypes.POINTER(ctypes.c_int), _u32p]),
    "cpm_fptree_download": (ctypes.c_int, [_vp, _i64p, _i64p, _u64p, _u32p, _u64p]),
    "cpm_fptree_destroy": (ctypes.c_int, [_vp]),
    "cpm_fpgrowth_mine": (ctypes.c_int, [_vp, _u64p, _u32p, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "cpm_eclat_mine_batch": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _u64p, ctypes.c_uint32,
                                            ctypes.c_int, ctypes.POINTER(_vp)]),
    "cpm_set_search_options": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                              ctypes.c_int]),
    "cpm_patterns_databases": (ctypes.c_int, [_vp, _u32p]),
    "cpm_patterns_view": (ctypes.c_int, [_vp, ctypes.POINTER(_u32p), ctypes.POINTER(_u32p),
                                         ctypes.POINTER(_u64p), ctypes.POINTER(_i64p),
                                         ctypes.POINTER(_u32p)]),
    "cpm_patterns_count": (ctypes.c_int, [_vp, _u64p]),
    "cpm_patterns_download": (ctypes.c_int, [_vp, _u32p, _u32p, _u64p, _i64p]),
    "cpm_patterns_destroy": (ctypes.c_int, [_vp]),
    "cpm_roof_stream": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, _vp,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "cpm_roof_mfma": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint32, _vp, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]),
    "cpm_roof_phases": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "cpm_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "cpm_stats_reset": (ctypes.c_int, [_vp]),
    "cpm_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
}

_lib = None
_lib_lock = threading.Lock()


def load():
    """Load the library (ImportError when it is missing or unloadable)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(
                    f"GPU mining requires the HIP library {LIB_PATH}; build it with "
                    "`python -m cupami_amd.build` (hipcc, gfx950)")
            try:
                lib = ctypes.CDLL(str(LIB_PATH))
            except OSError as error:
                raise ImportError(f"cannot load {LIB_PATH}: {error}") from error
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status):
    if status == OK:
        return
    message = load().cpm_last_error().decode("utf-8", "replace")
    if status == ERR_INVALID:
        raise ValueError(message)
    if status == ERR_NOMEM:
        raise MemoryError(message)
    raise RuntimeError(message)


def _p(array, ptype):
    return array.ctypes.data_as(ptype)


class Context:
    """One device and its stream (``cpm_ctx``)."""

    def __init__(self, device=0):
        lib = load()
        handle = _vp()
        status = lib.cpm_ctx_create(int(device), ctypes.byref(handle))
        if status == ERR_NODEVICE:
            raise RuntimeError("GPU mining requires a HIP device; " +
                               lib.cpm_last_error().decode())
        check(status)
        self.handle = handle
        self.device = int(device)

    def close(self):
        if self.handle:
            load().cpm_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(load().cpm_ctx_synchronize(self.handle))

    def stream(self):
        return load().cpm_ctx_stream(self.handle)

    def stats(self):
        out = (ctypes.c_double * 8)()
        check(load().cpm_stats(self.handle, out))
        keys = ("launches", "kernel_ms", "bytes", "candidates", "gram_launches", "gram_ms",
                "gram_ops", "levels")
        return dict(zip(keys, list(out)))

    def ingest_stats(self):
        out = (ctypes.c_double * 4)()
        check(load().cpm_ingest_stats(self.handle, out))
        return dict(zip(("builds", "tile_ms", "tile_bytes", "build_ms"), list(out)))

    def reset_stats(self):
        check(load().cpm_stats_reset(self.handle))

    def set_timing(self, on):
        check(load().cpm_set_timing(self.handle, 1 if on else 0))

    # ---- measured roofs --------------------------------------------------------
    def roof_stream(self, d_buf, nbytes, reps, xcd_sliced, d_sink):
        """AND + popcount streaming microbenchmark; returns (ms, bytes read)."""
        ms, nb = ctypes.c_double(), ctypes.c_double()
        check(load().cpm_roof_stream(self.handle, _vp(int(d_buf)), int(nbytes), int(reps), int(bool(xcd_sliced)),
                                     _vp(int(d_sink)), ctypes.byref(ms), ctypes.byref(nb)))
        return ms.value, nb.value

    def roof_phases(self, blocks, rounds, dirty_bytes):
        """(us per grid-wide barrier in one launch, us per launch in a graph)."""
        ub, ul = ctypes.c_double(), ctypes.c_double()
        check(load().cpm_roof_phases(self.handle, int(blocks), int(rounds), int(dirty_bytes),
                                     ctypes.byref(ub), ctypes.byref(ul)))
        return ub.value, ul.value

    def roof_mfma(self, iters, d_sink, shape=16):
        """FP4 MFMA loop microbenchmark; returns (ms, ops)."""
        ms, ops = ctypes.c_double(), ctypes.c_double()
        check(load().cpm_roof_mfma(self.handle, int(shape), int(iters), _vp(int(d_sink)), ctypes.byref(ms), ctypes.byref(ops)))
        return ms.value, ops.value

    # ---- store -------------------------------------------------------------
    def build_bits(self, rows, cols, n_transactions, n_items):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        handle = _vp()
        check(load().cpm_bits_build(self.handle, _p(rows, _u64p), _p(cols, _u32p), rows.size,
                                    int(n_transactions), int(n_items), ctypes.byref(handle)))
        return Bits(self, handle)

    def build_bits_device(self, d_rows, d_cols, nnz, n_transactions, n_items):
        """rows/cols given as device addresses (ints)."""
        handle = _vp()
        check(load().cpm_bits_build_device(self.handle, _vp(d_rows), _vp(d_cols), int(nnz),
                                           int(n_transactions), int(n_items),
                                           ctypes.byref(handle)))
        return Bits(self, handle)

    def build_bits_csr(self, offsets, ids, universe, keep_empty=False, labels=True):
        """Store straight from a basket file's arrays (``offsets`` u64[n+1],
        ``ids`` u32[m]).  Returns the store and the item id of each column
        (ids in decimal-string order; None with ``labels=False``, which
        saves the copy and the wait for them)."""
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        if offsets.size < 1:
            raise ValueError("Invalid BIN file: invalid transaction offsets")
        handle = _vp()
        check(load().cpm_bits_build_csr(self.handle, _p(offsets, _u64p), offsets.size - 1, _p(ids, _u32p),
                                        ids.size, int(universe), int(bool(keep_empty)), ctypes.byref(handle)))
        bits = Bits(self, handle)
        return bits, (bits.labels() if labels else None)

    def build_bits_csr_device(self, d_offsets, n_baskets, d_ids, n_ids, universe, keep_empty=False):
        """``build_bits_csr`` with the arrays given as device addresses (ints)."""
        handle = _vp()
        check(load().cpm_bits_build_csr_device(self.handle, _vp(d_offsets), int(n_baskets), _vp(d_ids),
                                               int(n_ids), int(universe), int(bool(keep_empty)),
                                               ctypes.byref(handle)))
        bits = Bits(self, handle)
        return bits, bits.labels()

    def parse_text(self, text, sep):
        """Delimited text of decimal ids parsed on the device
        (``cpm_baskets_parse_text``): ``Baskets``, or None when the text is
        for the host reader (quotes, non-ASCII, fields that are not
        canonical decimal ids)."""
        data = bytes(text)
        sep = sep.encode("latin-1") if isinstance(sep, str) else bytes(sep)
        if len(sep) != 1:
            return None
        handle = _vp()
        status = load().cpm_baskets_parse_text(self.handle, data, len(data), sep, ctypes.byref(handle))
        if status == ERR_LAYOUT:
            return None
        check(status)
        return Baskets(self, handle)

    def build_bits_device_batch(self, specs):
        """Several stores in one call: ``specs`` is a sequence of
        ``(d_rows, d_cols, nnz, n_transactions, n_items)`` with device
        addresses as ints.  Returns one ``Bits`` per spec."""
        n = len(specs)
        if n == 0:
            return []
        rows = (_vp * n)(*[_vp(int(s[0])) for s in specs])
        cols = (_vp * n)(*[_vp(int(s[1])) for s in specs])
        nnz = np.ascontiguousarray([int(s[2]) for s in specs], dtype=np.uint64)
        nt = np.ascontiguousarray([int(s[3]) for s in specs], dtype=np.uint64)
        ni = np.ascontiguousarray([int(s[4]) for s in specs], dtype=np.uint32)
        out = (_vp * n)()
        check(load().cpm_bits_build_device_batch(self.handle, n, rows, cols, _p(nnz, _u64p), _p(nt, _u64p),
                                                 _p(ni, _u32p), out))
        return [Bits(self, _vp(out[i])) for i in range(n)]

    def build_bits_batch(self, dbs):
        """Several stores from host pairs: ``dbs`` is a sequence of
        ``(rows, cols, n_transactions, n_items)`` numpy arrays / ints.
        Returns one ``Bits`` per database."""
        n = len(dbs)
        if n == 0:
            return []
        keep = [(np.ascontiguousarray(r, dtype=np.uint64), np.ascontiguousarray(c, dtype=np.uint32))
                for r, c, _, _ in dbs]
        rows = (_vp * n)(*[_vp(r.ctypes.data) if r.size else None for r, _ in keep])
        cols = (_vp * n)(*[_vp(c.ctypes.data) if c.size else None for _, c in keep])
        nnz = np.ascontiguousarray([r.size for r, _ in keep], dtype=np.uint64)
        nt = np.ascontiguousarray([int(d[2]) for d in dbs], dtype=np.uint64)
        ni = np.ascontiguousarray([int(d[3]) for d in dbs], dtype=np.uint32)
        out = (_vp * n)()
        check(load().cpm_bits_build_batch(self.handle, n, rows, cols, _p(nnz, _u64p), _p(nt, _u64p),
                                          _p(ni, _u32p), out))
        return [Bits(self, _vp(out[i])) for i in range(n)]

    # ---- operators -----------------------------------------------------------
    def count_supports(self, bits, candidates):
        cand = np.ascontiguousarray(candidates, dtype=np.uint32)
        if cand.ndim != 2:
            raise ValueError("candidates must be a 2-D array of item columns")
        out = np.zeros(cand.shape[0], dtype=np.uint64)
        if cand.shape[0]:
            check(load().cpm_count_supports(self.handle, bits.handle, _p(cand, _u32p),
                                            cand.shape[0], cand.shape[1], _p(out, _u64p)))
        return out

    def count_supports_noalloc(self, bits, d_cand, n_cand, width, d_supports):
        """``count_supports`` on device arrays (addresses as ints: candidates
        u32[n_cand x width], supports u64[n_cand]) without device allocation."""
        check(load().cpm_count_supports_noalloc(self.handle, bits.handle, _vp(int(d_cand)), int(n_cand),
                                                int(width), _vp(int(d_supports))))

    def intersect_extensions(self, prefix, extensions, keep_min=0, want_rows=True):
        prefix = np.ascontiguousarray(prefix, dtype=np.uint64)
        ext = np.ascontiguousarray(extensions, dtype=np.uint64)
        k = ext.shape[0]
        words = prefix.shape[0]
        if ext.ndim != 2 or ext.shape[1] != words:
            raise ValueError("extensions must be (k, words) with the prefix's word count")
        sup = np.zeros(k, dtype=np.uint64)
        rows = np.zeros((k, words), dtype=np.uint64) if want_rows else None
        idx = np.zeros(max(k, 1), dtype=np.uint32)
        nk = ctypes.c_uint32(0)
        if k:
            check(load().cpm_intersect_extensions(
                self.handle, _p(prefix, _u64p), _p(ext, _u64p), k, words, int(keep_min),
                _p(sup, _u64p), _p(rows, _u64p) if want_rows else None, _p(idx, _u32p),
                ctypes.byref(nk)))
        n = nk.value
        return sup, (rows[:n] if want_rows else None), idx[:n]

    def pair_gram(self, bits, cols):
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        n = cols.size
        out = np.zeros((n, n), dtype=np.uint32)
        if n:
            check(load().cpm_pair_gram(self.handle, bits.handle, _p(cols, _u32p), n,
                                          _p(out, _u32p)))
        return out

    def periodic_eval(self, bits, timestamps, end, candidates):
        cand = np.ascontiguousarray(candidates, dtype=np.uint32)
        ts = np.ascontiguousarray(timestamps, dtype=np.int64)
        if ts.size == 0:
            ts = np.zeros(1, dtype=np.int64)
        sup = np.zeros(cand.shape[0], dtype=np.uint64)
        per = np.zeros(cand.shape[0], dtype=np.int64)
        if cand.shape[0]:
            check(load().cpm_periodic_eval(self.handle, bits.handle, _p(ts, _i64p), int(end),
                                           _p(cand, _u32p), cand.shape[0], cand.shape[1],
                                           _p(sup, _u64p), _p(per, _i64p)))
        return sup, per

    # ---- searches -------------------------------------------------------------
    def eclat_mine(self, bits, minsup, use_gram=True):
        handle = _vp()
        check(load().cpm_eclat_mine(self.handle, bits.handle, clamp_u64(minsup),
                                    1 if use_gram else 0, ctypes.byref(handle)))
        return Patterns(handle)

    def eclat_mine_batch(self, bits_list, minsups, use_gram=True):
        """One search over several databases; ``Patterns.db`` tells them apart."""
        n = len(bits_list)
        handles = (_vp * max(n, 1))(*[b.handle for b in bits_list])
        ms = np.ascontiguousarray([clamp_u64(m) for m in minsups], dtype=np.uint64)
        handle = _vp()
        check(load().cpm_eclat_mine_batch(self.handle, handles, _p(ms, _u64p) if n else None, n,
                                          1 if use_gram else 0, ctypes.byref(handle)))
        return Patterns(handle, databases=True)

    def set_search_options(self, group=32, tile_words=1024, use_graph=True, subset_prune=True):
        check(load().cpm_set_search_options(self.handle, int(group), int(tile_words),
                                            1 if use_graph else 0, 1 if subset_prune else 0))

    def pfeclat_mine(self, bits, timestamps, end, minsup, maxper_floor):
        ts = np.ascontiguousarray(timestamps, dtype=np.int64)
        if ts.size == 0:
            ts = np.zeros(1, dtype=np.int64)
        handle = _vp()
        check(load().cpm_pfeclat_mine(self.handle, bits.handle, _p(ts, _i64p), int(end),
                                      clamp_u64(minsup), int(maxper_floor), ctypes.byref(handle)))
        return Patterns(handle, periodic=True)

    def pfeclat_mine_device(self, bits, timestamps_ptr, end, minsup, maxper_floor):
        """cpm_pfeclat_mine with device-resident timestamps (an int64 device
        pointer, e.g. ``tensor.data_ptr()`` of a torch tensor on this device)."""
        handle = _vp()
        check(load().cpm_pfeclat_mine_device(self.handle, bits.handle, ctypes.c_void_p(int(timestamps_ptr)),
                                             int(end), clamp_u64(minsup), int(maxper_floor),
                                             ctypes.byref(handle)))
        return Patterns(handle, periodic=True)


    # ---- FP-tree ----------------------------------------------------------------
    def unique_pairs(self, rows, cols):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        out_r = np.zeros(max(rows.size, 1), dtype=np.uint64)
        out_c = np.zeros(max(rows.size, 1), dtype=np.uint32)
        n = ctypes.c_uint64(0)
        if rows.size:
            check(load().cpm_unique_pairs(self.handle, _p(rows, _u64p), _p(cols, _u32p), rows.size,
                                          _p(out_r, _u64p), _p(out_c, _u32p), ctypes.byref(n)))
        return out_r[:n.value], out_c[:n.value]

    def fptree_build(self, rows, cols, weights, item_count, min_support, supports=None):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        weights = np.ascontiguousarray(weights, dtype=np.uint64)
        sup = None if supports is None else np.ascontiguousarray(supports, dtype=np.uint64)
        handle = _vp()
        check(load().cpm_fptree_build(self.handle, _p(rows, _u64p), _p(cols, _u32p), rows.size,
                                      _p(weights, _u64p), weights.size, int(item_count),
                                      clamp_u64(min_support),
                                      None if sup is None else _p(sup, _u64p),
                                      ctypes.byref(handle)))
        return FPTreeHandle(self, handle) if handle.value else None

    def fptree_conditional(self, tree, item, min_support):
        handle = _vp()
        check(load().cpm_fptree_conditional(self.handle, tree.handle, int(item),
                                            clamp_u64(min_support), ctypes.byref(handle)))
        return FPTreeHandle(self, handle) if handle.value else None

    def fpgrowth_mine(self, rows, cols, n_transactions, n_items, minsup):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        handle = _vp()
        check(load().cpm_fpgrowth_mine(self.handle, _p(rows, _u64p), _p(cols, _u32p), rows.size,
                                       int(n_transactions), int(n_items), clamp_u64(minsup),
                                       ctypes.byref(handle)))
        return Patterns(handle)


class FPTreeHandle:
    """Device FP-tree (``cpm_fptree``) with the reference's array view."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    def close(self):
        if self.handle:
            load().cpm_fptree_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        """``(parents, items, counts, header, single_path, depth)``."""
        n = ctypes.c_uint64()
        depth = ctypes.c_uint32()
        single = ctypes.c_int()
        hlen = ctypes.c_uint32()
        lib = load()
        check(lib.cpm_fptree_info(self.handle, ctypes.byref(n), ctypes.byref(depth),
                                  ctypes.byref(single), ctypes.byref(hlen)))
        parents = np.zeros(n.value, dtype=np.int64)
        items = np.zeros(n.value, dtype=np.int64)
        counts = np.zeros(n.value, dtype=np.uint64)
        hi = np.zeros(max(hlen.value, 1), dtype=np.uint32)
        hs = np.zeros(max(hlen.value, 1), dtype=np.uint64)
        check(lib.cpm_fptree_download(self.handle, _p(parents, _i64p), _p(items, _i64p),
                                      _p(counts, _u64p), _p(hi, _u32p), _p(hs, _u64p)))
        header = list(zip(hi[:hlen.value].tolist(), hs[:hlen.value].tolist()))
        return parents, items, counts, header, bool(single.value), int(depth.value)


def clamp_u64(value):
    value = int(value)
    return 0 if value < 0 else min(value, 2**64 - 1)


class Bits:
    """Device vertical store (``cpm_bits``)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    def close(self):
        if self.handle:
            load().cpm_bits_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self):
        n = ctypes.c_uint64()
        i = ctypes.c_uint32()
        w = ctypes.c_uint64()
        s = ctypes.c_uint64()
        check(load().cpm_bits_shape(self.handle, ctypes.byref(n), ctypes.byref(i),
                                    ctypes.byref(w), ctypes.byref(s)))
        return n.value, i.value, w.value, s.value

    def labels(self):
        """Item id of each column (stores built from basket arrays)."""
        _, items, _, _ = self.shape()
        out = np.zeros(items, dtype=np.uint32)
        check(load().cpm_bits_labels(self.ctx.handle, self.handle, _p(out, _u32p)))
        return out

    def download(self):
        _, items, words, _ = self.shape()
        out = np.zeros((items, words), dtype=np.uint64)
        if out.size:
            check(load().cpm_bits_download(self.ctx.handle, self.handle, _p(out, _u64p)))
        return out


class Baskets:
    """Basket arrays parsed on the device (``cpm_baskets``)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        nb, ni, ne = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        mx = ctypes.c_uint32()
        check(load().cpm_baskets_info(handle, ctypes.byref(nb), ctypes.byref(ni), ctypes.byref(ne),
                                      ctypes.byref(mx)))
        self.n_baskets, self.n_ids, self.n_nonempty, self.max_id = nb.value, ni.value, ne.value, mx.value

    def close(self):
        if self.handle:
            load().cpm_baskets_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        """``(offsets u64[n_baskets + 1], ids u32[n_ids])`` on the host."""
        off = np.zeros(self.n_baskets + 1, dtype=np.uint64)
        ids = np.zeros(max(self.n_ids, 1), dtype=np.uint32)
        check(load().cpm_baskets_download(self.ctx.handle, self.handle, _p(off, _u64p), _p(ids, _u32p)))
        return off, ids[: self.n_ids]

    def build_bits(self, keep_empty=False):
        """The store and each column's id (``cpm_bits_build_baskets``)."""
        handle = _vp()
        check(load().cpm_bits_build_baskets(self.ctx.handle, self.handle, int(bool(keep_empty)),
                                            ctypes.byref(handle)))
        bits = Bits(self.ctx, handle)
        return bits, bits.labels()


class _PatternsHandle:
    """Owns a ``cpm_patterns`` result; the arrays viewed from it keep it alive."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        if self.handle and _lib is not None:
            _lib.cpm_patterns_destroy(self.handle)
            self.handle = None


def _label_buffer(labels):
    """UTF-8 labels concatenated, and their byte offsets (n + 1)."""
    encoded = [label.encode("utf-8") for label in labels]
    off = np.zeros(len(encoded) + 1, dtype=np.uint64)
    if encoded:
        off[1:] = np.cumsum([len(e) for e in encoded], dtype=np.uint64)
    return b"".join(encoded), off


def _view(owner, address, n, ctype, dtype):
    """numpy view of n elements at `address`, keeping `owner` alive."""
    buf = (ctype * n).from_address(address)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype)


class Patterns:
    """Pattern forest returned by a search (``cpm_patterns``): numpy views of
    the library's result arrays (no copy; the result is released when the
    last view goes)."""

    def __init__(self, handle, periodic=False, databases=False):
        lib = load()
        owner = _PatternsHandle(handle)
        self._owner = owner
        n = ctypes.c_uint64()
        check(lib.cpm_patterns_count(handle, ctypes.byref(n)))
        self.n = n.value
        self.periodic = periodic
        parent, item, db = _u32p(), _u32p(), _u32p()
        support, period = _u64p(), _i64p()
        check(lib.cpm_patterns_view(handle, ctypes.byref(parent), ctypes.byref(item),
                                    ctypes.byref(support), ctypes.byref(period),
                                    ctypes.byref(db)))
        if self.n:
            addr = lambda p: ctypes.cast(p, ctypes.c_void_p).value
            self.parent = _view(owner, addr(parent), self.n, ctypes.c_uint32, np.uint32)
            self.item = _view(owner, addr(item), self.n, ctypes.c_uint32, np.uint32)
            self.support = _view(owner, addr(support), self.n, ctypes.c_uint64, np.uint64)
            self.period = (_view(owner, addr(period), self.n, ctypes.c_int64, np.int64)
                           if periodic else None)
            self.db = (_view(owner, addr(db), self.n, ctypes.c_uint32, np.uint32)
                       if databases and db else np.zeros(self.n, dtype=np.uint32))
        else:
            self.parent = np.zeros(0, dtype=np.uint32)
            self.item = np.zeros(0, dtype=np.uint32)
            self.support = np.zeros(0, dtype=np.uint64)
            self.period = np.zeros(0, dtype=np.int64) if periodic else None
            self.db = np.zeros(0, dtype=np.uint32)

    def __len__(self):
        return self.n

    def itemsets(self):
        """Column tuples, one per pattern (parents precede children)."""
        out = [None] * self.n
        for i, p, c in zip(range(self.n), self.parent.tolist(), self.item.tolist()):
            out[i] = (c,) if p == NO_PARENT else out[p] + (c,)
        return out

    def split_arrays(self, n_db):
        """Per-database results as ``PatternArrays`` (the forest cut by
        database with numpy: a stable sort by database keeps parents before
        children, and parents are renumbered within each database)."""
        counts = np.bincount(self.db, minlength=n_db).astype(np.int64)
        starts = np.zeros(n_db + 1, dtype=np.int64)
        starts[1:] = np.cumsum(counts)
        order = np.argsort(self.db, kind="stable")
        local = np.empty(self.n, dtype=np.int64)
        local[order] = np.arange(self.n, dtype=np.int64) - np.repeat(starts[:-1], counts)
        parent = self.parent[order]
        has = parent != NO_PARENT
        parent = parent.copy()
        parent[has] = local[parent[has]].astype(np.uint32)
        item, support = self.item[order], self.support[order]
        return [PatternArrays(parent[a:b], item[a:b], support[a:b]) for a, b in zip(starts[:-1], starts[1:])]

    def split(self, n_db):
        """Per-database ``{column tuple: support}`` dictionaries."""
        sets = self.itemsets()
        out = [dict() for _ in range(n_db)]
        for s, v, d in zip(sets, self.support.tolist(), self.db.tolist()):
            out[d][s] = v
        return out

    def keys(self, labels):
        """The key of every pattern (its labels joined by a tab, in column
        order), in result order, rendered by the library."""
        if self.n == 0:
            return []
        labels = [str(label) for label in labels]
        sep = next((c for c in ("\n", "\x00", "\x01", "\x1f") if not any(c in lab for lab in labels)), None)
        if sep is None:  # every candidate separator occurs in some label
            return ["\t".join(labels[c] for c in s) for s in self.itemsets()]
        text, off = _label_buffer(labels)
        size = ctypes.c_uint64()
        lib = load()
        check(lib.cpm_patterns_keys(self._owner.handle, text, _p(off, _u64p), len(labels), sep.encode(), None, 0,
                                    ctypes.byref(size)))
        buf = ctypes.create_string_buffer(max(size.value, 1))
        check(lib.cpm_patterns_keys(self._owner.handle, text, _p(off, _u64p), len(labels), sep.encode(), buf,
                                    size.value, ctypes.byref(size)))
        return buf.raw[:size.value].decode("utf-8").split(sep)

    def save(self, path, labels):
        """The reference's save file for this result: one line per pattern,
        sorted by label tuple, written by the library."""
        text, off = _label_buffer([str(label) for label in labels])
        check(load().cpm_patterns_save(self._owner.handle, text, _p(off, _u64p), len(off) - 1,
                                       os.fsencode(str(path))))

    def to_dict(self, labels=None):
        """``{itemset: support}`` (or ``(support, period)``), keyed by column tuples
        or, with ``labels``, by label tuples."""
        sets = self.itemsets()
        if labels is not None:
            lab = list(labels)
            sets = [tuple(lab[c] for c in s) for s in sets]
        sup = self.support.tolist()
        if self.periodic:
            return dict(zip(sets, zip(sup, self.period.tolist())))
        return dict(zip(sets, sup))


def _mix64(x):
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    x = x.astype(np.uint64, copy=True)
    x ^= x >> np.uint64(30)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(27)
    x *= np.uint64(0x94D049BB133111EB)
    x ^= x >> np.uint64(31)
    return x


def forest_digest(parent, item, support):
    """``PatternArrays.digest`` of a forest given as arrays."""
    n = len(item)
    if n == 0:
        return (0, 0)
    parent = np.asarray(parent, dtype=np.int64)
    parent = np.where(parent == NO_PARENT, -1, parent)
    hi = _mix64(np.asarray(item, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    with np.errstate(over="ignore"):
        h = hi.copy()
        p = parent.copy()
        live = p >= 0
        while live.any():
            idx = np.nonzero(live)[0]
            h[idx] += hi[p[idx]]
            p[idx] = parent[p[idx]]
            live[idx] = p[idx] >= 0
        m = _mix64(h ^ _mix64(np.asarray(support, dtype=np.uint64)))
        total = int(m.sum(dtype=np.uint64))
    return (n, total)


class PatternArrays(Mapping):
    """One database's result as arrays (parent, item, support: the forest
    of ``Patterns``), read as the mapping ``{column tuple: support}``.  The
    dict is built on first use; ``len`` needs only the arrays."""

    def __init__(self, parent, item, support):
        self.parent, self.item, self.support = parent, item, support
        self._dict = None

    @classmethod
    def of(cls, patterns):
        return cls(patterns.parent, patterns.item, patterns.support)

    def _materialise(self):
        if self._dict is None:
            sets = [None] * len(self.item)
            for i, p, c in zip(range(len(sets)), self.parent.tolist(), self.item.tolist()):
                sets[i] = (c,) if p == NO_PARENT else sets[p] + (c,)
            self._dict = dict(zip(sets, self.support.tolist()))
        return self._dict

    def __len__(self):
        return len(self.item)

    def digest(self):
        """Order-free fingerprint ``(count, u64)`` of the result: each
        pattern's itemset hashes to the sum of its items' hashes (walked
        through its parents), mixed with its support, and the mixes are
        summed.  Equal results give equal digests whatever order their
        patterns are in; for comparing large results without building dicts."""
        return forest_digest(self.parent, self.item, self.support)

    def __getitem__(self, key):
        return self._materialise()[key]

    def __iter__(self):
        return iter(self._materialise())

    def __eq__(self, other):
        if isinstance(other, Mapping):
            return self._materialise() == dict(other.items())
        return NotImplemented

    __hash__ = None

    def __repr__(self):
        return f"PatternArrays({len(self)} patterns)"


_contexts = {}
_ctx_lock = threading.Lock()


def device_count():
    try:
        return load().cpm_device_count()
    except ImportError:
        return 0


def context(device=0):
    """Shared per-device context (created on first use)."""
    with _ctx_lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = Context(device)
            _contexts[device] = ctx
        return ctx
