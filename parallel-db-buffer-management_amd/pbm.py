"""ctypes binding of libpbm.so (the C-ABI declared in include/pbm.h).

Device buffers passed to the batch entry points are plain device pointers;
callers here use torch tensors for them (torch is only the allocator and the
stream provider).  Loading fails loudly when libpbm.so has not been built:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libpbm.so"

# pbm_strategy
(RW_BM, COLLAB_RW_BM, CRW_BM, RW_MALLOC, LINKED_LIST, SINGLE_CLOCK, DIST_CLOCK, RW_BASIC,
 CRW_BASIC, COLLAB_RW_BASIC, OUROBOROS) = range(11)
STRATEGY_NAMES = {
    RW_BM: "RandomWalkBasic_BM", COLLAB_RW_BM: "CollabRW_BM", CRW_BM: "CRW_BM",
    RW_MALLOC: "RWMalloc", LINKED_LIST: "LinkedList", SINGLE_CLOCK: "SingleClock",
    DIST_CLOCK: "DistClock", RW_BASIC: "RandomWalkBasic",
    CRW_BASIC: "ClusteredRandomWalk", COLLAB_RW_BASIC: "CollabRW", OUROBOROS: "Ouroboros",
}
# pbm_order
ORDER_PACKED, ORDER_REFERENCE = 0, 1
# pbm_mode
MODE_PLAIN, MODE_BATCHED, MODE_SERIAL = 0, 1, 2
LAUNCH_NO_COUNTERS = 1
LAUNCH_ONE_WAVE = 2
LAUNCH_RANK = 4

PHASE_ALLOC, PHASE_FREE = 1, 2

ERRORS = {0: "PBM_OK", -1: "PBM_ERR_INVALID", -2: "PBM_ERR_HIP", -3: "PBM_ERR_OOM",
          -4: "PBM_ERR_UNSUPPORTED", -5: "PBM_ERR_RANGE"}


class Config(C.Structure):
    _fields_ = [("n_pages", C.c_uint32), ("page_size", C.c_uint32), ("strategy", C.c_int32),
                ("backing", C.c_int32), ("seed", C.c_uint64), ("walk_budget", C.c_uint32),
                ("max_threads", C.c_uint32), ("order", C.c_int32), ("device", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_ulonglong) for n in
                ("allocs", "frees", "failures", "steps", "probes", "atomics", "fallbacks",
                 "errors")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Ctx(C.Structure):
    _fields_ = [("bitmap", C.c_void_p), ("free_count", C.c_void_p), ("counters", C.c_void_p),
                ("pool", C.c_void_p), ("hint", C.c_void_p), ("arms", C.c_void_p),
                ("clock_arm", C.c_void_p), ("ring", C.c_void_p), ("ring_head", C.c_void_p),
                ("ring_tail", C.c_void_p), ("phase", C.c_void_p), ("scan", C.c_void_p),
                ("scan_ctl", C.c_void_p),
                ("n_pages", C.c_uint32), ("n_words", C.c_uint32), ("page_size", C.c_uint32),
                ("walk_budget", C.c_uint32), ("max_threads", C.c_uint32), ("order", C.c_int32),
                ("seed", C.c_uint64), ("n_tiles", C.c_uint32), ("scan_wpt", C.c_uint32),
                ("fresh", C.c_void_p), ("carve", C.c_void_p), ("q_avail", C.c_void_p),
                ("chunk_pages", C.c_uint32), ("scan_threads", C.c_uint32)]


class Launch(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("walk_budget", C.c_uint32), ("cluster_words", C.c_uint32), ("flags", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("get_launch", Launch), ("free_launch", Launch), ("ms_alloc", C.c_float),
                ("ms_free", C.c_float), ("candidates", C.c_uint32), ("reserved", C.c_uint32)]


TUNE_QUICK, TUNE_WALK_ONLY, TUNE_RANK_ONLY, TUNE_NO_PERSISTENT = 1, 2, 4, 8


class ShadowReport(C.Structure):
    _fields_ = [("double_grants", C.c_uint64), ("bad_frees", C.c_uint64),
                ("grant_page", C.c_int64), ("grant_owner", C.c_uint64),
                ("grant_holder", C.c_uint64), ("free_page", C.c_int64)]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


SHADOW_FREE, SHADOW_HELD = 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFE


class ServeConfig(C.Structure):
    _fields_ = [("max_slots", C.c_uint32), ("max_pages", C.c_uint32), ("max_seqs", C.c_uint32),
                ("flags", C.c_uint32), ("lat_capacity", C.c_uint32), ("reserved", C.c_uint32),
                ("seed", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("step", "submitted", "admitted", "finished", "active",
                                          "reserved_pages", "allocs", "frees", "alloc_failures",
                                          "payload_errors", "lat_samples")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


SERVE_PAYLOAD, SERVE_LATENCY = 1, 2
LAUNCH_PERSISTENT = 16
LAUNCH_EXCLUSIVE = 32
LAUNCH_ORDERED = 64


class CacheConfig(C.Structure):
    _fields_ = [("n_logical", C.c_uint32), ("capacity", C.c_uint32), ("batch", C.c_uint32),
                ("reserved", C.c_uint32), ("log_capacity", C.c_uint64)]


class CacheStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("hits", "misses", "evictions", "resident", "arm",
                                          "batches", "alloc_failures", "scan_timeouts",
                                          "invalid_refs")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class BurstTrace(C.Structure):
    _fields_ = [("n_bursts", C.c_uint32), ("reserved", C.c_uint32),
                ("alloc_offset", C.c_void_p), ("free_offset", C.c_void_p),
                ("sizes", C.c_void_p), ("free_index", C.c_void_p)]


class PbmError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "pbm_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "pbm_destroy": (C.c_int, [C.c_void_p]),
    "pbm_reset": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pbm_prefill_bernoulli": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_void_p]),
    "pbm_prefill_requests": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p]),
    "pbm_count_used": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p]),
    "pbm_free_page_percentage": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_void_p]),
    "pbm_snapshot_bitmap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "pbm_restore_bitmap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "pbm_read_free_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "pbm_shadow_enable": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "pbm_shadow_note": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                  C.c_int, C.c_void_p]),
    "pbm_shadow_read": (C.c_int, [C.c_void_p, C.POINTER(ShadowReport), C.c_void_p]),
    "pbm_shadow_owners": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "pbm_get_ctx": (C.c_int, [C.c_void_p, C.POINTER(Ctx)]),
    "pbm_read_counters": (C.c_int, [C.c_void_p, C.POINTER(Counters), C.c_void_p]),
    "pbm_clear_counters": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pbm_get_pages": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                C.POINTER(Launch), C.c_void_p]),
    "pbm_free_pages": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(Launch),
                                 C.c_void_p]),
    "pbm_get_n_pages": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.POINTER(Launch), C.c_void_p]),
    "pbm_get_x_pages": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.POINTER(Launch), C.c_void_p]),
    "pbm_free_n_pages": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.POINTER(Launch), C.c_void_p]),
    "pbm_malloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.POINTER(Launch), C.c_void_p]),
    "pbm_free_ptrs": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(Launch),
                                C.c_void_p]),
    "pbm_page_address": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    "pbm_page_id": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]),
    "pbm_collect_warp_data": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p,
                                        C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(Launch),
                                        C.c_void_p]),
    "pbm_time_next_kernel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbm_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "pbm_event_destroy": (C.c_int, [C.c_void_p]),
    "pbm_event_elapsed_ms": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "pbm_read_pool": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]),
    "pbm_kmalloc_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "pbm_kmalloc_destroy": (C.c_int, [C.c_void_p]),
    "pbm_kmalloc_reset": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pbm_kmalloc_malloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.POINTER(Launch), C.c_void_p]),
    "pbm_kmalloc_free": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(Launch),
                                   C.c_void_p]),
    "pbm_kmalloc_current": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    "pbm_count_clusters": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p]),
    "pbm_monitor": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "pbm_wall_clock_hz": (C.c_int, [C.c_int, C.POINTER(C.c_uint64)]),
    "pbm_graph_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                   C.POINTER(Launch), C.POINTER(Launch), C.c_uint32,
                                   C.POINTER(C.c_void_p)]),
    "pbm_graph_create_steps": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.POINTER(Launch), C.POINTER(Launch),
                                         C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "pbm_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pbm_graph_destroy": (C.c_int, [C.c_void_p]),
    "pbm_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                               C.POINTER(TuneResult), C.c_void_p]),
    "pbm_atomic_peak": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_double)]),
    "pbm_serve_create": (C.c_int, [C.c_void_p, C.POINTER(ServeConfig), C.POINTER(C.c_void_p)]),
    "pbm_serve_destroy": (C.c_int, [C.c_void_p]),
    "pbm_serve_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "pbm_serve_read_waits": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p]),
    "pbm_serve_run": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "pbm_serve_run_timed": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "pbm_serve_run_loop": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "pbm_serve_graph_loop": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "pbm_serve_loop_timeouts": (C.c_int, [C.c_void_p, C.POINTER(C.c_ulonglong), C.c_void_p]),
    "pbm_serve_graph": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "pbm_serve_read_stats": (C.c_int, [C.c_void_p, C.POINTER(ServeStats), C.c_void_p]),
    "pbm_serve_read_results": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pbm_serve_read_latency": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32,
                                         C.POINTER(C.c_uint32), C.c_void_p]),
    "pbm_serve_clock_hz": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "pbm_cache_create": (C.c_int, [C.c_void_p, C.POINTER(CacheConfig), C.POINTER(C.c_void_p)]),
    "pbm_cache_destroy": (C.c_int, [C.c_void_p]),
    "pbm_cache_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "pbm_cache_run_timed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                      C.c_void_p]),
    "pbm_cache_read_stats": (C.c_int, [C.c_void_p, C.POINTER(CacheStats), C.c_void_p]),
    "pbm_cache_read_evictions": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p,
                                           C.c_void_p]),
    "pbm_cache_read_slots": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    "pbm_zipf_trace": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64,
                                 C.c_void_p, C.c_int, C.c_void_p]),
    "pbm_zipf_trace_noalloc": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, C.c_uint64,
                                         C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]),
    "pbm_bursts_create": (C.c_int, [C.c_void_p, C.POINTER(BurstTrace), C.POINTER(C.c_void_p)]),
    "pbm_bursts_destroy": (C.c_int, [C.c_void_p]),
    "pbm_bursts_run": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(Launch),
                                 C.c_void_p]),
    "pbm_bursts_graph": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(Launch),
                                   C.POINTER(C.c_void_p)]),
    "pbm_bursts_read_ptrs": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p]),
    "pbm_bursts_reset": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pbm_bursts_wait_timeouts": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p]),
    "pbm_bursts_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(TuneResult),
                                      C.c_void_p]),
    "pbm_last_error": (C.c_char_p, []),
    "pbm_version": (C.c_char_p, []),
}


def exported_names() -> list[str]:
    return sorted(_SIGS)


def lib() -> C.CDLL:
    """The loaded libpbm.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no non-HIP implementation)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != 0:
        raise PbmError(code, where, lib().pbm_last_error().decode())


def launch(mode: int = MODE_PLAIN, block: int = 0, waves_per_cu: int = 0, walk_budget: int = 0,
           cluster_words: int = 0, flags: int = 0) -> Launch:
    return Launch(mode, block, waves_per_cu, walk_budget, cluster_words, flags)


def _ptr(t) -> C.c_void_p:
    """Device pointer of a torch tensor (or an int address, or None)."""
    if t is None:
        return C.c_void_p(0)
    if isinstance(t, int):
        return C.c_void_p(t)
    return C.c_void_p(t.data_ptr())


def _stream(stream) -> C.c_void_p:
    if stream is None:
        return C.c_void_p(0)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


class Pool:
    """One page pool (a pbm_handle)."""

    def __init__(self, n_pages: int, page_size: int = 4096, strategy: int = RW_BM,
                 backing: bool = False, seed: int = 42, walk_budget: int = 0,
                 max_threads: int = 0, order: int = ORDER_PACKED, device: int = 0):
        self.cfg = Config(n_pages, page_size, strategy, int(backing), seed, walk_budget,
                          max_threads, order, device)
        h = C.c_void_p()
        check(lib().pbm_create(C.byref(self.cfg), C.byref(h)), "pbm_create")
        self.h = h
        self.n_pages = n_pages
        self.n_words = (n_pages + 63) // 64

    def close(self) -> None:
        if self.h:
            check(lib().pbm_destroy(self.h), "pbm_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state
    def reset(self, stream=None) -> None:
        check(lib().pbm_reset(self.h, _stream(stream)), "pbm_reset")

    def prefill_bernoulli(self, fill: float, seed: int, stream=None) -> None:
        check(lib().pbm_prefill_bernoulli(self.h, fill, seed, _stream(stream)),
              "pbm_prefill_bernoulli")

    def prefill_requests(self, free_percentage: float, stream=None) -> None:
        check(lib().pbm_prefill_requests(self.h, free_percentage, _stream(stream)),
              "pbm_prefill_requests")

    def count_used(self, stream=None) -> int:
        v = C.c_uint64()
        check(lib().pbm_count_used(self.h, C.byref(v), _stream(stream)), "pbm_count_used")
        return int(v.value)

    def free_page_percentage(self, stream=None) -> float:
        v = C.c_float()
        check(lib().pbm_free_page_percentage(self.h, C.byref(v), _stream(stream)),
              "pbm_free_page_percentage")
        return float(v.value)

    def snapshot(self, stream=None) -> np.ndarray:
        out = np.empty(self.n_words, dtype=np.uint64)
        check(lib().pbm_snapshot_bitmap(self.h, out.ctypes.data_as(C.c_void_p), self.n_words,
                                        _stream(stream)), "pbm_snapshot_bitmap")
        return out

    def shadow_enable(self, on: bool = True, stream=None) -> None:
        """Debug owner shadow on (set from the map, report cleared) or off."""
        check(lib().pbm_shadow_enable(self.h, int(bool(on)), _stream(stream)),
              "pbm_shadow_enable")

    def shadow_note(self, n: int, ids, counts=None, req_base: int = 0, freeing: bool = False,
                    stream=None) -> None:
        check(lib().pbm_shadow_note(self.h, n, req_base, _ptr(ids), _ptr(counts),
                                    int(bool(freeing)), _stream(stream)), "pbm_shadow_note")

    def shadow_report(self, stream=None) -> dict:
        r = ShadowReport()
        check(lib().pbm_shadow_read(self.h, C.byref(r), _stream(stream)), "pbm_shadow_read")
        return r.as_dict()

    def shadow_owners(self, stream=None) -> np.ndarray:
        out = np.empty(self.n_pages, dtype=np.uint64)
        check(lib().pbm_shadow_owners(self.h, out.ctypes.data_as(C.c_void_p), self.n_pages,
                                      _stream(stream)), "pbm_shadow_owners")
        return out

    def restore(self, words: np.ndarray, stream=None) -> None:
        w = np.ascontiguousarray(words, dtype=np.uint64)
        check(lib().pbm_restore_bitmap(self.h, w.ctypes.data_as(C.c_void_p), w.size,
                                       _stream(stream)), "pbm_restore_bitmap")

    def free_count(self, stream=None) -> int:
        v = C.c_int64()
        check(lib().pbm_read_free_count(self.h, C.byref(v), _stream(stream)),
              "pbm_read_free_count")
        return int(v.value)

    def ctx(self) -> Ctx:
        c = Ctx()
        check(lib().pbm_get_ctx(self.h, C.byref(c)), "pbm_get_ctx")
        return c

    def counters(self, stream=None) -> dict:
        c = Counters()
        check(lib().pbm_read_counters(self.h, C.byref(c), _stream(stream)), "pbm_read_counters")
        return c.as_dict()

    def clear_counters(self, stream=None) -> None:
        check(lib().pbm_clear_counters(self.h, _stream(stream)), "pbm_clear_counters")

    # ---- batch entry points (device pointers)
    def get_pages(self, n: int, ids, steps=None, req_base: int = 0, lnch: Launch | None = None,
                  stream=None) -> None:
        check(lib().pbm_get_pages(self.h, n, req_base, _ptr(ids), _ptr(steps),
                                  C.byref(lnch or launch()), _stream(stream)), "pbm_get_pages")

    def free_pages(self, n: int, ids, lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_free_pages(self.h, n, _ptr(ids), C.byref(lnch or launch()),
                                   _stream(stream)), "pbm_free_pages")

    def count_clusters(self, stream=None) -> int:
        v = C.c_uint64()
        check(lib().pbm_count_clusters(self.h, C.byref(v), _stream(stream)), "pbm_count_clusters")
        return int(v.value)

    def monitor(self, gap_us: int, count: int, samples, stream=None) -> None:
        """Launches the sampling monitor; samples: device int64 tensor of
        2*count entries (tick, free pages)."""
        check(lib().pbm_monitor(self.h, gap_us, count, _ptr(samples), _stream(stream)),
              "pbm_monitor")

    def get_n_pages(self, n: int, counts, ids, steps=None, req_base: int = 0,
                    lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_get_n_pages(self.h, n, req_base, _ptr(counts), _ptr(ids), _ptr(steps),
                                    C.byref(lnch or launch()), _stream(stream)),
              "pbm_get_n_pages")

    def get_x_pages(self, n: int, counts, ids, steps=None, req_base: int = 0,
                    lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_get_x_pages(self.h, n, req_base, _ptr(counts), _ptr(ids), _ptr(steps),
                                    C.byref(lnch or launch()), _stream(stream)),
              "pbm_get_x_pages")

    def free_n_pages(self, n: int, ids, counts, lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_free_n_pages(self.h, n, _ptr(ids), _ptr(counts),
                                     C.byref(lnch or launch()), _stream(stream)),
              "pbm_free_n_pages")

    def malloc(self, n: int, sizes, ptrs, steps=None, req_base: int = 0,
               lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_malloc(self.h, n, req_base, _ptr(sizes), _ptr(ptrs), _ptr(steps),
                               C.byref(lnch or launch()), _stream(stream)), "pbm_malloc")

    def free_ptrs(self, n: int, ptrs, lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_free_ptrs(self.h, n, _ptr(ptrs), C.byref(lnch or launch()),
                                  _stream(stream)), "pbm_free_ptrs")

    def page_address(self, page: int) -> int:
        v = C.c_uint64()
        check(lib().pbm_page_address(self.h, page, C.byref(v)), "pbm_page_address")
        return int(v.value)

    def page_id(self, addr: int) -> int:
        v = C.c_int32()
        check(lib().pbm_page_id(self.h, addr, C.byref(v)), "pbm_page_id")
        return int(v.value)

    def collect_warp_data(self, n: int, ids, yields, max_steps: int, n_steps, req_base: int = 0,
                          lnch: Launch | None = None, stream=None) -> None:
        """CollabRW_BM getPage with collectWarpData (pbm_collect_warp_data)."""
        check(lib().pbm_collect_warp_data(self.h, n, req_base, _ptr(ids), _ptr(yields), max_steps,
                                          _ptr(n_steps), C.byref(lnch or launch()),
                                          _stream(stream)), "pbm_collect_warp_data")

    def time_next_kernel(self, start: "KernelEvent", stop: "KernelEvent") -> None:
        """The next get_pages / free_pages call records start and stop at its
        main kernel's own start and end (pbm_time_next_kernel)."""
        check(lib().pbm_time_next_kernel(self.h, start.e, stop.e), "pbm_time_next_kernel")

    def read_pool(self, offset: int, nbytes: int, stream=None) -> np.ndarray:
        """Bytes [offset, offset+nbytes) of page memory."""
        out = np.empty(nbytes, dtype=np.uint8)
        check(lib().pbm_read_pool(self.h, offset, nbytes, out.ctypes.data_as(C.c_void_p),
                                  _stream(stream)), "pbm_read_pool")
        return out


class KernelEvent:
    """A hipEvent_t for pbm_time_next_kernel."""

    def __init__(self):
        e = C.c_void_p()
        check(lib().pbm_event_create(C.byref(e)), "pbm_event_create")
        self.e = e

    def elapsed_ms(self, stop: "KernelEvent") -> float:
        v = C.c_float()
        check(lib().pbm_event_elapsed_ms(self.e, stop.e, C.byref(v)), "pbm_event_elapsed_ms")
        return float(v.value)

    def close(self) -> None:
        if self.e:
            check(lib().pbm_event_destroy(self.e), "pbm_event_destroy")
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


KMALLOC_SIMPLE, KMALLOC_COLLAB = 0, 1


class KMalloc:
    """KernelMallocSimple / KernelMallocCollab (malloc.cuh:14-37) over a
    pool with backing."""

    def __init__(self, pool: Pool, kind: int, max_threads: int = 0):
        self.pool = pool
        h = C.c_void_p()
        check(lib().pbm_kmalloc_create(pool.h, kind, max_threads, C.byref(h)),
              "pbm_kmalloc_create")
        self.k = h

    def reset(self, stream=None) -> None:
        check(lib().pbm_kmalloc_reset(self.k, _stream(stream)), "pbm_kmalloc_reset")

    def malloc(self, n: int, sizes, ptrs, steps=None, req_base: int = 0,
               lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_kmalloc_malloc(self.k, n, req_base, _ptr(sizes), _ptr(ptrs), _ptr(steps),
                                       C.byref(lnch or launch()), _stream(stream)),
              "pbm_kmalloc_malloc")

    def free(self, n: int, ptrs, lnch: Launch | None = None, stream=None) -> None:
        check(lib().pbm_kmalloc_free(self.k, n, _ptr(ptrs), C.byref(lnch or launch()),
                                     _stream(stream)), "pbm_kmalloc_free")

    def current(self) -> np.ndarray:
        m = C.c_uint32()
        check(lib().pbm_kmalloc_current(self.k, None, C.byref(m)), "pbm_kmalloc_current")
        out = np.empty(m.value, dtype=np.int32)
        check(lib().pbm_kmalloc_current(self.k, out.ctypes.data_as(C.c_void_p), None),
              "pbm_kmalloc_current")
        return out

    def close(self) -> None:
        if self.k:
            check(lib().pbm_kmalloc_destroy(self.k), "pbm_kmalloc_destroy")
            self.k = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured allocation and/or free phase (pbm_graph)."""

    def __init__(self, pool: Pool, n: int, ids, steps=None, req_base: int = 0,
                 get_launch: Launch | None = None, free_launch: Launch | None = None,
                 phases: int = 3, n_steps: int = 1):
        g = C.c_void_p()
        check(lib().pbm_graph_create_steps(pool.h, n, req_base, _ptr(ids), _ptr(steps),
                                           C.byref(get_launch or launch()),
                                           C.byref(free_launch or launch()), phases, n_steps,
                                           C.byref(g)),
              "pbm_graph_create_steps")
        self.g = g
        self.pool = pool  # keeps the pool alive while the graph exists

    def launch(self, stream=None) -> None:
        check(lib().pbm_graph_launch(self.g, _stream(stream)), "pbm_graph_launch")

    def close(self) -> None:
        if self.g:
            check(lib().pbm_graph_destroy(self.g), "pbm_graph_destroy")
            self.g = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _StepGraph(Graph):
    """A graph captured by another entry point (pbm_serve_graph)."""

    def __init__(self, g: C.c_void_p, owner):
        self.g = g
        self.pool = owner


class Server:
    """Continuous-batching serving loop over a pool (pbm_serve_*)."""

    def __init__(self, pool: Pool, max_slots: int, max_pages: int, max_seqs: int,
                 flags: int = 0, lat_capacity: int = 0, seed: int = 0):
        cfg = ServeConfig(max_slots, max_pages, max_seqs, flags, lat_capacity, 0, seed)
        s = C.c_void_p()
        check(lib().pbm_serve_create(pool.h, C.byref(cfg), C.byref(s)), "pbm_serve_create")
        self.s = s
        self.pool = pool
        self.cfg = cfg

    def submit(self, arrival_step, target, key=None, stream=None) -> None:
        a = np.ascontiguousarray(arrival_step, dtype=np.uint32)
        t = np.ascontiguousarray(target, dtype=np.uint32)
        k = None if key is None else np.ascontiguousarray(key, dtype=np.uint32)
        if a.shape != t.shape or (k is not None and k.shape != a.shape):
            raise ValueError("arrival_step, target and key differ in length")
        check(lib().pbm_serve_submit(self.s, a.size, a.ctypes.data_as(C.c_void_p),
                                     t.ctypes.data_as(C.c_void_p),
                                     None if k is None else k.ctypes.data_as(C.c_void_p),
                                     _stream(stream)), "pbm_serve_submit")

    def run(self, n_steps: int, stream=None) -> None:
        check(lib().pbm_serve_run(self.s, n_steps, _stream(stream)), "pbm_serve_run")

    def run_loop(self, n_steps: int, stream=None) -> None:
        """n_steps steps in one kernel (the device-resident loop)."""
        check(lib().pbm_serve_run_loop(self.s, n_steps, _stream(stream)), "pbm_serve_run_loop")

    def loop_timeouts(self, stream=None) -> int:
        v = C.c_ulonglong()
        check(lib().pbm_serve_loop_timeouts(self.s, C.byref(v), _stream(stream)),
              "pbm_serve_loop_timeouts")
        return int(v.value)

    def run_timed(self, n_steps: int, stream=None) -> dict:
        """Runs n_steps steps; the mean kernel-own ms of admit, grow, release."""
        out = (C.c_float * 3)()
        check(lib().pbm_serve_run_timed(self.s, n_steps, out, _stream(stream)),
              "pbm_serve_run_timed")
        return {"admit": float(out[0]), "grow": float(out[1]), "release": float(out[2])}

    def graph(self, n_steps: int) -> Graph:
        g = C.c_void_p()
        check(lib().pbm_serve_graph(self.s, n_steps, C.byref(g)), "pbm_serve_graph")
        return _StepGraph(g, self)

    def graph_loop(self, n_steps: int) -> Graph:
        """One device-resident loop launch of n_steps steps as a HIP graph."""
        g = C.c_void_p()
        check(lib().pbm_serve_graph_loop(self.s, n_steps, C.byref(g)), "pbm_serve_graph_loop")
        return _StepGraph(g, self)

    def stats(self, stream=None) -> dict:
        st = ServeStats()
        check(lib().pbm_serve_read_stats(self.s, C.byref(st), _stream(stream)),
              "pbm_serve_read_stats")
        return st.as_dict()

    def results(self, first: int = 0, n: int | None = None, stream=None) -> dict:
        if n is None:
            n = self.cfg.max_seqs - first
        out = {k: np.zeros(n, dtype=np.uint32) for k in ("admit_step", "finish_step", "length")}
        out["checksum"] = np.zeros(n, dtype=np.uint64)
        check(lib().pbm_serve_read_results(
            self.s, first, n, *(out[k].ctypes.data_as(C.c_void_p)
                                for k in ("admit_step", "finish_step", "length", "checksum")),
            _stream(stream)), "pbm_serve_read_results")
        return out

    def wait_ticks(self, first: int = 0, n: int | None = None, stream=None) -> np.ndarray:
        """Per sequence: wall-clock ticks from arrival to admission
        (0xFFFFFFFF: not admitted)."""
        if n is None:
            n = self.cfg.max_seqs - first
        buf = np.zeros(max(n, 1), dtype=np.uint32)
        check(lib().pbm_serve_read_waits(self.s, first, n, buf.ctypes.data_as(C.c_void_p),
                                         _stream(stream)), "pbm_serve_read_waits")
        return buf[:n]

    def latency_ticks(self, n: int, stream=None) -> np.ndarray:
        buf = np.zeros(max(n, 1), dtype=np.uint32)
        got = C.c_uint32()
        check(lib().pbm_serve_read_latency(self.s, buf.ctypes.data_as(C.c_void_p), n,
                                           C.byref(got), _stream(stream)),
              "pbm_serve_read_latency")
        return buf[:got.value]

    def clock_hz(self) -> int:
        hz = C.c_uint64()
        check(lib().pbm_serve_clock_hz(self.s, C.byref(hz)), "pbm_serve_clock_hz")
        return int(hz.value)

    def close(self) -> None:
        if self.s:
            check(lib().pbm_serve_destroy(self.s), "pbm_serve_destroy")
            self.s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cache:
    """Buffer cache with batched CLOCK replacement over a pool (pbm_cache_*)."""

    def __init__(self, pool: Pool, n_logical: int, capacity: int, batch: int,
                 log_capacity: int = 0):
        cfg = CacheConfig(n_logical, capacity, batch, 0, log_capacity)
        k = C.c_void_p()
        check(lib().pbm_cache_create(pool.h, C.byref(cfg), C.byref(k)), "pbm_cache_create")
        self.k = k
        self.pool = pool
        self.cfg = cfg

    def run(self, refs, n: int | None = None, stream=None) -> None:
        """refs: a device int32 tensor (or address) of logical page ids."""
        n = refs.numel() if n is None else n
        check(lib().pbm_cache_run(self.k, _ptr(refs), n, _stream(stream)), "pbm_cache_run")

    def run_timed(self, refs, n: int | None = None, stream=None) -> dict:
        """run(), with the mean kernel-own ms of hits, plan and load per batch."""
        n = refs.numel() if n is None else n
        out = (C.c_float * 3)()
        check(lib().pbm_cache_run_timed(self.k, _ptr(refs), n, out, _stream(stream)),
              "pbm_cache_run_timed")
        return {"hits": float(out[0]), "plan": float(out[1]), "load": float(out[2])}

    def stats(self, stream=None) -> dict:
        st = CacheStats()
        check(lib().pbm_cache_read_stats(self.k, C.byref(st), _stream(stream)),
              "pbm_cache_read_stats")
        return st.as_dict()

    def evictions(self, first: int = 0, n: int | None = None, stream=None) -> np.ndarray:
        if n is None:
            n = min(self.stats(stream)["evictions"], self.cfg.log_capacity) - first
        out = np.zeros(max(n, 1), dtype=np.int32)
        check(lib().pbm_cache_read_evictions(self.k, first, n, out.ctypes.data_as(C.c_void_p),
                                             _stream(stream)), "pbm_cache_read_evictions")
        return out[:n]

    def slots(self, stream=None):
        cap = self.cfg.capacity
        owner = np.zeros(cap, dtype=np.int32)
        frame = np.zeros(cap, dtype=np.int32)
        bit = np.zeros(cap, dtype=np.uint32)
        check(lib().pbm_cache_read_slots(self.k, owner.ctypes.data_as(C.c_void_p),
                                         frame.ctypes.data_as(C.c_void_p),
                                         bit.ctypes.data_as(C.c_void_p), _stream(stream)),
              "pbm_cache_read_slots")
        return owner, frame, bit

    def close(self) -> None:
        if self.k:
            check(lib().pbm_cache_destroy(self.k), "pbm_cache_destroy")
            self.k = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def zipf_trace(out, n_logical: int, s: float, seed: int, first: int = 0, device: int = 0,
               stream=None) -> None:
    """Fills the device int32 tensor `out` with a Zipf(s) trace over
    n_logical pages (pbm_zipf_trace)."""
    check(lib().pbm_zipf_trace(n_logical, s, seed, first, out.numel(), _ptr(out), device,
                               _stream(stream)), "pbm_zipf_trace")


def zipf_trace_noalloc(out, n_logical: int, s: float, seed: int, first: int = 0,
                       device: int = 0, stream=None) -> None:
    """zipf_trace with no device allocation (pbm_zipf_trace_noalloc): the
    CDF of at most 4,194,304 pages in the library's static scratch."""
    check(lib().pbm_zipf_trace_noalloc(n_logical, s, seed, first, out.numel(), _ptr(out),
                                       device, _stream(stream)), "pbm_zipf_trace_noalloc")


class Bursts:
    """A trace of mallocRW/freeRW bursts on a backed pool (pbm_bursts_*)."""

    def __init__(self, pool: Pool, trace: dict):
        self._keep = {k: np.ascontiguousarray(trace[k], dtype=t) for k, t in
                      (("alloc_offset", np.uint32), ("free_offset", np.uint32),
                       ("sizes", np.uint64), ("free_index", np.uint32))}
        k = self._keep
        tr = BurstTrace(len(k["alloc_offset"]) - 1, 0,
                        *(k[n].ctypes.data_as(C.c_void_p) for n in
                          ("alloc_offset", "free_offset", "sizes", "free_index")))
        b = C.c_void_p()
        check(lib().pbm_bursts_create(pool.h, C.byref(tr), C.byref(b)), "pbm_bursts_create")
        self.b = b
        self.pool = pool
        self.n_bursts = tr.n_bursts
        self.n_allocs = int(k["alloc_offset"][-1])

    def run(self, first: int, count: int, launch_: Launch | None = None, stream=None) -> None:
        check(lib().pbm_bursts_run(self.b, first, count, C.byref(launch_ or launch()),
                                   _stream(stream)), "pbm_bursts_run")

    def graph(self, first: int, count: int, launch_: Launch | None = None) -> Graph:
        g = C.c_void_p()
        check(lib().pbm_bursts_graph(self.b, first, count, C.byref(launch_ or launch()),
                                     C.byref(g)), "pbm_bursts_graph")
        return _StepGraph(g, self)

    def ptrs(self, first: int = 0, n: int | None = None, stream=None) -> np.ndarray:
        n = self.n_allocs - first if n is None else n
        out = np.zeros(max(n, 1), dtype=np.uint64)
        check(lib().pbm_bursts_read_ptrs(self.b, first, n, out.ctypes.data_as(C.c_void_p),
                                         _stream(stream)), "pbm_bursts_read_ptrs")
        return out[:n]

    def reset(self, stream=None) -> None:
        check(lib().pbm_bursts_reset(self.b, _stream(stream)), "pbm_bursts_reset")

    def wait_timeouts(self, stream=None) -> int:
        """Dependency waits of persistent runs that gave up since the last reset."""
        v = C.c_uint64()
        check(lib().pbm_bursts_wait_timeouts(self.b, C.byref(v), _stream(stream)),
              "pbm_bursts_wait_timeouts")
        return int(v.value)

    def autotune(self, count: int, flags: int = 0, stream=None) -> dict:
        r = TuneResult()
        check(lib().pbm_bursts_autotune(self.b, count, flags, C.byref(r), _stream(stream)),
              "pbm_bursts_autotune")
        g = r.get_launch
        return {"launch": g, "waves_per_cu": int(g.waves_per_cu), "walk_budget": int(g.walk_budget),
                "cluster_words": int(g.cluster_words) or 64,
                "persistent": bool(g.flags & LAUNCH_PERSISTENT), "ms": float(r.ms_alloc),
                "ms_longer_run": float(r.ms_free), "candidates": int(r.candidates)}

    def close(self) -> None:
        if self.b:
            check(lib().pbm_bursts_destroy(self.b), "pbm_bursts_destroy")
            self.b = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def burst_trace(n_bursts: int, median: float, sigma: float, min_size: int, max_size: int,
                page_size: int, seed: int) -> dict:
    """Bursts with lognormal request counts (median, sigma; at least 1),
    each request an alloc or a free with probability 1/2 (a free needs a
    live allocation of an earlier burst), alloc sizes log-uniform in
    [min_size, max_size] bytes.  mallocRW packs the small requests
    (size+4 rounded to 4 < page_size) of each 64-request chunk of a burst
    behind one leader, and freeing the leader returns them all, so a free
    picks a live packed group or a live large request and frees all of it.
    Returns alloc_offset, free_offset, sizes, free_index (numpy arrays) and
    the request counts."""
    rng = np.random.default_rng(seed)
    counts = np.maximum(1, np.rint(np.exp(np.log(median) + sigma * rng.standard_normal(n_bursts))))
    counts = counts.astype(np.int64)
    lo, hi = np.log(min_size), np.log(max_size)
    units: list[np.ndarray] = []        # live units (arrays of request indices)
    alloc_off, free_off = [0], [0]
    sizes_l, frees_l = [], []
    n_alloc = 0
    for b in range(n_bursts):
        c = int(counts[b])
        want_free = int((rng.random(c) < 0.5).sum())
        freed = 0
        if want_free and units:
            order = rng.permutation(len(units))
            take = []
            for u in order:
                if freed >= want_free:
                    break
                take.append(u)
                freed += len(units[u])
            chosen = set(take)
            frees_l.extend(np.concatenate([units[u] for u in take]).tolist())
            units = [u for i, u in enumerate(units) if i not in chosen]
        k = max(c - freed, 0)
        sz = np.exp(rng.uniform(lo, hi, size=k)).astype(np.uint64)
        sizes_l.append(sz)
        small = ((sz + 4 + 3) & ~np.uint64(3)) < page_size
        idx = np.arange(n_alloc, n_alloc + k)
        for ch in range(0, k, 64):
            s = small[ch:ch + 64]
            if s.any():
                units.append(idx[ch:ch + 64][s])
            units.extend(idx[ch:ch + 64][~s].reshape(-1, 1))
        n_alloc += k
        alloc_off.append(n_alloc)
        free_off.append(len(frees_l))
    return {"alloc_offset": np.array(alloc_off, dtype=np.uint32),
            "free_offset": np.array(free_off, dtype=np.uint32),
            "sizes": np.concatenate(sizes_l) if sizes_l else np.zeros(0, np.uint64),
            "free_index": np.array(frees_l, dtype=np.uint32), "counts": counts}


def burst_live_runs(trace: dict, page_size: int) -> dict:
    """The runs a burst trace leaves allocated, from its requests alone:
    {request that owns the run: (pages, requests packed into it)}.
    mallocRW packs the small requests (size + 4, rounded up to 4, below the
    page size) of each 64-request chunk of a burst into one run owned by
    the first of them (ceil(bytes / page_size) pages); a larger request
    owns ceil((size + 4) / page_size) pages.  A run is live while the
    request that owns it is not freed (freeing a packed member frees
    nothing)."""
    a_off, sizes = trace["alloc_offset"], np.asarray(trace["sizes"], dtype=np.uint64)
    freed = np.zeros(len(sizes), dtype=bool)
    freed[np.asarray(trace["free_index"], dtype=np.int64)] = True
    need = (sizes + np.uint64(4 + 3)) & ~np.uint64(3)
    small = need < np.uint64(page_size)
    # chunk start of every request
    burst = np.repeat(np.arange(len(a_off) - 1), np.diff(a_off.astype(np.int64)))
    idx = np.arange(len(sizes), dtype=np.int64)
    start = a_off[burst].astype(np.int64)
    chunk = start + (idx - start) // 64 * 64
    runs = {}
    # packed runs: per chunk, the small requests' bytes
    cs = chunk[small]
    if cs.size:
        uniq, first_pos, inv = np.unique(cs, return_index=True, return_inverse=True)
        tot = np.bincount(inv, weights=need[small].astype(np.float64)).astype(np.int64)
        members = np.split(idx[small][np.argsort(inv, kind="stable")],
                           np.cumsum(np.bincount(inv))[:-1])
        leaders = idx[small][first_pos]
        for u in np.nonzero(~freed[leaders])[0]:
            runs[int(leaders[u])] = (int((tot[u] + page_size - 1) // page_size), members[u])
    big = np.nonzero(~small & ~freed)[0]
    pages = ((need[big] + np.uint64(page_size - 1)) // np.uint64(page_size)).astype(np.int64)
    for r, n in zip(big.tolist(), pages.tolist()):
        runs[r] = (n, np.array([r]))
    return runs


def check_burst_end_state(trace: dict, ptrs: np.ndarray, base: int, page_size: int,
                          before: np.ndarray, after: np.ndarray) -> dict:
    """Checks the end state of a burst run against the trace: every live
    run starts on a page inside the pool, live runs are disjoint, each of
    their pages is used now and was free in `before`, and the pool holds
    exactly `before` plus the live runs' pages.  Returns the live runs'
    leader pointers and page count; raises AssertionError on a violation."""
    runs = burst_live_runs(trace, page_size)
    lead = np.array(sorted(runs), dtype=np.int64)
    n_pages = np.array([runs[k][0] for k in lead], dtype=np.int64)
    p = ptrs[lead].astype(np.int64) if lead.size else np.zeros(0, np.int64)
    ok = p > 0
    start = (p[ok] - 4 - base) // page_size
    assert ((p[ok] - 4 - base) % page_size == 0).all(), "a live run does not start on a page"
    cnt = n_pages[ok]
    used = np.zeros(before.size * 64, dtype=bool)
    order = np.argsort(start)
    s, c = start[order], cnt[order]
    assert (s >= 0).all() and (s + c <= before.size * 64).all(), "a live run leaves the pool"
    assert (s[1:] >= s[:-1] + c[:-1]).all(), "two live runs overlap"
    for a, n in zip(s.tolist(), c.tolist()):
        used[a:a + n] = True
    bits_b = np.unpackbits(before.view(np.uint8), bitorder="little").astype(bool)
    bits_a = np.unpackbits(after.view(np.uint8), bitorder="little").astype(bool)
    assert not (used & bits_b).any(), "a live run holds a page used before the trace"
    assert np.array_equal(bits_a, bits_b | used), "the map is not the pre-fill plus the live runs"
    return {"leaders": p[ok].astype(np.uint64), "live_runs": int(ok.sum()),
            "live_pages": int(c.sum()), "failed_runs": int((~ok).sum())}


def serve_trace(n: int, rate_per_s: float, step_s: float, median: float, sigma: float,
                seed: int, max_pages: int):
    """Poisson arrivals at rate_per_s (exponential gaps from a seeded numpy
    generator), binned into steps of step_s seconds (a sequence arriving in
    [ (k-1) step_s, k step_s ) may be admitted from step k), and lognormal
    target lengths (median, sigma) rounded and clipped to [1, max_pages]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))
    arrival = (np.floor(t / step_s) + 1).astype(np.uint32)
    target = np.rint(np.exp(np.log(median) + sigma * rng.standard_normal(n)))
    target = np.clip(target, 1, max_pages).astype(np.uint32)
    return arrival, target


def wall_clock_hz(device: int = 0) -> int:
    hz = C.c_uint64()
    check(lib().pbm_wall_clock_hz(device, C.byref(hz)), "pbm_wall_clock_hz")
    return int(hz.value)


def atomic_peak(words: int = 16384, rounds: int = 64, device: int = 0) -> float:
    """Device-scope u64 atomicOr rate (ops/s) over `words` spread words."""
    v = C.c_double()
    check(lib().pbm_atomic_peak(words, rounds, device, C.byref(v)), "pbm_atomic_peak")
    return float(v.value)


def autotune(pool: Pool, n: int, ids, flags: int = 0, stream=None) -> dict:
    """Fastest batched alloc launch for n requests (ids: device buffer of n
    int32) and the fastest free launch after it."""
    r = TuneResult()
    check(lib().pbm_autotune(pool.h, n, _ptr(ids), flags, C.byref(r), _stream(stream)),
          "pbm_autotune")
    g = r.get_launch
    return {"get_launch": g, "free_launch": r.free_launch,
            "rank": bool(g.flags & LAUNCH_RANK), "cluster_words": int(g.cluster_words),
            "waves_per_cu": int(g.waves_per_cu), "walk_budget": int(g.walk_budget),
            "free_waves_per_cu": int(r.free_launch.waves_per_cu),
            "ms_alloc": float(r.ms_alloc), "ms_free": float(r.ms_free),
            "candidates": int(r.candidates)}
