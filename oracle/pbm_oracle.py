"""ctypes binding of liborc.so, the serial CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker or the CPU baseline,
never as a path of the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

_lib = None
_P32 = C.POINTER(C.c_int32)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)
_PI64 = C.POINTER(C.c_int64)
_PU8 = C.POINTER(C.c_uint8)

_SIGS = {
    "orc_lcg": (C.c_int32, [C.c_int32]),
    "orc_mix64": (C.c_uint64, [C.c_uint64]),
    "orc_request_seed": (C.c_int32, [C.c_uint64, C.c_uint64]),
    "orc_page_hash": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "orc_n_words": (C.c_uint32, [C.c_uint32]),
    "orc_clear": (None, [_PU64, C.c_uint32]),
    "orc_prefill_bernoulli": (None, [_PU64, C.c_uint32, C.c_double, C.c_uint64]),
    "orc_count_free": (C.c_int64, [_PU64, C.c_uint32]),
    "orc_rwbm_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                     C.c_uint64, C.c_uint32, _P32, _PU32]),
    "orc_rwbm_threads": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                      C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, _P32,
                                      _PU32]),
    "orc_collab_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                       C.c_uint64, C.c_uint32, C.c_int, _P32, _PU32]),
    "orc_batched_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                        C.c_uint64, C.c_uint32, C.c_int, C.c_uint32, _P32,
                                        _PU32]),
    "orc_free": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _P32]),
    "orc_batched_rank": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _P32, _PU32]),
    "orc_crw_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                    C.c_uint64, C.c_uint32, _P32, _P32, _PU32]),
    "orc_single_clock_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _PU64,
                                             C.c_uint32, _P32, _PU32]),
    "orc_dist_clock_arm": (C.c_uint32, [C.c_uint64, C.c_uint32, C.c_uint32]),
    "orc_dist_clock_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _PU32,
                                           C.c_uint32, _P32, _PU32]),
    "orc_list_build": (C.c_uint32, [_PU64, C.c_uint32, _P32]),
    "orc_list_get": (None, [_PU64, C.c_uint32, _P32, _PU64, _PU64, C.c_uint32, _P32]),
    "orc_list_put": (None, [_PU64, C.c_uint32, _P32, _PU64, _PU64, C.c_uint32, _P32]),
    "orc_rwmalloc_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                         C.c_uint64, C.c_uint32, C.c_int, _P32, _PU32]),
    "orc_getnpage_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _PU32,
                                         C.c_uint64, C.c_uint64, C.c_uint32, _P32, _PU32]),
    "orc_getnpage_spans": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _PU32,
                                        C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _P32,
                                        _PU32]),
    "orc_free_n": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _P32, _PU32]),
    "orc_malloc_serial": (C.c_uint32, [_PU64, C.c_uint32, C.c_uint32, _PI64, C.c_uint32, _PU64,
                                       C.c_uint64, C.c_uint64, C.c_uint32, _P32, _PU64, _P32,
                                       _PU32]),
    "orc_malloc_spans": (C.c_uint32, [_PU64, C.c_uint32, C.c_uint32, _PI64, C.c_uint32, _PU64,
                                      C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _P32,
                                      _PU64, _P32, _PU32]),
    "orc_check_alloc": (C.c_int, [_PU64, _PU64, C.c_uint32, C.c_uint32, _P32, _PU32]),
    "orc_serve_payload": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32]),
    "orc_rwbasic_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                        C.c_uint64, C.c_uint32, _P32, _PU32]),
    "orc_crw_basic_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                          C.c_uint64, C.c_uint32, _P32, _P32, _PU32]),
    "orc_collab_basic_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                             C.c_uint64, C.c_uint32, _P32, _PU32]),
    "orc_kmalloc_collab": (C.c_uint32, [_PU64, C.c_uint32, C.c_uint32, _PI64, _PU8, _PU8,
                                        C.c_uint32, _PU32, C.c_uint64, C.c_uint64, C.c_uint32,
                                        _PU64, _PU32]),
    "orc_kfree_collab": (C.c_uint32, [_PU64, C.c_uint32, _PI64, _PU8, C.c_uint32, _PU64]),
    "orc_kmalloc_simple": (C.c_uint32, [_PU64, C.c_uint32, C.c_uint32, _PI64, _PU8, _PU8,
                                        C.c_uint32, _PU32, C.c_uint64, C.c_uint64, C.c_uint32,
                                        _P32, _PU64, _PU32]),
    "orc_ouroboros_chunk_pages": (C.c_uint32, [C.c_uint32]),
    "orc_queue_reset": (C.c_uint32, [_PU64, C.c_uint32, _PU32, _PI64]),
    "orc_queue_get": (C.c_uint32, [_PU64, C.c_uint32, _PI64, _PU32, _P32, _PI64, C.c_uint32,
                                   C.c_uint32, _P32, _PU32]),
    "orc_queue_put": (C.c_uint32, [_PU64, C.c_uint32, _PI64, _P32, _PI64, C.c_uint32, _P32]),
    "orc_collab_yields": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, C.c_uint64,
                                       C.c_uint64, C.c_uint32, C.c_int, _P32, _PU32, C.c_uint32,
                                       _PU32]),
    "orc_rw_tas": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "orc_count_clusters": (C.c_uint64, [_PU64, C.c_uint32]),
    "orc_getx_serial": (C.c_uint32, [_PU64, C.c_uint32, _PI64, C.c_uint32, _PU32, C.c_uint64,
                                     C.c_uint64, C.c_uint32, _P32, _PU32]),
    "orc_zipf_cdf": (None, [C.c_uint32, C.c_double, C.POINTER(C.c_double)]),
    "orc_zipf_trace": (None, [C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, _P32]),
    "orc_clock_run": (None, [C.c_uint32, C.c_uint32, C.c_uint32, _P32, C.c_uint64, _P32, _P32,
                             C.POINTER(C.c_uint8), _PU32, _P32, C.c_uint64, _PU64]),
    "orc_clock_textbook": (None, [C.c_uint32, C.c_uint32, _P32, C.c_uint64, _P32, _P32,
                                  C.POINTER(C.c_uint8), _PU32, _P32, C.c_uint64, _PU64]),
    "orc_serve_checksum": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32]),
    "orc_serve_sim": (None, [C.c_int64, C.c_uint32, C.c_uint32, _PU32, _PU32, C.c_uint32, _PU32,
                             _PU32, _PU64]),
    "orc_rwbm_expected_steps": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def n_words(n_pages: int) -> int:
    return (n_pages + 63) // 64


def lcg(x: int) -> int:
    return lib().orc_lcg(x)


def mix64(z: int) -> int:
    return lib().orc_mix64(z)


def request_seed(seed: int, r: int) -> int:
    return lib().orc_request_seed(seed, r)


def page_hash(seed: int, i: int) -> int:
    return lib().orc_page_hash(seed, i)


def cleared(n_pages: int) -> np.ndarray:
    bm = np.zeros(n_words(n_pages), dtype=np.uint64)
    lib().orc_clear(_p(bm, _PU64), n_pages)
    return bm


def prefill(n_pages: int, fill: float, seed: int) -> np.ndarray:
    bm = np.zeros(n_words(n_pages), dtype=np.uint64)
    lib().orc_prefill_bernoulli(_p(bm, _PU64), n_pages, fill, seed)
    return bm


def count_free(bm: np.ndarray, n_pages: int) -> int:
    return int(lib().orc_count_free(_p(bm, _PU64), n_pages))


class State:
    """A page map plus its reservation counter, mutated by the serial policies."""

    def __init__(self, bm: np.ndarray, n_pages: int, free_count: int | None = None):
        self.bm = np.ascontiguousarray(bm, dtype=np.uint64).copy()
        self.n_pages = n_pages
        fc = count_free(self.bm, n_pages) if free_count is None else free_count
        self.free = np.array([fc], dtype=np.int64)

    def _out(self, n):
        return np.full(n, -1, dtype=np.int32), np.zeros(n, dtype=np.uint32)

    def rwbm(self, n: int, req_base: int, seed: int, budget: int = 1 << 20):
        ids, steps = self._out(n)
        lib().orc_rwbm_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                              req_base, seed, budget, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def rwbm_threads(self, n: int, req_base: int, seed: int, threads: int,
                     free_after: bool = False, budget: int = 1 << 20):
        """The RandomWalkBasic_BM policy on `threads` host threads sharing
        the map (a CPU baseline; one thread equals rwbm)."""
        ids, steps = self._out(n)
        g = lib().orc_rwbm_threads(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                   req_base, seed, budget, threads, int(free_after),
                                   _p(ids, _P32), _p(steps, _PU32))
        return ids, steps, int(g)

    def rwbasic(self, n: int, req_base: int, seed: int, budget: int = 1 << 20):
        ids, steps = self._out(n)
        lib().orc_rwbasic_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                 req_base, seed, budget, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def crw_basic(self, n: int, req_base: int, seed: int, hint: int = -1,
                  budget: int = 1 << 20):
        """Legacy ClusteredRandomWalk for one thread; returns ids, steps, final hint."""
        ids, steps = self._out(n)
        h = np.array([hint], dtype=np.int32)
        lib().orc_crw_basic_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                   req_base, seed, budget, _p(h, _P32), _p(ids, _P32),
                                   _p(steps, _PU32))
        return ids, steps, int(h[0])

    def collab_basic(self, n: int, req_base: int, seed: int, budget: int = 1 << 20):
        """Legacy CollabRW, one wave per chunk of 64 requests."""
        ids, steps = self._out(n)
        lib().orc_collab_basic_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                      req_base, seed, budget, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def kmalloc_collab(self, mem: np.ndarray, written: np.ndarray, sizes: np.ndarray,
                       page_size: int, req_base: int, seed: int, budget: int = 1 << 20):
        """KernelMallocCollab::malloc; returns (pool offsets, ~0 for none; steps)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        n = sizes.size
        off = np.zeros(n, dtype=np.uint64)
        steps = np.zeros(n, dtype=np.uint32)
        lib().orc_kmalloc_collab(_p(self.bm, _PU64), self.n_pages, page_size, _p(self.free, _PI64),
                                 _p(mem, _PU8), _p(written, _PU8), n, _p(sizes, _PU32), req_base,
                                 seed, budget, _p(off, _PU64), _p(steps, _PU32))
        return off, steps

    def kfree_collab(self, mem: np.ndarray, off: np.ndarray) -> int:
        off = np.ascontiguousarray(off, dtype=np.uint64)
        return int(lib().orc_kfree_collab(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64),
                                          _p(mem, _PU8), off.size, _p(off, _PU64)))

    def kmalloc_simple(self, mem: np.ndarray, written: np.ndarray, sizes: np.ndarray,
                       page_size: int, req_base: int, seed: int, current: int = -1,
                       budget: int = 1 << 20):
        """KernelMallocSimple::malloc for one thread; returns (offsets, steps, current)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        n = sizes.size
        off = np.zeros(n, dtype=np.uint64)
        steps = np.zeros(n, dtype=np.uint32)
        cur = np.array([current], dtype=np.int32)
        lib().orc_kmalloc_simple(_p(self.bm, _PU64), self.n_pages, page_size, _p(self.free, _PI64),
                                 _p(mem, _PU8), _p(written, _PU8), n, _p(sizes, _PU32), req_base,
                                 seed, budget, _p(cur, _P32), _p(off, _PU64), _p(steps, _PU32))
        return off, steps, int(cur[0])

    def collab_yields(self, n: int, req_base: int, seed: int, order: int, max_steps: int,
                      budget: int = 1 << 20):
        """CollabRW_BM with collectWarpData: (ids, yields[waves, max_steps],
        steps per wave)."""
        waves = (n + 63) // 64
        ids = np.full(n, -1, dtype=np.int32)
        yields = np.zeros(waves * max_steps, dtype=np.uint32)
        n_steps = np.zeros(waves, dtype=np.uint32)
        lib().orc_collab_yields(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                req_base, seed, budget, order, _p(ids, _P32), _p(yields, _PU32),
                                max_steps, _p(n_steps, _PU32))
        return ids, yields.reshape(waves, max_steps), n_steps

    def collab(self, n: int, req_base: int, seed: int, order: int, budget: int = 1 << 20):
        ids, steps = self._out(n)
        lib().orc_collab_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                req_base, seed, budget, order, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def batched(self, n: int, req_base: int, seed: int, order: int, cluster_words: int = 64,
                budget: int = 1 << 20):
        ids, steps = self._out(n)
        lib().orc_batched_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                 req_base, seed, budget, order, cluster_words, _p(ids, _P32),
                                 _p(steps, _PU32))
        return ids, steps

    def batched_rank(self, n: int):
        ids, steps = self._out(n)
        lib().orc_batched_rank(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                               _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def crw(self, n: int, req_base: int, seed: int, hint: int = -1, budget: int = 1 << 20):
        """CRW_BM for one thread; returns ids, steps, final hint."""
        ids, steps = self._out(n)
        h = np.array([hint], dtype=np.int32)
        lib().orc_crw_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n, req_base,
                             seed, budget, _p(h, _P32), _p(ids, _P32), _p(steps, _PU32))
        return ids, steps, int(h[0])

    def single_clock(self, n: int, arm: int = 0, budget: int = 1 << 20):
        ids, steps = self._out(n)
        a = np.array([arm], dtype=np.uint64)
        lib().orc_single_clock_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                      _p(a, _PU64), budget, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps, int(a[0])

    def dist_clock(self, n: int, arm: int, budget: int = 1 << 20):
        ids, steps = self._out(n)
        a = np.array([arm], dtype=np.uint32)
        lib().orc_dist_clock_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                    _p(a, _PU32), budget, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps, int(a[0])

    def rwmalloc(self, n: int, req_base: int, seed: int, order: int, budget: int = 1 << 20):
        ids, steps = self._out(n)
        lib().orc_rwmalloc_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                  req_base, seed, budget, order, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def get_n(self, counts: np.ndarray, req_base: int, seed: int, budget: int = 1 << 20,
              spans: int = 1):
        """getNPage; windows of `spans` x 64 words (1 = the reference's)."""
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        n = counts.size
        ids, steps = self._out(n)
        lib().orc_getnpage_spans(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), n,
                                 _p(counts, _PU32), req_base, seed, budget, spans, _p(ids, _P32),
                                 _p(steps, _PU32))
        return ids, steps

    def get_x(self, counts: np.ndarray, req_base: int, seed: int, budget: int = 1 << 20):
        """Per-thread getXPage, requests in index order."""
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        ids, steps = self._out(counts.size)
        lib().orc_getx_serial(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64), counts.size,
                              _p(counts, _PU32), req_base, seed, budget, _p(ids, _P32),
                              _p(steps, _PU32))
        return ids, steps

    def free_n(self, ids: np.ndarray, counts: np.ndarray) -> int:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        return int(lib().orc_free_n(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64),
                                    ids.size, _p(ids, _P32), _p(counts, _PU32)))

    def malloc(self, sizes: np.ndarray, page_size: int, req_base: int, seed: int,
               budget: int = 1 << 20, spans: int = 1):
        """(page, offset, header, steps) per request; getNPage windows of
        `spans` x 64 words."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        n = sizes.size
        page, steps = self._out(n)
        off = np.zeros(n, dtype=np.uint64)
        hdr = np.zeros(n, dtype=np.int32)
        lib().orc_malloc_spans(_p(self.bm, _PU64), self.n_pages, page_size, _p(self.free, _PI64),
                               n, _p(sizes, _PU64), req_base, seed, budget, spans, _p(page, _P32),
                               _p(off, _PU64), _p(hdr, _P32), _p(steps, _PU32))
        return page, off, hdr, steps

    def free_ids(self, ids: np.ndarray) -> int:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        return int(lib().orc_free(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64),
                                  ids.size, _p(ids, _P32)))


def ouroboros_chunk_pages(page_size: int) -> int:
    return int(lib().orc_ouroboros_chunk_pages(page_size))


class QueueState(State):
    """The Ouroboros comparator's queue over a page map, one request at a
    time: pages free at construction are carved in chunks of chunk_pages."""

    def __init__(self, bm: np.ndarray, n_pages: int, chunk_pages: int):
        super().__init__(bm, n_pages)
        self.cp = chunk_pages
        self.fresh = np.zeros(n_pages, dtype=np.uint32)
        self.ring = np.zeros(n_pages, dtype=np.int32)
        self.st = np.zeros(5, dtype=np.int64)
        lib().orc_queue_reset(_p(self.bm, _PU64), n_pages, _p(self.fresh, _PU32),
                              _p(self.st, _PI64))

    def get(self, n: int):
        ids, steps = self._out(n)
        lib().orc_queue_get(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64),
                            _p(self.fresh, _PU32), _p(self.ring, _P32), _p(self.st, _PI64),
                            self.cp, n, _p(ids, _P32), _p(steps, _PU32))
        return ids, steps

    def put(self, ids: np.ndarray) -> int:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        return int(lib().orc_queue_put(_p(self.bm, _PU64), self.n_pages, _p(self.free, _PI64),
                                       _p(self.ring, _P32), _p(self.st, _PI64), ids.size,
                                       _p(ids, _P32)))

    def queued(self) -> list:
        h, t = int(self.st[2]), int(self.st[3])
        return [int(self.ring[k % self.n_pages]) for k in range(h, t)]


def check_alloc(before: np.ndarray, after: np.ndarray, n_pages: int, ids: np.ndarray):
    """(code, granted): code 0 when every invariant holds."""
    before = np.ascontiguousarray(before, dtype=np.uint64)
    after = np.ascontiguousarray(after, dtype=np.uint64)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    g = C.c_uint32()
    rc = lib().orc_check_alloc(_p(before, _PU64), _p(after, _PU64), n_pages, ids.size,
                               _p(ids, _P32), C.byref(g))
    return int(rc), int(g.value)


def count_clusters(bm: np.ndarray, n_pages: int) -> int:
    bm = np.ascontiguousarray(bm, dtype=np.uint64)
    return int(lib().orc_count_clusters(_p(bm, _PU64), n_pages))


def rw_tas(T: float, A: float, n: float) -> float:
    return float(lib().orc_rw_tas(T, A, n))


def expected_steps(T: float, A: float, n: float, W: int = 64) -> float:
    return float(lib().orc_rwbm_expected_steps(T, A, n, W))


def serve_payload(seed: int, seq: int, j: int) -> int:
    return int(lib().orc_serve_payload(seed, seq, j))


def serve_checksum(seed: int, seq: int, length: int) -> int:
    return int(lib().orc_serve_checksum(seed, seq, length))


def serve_sim(free_pages: int, max_slots: int, arrival: np.ndarray, target: np.ndarray,
              n_steps: int):
    """(admit_step, finish_step, totals) of the serving loop's schedule;
    totals = allocs, frees, admitted, finished."""
    a = np.ascontiguousarray(arrival, dtype=np.uint32)
    t = np.ascontiguousarray(target, dtype=np.uint32)
    adm = np.zeros(a.size, dtype=np.uint32)
    fin = np.zeros(a.size, dtype=np.uint32)
    tot = np.zeros(4, dtype=np.uint64)
    lib().orc_serve_sim(free_pages, max_slots, a.size, _p(a, _PU32), _p(t, _PU32), n_steps,
                        _p(adm, _PU32), _p(fin, _PU32), _p(tot, _PU64))
    return adm, fin, tot


def zipf_cdf(n: int, s: float) -> np.ndarray:
    cdf = np.zeros(n, dtype=np.float64)
    lib().orc_zipf_cdf(n, s, cdf.ctypes.data_as(C.POINTER(C.c_double)))
    return cdf


def zipf_trace(n: int, s: float, seed: int, first: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.int32)
    lib().orc_zipf_trace(n, s, seed, first, count, _p(out, _P32))
    return out


class ClockCache:
    """The buffer cache's batched CLOCK schedule (orc_clock_run)."""

    def __init__(self, n_logical: int, capacity: int, batch: int, log_cap: int = 0):
        self.n_logical, self.capacity, self.batch = n_logical, capacity, batch
        self.slot_of = np.full(n_logical, -1, dtype=np.int32)
        self.owner = np.full(capacity, -1, dtype=np.int32)
        self.refbit = np.zeros(capacity, dtype=np.uint8)
        self.arm_res = np.zeros(2, dtype=np.uint32)
        self.stats = np.zeros(3, dtype=np.uint64)      # hits, misses, evictions
        self.log = np.zeros(max(log_cap, 1), dtype=np.int32)
        self.log_cap = log_cap

    def run(self, refs: np.ndarray) -> None:
        refs = np.ascontiguousarray(refs, dtype=np.int32)
        lib().orc_clock_run(self.n_logical, self.capacity, self.batch, _p(refs, _P32), refs.size,
                            _p(self.slot_of, _P32), _p(self.owner, _P32),
                            self.refbit.ctypes.data_as(C.POINTER(C.c_uint8)),
                            _p(self.arm_res, _PU32), _p(self.log, _P32) if self.log_cap else None,
                            self.log_cap, _p(self.stats, _PU64))

    def run_textbook(self, refs: np.ndarray) -> None:
        """The references one at a time by textbook second-chance CLOCK
        (orc_clock_textbook): the batch rule at batch 1, in linear time."""
        refs = np.ascontiguousarray(refs, dtype=np.int32)
        lib().orc_clock_textbook(self.n_logical, self.capacity, _p(refs, _P32), refs.size,
                                 _p(self.slot_of, _P32), _p(self.owner, _P32),
                                 self.refbit.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 _p(self.arm_res, _PU32),
                                 _p(self.log, _P32) if self.log_cap else None, self.log_cap,
                                 _p(self.stats, _PU64))

    @property
    def evictions(self) -> np.ndarray:
        return self.log[:min(int(self.stats[2]), self.log_cap)]


def dist_clock_arm(seed: int, slot: int, n_pages: int) -> int:
    return int(lib().orc_dist_clock_arm(seed, slot, n_pages))


class ListState:
    """The LinkedList FIFO (a ring of n_pages slots) plus its page map."""

    def __init__(self, bm: np.ndarray, n_pages: int):
        self.bm = np.ascontiguousarray(bm, dtype=np.uint64).copy()
        self.n_pages = n_pages
        self.q = np.zeros(n_pages, dtype=np.int32)
        self.ht = np.zeros(2, dtype=np.uint64)
        self.ht[1] = lib().orc_list_build(_p(self.bm, _PU64), n_pages, _p(self.q, _P32))

    def get(self, n: int) -> np.ndarray:
        ids = np.full(n, -1, dtype=np.int32)
        h, t = self.ht[0:1], self.ht[1:2]
        lib().orc_list_get(_p(self.bm, _PU64), self.n_pages, _p(self.q, _P32), _p(h, _PU64),
                           _p(t, _PU64), n, _p(ids, _P32))
        self.ht[0], self.ht[1] = h[0], t[0]
        return ids

    def put(self, ids: np.ndarray) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        h, t = self.ht[0:1].copy(), self.ht[1:2].copy()
        lib().orc_list_put(_p(self.bm, _PU64), self.n_pages, _p(self.q, _P32), _p(h, _PU64),
                           _p(t, _PU64), ids.size, _p(ids, _P32))
        self.ht[0], self.ht[1] = h[0], t[0]
