"""mallocRW/freeRW bursts through the C-ABI: a serial launch reproduces the
serial oracle burst by burst; concurrent launches (plain, persistent
batched, graph-captured, one cooperative kernel) keep every live
allocation's pages disjoint and in use, and hand every page back."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu

PS = 4096


def _units(tr):
    """Live units left by the trace: {leader request: pages needed}."""
    a_off, sizes = tr["alloc_offset"], tr["sizes"]
    freed = np.zeros(len(sizes), dtype=bool)
    freed[tr["free_index"]] = True
    need = (sizes + 4 + 3) & ~np.uint64(3)
    units = {}
    for b in range(len(a_off) - 1):
        a0, a1 = int(a_off[b]), int(a_off[b + 1])
        for ch in range(a0, a1, 64):
            idx = np.arange(ch, min(ch + 64, a1))
            small = need[idx] < PS
            if small.any():
                lead = int(idx[small][0])
                units[lead] = (int((need[idx][small].sum() + PS - 1) // PS), idx[small])
            for r in idx[~small]:
                units[int(r)] = (int((need[r] + PS - 1) // PS), np.array([r]))
    return {k: v for k, v in units.items() if not freed[v[1]].any()}


def _oracle(tr, T, fill, pseed, seed):
    st = orc.State(orc.prefill(T, fill, pseed), T)
    n = int(tr["alloc_offset"][-1])
    page = np.full(n, -1, dtype=np.int64)
    off = np.zeros(n, dtype=np.int64)
    hdr = np.zeros(n, dtype=np.int64)
    for b in range(len(tr["alloc_offset"]) - 1):
        for i in tr["free_index"][tr["free_offset"][b]:tr["free_offset"][b + 1]]:
            if page[i] >= 0 and hdr[i] > 0:
                st.free_n(np.array([page[i]]), np.array([hdr[i]]))
        a0, a1 = int(tr["alloc_offset"][b]), int(tr["alloc_offset"][b + 1])
        if a1 > a0:
            p, o, h, _ = st.malloc(tr["sizes"][a0:a1], PS, a0, seed)
            page[a0:a1], off[a0:a1], hdr[a0:a1] = p, o, h
    return page, off, st


def test_bursts_serial_matches_oracle(gpu):
    T, fill = 1 << 14, 0.5
    log_case("bursts serial")
    tr = pbm.burst_trace(48, 150, 1.0, 16, 65536, PS, 3)
    page, off, st = _oracle(tr, T, fill, 2, 31)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=31) as pool:
        pool.prefill_bernoulli(fill, 2)
        with pbm.Bursts(pool, tr) as bs:
            for b in range(bs.n_bursts):
                bs.run(b, 1, pbm.launch(mode=pbm.MODE_SERIAL))
            p = bs.ptrs().astype(np.int64)
        base = pool.page_address(0)
        ok = page >= 0
        assert np.array_equal(p > 0, ok)
        assert np.array_equal(p[ok] - 4 - base, page[ok] * PS + off[ok])
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


def _modes():
    b = pbm.MODE_BATCHED
    return {"plain": (pbm.launch(mode=pbm.MODE_PLAIN), False),
            "batched": (pbm.launch(mode=b, waves_per_cu=4, walk_budget=16), False),
            "graph": (pbm.launch(mode=b, waves_per_cu=8, walk_budget=16), True),
            "cooperative": (pbm.launch(mode=b, waves_per_cu=8, walk_budget=16,
                                       flags=pbm.LAUNCH_COOPERATIVE), False)}


@pytest.mark.parametrize("mode", ["plain", "batched", "graph", "cooperative"])
def test_bursts_concurrent_invariants(gpu, mode):
    T, fill = 1 << 18, 0.5
    log_case(f"bursts concurrent {mode}")
    tr = pbm.burst_trace(160, 2048, 1.0, 16, 65536, PS, 5)
    units = _units(tr)
    launch_, use_graph = _modes()[mode]
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=9) as pool:
        pool.prefill_bernoulli(fill, 4)
        before = pool.snapshot()
        used0 = pool.count_used()
        with pbm.Bursts(pool, tr) as bs:
            if use_graph:
                for first in range(0, bs.n_bursts, 40):
                    g = bs.graph(first, min(40, bs.n_bursts - first), launch_)
                    g.launch()
                    gpu.cuda.synchronize()
                    g.close()
            else:
                bs.run(0, bs.n_bursts, launch_)
            gpu.cuda.synchronize()
            p = bs.ptrs().astype(np.int64)
            assert bs.barrier_timeouts() == 0
        base = pool.page_address(0)
        live = [(int(p[k]) - 4 - base, n) for k, (n, _) in units.items() if p[k] > 0]
        # every live run: page-aligned start, inside the pool, disjoint from the rest
        starts = np.array(sorted(s for s, _ in live), dtype=np.int64)
        assert (starts % PS == 0).all() and (starts >= 0).all()
        spans = sorted((s // PS, s // PS + n) for s, n in live)
        assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:]))
        assert spans[-1][1] <= T
        bm = pool.snapshot()
        used_pages = sum(n for _, n in live)
        assert pool.count_used() == used0 + used_pages
        for s, e in spans[:2000]:
            for q in range(s, e):
                assert (int(bm[q // 64]) >> (q % 64)) & 1 and not (int(before[q // 64]) >> (q % 64)) & 1
        # hand every live run back: the pool returns to its pre-fill
        leaders = np.array([p[k] for k in units if p[k] > 0], dtype=np.uint64)
        d = gpu.from_numpy(leaders.view(np.int64)).cuda()
        pool.free_ptrs(len(leaders), d)
        gpu.cuda.synchronize()
        assert np.array_equal(pool.snapshot(), before)
        assert pool.free_count() == orc.count_free(before, T)


def test_bursts_autotune_leaves_the_pool(gpu):
    T = 1 << 17
    log_case("bursts autotune")
    tr = pbm.burst_trace(24, 1024, 1.0, 16, 65536, PS, 8)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=1) as pool:
        pool.prefill_bernoulli(0.3, 1)
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            r = bs.autotune(24, pbm.TUNE_QUICK)
            assert r["candidates"] == 30 and r["ms"] > 0
            assert 1 <= r["waves_per_cu"] <= 16
            assert not bs.ptrs().any()
            assert np.array_equal(pool.snapshot(), before)
            bs.run(0, 24, r["launch"])
            gpu.cuda.synchronize()
            assert bs.barrier_timeouts() == 0
            assert bs.ptrs().any()


def test_bursts_autotune_phase_kernels_only(gpu):
    # PBM_TUNE_NO_COOPERATIVE keeps the tuner to the phase kernels, whose
    # launch a HIP graph can capture
    T = 1 << 17
    log_case("bursts autotune, no cooperative candidates")
    tr = pbm.burst_trace(24, 1024, 1.0, 16, 65536, PS, 9)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=1) as pool:
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            r = bs.autotune(24, pbm.TUNE_QUICK | pbm.TUNE_NO_COOPERATIVE)
            assert r["candidates"] == 15 and not r["cooperative"]
            assert np.array_equal(pool.snapshot(), before)
            with bs.graph(0, 24, r["launch"]) as g:
                g.launch()
                gpu.cuda.synchronize()
            assert bs.ptrs().any()
