"""mallocRW/freeRW bursts through the C-ABI.

A serial launch reproduces the serial oracle (pointers, map, reservation
counter), on a small trace and on the whole 4,096-burst trace of the
benchmark (lognormal counts, median 2,048, sigma 1.0, sizes 16 B .. 64 KiB,
262,144 empty pages), for each getNPage window width.  Concurrent launches
(plain, batched, graph-captured phase kernels, the persistent kernel and
its graph) leave exactly the trace's live runs: disjoint, in use, the map
equal to the pre-fill plus those runs, and every page handed back when
they are freed."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu

PS = 4096
C4_PAGES = 1 << 18


@pytest.fixture(scope="module")
def c4_trace():
    return pbm.burst_trace(4096, 2048, 1.0, 16, 65536, PS, 7)


def _oracle(tr, T, fill, pseed, seed, spans=1, n_bursts=None):
    """The trace run serially by the oracle: pages, offsets, final state."""
    st = orc.State(orc.prefill(T, fill, pseed), T)
    n = int(tr["alloc_offset"][-1])
    page = np.full(n, -1, dtype=np.int64)
    off = np.zeros(n, dtype=np.int64)
    hdr = np.zeros(n, dtype=np.int64)
    nb = len(tr["alloc_offset"]) - 1 if n_bursts is None else n_bursts
    for b in range(nb):
        fi = tr["free_index"][tr["free_offset"][b]:tr["free_offset"][b + 1]]
        live = (page[fi] >= 0) & (hdr[fi] > 0)
        if live.any():
            st.free_n(page[fi][live], hdr[fi][live])
        a0, a1 = int(tr["alloc_offset"][b]), int(tr["alloc_offset"][b + 1])
        if a1 > a0:
            p, o, h, _ = st.malloc(tr["sizes"][a0:a1], PS, a0, seed, spans=spans)
            page[a0:a1], off[a0:a1], hdr[a0:a1] = p, o, h
    return page, off, st


def _serial_matches(gpu, tr, T, fill, pseed, seed, spans, n_bursts=None):
    page, off, st = _oracle(tr, T, fill, pseed, seed, spans, n_bursts)
    nb = len(tr["alloc_offset"]) - 1 if n_bursts is None else n_bursts
    n_req = int(tr["alloc_offset"][nb])
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=seed) as pool:
        if fill > 0:
            pool.prefill_bernoulli(fill, pseed)
        with pbm.Bursts(pool, tr) as bs:
            bs.run(0, nb, pbm.launch(mode=pbm.MODE_SERIAL, cluster_words=64 * spans))
            gpu.cuda.synchronize()
            p = bs.ptrs(0, n_req).astype(np.int64)
        base = pool.page_address(0)
        ok = page[:n_req] >= 0
        assert np.array_equal(p > 0, ok)
        assert np.array_equal(p[ok] - 4 - base, page[:n_req][ok] * PS + off[:n_req][ok])
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]
        return int((~ok).sum())


@pytest.mark.parametrize("spans", [1, 2, 4])
def test_bursts_serial_matches_oracle(gpu, spans):
    log_case(f"bursts serial spans={spans}")
    tr = pbm.burst_trace(48, 150, 1.0, 16, 65536, PS, 3)
    _serial_matches(gpu, tr, 1 << 14, 0.5, 2, 31, spans)


@pytest.mark.parametrize("spans", [1, 4])
def test_c4_trace_serial_matches_oracle(gpu, c4_trace, spans):
    # the benchmark's whole trace, pointer for pointer
    log_case(f"bursts C4 trace serial, 4096 bursts, spans={spans}")
    fails = _serial_matches(gpu, c4_trace, C4_PAGES, 0.0, 0, 7, spans)
    assert fails == 0


def _run(gpu, bs, mode, n_bursts, launch_, stride=256):
    if mode.startswith("graph"):
        for first in range(0, n_bursts, stride):
            g = bs.graph(first, min(stride, n_bursts - first), launch_)
            g.launch()
            gpu.cuda.synchronize()
            g.close()
    else:
        bs.run(0, n_bursts, launch_)
    gpu.cuda.synchronize()


def _modes():
    b = pbm.MODE_BATCHED
    per = pbm.LAUNCH_PERSISTENT
    return {"plain": pbm.launch(mode=pbm.MODE_PLAIN),
            "batched": pbm.launch(mode=b, waves_per_cu=4, walk_budget=16),
            "graph_phases": pbm.launch(mode=b, waves_per_cu=8, walk_budget=16),
            "persistent": pbm.launch(mode=b, waves_per_cu=8, walk_budget=16, flags=per),
            "persistent_ordered": pbm.launch(mode=b, waves_per_cu=8, walk_budget=16,
                                             flags=per | pbm.LAUNCH_ORDERED),
            "persistent_wide": pbm.launch(mode=b, waves_per_cu=2, walk_budget=4,
                                          cluster_words=256, flags=per),
            "graph_persistent": pbm.launch(mode=b, waves_per_cu=16, walk_budget=16, flags=per)}


def _end_state(gpu, pool, bs, tr, before):
    st = pbm.check_burst_end_state(tr, bs.ptrs(), pool.page_address(0), PS, before,
                                   pool.snapshot())
    assert bs.wait_timeouts() == 0
    # hand every live run back: the pool returns to its pre-fill
    d = gpu.from_numpy(st["leaders"].view(np.int64)).cuda()
    pool.free_ptrs(len(st["leaders"]), d)
    gpu.cuda.synchronize()
    assert np.array_equal(pool.snapshot(), before)
    return st


@pytest.mark.parametrize("mode", list(_modes()))
def test_bursts_concurrent_end_state(gpu, mode):
    T, fill = 1 << 18, 0.5
    log_case(f"bursts concurrent {mode}")
    tr = pbm.burst_trace(160, 2048, 1.0, 16, 65536, PS, 5)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=9) as pool:
        pool.prefill_bernoulli(fill, 4)
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            _run(gpu, bs, mode, bs.n_bursts, _modes()[mode], stride=40)
            st = _end_state(gpu, pool, bs, tr, before)
        assert pool.free_count() == orc.count_free(before, T)
        assert st["live_runs"] > 0


@pytest.mark.parametrize("mode", ["graph_phases", "persistent", "persistent_ordered",
                                  "graph_persistent"])
def test_c4_trace_concurrent_end_state(gpu, c4_trace, mode):
    # the benchmark's configuration: 4,096 bursts on 262,144 empty pages
    log_case(f"bursts C4 trace concurrent {mode}")
    with pbm.Pool(C4_PAGES, PS, strategy=pbm.RW_MALLOC, backing=True, seed=7) as pool:
        before = pool.snapshot()
        with pbm.Bursts(pool, c4_trace) as bs:
            _run(gpu, bs, mode, bs.n_bursts, _modes()[mode])
            # the serial oracle fails no request on this trace, and the pool
            # never holds more than a fraction of its pages
            assert int((bs.ptrs() == 0).sum()) == 0
            st = _end_state(gpu, pool, bs, c4_trace, before)
        assert st["failed_runs"] == 0 and st["live_pages"] == 38875
        assert pool.free_count() == C4_PAGES


@pytest.mark.parametrize("ordered", [False, True])
def test_persistent_on_a_fragmented_pool(gpu, ordered):
    # 30% of the pages free, scattered: alloc items run ahead of their
    # burst's frees (more than a quarter of the pool is unreserved), most
    # multi-page runs find no room, wait for the burst's frees and try again;
    # requests that still find none fail.  Live runs stay disjoint and in
    # use, and every page comes back.
    T = 1 << 16
    log_case(f"bursts persistent, fragmented pool, ordered={ordered}")
    tr = pbm.burst_trace(64, 512, 1.0, 16, 65536, PS, 21)
    flags = pbm.LAUNCH_PERSISTENT | (pbm.LAUNCH_ORDERED if ordered else 0)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=5) as pool:
        pool.prefill_bernoulli(0.7, 6)
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            bs.run(0, bs.n_bursts, pbm.launch(mode=pbm.MODE_BATCHED, waves_per_cu=8,
                                              walk_budget=16, flags=flags))
            gpu.cuda.synchronize()
            fails = int((bs.ptrs() == 0).sum())
            st = _end_state(gpu, pool, bs, tr, before)
        assert fails > 0 and st["failed_runs"] > 0
        assert pool.free_count() == orc.count_free(before, T)


def test_persistent_runs_in_pieces_match_one_run(gpu):
    # bursts 0..99 then 100..159 in two persistent launches: the second
    # launch's frees of the first launch's pointers find them published
    T = 1 << 17
    log_case("bursts persistent in two pieces")
    tr = pbm.burst_trace(160, 1024, 1.0, 16, 65536, PS, 12)
    lp = _modes()["persistent"]
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=3) as pool:
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            bs.run(0, 100, lp)
            bs.run(100, 60, lp)
            gpu.cuda.synchronize()
            _end_state(gpu, pool, bs, tr, before)


def test_bursts_create_rejects_traces_that_free_twice_or_split_a_packed_run(gpu):
    with pbm.Pool(1 << 12, PS, strategy=pbm.RW_MALLOC, backing=True) as pool:
        sizes = np.array([100, 200, 300, 50000], dtype=np.uint64)
        good = {"alloc_offset": np.array([0, 4, 4, 4]), "free_offset": np.array([0, 0, 3, 4]),
                "sizes": sizes, "free_index": np.array([0, 1, 2, 3])}
        pbm.Bursts(pool, good).close()
        twice = dict(good, free_index=np.array([0, 1, 3, 3]))
        with pytest.raises(pbm.PbmError, match="freed twice"):
            pbm.Bursts(pool, twice)
        # requests 0..2 are one packed run: freeing 0 and 1 in burst 1 and
        # 2 in burst 2 splits it
        split = dict(good, free_index=np.array([0, 1, 3, 2]))
        with pytest.raises(pbm.PbmError, match="one burst"):
            pbm.Bursts(pool, split)


def test_bursts_autotune_leaves_the_pool(gpu):
    T = 1 << 17
    log_case("bursts autotune")
    tr = pbm.burst_trace(24, 1024, 1.0, 16, 65536, PS, 8)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=1) as pool:
        pool.prefill_bernoulli(0.3, 1)
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            r = bs.autotune(24, pbm.TUNE_QUICK)
            assert r["candidates"] == 90 and r["ms"] > 0
            assert 1 <= r["waves_per_cu"] <= 16 and r["cluster_words"] in (64, 128, 256)
            assert not bs.ptrs().any()
            assert np.array_equal(pool.snapshot(), before)
            bs.run(0, 24, r["launch"])
            gpu.cuda.synchronize()
            assert bs.wait_timeouts() == 0
            assert bs.ptrs().any()


def test_bursts_autotune_phase_kernels_only(gpu):
    T = 1 << 17
    log_case("bursts autotune, phase kernels only")
    tr = pbm.burst_trace(24, 1024, 1.0, 16, 65536, PS, 9)
    with pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=1) as pool:
        before = pool.snapshot()
        with pbm.Bursts(pool, tr) as bs:
            r = bs.autotune(24, pbm.TUNE_QUICK | pbm.TUNE_NO_PERSISTENT)
            assert r["candidates"] == 45 and not r["persistent"]
            assert np.array_equal(pool.snapshot(), before)
            with bs.graph(0, 24, r["launch"]) as g:
                g.launch()
                gpu.cuda.synchronize()
            assert bs.ptrs().any()
