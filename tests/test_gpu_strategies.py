"""Parity of CRW_BM, SingleClock, DistClock, LinkedList, RWMalloc
(getPage, getNPage, mallocRW/freeRW), RandomWalkBasic and the legacy
clustered and collaborative page walks with the serial oracle, through the
C-ABI.  Serial launches must match the oracle bit for bit; concurrent ones
are checked on the schedule-independent invariants."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu


def _get(torch, pool, n, mode, req_base=0):
    ids = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    steps = torch.zeros((max(n, 1),), dtype=torch.int32, device="cuda")
    pool.get_pages(n, ids, steps, req_base, pbm.launch(mode=mode))
    torch.cuda.synchronize()
    return ids.cpu().numpy()[:n], steps.cpu().numpy().view(np.uint32)[:n]


def _free(torch, pool, ids_np, mode):
    ids = torch.from_numpy(np.ascontiguousarray(ids_np, dtype=np.int32)).cuda()
    pool.free_pages(len(ids_np), ids, pbm.launch(mode=mode))
    torch.cuda.synchronize()


# ------------------------------------------------------------------ serial = oracle

def test_serial_crw_matches_oracle(gpu):
    T, n = 1 << 16, 6000
    log_case("serial crw")
    with pbm.Pool(T, strategy=pbm.CRW_BM, seed=13) as pool:
        pool.prefill_bernoulli(0.9, 2)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=7)
        st = orc.State(orc.prefill(T, 0.9, 2), T)
        eids, esteps, _ = st.crw(n, 7, 13)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        # the hint of slot 0 persists into the next call
        ids2, steps2 = _get(gpu, pool, 500, pbm.MODE_SERIAL, req_base=n + 7)
        eids2, esteps2, _ = st.crw(500, n + 7, 13, hint=int(eids[-1]))
        assert np.array_equal(ids2, eids2) and np.array_equal(steps2, esteps2)
        assert np.array_equal(pool.snapshot(), st.bm)


@pytest.mark.parametrize("T,fill", [(1 << 16, 0.9), (1_000_003, 0.99), (777, 0.3)])
def test_serial_rwbasic_matches_oracle(gpu, T, fill):
    n = 5000 if T > 1000 else 400
    log_case(f"serial RandomWalkBasic T={T} fill={fill}")
    with pbm.Pool(T, strategy=pbm.RW_BASIC, seed=19) as pool:
        pool.prefill_bernoulli(fill, 3)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=11)
        st = orc.State(orc.prefill(T, fill, 3), T)
        eids, esteps = st.rwbasic(n, 11, 19)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


@pytest.mark.parametrize("T,fill", [(1 << 16, 0.9), (777, 0.3), (1_000_003, 0.99)])
def test_serial_crw_basic_matches_oracle(gpu, T, fill):
    n = 5000 if T > 1000 else 400
    log_case(f"serial ClusteredRandomWalk T={T} fill={fill}")
    with pbm.Pool(T, strategy=pbm.CRW_BASIC, seed=37) as pool:
        pool.prefill_bernoulli(fill, 4)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=5)
        st = orc.State(orc.prefill(T, fill, 4), T)
        eids, esteps, hint = st.crw_basic(n, 5, 37)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        # the hint of slot 0 carries into the next call
        ids2, steps2 = _get(gpu, pool, 300, pbm.MODE_SERIAL, req_base=n + 5)
        eids2, esteps2, _ = st.crw_basic(300, n + 5, 37, hint=hint)
        assert np.array_equal(ids2, eids2) and np.array_equal(steps2, esteps2)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


@pytest.mark.parametrize("T,fill,n", [(1 << 16, 0.9, 6000), (4096, 0.5, 3000), (100, 0.0, 130),
                                      (1_000_003, 0.99, 20000)])
def test_serial_collab_basic_matches_oracle(gpu, T, fill, n):
    log_case(f"serial CollabRW T={T} fill={fill} n={n}")
    with pbm.Pool(T, strategy=pbm.COLLAB_RW_BASIC, seed=41) as pool:
        pool.prefill_bernoulli(fill, 6)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=9)
        st = orc.State(orc.prefill(T, fill, 6), T)
        eids, esteps = st.collab_basic(n, 9, 41)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


def test_serial_collab_basic_fallback_matches_oracle(gpu):
    # a budget of 2 rounds on a 99%-full pool leaves most lanes to the sweep
    T, n = 1 << 14, 640
    log_case("serial CollabRW sweep fallback")
    with pbm.Pool(T, strategy=pbm.COLLAB_RW_BASIC, seed=43) as pool:
        pool.prefill_bernoulli(0.99, 8)
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.get_pages(n, ids, steps, 0, pbm.launch(mode=pbm.MODE_SERIAL, walk_budget=2))
        gpu.cuda.synchronize()
        st = orc.State(orc.prefill(T, 0.99, 8), T)
        eids, esteps = st.collab_basic(n, 0, 43, budget=2)
        assert np.array_equal(ids.cpu().numpy(), eids)
        assert np.array_equal(steps.cpu().numpy().view(np.uint32), esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.counters()["fallbacks"] > 0


def test_serial_single_clock_matches_oracle(gpu):
    T, n = 1 << 16, 20000
    log_case("serial single clock")
    with pbm.Pool(T, strategy=pbm.SINGLE_CLOCK, seed=1) as pool:
        pool.prefill_bernoulli(0.5, 8)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL)
        st = orc.State(orc.prefill(T, 0.5, 8), T)
        eids, esteps, _ = st.single_clock(n)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)


def test_serial_dist_clock_matches_oracle(gpu):
    T, n = 1 << 16, 20000
    log_case("serial dist clock")
    with pbm.Pool(T, strategy=pbm.DIST_CLOCK, seed=21) as pool:
        pool.prefill_bernoulli(0.7, 8)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL)
        st = orc.State(orc.prefill(T, 0.7, 8), T)
        eids, esteps, _ = st.dist_clock(n, orc.dist_clock_arm(21, 0, T))
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)


def test_serial_list_is_fifo(gpu):
    T = 1 << 14
    log_case("serial linked list fifo")
    with pbm.Pool(T, strategy=pbm.LINKED_LIST) as pool:
        pool.prefill_bernoulli(0.5, 4)
        ls = orc.ListState(orc.prefill(T, 0.5, 4), T)
        ids, _ = _get(gpu, pool, 3000, pbm.MODE_SERIAL)
        assert np.array_equal(ids, ls.get(3000))
        rng = np.random.default_rng(0)
        back = rng.permutation(ids[ids >= 0])[:1500]
        _free(gpu, pool, back, pbm.MODE_SERIAL)
        ls.put(back)
        n = int(ls.ht[1] - ls.ht[0]) + 10      # drain it and ask 10 more
        ids, _ = _get(gpu, pool, n, pbm.MODE_SERIAL)
        assert np.array_equal(ids, ls.get(n))
        assert np.array_equal(pool.snapshot(), ls.bm)
        assert pool.free_count() == 0


@pytest.mark.parametrize("order", [pbm.ORDER_PACKED, pbm.ORDER_REFERENCE])
def test_serial_rwmalloc_getpage_matches_oracle(gpu, order):
    T, n = 1 << 16, 6000
    log_case(f"serial rwmalloc getPage order={order}")
    with pbm.Pool(T, strategy=pbm.RW_MALLOC, seed=17, order=order) as pool:
        pool.prefill_bernoulli(0.9, 5)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=3)
        st = orc.State(orc.prefill(T, 0.9, 5), T)
        eids, esteps = st.rwmalloc(n, 3, 17, order)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)


def _counts(rng, n, hi):
    c = rng.integers(1, hi + 1, size=n).astype(np.uint32)
    c[rng.random(n) < 0.1] = 0
    return c


@pytest.mark.parametrize("fill,hi,spans", [(0.5, 8, 1), (0.3, 40, 1), (0.8, 3, 1),
                                           (0.5, 8, 4), (0.8, 3, 2), (0.3, 40, 4)])
def test_serial_getnpage_matches_oracle(gpu, fill, hi, spans):
    T, n = 1 << 15, 3000
    log_case(f"serial getNPage fill={fill} hi={hi} spans={spans}")
    counts = _counts(np.random.default_rng(1), n, hi)
    with pbm.Pool(T, strategy=pbm.RW_MALLOC, seed=23) as pool:
        pool.prefill_bernoulli(fill, 6)
        d_counts = gpu.from_numpy(counts.view(np.int32)).cuda()
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.get_n_pages(n, d_counts, ids, steps, 11,
                         pbm.launch(mode=pbm.MODE_SERIAL, cluster_words=64 * spans))
        gpu.cuda.synchronize()
        st = orc.State(orc.prefill(T, fill, 6), T)
        eids, esteps = st.get_n(counts, 11, 23, spans=spans)
        assert np.array_equal(ids.cpu().numpy(), eids)
        assert np.array_equal(steps.cpu().numpy().view(np.uint32), esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]
        pool.free_n_pages(n, ids, d_counts)
        gpu.cuda.synchronize()
        st.free_n(eids, counts)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


@pytest.mark.parametrize("fill,hi", [(0.3, 16), (0.7, 4), (0.0, 64)])
def test_serial_getxpage_matches_oracle(gpu, fill, hi):
    T, n = 1 << 15, 2000
    log_case(f"serial getXPage fill={fill} hi={hi}")
    counts = _counts(np.random.default_rng(5), n, hi)
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=29) as pool:
        pool.prefill_bernoulli(fill, 7)
        d_counts = gpu.from_numpy(counts.view(np.int32)).cuda()
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.get_x_pages(n, d_counts, ids, steps, 13, pbm.launch(mode=pbm.MODE_SERIAL))
        gpu.cuda.synchronize()
        st = orc.State(orc.prefill(T, fill, 7), T)
        eids, esteps = st.get_x(counts, 13, 29)
        assert np.array_equal(ids.cpu().numpy(), eids)
        assert np.array_equal(steps.cpu().numpy().view(np.uint32), esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]
        pool.free_n_pages(n, ids, d_counts)
        gpu.cuda.synchronize()
        assert np.array_equal(pool.snapshot(), orc.prefill(T, fill, 7))


def test_concurrent_getxpage_runs_are_disjoint(gpu):
    T, n = 1 << 18, 1 << 16
    log_case("concurrent getXPage")
    counts = _counts(np.random.default_rng(6), n, 12)
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=30) as pool:
        pool.prefill_bernoulli(0.5, 8)
        before = pool.snapshot()
        d_counts = gpu.from_numpy(counts.view(np.int32)).cuda()
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        pool.get_x_pages(n, d_counts, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN))
        gpu.cuda.synchronize()
        got = ids.cpu().numpy()
        used = np.unpackbits(before.view(np.uint8), bitorder="little").astype(bool)[:T]
        seen = np.zeros(T, dtype=bool)
        for i, x in zip(got, counts):
            if x == 0 or i < 0:
                continue
            assert i // 64 == (i + int(x) - 1) // 64
            assert not seen[i:i + x].any() and not used[i:i + x].any()
            seen[i:i + x] = True
        after = np.unpackbits(pool.snapshot().view(np.uint8), bitorder="little").astype(bool)[:T]
        assert np.array_equal(after, used | seen)
        assert pool.free_count() == orc.count_free(before, T) - int(seen.sum())
        pool.free_n_pages(n, ids, d_counts)
        gpu.cuda.synchronize()
        assert np.array_equal(pool.snapshot(), before)


def _sizes(rng, n):
    # log-uniform 16 B .. 64 KiB
    return np.exp(rng.uniform(np.log(16), np.log(65536), size=n)).astype(np.uint64)


@pytest.mark.parametrize("spans", [1, 4])
def test_serial_malloc_matches_oracle(gpu, spans):
    T, ps, n = 1 << 15, 4096, 3000
    log_case(f"serial mallocRW/freeRW spans={spans}")
    sizes = _sizes(np.random.default_rng(2), n)
    with pbm.Pool(T, ps, strategy=pbm.RW_MALLOC, backing=True, seed=29) as pool:
        pool.prefill_bernoulli(0.4, 9)
        d_sizes = gpu.from_numpy(sizes.view(np.int64)).cuda()
        ptrs = gpu.zeros((n,), dtype=gpu.int64, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.malloc(n, d_sizes, ptrs, steps, 0,
                    pbm.launch(mode=pbm.MODE_SERIAL, cluster_words=64 * spans))
        gpu.cuda.synchronize()
        st = orc.State(orc.prefill(T, 0.4, 9), T)
        page, off, hdr, esteps = st.malloc(sizes, ps, 0, 29, spans=spans)
        p = ptrs.cpu().numpy().view(np.uint64).astype(np.int64)
        base = pool.page_address(0)
        # a packed group may span several pages: compare byte positions
        assert np.array_equal(p > 0, page >= 0)
        ok = page >= 0
        exp_pos = page[ok].astype(np.int64) * ps + off[ok].astype(np.int64)
        assert np.array_equal(p[ok] - 4 - base, exp_pos)
        assert np.array_equal(steps.cpu().numpy().view(np.uint32), esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        pool.free_ptrs(n, ptrs)
        gpu.cuda.synchronize()
        assert np.array_equal(pool.snapshot(), orc.prefill(T, 0.4, 9))


# ------------------------------------------------------------------ concurrent

@pytest.mark.parametrize("strategy", [pbm.CRW_BM, pbm.SINGLE_CLOCK, pbm.DIST_CLOCK,
                                      pbm.LINKED_LIST, pbm.RW_MALLOC, pbm.RW_BASIC,
                                      pbm.CRW_BASIC, pbm.COLLAB_RW_BASIC])
@pytest.mark.parametrize("fill", [0.5, 0.9])
def test_concurrent_cycle_invariants(gpu, strategy, fill):
    T, n = 1 << 18, 65536
    log_case(f"concurrent {pbm.STRATEGY_NAMES[strategy]} fill={fill}")
    with pbm.Pool(T, strategy=strategy, seed=31) as pool:
        pool.prefill_bernoulli(fill, 42)
        before = orc.prefill(T, fill, 42)
        A = orc.count_free(before, T)
        ids, _ = _get(gpu, pool, n, pbm.MODE_PLAIN)
        rc, g = orc.check_alloc(before, pool.snapshot(), T, ids)
        assert rc == 0 and g == min(n, A)
        assert pool.free_count() == A - g
        _free(gpu, pool, ids, pbm.MODE_PLAIN)
        assert np.array_equal(pool.snapshot(), before)
        assert pool.free_count() == A
        c = pool.counters()
        assert c["errors"] == 0


@pytest.mark.parametrize("strategy", [pbm.SINGLE_CLOCK, pbm.LINKED_LIST])
def test_fresh_pool_concurrent_gets_return_the_first_k_pages(gpu, strategy):
    # unique arm tickets / FIFO head tickets: k concurrent gets on an empty
    # pool return exactly pages 0..k-1
    T, k = 1 << 16, 40000
    with pbm.Pool(T, strategy=strategy) as pool:
        ids, _ = _get(gpu, pool, k, pbm.MODE_PLAIN)
        assert np.array_equal(np.sort(ids), np.arange(k))


def test_list_refills_after_running_empty(gpu):
    T = 4096
    with pbm.Pool(T, strategy=pbm.LINKED_LIST) as pool:
        ids, _ = _get(gpu, pool, T + 100, pbm.MODE_PLAIN)
        assert (ids >= 0).sum() == T
        # which requests got the T pages depends on the wave schedule
        back = ids[ids >= 0][:50]
        _free(gpu, pool, back, pbm.MODE_PLAIN)
        again, _ = _get(gpu, pool, 60, pbm.MODE_PLAIN)
        assert np.array_equal(np.sort(again[again >= 0]), np.sort(back))


def test_multipage_calls_reject_other_window_widths(gpu):
    with pbm.Pool(4096, strategy=pbm.RW_MALLOC) as pool:
        c = gpu.ones((64,), dtype=gpu.int32, device="cuda")
        ids = gpu.full((64,), -1, dtype=gpu.int32, device="cuda")
        with pytest.raises(pbm.PbmError, match="cluster_words"):
            pool.get_n_pages(64, c, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN,
                                                             cluster_words=192))
        assert (ids.cpu().numpy() == -1).all()


@pytest.mark.parametrize("cw", [64, 256])
def test_concurrent_getnpage_runs_are_disjoint(gpu, cw):
    T, n = 1 << 18, 30000
    log_case(f"concurrent getNPage cluster_words={cw}")
    counts = _counts(np.random.default_rng(7), n, 17)
    with pbm.Pool(T, strategy=pbm.RW_MALLOC, seed=3) as pool:
        pool.prefill_bernoulli(0.5, 1)
        before = orc.prefill(T, 0.5, 1)
        d_counts = gpu.from_numpy(counts.view(np.int32)).cuda()
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        pool.get_n_pages(n, d_counts, ids, None, 0,
                         pbm.launch(mode=pbm.MODE_PLAIN, cluster_words=cw))
        gpu.cuda.synchronize()
        got = ids.cpu().numpy()
        after = pool.snapshot()
        used = np.zeros(T, dtype=np.int64)
        for i, c in zip(got, counts):
            if i >= 0 and c > 0:
                assert i % 64 + c <= 64
                used[i:i + c] += 1
        assert used.max() <= 1
        bits = np.unpackbits(before.view(np.uint8), bitorder="little")[:T].astype(bool)
        assert not (used.astype(bool) & bits).any()
        newbits = np.unpackbits(after.view(np.uint8), bitorder="little")[:T].astype(bool)
        assert np.array_equal(newbits, bits | used.astype(bool))
        pool.free_n_pages(n, ids, d_counts)
        gpu.cuda.synchronize()
        assert np.array_equal(pool.snapshot(), before)
        assert pool.free_count() == orc.count_free(before, T)


def test_concurrent_malloc_regions_are_disjoint(gpu):
    T, ps, n = 1 << 16, 4096, 20000
    log_case("concurrent mallocRW")
    sizes = _sizes(np.random.default_rng(4), n)
    with pbm.Pool(T, ps, strategy=pbm.RW_MALLOC, backing=True, seed=5) as pool:
        d_sizes = gpu.from_numpy(sizes.view(np.int64)).cuda()
        ptrs = gpu.zeros((n,), dtype=gpu.int64, device="cuda")
        pool.malloc(n, d_sizes, ptrs)
        gpu.cuda.synchronize()
        p = ptrs.cpu().numpy().view(np.uint64).astype(np.int64)
        base = pool.page_address(0)
        ok = p > 0
        lo = p[ok] - 4 - base
        hi = p[ok] + sizes[ok].astype(np.int64) - base
        order = np.argsort(lo)
        assert (lo >= 0).all() and (hi <= T * ps).all()
        assert (lo[order][1:] >= hi[order][:-1]).all()   # no two regions overlap
        pool.free_ptrs(n, ptrs)
        gpu.cuda.synchronize()
        assert pool.count_used() == 0 and pool.free_count() == T


@pytest.mark.parametrize("order", [pbm.ORDER_PACKED, pbm.ORDER_REFERENCE])
def test_serial_collect_warp_data_matches_oracle(gpu, order):
    # collectWarpData (CollabRW_BM.cuh:131-138): per wave, the free bits its
    # lanes found in each step
    T, n, K = 1 << 16, 64 * 120 + 17, 24
    log_case(f"serial collectWarpData order={order}")
    with pbm.Pool(T, strategy=pbm.COLLAB_RW_BM, seed=47, order=order) as pool:
        pool.prefill_bernoulli(0.95, 9)
        waves = -(-n // 64)
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        yields = gpu.zeros((waves * K,), dtype=gpu.int32, device="cuda")
        n_steps = gpu.zeros((waves,), dtype=gpu.int32, device="cuda")
        pool.collect_warp_data(n, ids, yields, K, n_steps, 3, pbm.launch(mode=pbm.MODE_SERIAL))
        gpu.cuda.synchronize()
        st = orc.State(orc.prefill(T, 0.95, 9), T)
        eids, ey, ens = st.collab_yields(n, 3, 47, order, K)
        assert np.array_equal(ids.cpu().numpy(), eids)
        assert np.array_equal(n_steps.cpu().numpy().view(np.uint32), ens)
        got = yields.cpu().numpy().view(np.uint32).reshape(waves, K)
        for w in range(waves):
            k = min(int(ens[w]), K)
            assert np.array_equal(got[w, :k], ey[w, :k]), w
        assert np.array_equal(pool.snapshot(), st.bm)


def test_collect_warp_data_needs_collab(gpu):
    with pbm.Pool(4096, strategy=pbm.RW_BM) as pool:
        ids = gpu.zeros((64,), dtype=gpu.int32, device="cuda")
        with pytest.raises(pbm.PbmError):
            pool.collect_warp_data(64, ids, ids, 1, ids)
