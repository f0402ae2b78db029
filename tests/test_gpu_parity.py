"""Parity of the HIP allocators with the serial oracle, through the C-ABI.

Serial schedules (PBM_MODE_SERIAL, PBM_LAUNCH_ONE_WAVE) must reproduce the
oracle's page ids and step counts bit for bit.  Concurrent schedules are
checked on what does not depend on the schedule: every granted page unique
and free before, the map after the alloc phase equal to the map before plus
exactly the granted bits, grants = min(requests, free pages), and the map
after freeing everything equal to the map before.
"""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu

BIG = 1 << 20


def _bufs(torch, n):
    ids = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    steps = torch.zeros((max(n, 1),), dtype=torch.int32, device="cuda")
    return ids, steps


def _get(torch, pool, n, mode, req_base=0, cw=0, flags=0, budget=0):
    ids, steps = _bufs(torch, n)
    pool.get_pages(n, ids, steps, req_base,
                   pbm.launch(mode=mode, cluster_words=cw, flags=flags, walk_budget=budget))
    torch.cuda.synchronize()
    return ids.cpu().numpy()[:n], steps.cpu().numpy().view(np.uint32)[:n]


def _free(torch, pool, ids_np, mode):
    ids = torch.from_numpy(np.ascontiguousarray(ids_np, dtype=np.int32)).cuda()
    pool.free_pages(len(ids_np), ids, pbm.launch(mode=mode))
    torch.cuda.synchronize()


# ------------------------------------------------------------------ prefill

@pytest.mark.parametrize("T,fill,seed", [(65536, 0.0, 42), (1000, 0.5, 3), (BIG, 0.9, 42),
                                         (BIG, 0.99, 7), (4096, 1.0, 1)])
def test_prefill_matches_oracle(gpu, T, fill, seed):
    log_case(f"prefill T={T} fill={fill} seed={seed}")
    with pbm.Pool(T) as pool:
        pool.prefill_bernoulli(fill, seed)
        got = pool.snapshot()
        exp = orc.prefill(T, fill, seed)
        assert np.array_equal(got, exp)
        assert pool.free_count() == orc.count_free(exp, T)
        assert pool.count_used() == T - orc.count_free(exp, T)


# ------------------------------------------------------------------ serial = oracle

@pytest.mark.parametrize("T,fill,n", [(65536, 0.0, 32768), (1 << 16, 0.9, 8000),
                                      (1000, 0.3, 900)])
def test_serial_rwbm_matches_oracle(gpu, T, fill, n):
    log_case(f"serial rwbm T={T} fill={fill} n={n}")
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=42) as pool:
        pool.prefill_bernoulli(fill, 42)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=5)
        st = orc.State(orc.prefill(T, fill, 42), T)
        eids, esteps = st.rwbm(n, 5, 42)
        assert np.array_equal(ids, eids)
        assert np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


@pytest.mark.parametrize("order", [pbm.ORDER_PACKED, pbm.ORDER_REFERENCE])
@pytest.mark.parametrize("T,fill,n", [(1 << 16, 0.9, 8000), (1 << 14, 0.99, 500),
                                      (64 * 40, 0.5, 1500)])
def test_serial_collab_matches_oracle(gpu, order, T, fill, n):
    log_case(f"serial collab order={order} T={T} fill={fill} n={n}")
    with pbm.Pool(T, strategy=pbm.COLLAB_RW_BM, seed=11, order=order) as pool:
        pool.prefill_bernoulli(fill, 3)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=64)
        st = orc.State(orc.prefill(T, fill, 3), T)
        eids, esteps = st.collab(n, 64, 11, order)
        assert np.array_equal(ids, eids)
        assert np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)


@pytest.mark.parametrize("cw", [64, 128, 256])
@pytest.mark.parametrize("order", [pbm.ORDER_PACKED, pbm.ORDER_REFERENCE])
def test_one_wave_batched_matches_oracle(gpu, cw, order):
    T, fill, n = 1 << 16, 0.9, 8000
    log_case(f"one-wave batched cw={cw} order={order}")
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=9, order=order) as pool:
        pool.prefill_bernoulli(fill, 42)
        ids, steps = _get(gpu, pool, n, pbm.MODE_BATCHED, req_base=3, cw=cw,
                          flags=pbm.LAUNCH_ONE_WAVE)
        st = orc.State(orc.prefill(T, fill, 42), T)
        eids, esteps = st.batched(n, 3, 9, order, cw)
        assert np.array_equal(ids, eids)
        assert np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]


@pytest.mark.parametrize("T,fill,n,budget", [(1 << 14, 0.99, 600, 1), (1 << 18, 0.999, 400, 2),
                                             (1 << 20, 0.9, 120000, 4)])
@pytest.mark.parametrize("order", [pbm.ORDER_PACKED, pbm.ORDER_REFERENCE])
def test_one_wave_batched_budget_sweep_matches_oracle(gpu, T, fill, n, budget, order):
    # small walk budget: after it the wave sweeps 1024 words per step
    log_case(f"one-wave batched sweep T={T} fill={fill} budget={budget} order={order}")
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=2, order=order) as pool:
        pool.prefill_bernoulli(fill, 5)
        ids, steps = _get(gpu, pool, n, pbm.MODE_BATCHED, cw=64, flags=pbm.LAUNCH_ONE_WAVE,
                          budget=budget)
        st = orc.State(orc.prefill(T, fill, 5), T)
        eids, esteps = st.batched(n, 0, 2, order, 64, budget=budget)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)


@pytest.mark.parametrize("T,fill,n", [(BIG, 0.9, 262144), (BIG, 0.5, 262144), (BIG, 0.99, 262144),
                                      (65536, 0.0, 32768), (1000, 0.4, 700), (1 << 16, 0.999, 100),
                                      (1 << 22, 0.9, 300000), (1 << 24, 0.97, 200000),
                                      (64 * 256 * 65 + 17, 0.6, 1 << 20),
                                      # more than 256 tiles: two-level look-back,
                                      # a last group of 45 tiles
                                      (301 * 256 * 64 - 4000, 0.7, 1 << 20),
                                      # 2^20 map words and more: four words per thread
                                      (1 << 26, 0.9, 300000), (1 << 26, 0.999, 100000)])
@pytest.mark.parametrize("one_wave,excl", [(False, False), (True, False), (False, True),
                                           (True, True)])
def test_batched_rank_matches_oracle_exactly(gpu, T, fill, n, one_wave, excl):
    # by rank, request r takes the r-th free page whatever the schedule, so
    # even the full concurrent launch matches the oracle bit for bit, with
    # atomic claims or with the plain stores of an exclusive call
    log_case(f"batched rank T={T} fill={fill} n={n} one_wave={one_wave} exclusive={excl}")
    flags = (pbm.LAUNCH_RANK | (pbm.LAUNCH_ONE_WAVE if one_wave else 0)
             | (pbm.LAUNCH_EXCLUSIVE if excl else 0))
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=3) as pool:
        pool.prefill_bernoulli(fill, 42)
        ids, steps = _get(gpu, pool, n, pbm.MODE_BATCHED, flags=flags)
        st = orc.State(orc.prefill(T, fill, 42), T)
        eids, esteps = st.batched_rank(n)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), st.bm)
        assert pool.free_count() == st.free[0]
        _free(gpu, pool, ids, pbm.MODE_BATCHED)
        assert np.array_equal(pool.snapshot(), orc.prefill(T, fill, 42))


@pytest.mark.parametrize("T", [1 << 21, 1 << 26])
def test_batched_rank_calls_in_a_row(gpu, T):
    # each call flips every tile's look-back parity; calls of different
    # sizes and kinds (atomic or exclusive claims, one workgroup or all) on
    # the same pool must each match the oracle continuing from the last state
    log_case(f"batched rank calls in a row T={T}")
    kinds = [pbm.LAUNCH_RANK, pbm.LAUNCH_RANK | pbm.LAUNCH_EXCLUSIVE,
             pbm.LAUNCH_RANK | pbm.LAUNCH_ONE_WAVE]
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=4) as pool:
        pool.prefill_bernoulli(0.8, 9)
        st = orc.State(orc.prefill(T, 0.8, 9), T)
        for k, n in enumerate([1, 64, 5000, 100000, 777, 300000, 3, 150000]):
            ids, steps = _get(gpu, pool, n, pbm.MODE_BATCHED, flags=kinds[k % 3])
            eids, esteps = st.batched_rank(n)
            assert np.array_equal(ids, eids), f"call {k} n={n}"
            assert np.array_equal(steps, esteps), f"call {k} n={n}"
            assert pool.free_count() == st.free[0]
            if k % 3 == 2:
                _free(gpu, pool, ids, pbm.MODE_BATCHED)
                st.free_ids(ids)
        assert np.array_equal(pool.snapshot(), st.bm)


# ------------------------------------------------------------------ concurrent invariants


def _with_bits(words, ids):
    out = words.copy()
    np.bitwise_or.at(out, ids >> 6, np.left_shift(np.uint64(1), (ids & 63).astype(np.uint64)))
    return out


@pytest.mark.parametrize("fill,n,drained", [(0.5, 1 << 17, False), (0.9, 1 << 18, True)])
def test_rank_beside_plain_getpage_on_another_stream(gpu, fill, n, drained):
    # the rank allocation while plain RandomWalkBasic_BM getPage kernels
    # claim pages of the same map from another stream: every page granted
    # once, the map is the pre-fill plus every grant, and the reservation
    # counter still counts the map's free pages.  Below capacity every
    # request is granted and no reservation is lost; a drained pool leaves
    # no page free that a request went without
    torch = gpu
    T = 1 << 20
    log_case(f"rank beside plain getPage fill={fill} n={n}")
    with pbm.Pool(T, 4096, strategy=pbm.RW_BM, seed=11) as pool:
        pool.prefill_bernoulli(fill, 5)
        before = pool.snapshot()
        A = pool.free_count()
        sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
        rank = pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK)
        plain = pbm.launch(pbm.MODE_PLAIN)
        grants = []
        fallbacks = 0
        for rep in range(3):
            ids_a = torch.full((n,), -7, dtype=torch.int32, device="cuda")
            ids_b = torch.full((n,), -7, dtype=torch.int32, device="cuda")
            pool.clear_counters()
            torch.cuda.synchronize()
            if rep % 2:
                pool.get_pages(n, ids_a, None, 0, rank, sa)
                pool.get_pages(n, ids_b, None, 1 << 28, plain, sb)
            else:
                pool.get_pages(n, ids_b, None, 1 << 28, plain, sb)
                pool.get_pages(n, ids_a, None, 0, rank, sa)
            torch.cuda.synchronize()
            a, b = ids_a.cpu().numpy(), ids_b.cpu().numpy()
            assert (a >= -1).all() and (b >= -1).all()  # every request answered
            g = np.concatenate([a[a >= 0], b[b >= 0]]).astype(np.int64)
            cnt = pool.counters()
            fallbacks += cnt["fallbacks"]
            assert len(np.unique(g)) == len(g) and (g < T).all()
            snap = pool.snapshot()
            assert np.array_equal(snap, _with_bits(before, g))
            used = int(np.unpackbits(snap.view(np.uint8)).sum())
            assert pool.free_count() == T - used
            if drained:
                assert used == T
            else:
                assert len(g) == 2 * n and cnt["errors"] == 0
            grants.append(len(g))
            _free(torch, pool, g.astype(np.int32), pbm.MODE_BATCHED)
            assert np.array_equal(pool.snapshot(), before) and pool.free_count() == A
        print(f"grants per run {grants}, rank fallbacks over the runs {fallbacks}", flush=True)

def _check_cycle(torch, pool, T, before, n, get_mode, free_mode, cw=0, expect_prefix=False,
                 budget=0):
    A = orc.count_free(before, T)
    ids, steps = _get(torch, pool, n, get_mode, cw=cw, budget=budget)
    after = pool.snapshot()
    rc, g = orc.check_alloc(before, after, T, ids)
    assert rc == 0, f"invariant {rc} violated"
    assert g == min(n, A)
    assert pool.free_count() == A - g
    if expect_prefix:
        assert (ids[:g] >= 0).all() and (ids[g:] == -1).all()
    _free(torch, pool, ids, free_mode)
    assert np.array_equal(pool.snapshot(), before)
    assert pool.free_count() == A
    return ids, steps


def test_c1_plain_rwbm_cycle(gpu):
    # 65,536 pages empty, 32,768 threads each getPage then freePage
    T, n = 65536, 32768
    log_case("C1 plain rwbm")
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=42) as pool:
        pool.clear_counters()
        ids, steps = _check_cycle(gpu, pool, T, orc.cleared(T), n, pbm.MODE_PLAIN,
                                  pbm.MODE_PLAIN)
        c = pool.counters()
        assert c["allocs"] == n and c["frees"] == n and c["failures"] == 0
        assert c["errors"] == 0
        assert c["steps"] == int(steps.sum())


@pytest.mark.parametrize("fill", [0.5, 0.9, 0.99])
@pytest.mark.parametrize("kind", ["plain_rwbm", "plain_collab", "batched64", "batched128",
                                  "batched256", "batched64_b2"])
def test_c2_cycle_invariants(gpu, fill, kind):
    # 1,048,576 pages, Bernoulli pre-fill, 262,144 requests then frees
    T, n = BIG, 262144
    log_case(f"C2 {kind} fill={fill}")
    strat = pbm.COLLAB_RW_BM if kind == "plain_collab" else pbm.RW_BM
    with pbm.Pool(T, strategy=strat, seed=42) as pool:
        pool.prefill_bernoulli(fill, 42)
        before = orc.prefill(T, fill, 42)
        if kind.startswith("batched"):
            cw, _, b = kind[7:].partition("_b")
            _check_cycle(gpu, pool, T, before, n, pbm.MODE_BATCHED, pbm.MODE_BATCHED,
                         cw=int(cw), expect_prefix=True, budget=int(b or 0))
        else:
            _check_cycle(gpu, pool, T, before, n, pbm.MODE_PLAIN, pbm.MODE_PLAIN)
        c = pool.counters()
        A = orc.count_free(before, T)
        assert c["allocs"] == min(n, A) and c["failures"] == n - min(n, A)
        assert c["frees"] == min(n, A) and c["errors"] == 0


def test_repeated_batched_cycles_restore_the_map(gpu):
    T, n = BIG, 262144
    log_case("C2 batched x5 cycles")
    with pbm.Pool(T, seed=1) as pool:
        pool.prefill_bernoulli(0.9, 42)
        before = orc.prefill(T, 0.9, 42)
        for k in range(5):
            ids, _ = _get(gpu, pool, n, pbm.MODE_BATCHED, req_base=k * n)
            assert orc.check_alloc(before, pool.snapshot(), T, ids)[0] == 0
            _free(gpu, pool, ids, pbm.MODE_BATCHED)
            assert np.array_equal(pool.snapshot(), before)


# ------------------------------------------------------------------ edge cases

def test_zero_requests_is_a_no_op(gpu):
    with pbm.Pool(4096) as pool:
        pool.get_pages(0, None, None)
        pool.free_pages(0, None)
        assert pool.count_used() == 0


@pytest.mark.parametrize("mode", [pbm.MODE_PLAIN, pbm.MODE_SERIAL, pbm.MODE_BATCHED])
def test_one_page_pool(gpu, mode):
    log_case(f"one-page pool mode={mode}")
    with pbm.Pool(1, seed=3) as pool:
        ids, _ = _get(gpu, pool, 5, mode)
        assert sorted(ids.tolist()) == [-1, -1, -1, -1, 0]
        assert pool.count_used() == 1


@pytest.mark.parametrize("mode", [pbm.MODE_PLAIN, pbm.MODE_BATCHED])
def test_full_pool_fails_every_request(gpu, mode):
    with pbm.Pool(65536) as pool:
        pool.prefill_bernoulli(1.0, 1)
        pool.clear_counters()
        ids, steps = _get(gpu, pool, 10000, mode)
        assert (ids == -1).all() and (steps == 0).all()
        assert pool.counters()["failures"] == 10000


@pytest.mark.parametrize("strategy", [pbm.RW_BM, pbm.COLLAB_RW_BM])
def test_padding_pages_are_never_granted(gpu, strategy):
    T = 1000  # last word holds 40 padding bits
    with pbm.Pool(T, strategy=strategy, seed=4) as pool:
        ids, _ = _get(gpu, pool, 1200, pbm.MODE_PLAIN)
        got = ids[ids >= 0]
        assert len(got) == T and len(set(got.tolist())) == T and got.max() < T
        assert pool.free_count() == 0


def test_double_free_and_bad_ids_count_errors(gpu):
    T = 4096
    with pbm.Pool(T) as pool:
        ids, _ = _get(gpu, pool, 100, pbm.MODE_PLAIN)
        _free(gpu, pool, ids, pbm.MODE_PLAIN)
        pool.clear_counters()
        before = pool.snapshot()
        _free(gpu, pool, ids[:10], pbm.MODE_PLAIN)                  # double free
        _free(gpu, pool, np.array([T, T + 5, 1 << 30]), pbm.MODE_PLAIN)  # out of range
        _free(gpu, pool, ids[10:20], pbm.MODE_BATCHED)
        assert np.array_equal(pool.snapshot(), before)
        c = pool.counters()
        assert c["errors"] == 23 and c["frees"] == 0
        assert pool.free_count() == T


def test_reference_style_prefill_requests(gpu):
    T = 1 << 16
    with pbm.Pool(T, seed=8) as pool:
        pool.prefill_requests(0.1)
        # (int)((1-0.1f)*T) pages requested and kept, as prefillBuffer does
        used = int(np.float32(1.0 - np.float32(0.1)) * np.float32(T))
        assert pool.count_used() == used
        assert abs(pool.free_page_percentage() - (T - used) / T) < 1e-6


def test_snapshot_restore_round_trip(gpu):
    T = 1 << 16
    with pbm.Pool(T) as pool:
        bm = orc.prefill(T, 0.7, 99)
        pool.restore(bm)
        assert np.array_equal(pool.snapshot(), bm)
        assert pool.free_count() == orc.count_free(bm, T)
        pool.reset()
        assert np.array_equal(pool.snapshot(), orc.cleared(T))


def test_free_n_pages_run_crossing_n_pages_is_an_error(gpu):
    # 1,000 pages: the last word holds 40 pages and 24 padding bits, which
    # must stay set whatever a free asks for
    T = 1000
    with pbm.Pool(T, strategy=pbm.RW_MALLOC) as pool:
        pool.prefill_bernoulli(1.0, 1)
        before = pool.snapshot()
        pool.clear_counters()
        ids = gpu.tensor([995, 990, 1 << 30], dtype=gpu.int32, device="cuda")
        counts = gpu.tensor([10, 10, 1], dtype=gpu.int32, device="cuda")
        pool.free_n_pages(1, ids[:1], counts[:1])   # 995..1004 crosses n_pages
        gpu.cuda.synchronize()
        assert np.array_equal(pool.snapshot(), before)
        assert pool.free_count() == 0
        c = pool.counters()
        assert c["errors"] == 1 and c["frees"] == 0
        pool.free_n_pages(1, ids[1:2], counts[1:2])  # 990..999 is the pool's last 10 pages
        gpu.cuda.synchronize()
        after = pool.snapshot()
        assert int(after[-1]) == int(before[-1]) & ~(((1 << 10) - 1) << (990 - 960))
        assert pool.free_count() == 10 and pool.counters()["frees"] == 10
        # nothing past n_pages is ever granted afterwards
        got, _ = _get(gpu, pool, 64, pbm.MODE_PLAIN)
        assert sorted(got[got >= 0].tolist()) == list(range(990, 1000))


def test_restore_sets_the_padding_bits(gpu):
    T = 1000
    with pbm.Pool(T) as pool:
        pool.restore(np.zeros(orc.n_words(T), dtype=np.uint64))  # padding bits clear
        assert np.array_equal(pool.snapshot(), orc.cleared(T))
        assert pool.free_count() == T
        ids, _ = _get(gpu, pool, 1200, pbm.MODE_PLAIN)
        got = ids[ids >= 0]
        assert len(got) == T and got.max() < T and len(set(got.tolist())) == T


def test_autotune_returns_a_working_launch_and_leaves_the_pool(gpu):
    T, n = 1 << 18, 65536
    log_case("autotune quick")
    with pbm.Pool(T, seed=6) as pool:
        pool.prefill_bernoulli(0.9, 42)
        before = pool.snapshot()
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        r = pbm.autotune(pool, n, ids, pbm.TUNE_QUICK)
        assert r["candidates"] == 52 and r["ms_alloc"] > 0 and r["ms_free"] > 0
        assert np.array_equal(pool.snapshot(), before)
        assert pool.free_count() == orc.count_free(before, T)
        pool.get_pages(n, ids, None, 0, r["get_launch"])
        gpu.cuda.synchronize()
        got = ids.cpu().numpy()
        rc, g = orc.check_alloc(before, pool.snapshot(), T, got)
        assert rc == 0 and g == min(n, orc.count_free(before, T))
