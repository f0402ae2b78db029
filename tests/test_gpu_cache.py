"""The buffer cache (batched CLOCK over pool frames) through the C-ABI: the
device Zipf trace equals the oracle's; hits, misses, the eviction order,
the arm and every slot's page and bit equal the serial oracle's for
batches of 1 (the textbook CLOCK) up to 65,536; frames are distinct pool
pages, exactly as many as pages held."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu


def _trace(gpu, n_logical, s, seed, n):
    t = gpu.empty(n, dtype=gpu.int32, device="cuda")
    pbm.zipf_trace(t, n_logical, s, seed)
    gpu.cuda.synchronize()
    return t


def test_zipf_trace_matches_oracle(gpu):
    log_case("zipf trace")
    t = _trace(gpu, 1 << 20, 1.1, 7, 1 << 20).cpu().numpy()
    assert np.array_equal(t, orc.zipf_trace(1 << 20, 1.1, 7, 0, 1 << 20))


def _check(gpu, pool, cache, cc, T):
    st = cache.stats()
    assert st["hits"] == int(cc.stats[0]) and st["misses"] == int(cc.stats[1])
    assert st["evictions"] == int(cc.stats[2]) and st["alloc_failures"] == 0
    assert st["scan_timeouts"] == 0
    assert st["arm"] == int(cc.arm_res[0]) and st["resident"] == int(cc.arm_res[1])
    assert np.array_equal(cache.evictions(), cc.evictions)
    owner, frame, bit = cache.slots()
    assert np.array_equal(owner, cc.owner)
    assert np.array_equal(bit.astype(np.uint8), cc.refbit)
    held = frame[owner >= 0]
    assert (held >= 0).all() and len(np.unique(held)) == held.size
    assert pool.count_used() == held.size == st["resident"]
    bm = pool.snapshot()
    assert all((int(bm[int(f) // 64]) >> (int(f) % 64)) & 1 for f in held[:5000])


@pytest.mark.parametrize("B,n", [(1, 20000), (7, 60000), (256, 300000), (2048, 1000000)])
def test_cache_matches_oracle(gpu, B, n):
    T, L, C = 1 << 16, 1 << 14, 2048
    log_case(f"cache batch={B} refs={n}")
    refs = _trace(gpu, L, 1.1, 3, n)
    cc = orc.ClockCache(L, C, B, log_cap=n)
    cc.run(refs.cpu().numpy())
    with pbm.Pool(T, strategy=pbm.CRW_BM, seed=5) as pool:
        with pbm.Cache(pool, L, C, B, log_capacity=n) as cache:
            cache.run(refs)
            gpu.cuda.synchronize()
            _check(gpu, pool, cache, cc, T)


def test_cache_runs_continue(gpu):
    # a trace run in three calls (split at batch boundaries) ends where one
    # call over the whole trace ends
    T, L, C, B, n = 1 << 16, 1 << 14, 2048, 256, 300000
    log_case("cache run in three calls")
    refs = _trace(gpu, L, 1.1, 17, n)
    cc = orc.ClockCache(L, C, B, log_cap=n)
    cc.run(refs.cpu().numpy())
    with pbm.Pool(T, strategy=pbm.CRW_BM, seed=5) as pool:
        with pbm.Cache(pool, L, C, B, log_capacity=n) as cache:
            for a, b in [(0, 256 * 300), (256 * 300, 256 * 301), (256 * 301, n)]:
                cache.run(refs[a:b])
                gpu.cuda.synchronize()
            assert cache.stats()["batches"] == -(-n // B)
            _check(gpu, pool, cache, cc, T)


def test_cache_skips_out_of_range_references(gpu):
    # ids outside 0 .. n_logical-1 are neither hits nor misses; the rest of
    # each batch runs as if they were absent
    T, L, C, B, n = 1 << 16, 1 << 14, 2048, 256, 200000
    log_case("cache with out-of-range references")
    refs = _trace(gpu, L, 1.1, 23, n).cpu().numpy()
    rng = np.random.default_rng(5)
    bad = rng.choice(n, n // 100, replace=False)
    refs[bad] = rng.choice(np.array([-1, -7, L, L + 5, 1 << 30], dtype=np.int32), bad.size)
    cc = orc.ClockCache(L, C, B, log_cap=n)
    cc.run(refs)
    with pbm.Pool(T, strategy=pbm.CRW_BM, seed=5) as pool:
        with pbm.Cache(pool, L, C, B, log_capacity=n) as cache:
            cache.run(gpu.from_numpy(refs).cuda())
            gpu.cuda.synchronize()
            assert cache.stats()["invalid_refs"] == bad.size
            _check(gpu, pool, cache, cc, T)


@pytest.mark.parametrize("C,B,n", [(3001, 1111, 150000), (1025, 1025, 150000), (5000, 1, 30000)])
def test_cache_ragged_tiles(gpu, C, B, n):
    # capacities and batches that end inside a look-back tile of 1,024
    T, L = 1 << 15, 1 << 13
    log_case(f"cache capacity={C} batch={B}")
    refs = _trace(gpu, L, 1.0, 13, n)
    cc = orc.ClockCache(L, C, B, log_cap=n)
    cc.run(refs.cpu().numpy())
    with pbm.Pool(T, strategy=pbm.CRW_BM, seed=4) as pool:
        with pbm.Cache(pool, L, C, B, log_capacity=n) as cache:
            cache.run(refs)
            gpu.cuda.synchronize()
            _check(gpu, pool, cache, cc, T)


@pytest.mark.parametrize("strategy", [pbm.RW_BM, pbm.DIST_CLOCK, pbm.COLLAB_RW_BM,
                                      pbm.CRW_BASIC, pbm.COLLAB_RW_BASIC])
def test_cache_frame_policies(gpu, strategy):
    T, L, C, B, n = 1 << 15, 1 << 13, 4096, 512, 200000
    log_case(f"cache frames strategy={strategy}")
    refs = _trace(gpu, L, 0.9, 11, n)
    cc = orc.ClockCache(L, C, B, log_cap=n)
    cc.run(refs.cpu().numpy())
    with pbm.Pool(T, strategy=strategy, seed=6) as pool:
        with pbm.Cache(pool, L, C, B, log_capacity=n) as cache:
            cache.run(refs)
            gpu.cuda.synchronize()
            _check(gpu, pool, cache, cc, T)


@pytest.mark.parametrize("B", [1 << 16, 1 << 18, 1 << 20])
def test_cache_c3_sizes(gpu, B):
    # the buffer-cache workload whole: 16,777,216 frames (page map only),
    # 4,194,304 logical pages, 1,048,576 slots, Zipf(1.1) seed 7, all
    # 67,108,864 references, in batches of 65,536 (the benchmark's) and of
    # 262,144 and 1,048,576 (as many as the slots)
    T, L, C, n = 1 << 24, 1 << 22, 1 << 20, 1 << 26
    log_case(f"cache C3 sizes, 64M references, batch {B}")
    refs = _trace(gpu, L, 1.1, 7, n)
    h_refs = refs.cpu().numpy()
    cc = orc.ClockCache(L, C, B, log_cap=n)
    cc.run(h_refs)
    with pbm.Pool(T, strategy=pbm.CRW_BM, seed=7) as pool:
        with pbm.Cache(pool, L, C, B, log_capacity=n) as cache:
            cache.run(refs)
            gpu.cuda.synchronize()
            _check(gpu, pool, cache, cc, T)
    # the textbook CLOCK (one reference at a time) on the same trace: the
    # batch rule's hit count lies within 0.01% of the references of it at
    # the benchmark's batch
    if B == 1 << 16:
        tb = orc.ClockCache(L, C, 1)
        tb.run_textbook(h_refs)
        assert int(tb.stats[0] + tb.stats[1]) == n
        assert abs(int(cc.stats[0]) - int(tb.stats[0])) < n // 10000


def test_cache_rejects_bad_configs(gpu):
    log_case("cache bad configs")
    with pbm.Pool(1024, strategy=pbm.CRW_BM) as pool:
        for args in [(0, 16, 4), (100, 2048, 4), (100, 16, 0), (100, 16, 17)]:
            with pytest.raises(pbm.PbmError):
                pbm.Cache(pool, *args)
    with pbm.Pool(1024, strategy=pbm.LINKED_LIST) as pool:
        with pytest.raises(pbm.PbmError):
            pbm.Cache(pool, 100, 16, 4)
