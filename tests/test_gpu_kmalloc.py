"""KernelMallocSimple / KernelMallocCollab (malloc.cu) through the C-ABI:
serial runs match the oracle bit for bit, on-page headers included;
concurrent runs keep spaces disjoint and give every page back."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu

NONE = np.uint64(2 ** 64 - 1)


def _sizes(rng, n, hi):
    s = np.exp(rng.uniform(0, np.log(hi), size=n)).astype(np.uint32)
    s[rng.random(n) < 0.02] = 0
    return s


def _offsets(pool, ptrs):
    p = ptrs.cpu().numpy().view(np.uint64)
    base = np.uint64(pool.page_address(0))
    return np.where(p == 0, NONE, p - base)


def _malloc(gpu, km, sizes, mode, req_base=0):
    n = sizes.size
    d_sizes = gpu.from_numpy(sizes.view(np.int32)).cuda()
    ptrs = gpu.zeros((n,), dtype=gpu.int64, device="cuda")
    steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
    km.malloc(n, d_sizes, ptrs, steps, req_base, pbm.launch(mode=mode))
    gpu.cuda.synchronize()
    return ptrs, steps.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("fill,hi", [(0.5, 6000), (0.0, 300), (0.9, 70000)])
def test_serial_kmalloc_collab_matches_oracle(gpu, fill, hi):
    T, ps, n = 4096, 4096, 1500
    log_case(f"serial KernelMallocCollab fill={fill} hi={hi}")
    sizes = _sizes(np.random.default_rng(3), n, hi)
    with pbm.Pool(T, ps, strategy=pbm.COLLAB_RW_BM, backing=True, seed=11) as pool:
        pool.prefill_bernoulli(fill, 2)
        with pbm.KMalloc(pool, pbm.KMALLOC_COLLAB) as km:
            ptrs, steps = _malloc(gpu, km, sizes, pbm.MODE_SERIAL, req_base=5)
            st = orc.State(orc.prefill(T, fill, 2), T)
            mem = np.zeros(T * ps, np.uint8)
            wr = np.zeros(T * ps, np.uint8)
            off, esteps = st.kmalloc_collab(mem, wr, sizes, ps, 5, 11)
            assert np.array_equal(_offsets(pool, ptrs), off)
            assert np.array_equal(steps, esteps)
            assert np.array_equal(pool.snapshot(), st.bm)
            dev = pool.read_pool(0, T * ps)
            w = wr.astype(bool)
            assert w.any() and np.array_equal(dev[w], mem[w])
            # free every other request, then the rest
            half = np.zeros(n, dtype=np.int64)
            half[::2] = ptrs.cpu().numpy()[::2]
            km.free(n, gpu.from_numpy(half).cuda(), pbm.launch(mode=pbm.MODE_SERIAL))
            gpu.cuda.synchronize()
            st.kfree_collab(mem, np.where(np.arange(n) % 2 == 0, off, NONE))
            assert np.array_equal(pool.snapshot(), st.bm)
            dev = pool.read_pool(0, T * ps)
            assert np.array_equal(dev[w], mem[w])      # occupied fields included
            rest = ptrs.cpu().numpy().copy()
            rest[::2] = 0
            km.free(n, gpu.from_numpy(rest).cuda(), pbm.launch(mode=pbm.MODE_SERIAL))
            gpu.cuda.synchronize()
            assert np.array_equal(pool.snapshot(), orc.prefill(T, fill, 2))
            assert pool.free_count() == orc.count_free(orc.prefill(T, fill, 2), T)


@pytest.mark.parametrize("fill,hi", [(0.5, 3000), (0.2, 40000)])
def test_serial_kmalloc_simple_matches_oracle(gpu, fill, hi):
    T, ps, n = 4096, 4096, 2000
    log_case(f"serial KernelMallocSimple fill={fill} hi={hi}")
    sizes = _sizes(np.random.default_rng(4), n, hi)
    with pbm.Pool(T, ps, strategy=pbm.RW_BM, backing=True, seed=13) as pool:
        pool.prefill_bernoulli(fill, 6)
        with pbm.KMalloc(pool, pbm.KMALLOC_SIMPLE, max_threads=64) as km:
            ptrs, steps = _malloc(gpu, km, sizes[:1200], pbm.MODE_SERIAL, req_base=1)
            ptrs2, steps2 = _malloc(gpu, km, sizes[1200:], pbm.MODE_SERIAL, req_base=1201)
            st = orc.State(orc.prefill(T, fill, 6), T)
            mem = np.zeros(T * ps, np.uint8)
            wr = np.zeros(T * ps, np.uint8)
            off, esteps, cur = st.kmalloc_simple(mem, wr, sizes[:1200], ps, 1, 13)
            off2, esteps2, cur2 = st.kmalloc_simple(mem, wr, sizes[1200:], ps, 1201, 13,
                                                    current=cur)
            assert np.array_equal(_offsets(pool, ptrs), off)
            assert np.array_equal(_offsets(pool, ptrs2), off2)
            assert np.array_equal(steps, esteps) and np.array_equal(steps2, esteps2)
            assert km.current()[0] == cur2
            assert np.array_equal(pool.snapshot(), st.bm)
            dev = pool.read_pool(0, T * ps)
            w = wr.astype(bool)
            assert np.array_equal(dev[w], mem[w])
            km.reset()
            gpu.cuda.synchronize()
            assert pool.count_used() == 0 and (km.current() == -1).all()


def _disjoint(off, sizes):
    ok = off != NONE
    lo = off[ok].astype(np.int64) - 8
    hi = off[ok].astype(np.int64) + ((sizes[ok].astype(np.int64) + 7) & ~7)
    order = np.argsort(lo)
    return bool((lo[order][1:] >= hi[order][:-1]).all())


def test_concurrent_kmalloc_collab_spaces_are_disjoint_and_freed(gpu):
    T, ps, n = 1 << 17, 4096, 1 << 17
    log_case("concurrent KernelMallocCollab")
    sizes = _sizes(np.random.default_rng(5), n, 9000)
    with pbm.Pool(T, ps, strategy=pbm.RW_BM, backing=True, seed=17) as pool:
        pool.prefill_bernoulli(0.3, 9)
        before = pool.snapshot()
        with pbm.KMalloc(pool, pbm.KMALLOC_COLLAB) as km:
            ptrs, _ = _malloc(gpu, km, sizes, pbm.MODE_PLAIN)
            off = _offsets(pool, ptrs)
            assert (off != NONE).all()
            assert _disjoint(off, sizes)
            # every space lies on pages the map shows used
            used = np.unpackbits(pool.snapshot().view(np.uint8), bitorder="little")[:T]
            first = ((off.astype(np.int64) - 8) // ps)
            assert used[first].all()
            km.free(n, ptrs)
            gpu.cuda.synchronize()
            assert np.array_equal(pool.snapshot(), before)
            assert pool.free_count() == orc.count_free(before, T)
            assert pool.counters()["errors"] == 0


def test_concurrent_kmalloc_simple_spaces_are_disjoint(gpu):
    T, ps, n = 1 << 15, 4096, 1 << 16
    log_case("concurrent KernelMallocSimple")
    sizes = _sizes(np.random.default_rng(6), n, 2000)
    with pbm.Pool(T, ps, strategy=pbm.RW_BM, backing=True, seed=19) as pool:
        with pbm.KMalloc(pool, pbm.KMALLOC_SIMPLE, max_threads=8192) as km:
            ptrs, _ = _malloc(gpu, km, sizes, pbm.MODE_PLAIN)
            off = _offsets(pool, ptrs)
            assert (off != NONE).all() and _disjoint(off, sizes)
            # thread t served requests t, t + 8192, ...: each chunk holds the
            # spaces of one thread only
            cur = km.current()
            assert (cur >= 0).all() and np.unique(cur).size == cur.size
            km.free(n, ptrs)                 # a no-op, as in the reference
            gpu.cuda.synchronize()
            assert pool.count_used() > 0
            km.reset()
            gpu.cuda.synchronize()
            assert pool.count_used() == 0


def test_kmalloc_rejects_pools_it_cannot_serve(gpu):
    with pbm.Pool(1024, 4096, backing=False) as pool:
        with pytest.raises(pbm.PbmError):
            pbm.KMalloc(pool, pbm.KMALLOC_COLLAB)
    with pbm.Pool(1024, 100, backing=True) as pool:
        with pytest.raises(pbm.PbmError):
            pbm.KMalloc(pool, pbm.KMALLOC_SIMPLE)
    with pbm.Pool(1024, 4096, backing=True) as pool:
        with pytest.raises(pbm.PbmError):
            pbm.KMalloc(pool, 7)


def test_frees_of_pointers_outside_the_pool_count_errors(gpu):
    # freeRW and KernelMallocCollab's free check a pointer against the pool
    # before they read the header in front of it
    T, ps = 4096, 4096
    log_case("frees of pointers outside the pool")
    with pbm.Pool(T, ps, strategy=pbm.RW_MALLOC, backing=True, seed=2) as pool:
        base = pool.page_address(0)
        bad = np.array([8, base - 4096, base + T * ps + 64], dtype=np.uint64)
        d_bad = gpu.from_numpy(bad.view(np.int64)).cuda()
        before = pool.snapshot()
        pool.clear_counters()
        pool.free_ptrs(3, d_bad)
        gpu.cuda.synchronize()
        assert pool.counters()["errors"] == 3
        with pbm.KMalloc(pool, pbm.KMALLOC_COLLAB) as km:
            km.free(3, d_bad)
            gpu.cuda.synchronize()
        assert pool.counters()["errors"] == 6
        assert np.array_equal(pool.snapshot(), before)
        assert pool.free_count() == T
