"""The Ouroboros comparator (Ouroboros.cuh:22-62): a FIFO queue of free
pages filled by lazily carved chunks.  Serial runs match the oracle bit
for bit; concurrent runs hand out each page once, carve in order and give
every page back."""
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


@pytest.mark.parametrize("ps", [8, 264, 4096, 10000, 65536])
def test_chunk_pages_match_the_size_classes(gpu, ps):
    with pbm.Pool(256, ps, strategy=pbm.OUROBOROS) as pool:
        assert pool.ctx().chunk_pages == orc.ouroboros_chunk_pages(ps)


@pytest.mark.parametrize("T,ps,fill", [(1 << 14, 4096, 0.5), (5000, 512, 0.0),
                                       (1 << 16, 264, 0.9)])
def test_serial_queue_matches_oracle(gpu, T, ps, fill):
    log_case(f"serial Ouroboros T={T} ps={ps} fill={fill}")
    cp = orc.ouroboros_chunk_pages(ps)
    rng = np.random.default_rng(1)
    with pbm.Pool(T, ps, strategy=pbm.OUROBOROS, seed=3) as pool:
        pool.prefill_bernoulli(fill, 5)
        q = orc.QueueState(orc.prefill(T, fill, 5), T, cp)
        n = min(3000, T)
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL)
        eids, esteps = q.get(n)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        back = rng.permutation(ids[ids >= 0])[: max(1, (ids >= 0).sum() // 2)]
        back = np.concatenate([back, back[:3]])          # three double frees
        _free(gpu, pool, back, pbm.MODE_SERIAL)
        q.put(back)
        assert pool.counters()["errors"] == 3
        ids, steps = _get(gpu, pool, n, pbm.MODE_SERIAL, req_base=n)
        eids, esteps = q.get(n)
        assert np.array_equal(ids, eids) and np.array_equal(steps, esteps)
        assert np.array_equal(pool.snapshot(), q.bm)
        assert pool.free_count() == q.free[0]


@pytest.mark.parametrize("fill", [0.5, 0.9])
def test_concurrent_queue_cycle(gpu, fill):
    T, n = 1 << 18, 65536
    log_case(f"concurrent Ouroboros fill={fill}")
    with pbm.Pool(T, strategy=pbm.OUROBOROS, seed=31) as pool:
        pool.prefill_bernoulli(fill, 42)
        before = orc.prefill(T, fill, 42)
        A = orc.count_free(before, T)
        for rnd in range(3):
            ids, _ = _get(gpu, pool, n, pbm.MODE_PLAIN, req_base=rnd * n)
            rc, g = orc.check_alloc(before, pool.snapshot(), T, ids)
            assert rc == 0 and g == min(n, A)
            assert pool.free_count() == A - g
            _free(gpu, pool, ids, pbm.MODE_PLAIN)
            assert np.array_equal(pool.snapshot(), before)
            assert pool.free_count() == A
        assert pool.counters()["errors"] == 0


def test_fresh_pool_carves_in_order(gpu):
    # k concurrent gets on an empty pool take pages from the front of the
    # carve order; each wave carves at most chunk_pages - 1 pages it leaves
    # to others
    T, k, ps = 1 << 16, 40000, 512
    cp = orc.ouroboros_chunk_pages(ps)
    with pbm.Pool(T, ps, strategy=pbm.OUROBOROS) as pool:
        ids, _ = _get(gpu, pool, k, pbm.MODE_PLAIN)
        assert np.unique(ids).size == k and ids.min() >= 0
        waves = -(-k // 64)
        assert ids.max() < k + waves * (cp - 1)


def test_prefill_requests_takes_the_first_carved_pages(gpu):
    # prefillBuffer through the allocator itself (Ouroboros.cuh:50-61):
    # the used pages lie in the carved prefix
    T, ps = 1 << 16, 4096
    cp = orc.ouroboros_chunk_pages(ps)
    with pbm.Pool(T, ps, strategy=pbm.OUROBOROS) as pool:
        pool.prefill_requests(0.25)
        used = np.unpackbits(pool.snapshot().view(np.uint8), bitorder="little")[:T]
        n_req = int((1.0 - np.float32(0.25)) * np.float32(T))
        assert used.sum() == n_req
        assert not used[n_req + (-(-n_req // 64)) * (cp - 1):].any()


def test_queue_rejects_bitmap_only_calls(gpu):
    with pbm.Pool(4096, strategy=pbm.OUROBOROS) as pool:
        ids = gpu.zeros((64,), dtype=gpu.int32, device="cuda")
        with pytest.raises(pbm.PbmError):
            pool.get_pages(64, ids, None, 0, pbm.launch(mode=pbm.MODE_BATCHED))
        counts = gpu.ones((64,), dtype=gpu.int32, device="cuda")
        with pytest.raises(pbm.PbmError):
            pool.get_n_pages(64, counts, ids)
