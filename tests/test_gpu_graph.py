"""Allocation and free phases captured as HIP graphs (pbm_graph_create):
every replay, with eager work and synchronisations between replays,
equals the eager result and the oracle's."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T", [1 << 20, 100_003])
def test_rank_graph_replays_match_eager_and_oracle(gpu, T):
    n, fill = 262144, 0.9
    log_case(f"rank graph replays T={T}")
    rank = pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK)
    free = pbm.launch(pbm.MODE_BATCHED)
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=42) as pool:
        pool.prefill_bernoulli(fill, 42)
        before = pool.snapshot()
        st = orc.State(before, T)
        eids, esteps = st.batched_rank(n)
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        # alloc phase alone: replay, check, free eagerly, replay again
        g = pbm.Graph(pool, n, ids, steps, 0, rank, free, pbm.PHASE_ALLOC)
        for _ in range(4):
            ids.fill_(-7)
            g.launch()
            gpu.cuda.synchronize()
            assert np.array_equal(ids.cpu().numpy(), eids)
            assert np.array_equal(steps.cpu().numpy().view(np.uint32), esteps)
            assert np.array_equal(pool.snapshot(), st.bm)
            pool.free_pages(n, ids, free)                 # eager work between replays
            gpu.cuda.synchronize()
            assert np.array_equal(pool.snapshot(), before)
        g.close()
        # both phases in one graph: the map comes back after every replay
        g = pbm.Graph(pool, n, ids, None, 0, rank, free, pbm.PHASE_ALLOC | pbm.PHASE_FREE)
        for k in range(5):
            g.launch()
            gpu.cuda.synchronize()
            assert np.array_equal(ids.cpu().numpy(), eids), k
            assert np.array_equal(pool.snapshot(), before), k
            assert pool.free_count() == orc.count_free(before, T)
            pool.get_pages(64, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN))
            pool.free_pages(64, ids, pbm.launch(mode=pbm.MODE_PLAIN))
            gpu.cuda.synchronize()
        g.close()
        assert pool.counters()["errors"] == 0


def test_walk_graph_replays_keep_the_invariants(gpu):
    T, n = 1 << 18, 65536
    log_case("batched walk graph replays")
    walk = pbm.launch(pbm.MODE_BATCHED, cluster_words=128)
    with pbm.Pool(T, strategy=pbm.COLLAB_RW_BM, seed=7) as pool:
        pool.prefill_bernoulli(0.5, 3)
        before = pool.snapshot()
        A = orc.count_free(before, T)
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        g = pbm.Graph(pool, n, ids, None, 0, walk, walk, pbm.PHASE_ALLOC)
        for k in range(4):
            g.launch()
            gpu.cuda.synchronize()
            got = ids.cpu().numpy()
            rc, granted = orc.check_alloc(before, pool.snapshot(), T, got)
            assert rc == 0 and granted == min(n, A), k
            pool.free_pages(n, ids, walk)
            gpu.cuda.synchronize()
            assert np.array_equal(pool.snapshot(), before), k
        g.close()


def test_multi_step_graph_replays_match_the_oracle(gpu):
    # 16 alloc+free steps in one graph: each replay runs all of them, the
    # ids are the last step's (the oracle's rank answer), the map comes back
    T, n, S = 1 << 20, 262144, 16
    log_case("rank graph of 16 steps")
    rank = pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK)
    free = pbm.launch(pbm.MODE_BATCHED)
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=42) as pool:
        pool.prefill_bernoulli(0.9, 42)
        before = pool.snapshot()
        eids, _ = orc.State(before, T).batched_rank(n)
        granted = int((eids >= 0).sum())
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        pool.clear_counters()
        with pbm.Graph(pool, n, ids, None, 0, rank, free, pbm.PHASE_ALLOC | pbm.PHASE_FREE,
                       n_steps=S) as g:
            for k in range(3):
                ids.fill_(-7)
                g.launch()
                gpu.cuda.synchronize()
                assert np.array_equal(ids.cpu().numpy(), eids), k
                assert np.array_equal(pool.snapshot(), before), k
        c = pool.counters()
        assert c["allocs"] == c["frees"] == 3 * S * granted and c["errors"] == 0
        with pytest.raises(pbm.PbmError):
            pbm.Graph(pool, n, ids, None, 0, rank, free, 3, n_steps=0)
