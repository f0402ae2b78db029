"""Cluster counts and the occupancy monitor through the C-ABI, and the
analytic models against measured plain-walk step counts."""
import numpy as np
import pytest

import models
import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T,fill", [(1, 0.0), (64, 0.5), (1000, 0.3), (1 << 20, 0.9),
                                    (70001, 0.999), (4096, 0.0), (4096, 1.0)])
def test_cluster_count_matches_oracle(gpu, T, fill):
    log_case(f"clusters T={T} fill={fill}")
    with pbm.Pool(T, strategy=pbm.RW_BM) as pool:
        pool.prefill_bernoulli(fill, 5)
        assert pool.count_clusters() == orc.count_clusters(orc.prefill(T, fill, 5), T)


def test_monitor_sees_an_allocation_phase(gpu):
    T, n = 1 << 20, 1 << 18
    log_case("monitor")
    with pbm.Pool(T, strategy=pbm.RW_BM, seed=3) as pool:
        pool.prefill_bernoulli(0.5, 1)
        A = pool.free_count()
        count = 40
        samples = gpu.zeros(2 * count, dtype=gpu.int64, device="cuda")
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        gpu.cuda.synchronize()
        side, work = gpu.cuda.Stream(), gpu.cuda.Stream()
        pool.monitor(200, count, samples, side)         # 40 samples, 200 us apart
        pool.get_pages(n, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN), work)
        gpu.cuda.synchronize()
        s = samples.cpu().numpy().reshape(-1, 2)
        ticks, free = s[:, 0], s[:, 1]
        assert (np.diff(ticks) > 0).all()
        hz = pbm.wall_clock_hz()
        gap = np.diff(ticks).mean() / hz
        assert 150e-6 < gap < 5e-3
        # an allocation-only phase: free pages never rise, and the last
        # sample (taken after the phase) sees every grant
        assert (free <= A).all() and (np.diff(free) <= 0).all()
        assert free[-1] == A - min(n, A) == pool.free_count()


def _tas(gpu, strategy, T, fill, n, mode):
    with pbm.Pool(T, strategy=strategy, seed=17) as pool:
        pool.prefill_bernoulli(fill, 2)
        A = pool.free_count()
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.get_pages(n, ids, steps, 0, pbm.launch(mode=mode))
        gpu.cuda.synchronize()
        return A, steps.cpu().numpy().astype(np.float64).mean()


def test_models_against_measured_walks(gpu):
    # The models count the k-th grant against T-A+k used pages, one grant
    # after another.  Pool sizes that are not powers of two keep the LCG's
    # low bits from folding walks onto one cycle.
    log_case("models vs measured")
    T, fill, n = 1_000_003, 0.9, 60000
    # single-page walks: 60,000 concurrent threads on a million pages
    # rarely meet, so even the plain launch follows the serial model
    A, tas = _tas(gpu, pbm.RW_BASIC, T, fill, n, pbm.MODE_PLAIN)
    assert abs(tas - models.RW_TAS(T, A, n)) / models.RW_TAS(T, A, n) < 0.03
    # word walks: the model takes a word's free bits to be independent,
    # which holds while grants are about one per word (with several per
    # word, grants drain the same words and the model underestimates);
    # serially it holds, and concurrently threads that land on one word
    # beyond its free bits jump again, so the plain launch takes more steps
    n = 15000
    A, tas = _tas(gpu, pbm.RW_BM, T, fill, n, pbm.MODE_SERIAL)
    m = models.RWBM_TAS_mean(T, A, n, 64)
    assert abs(tas - m) / m < 0.03
    A, tas_plain = _tas(gpu, pbm.RW_BM, T, fill, n, pbm.MODE_PLAIN)
    assert tas_plain >= m
