"""The analytic models against what fixes them independently: Monte Carlo
of the random variables they describe, exact sums, and the closed form of
the inverse-Gaussian mean the reference integrates."""
import math

import numpy as np
import pytest

import models
import pbm_oracle as orc


def test_queue_models_are_positions():
    rng = np.random.default_rng(1)
    N = 5000
    # WAS: the largest of 32 positions drawn without replacement from 1..N,
    # whose mean is 32 (N+1) / 33
    draws = np.array([rng.choice(N, 32, replace=False).max() + 1 for _ in range(4000)])
    assert abs(draws.mean() - models.Queue_WAS(0, 0, N)) < 4 * draws.std() / math.sqrt(draws.size) + 1
    assert models.Queue_TAS(0, 0, N) == np.arange(1, N + 1).mean()


@pytest.mark.parametrize("T,A,N", [(10000, 3000, 1500), (100000, 12000, 5000)])
def test_rw_was_is_the_expected_max_of_geometric_walks(T, A, N):
    q = (T - A + N) / T
    rng = np.random.default_rng(2)
    g = rng.geometric(1 - q, size=(200000, 32))   # probes until a free page
    m = g.max(axis=1)
    # sum_k 1-(1-q^k)^32 counts P(max > k) for k >= 0, i.e. E[max]
    assert abs(m.mean() - models.RW_WAS(T, A, N)) < 5 * m.std() / math.sqrt(m.size)


@pytest.mark.parametrize("W,lanes", [(32, 32), (64, 64)])
def test_rwbm_was_is_the_expected_max_of_word_walks(W, lanes):
    T, A, N = 100000, 3000, 1000
    q = ((T - A + N) / T) ** W
    rng = np.random.default_rng(3)
    g = rng.geometric(1 - q, size=(200000, lanes))
    m = g.max(axis=1)
    assert abs(m.mean() - models.RWBM_WAS(T, A, N, W, lanes)) < 5 * m.std() / math.sqrt(m.size)


def test_rw_tas_approximates_the_exact_mean():
    T, A, N = 1_000_003, 100_000, 60_000
    exact = np.mean(T / (A - np.arange(N)))
    assert abs(models.RW_TAS(T, A, N) - exact) / exact < 1e-4
    assert models.RW_TAS(T, A, N) == pytest.approx(orc.rw_tas(T, A, N), rel=1e-15)


def test_rwbm_tas_mean_is_the_oracle_expectation_and_the_reference_sum_differs_by_one_term():
    T, A, N = 1 << 20, 104858, 104000
    assert models.RWBM_TAS_mean(T, A, N, 64) == pytest.approx(orc.expected_steps(T, A, N, 64),
                                                              rel=1e-12)
    last = 1.0 / (1.0 - ((T - A + N - 1) / T) ** 32) / N
    assert models.RWBM_TAS(T, A, N) + last == pytest.approx(models.RWBM_TAS_mean(T, A, N), rel=1e-12)


@pytest.mark.parametrize("T,A,N", [(1 << 20, 104858, 50000), (65536, 20000, 5000),
                                   (1 << 20, 600000, 100000)])
def test_corwbm_was_closed_form_equals_the_reference_integral(T, A, N):
    assert models.CoRWBM_WAS_integral(T, A, N) == pytest.approx(models.CoRWBM_WAS(T, A, N),
                                                                rel=1e-6)


def test_sweep_rows():
    rows = models.sweep(1 << 20, 262144, [0.5, 0.9, 0.99])
    assert [r["fill"] for r in rows] == [0.5, 0.9, 0.99]
    assert rows[1]["A"] == 104858 and rows[1]["N"] == 104857
    assert all(r["RWBM_TAS"] >= 1 and r["RW_TAS"] >= 1 for r in rows)


# ------------------------------------------------------------------ metrics (unitTests/metrics.h)

def test_aggregate_metrics_known_answer():
    from metrics import aggregate_metrics, sample_average
    # 67 threads: two full warps and a partial third (metrics.h:21-30)
    steps = [5] * 32 + [1] * 32 + [9, 0, 0]
    avg, avg_max = aggregate_metrics(steps)
    assert avg == pytest.approx((160 + 32 + 9) / 67, rel=1e-7)
    assert avg_max == 5.0                                   # (5 + 1 + 9) / 3
    assert sample_average([(1.0, 2.0, 3.0), (3.0, 4.0, 5.0)]) == (2.0, 3.0, 4.0)


def test_walk_stats_equals_the_reference_metrics_when_all_granted():
    from metrics import aggregate_metrics, walk_stats
    rng = np.random.default_rng(1)
    steps = rng.integers(1, 40, size=1000)
    ws = walk_stats(steps, np.arange(1000))
    avg, avg_max = aggregate_metrics(steps)
    assert ws["tas"] == pytest.approx(avg, rel=1e-6)
    assert ws["was32"] == pytest.approx(avg_max, rel=1e-6)
    assert ws["max"] == steps.max()
    # a failed request's steps do not count, nor does a group without grants
    ids = np.arange(128)
    ids[64:] = -1
    st = np.ones(128, dtype=np.int64)
    st[70] = 99
    ws = walk_stats(st, ids)
    assert ws["tas"] == 1.0 and ws["was64"] == 1.0 and ws["was32"] == 1.0 and ws["max"] == 1
