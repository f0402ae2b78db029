"""Analytic step-count models of the allocation policies.

Restates theoretical_models.py:12-62 of the reference with the same names
and argument meaning: T pages in the pool, A of them free, N requesting
threads.  Each takes the word width W (bits per map word) and the lanes
that walk together (a warp or wave) as keywords; the defaults, W=32 and 32
lanes, are the reference's.  The build's map words and waves are 64 wide.

  Queue_TAS, Queue_WAS  linked-list FIFO: mean and per-warp worst position
  RW_TAS, RW_WAS        single-page random walk (RandomWalkBasic)
  RWBM_TAS, RWBM_WAS    random walk over bitmap words (RandomWalkBasic_BM)
  CoRWBM_WAS            collaborative walk (CollabRW_BM), steps per warp

TAS = mean steps per thread, WAS = mean over warps of the warp's largest
step count (unitTests/metrics.h:13-33).
"""
from __future__ import annotations

import math

_EPS = 1e-15
_MAX_TERMS = 10 ** 7     # the reference's INFINITY (theoretical_models.py:9)


def Queue_TAS(T, A, N):
    """Mean FIFO position of N serialised requests."""
    return (N + 1) / 2


def Queue_WAS(T, A, N, lanes: int = 32):
    """Expected largest of `lanes` positions among N: N * lanes / (lanes + 1)."""
    return lanes / (lanes + 1) * N


def RW_TAS(T, A, N):
    """T/N * ln(A / (A - N)): the integral of T/(A-k) over the N grants."""
    return T / N * math.log(A / (A - N))


def _tail_sum(term) -> float:
    """sum_{k>=0} term(k) for a non-increasing term, stopped when a term
    falls below 1e-15 or after the reference's 10^7 terms."""
    s = 0.0
    for k in range(_MAX_TERMS):
        t = term(k)
        s += t
        if k > 0 and t < _EPS:
            break
    return s


def RW_WAS(T, A, N, lanes: int = 32):
    """E[max of `lanes` walks], each probe failing with q = (T-A+N)/T:
    sum_k 1 - (1 - q^k)^lanes."""
    q = (T - A + N) / T
    return _tail_sum(lambda k: 1.0 - (1.0 - q ** k) ** lanes)


def RWBM_TAS(T, A, N, W: int = 32):
    """The reference's sum (theoretical_models.py:34-41): grant k sees
    T-A+k used pages, a word is full with probability ((T-A+k)/T)^W; k runs
    over N-1 grants and the sum is divided by N."""
    s = 0.0
    for k in range(N - 1):
        s += 1.0 / (1.0 - ((T - A + k) / T) ** W) / N
    return s


def RWBM_TAS_mean(T, A, N, W: int = 32):
    """The same sum over all N grants divided by N (the mean the build's
    serial oracle reproduces)."""
    s = 0.0
    for k in range(N):
        s += 1.0 / (1.0 - ((T - A + k) / T) ** W)
    return s / N


def RWBM_WAS(T, A, N, W: int = 32, lanes: int = 32):
    """E[max of `lanes` word walks], a word full with q^W, q = (T-A+N)/T:
    sum_k 1 - (1 - q^(W k))^lanes."""
    q = (T - A + N) / T
    return _tail_sum(lambda k: 1.0 - (1.0 - q ** (W * k)) ** lanes)


def CoRWBM_WAS(T, A, N, W: int = 32, lanes: int = 32):
    """Steps of a collaborative warp: the window of W*lanes bits yields on
    average mu = W*lanes*(A-N)/T free bits per step; the steps to gather
    `lanes` pages are the first passage of a drifted random walk to level
    `lanes`, whose mean is lanes/mu (the inverse-Gaussian mean the
    reference integrates numerically, theoretical_models.py:54-62); at
    least one step."""
    mu = W * lanes * (A - N) / T
    return max(1.0, lanes / mu)


def CoRWBM_WAS_integral(T, A, N, W: int = 32, lanes: int = 32):
    """The reference's integral, evaluated with scipy: int_0^inf lanes /
    (sigma sqrt(2 pi t)) exp(-(lanes - mu t)^2 / (2 sigma^2 t)) dt."""
    from scipy import integrate

    mu = W * lanes * (A - N) / T
    sigma = W * lanes * (A - N) / T * (T - A + N) / T

    def f(t):
        if t <= 0:
            return 0.0
        return lanes / (sigma * math.sqrt(2 * math.pi * t)) * math.exp(
            -(lanes - mu * t) ** 2 / (2 * sigma ** 2 * t))

    peak = lanes / mu
    a, _ = integrate.quad(f, 0, peak, limit=200)
    b, _ = integrate.quad(f, peak, math.inf, limit=200)
    return max(1.0, a + b)


def sweep(T: int, N: int, fills, W: int = 64, lanes: int = 64) -> list[dict]:
    """Every model at pool fill f (A = round((1-f) T)) for each f, with the
    given word width and lanes."""
    rows = []
    for f in fills:
        A = int(round((1 - f) * T))
        n = min(N, A - 1)
        rows.append({"fill": f, "T": T, "A": A, "N": n,
                     "RW_TAS": RW_TAS(T, A, n), "RW_WAS": RW_WAS(T, A, n, lanes),
                     "RWBM_TAS": RWBM_TAS_mean(T, A, n, W),
                     "RWBM_WAS": RWBM_WAS(T, A, n, W, lanes),
                     "CoRWBM_WAS": CoRWBM_WAS(T, A, n, W, lanes),
                     "Queue_TAS": Queue_TAS(T, A, n), "Queue_WAS": Queue_WAS(T, A, n, lanes)})
    return rows
