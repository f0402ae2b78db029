"""Walk-step metrics of an allocation batch (unitTests/metrics.h:13-47).

aggregate_metrics and sample_average follow the reference's definitions
exactly: TAS is the mean step count over all threads, and WAS is the mean
over consecutive groups of 32 thread indices (the last group may be
partial) of each group's largest step count, both in float32.
walk_stats is the form the benchmark reports for batches in which some
requests get no page: it restricts both means to granted requests, and
gives WAS for 64-lane waves as well as 32-lane warps.  When every request
is granted, its tas and was32 equal aggregate_metrics's two values.
"""
from __future__ import annotations

import numpy as np


def aggregate_metrics(step_counts, group: int = 32) -> tuple[float, float]:
    """(avgStep, avgMaxWarp) of metrics.h:13-33."""
    s = np.asarray(step_counts, dtype=np.int64)
    n = s.size
    if n == 0:
        return 0.0, 0.0
    avg = np.float32(np.float32(s.sum()) / np.float32(n))
    n_groups = -(-n // group)
    pad = np.zeros(n_groups * group, dtype=np.int64)
    pad[:n] = s
    maxes = pad.reshape(n_groups, group).max(axis=1)
    avg_max = np.float32(np.float32(maxes.sum()) / np.float32(n_groups))
    return float(avg), float(avg_max)


def sample_average(metrics) -> tuple[float, ...]:
    """Mean of per-sample metric tuples, term by term as metrics.h:37-47
    accumulates them (each divided by the count first), in float32."""
    rows = list(metrics)
    if not rows:
        return ()
    k = np.float32(len(rows))
    out = [np.float32(0)] * len(rows[0])
    for r in rows:
        out = [o + np.float32(v) / k for o, v in zip(out, r)]
    return tuple(float(o) for o in out)


def walk_stats(steps, ids) -> dict:
    """TAS, WAS over 64- and 32-lane groups, and the largest step count,
    over granted requests only (ids >= 0); groups with no grant are left
    out of WAS."""
    steps = np.asarray(steps, dtype=np.int64)
    g = np.asarray(ids) >= 0
    out = {"tas": float(steps[g].mean()) if g.any() else 0.0}
    for w in (64, 32):
        n_groups = -(-steps.size // w)
        s = np.zeros(n_groups * w, dtype=np.int64)
        gg = np.zeros(n_groups * w, dtype=bool)
        s[:steps.size] = steps
        gg[:g.size] = g
        s, gg = s.reshape(n_groups, w), gg.reshape(n_groups, w)
        rows = gg.any(axis=1)
        out[f"was{w}"] = (float(np.where(gg, s, 0).max(axis=1)[rows].mean())
                          if rows.any() else 0.0)
    out["max"] = int(steps[g].max()) if g.any() else 0
    return out
