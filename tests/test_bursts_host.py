"""Host logic of the mallocRW bursts: the trace's live runs
(pbm.burst_live_runs) and the end-state checker (pbm.check_burst_end_state)
against the serial oracle, with planted violations; and the prediction the
GPU tests use for the benchmark's 4,096-burst trace."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc

PS = 4096
BASE = 1 << 40  # any page-aligned pool base


def _oracle_run(tr, T, fill, seed=7):
    before = orc.prefill(T, fill, 3)
    st = orc.State(before.copy(), T)
    n = int(tr["alloc_offset"][-1])
    page = np.full(n, -1, dtype=np.int64)
    off = np.zeros(n, dtype=np.int64)
    hdr = np.zeros(n, dtype=np.int64)
    for b in range(len(tr["alloc_offset"]) - 1):
        fi = tr["free_index"][tr["free_offset"][b]:tr["free_offset"][b + 1]]
        live = (page[fi] >= 0) & (hdr[fi] > 0)
        if live.any():
            st.free_n(page[fi][live], hdr[fi][live])
        a0, a1 = int(tr["alloc_offset"][b]), int(tr["alloc_offset"][b + 1])
        if a1 > a0:
            p, o, h, _ = st.malloc(tr["sizes"][a0:a1], PS, a0, seed)
            page[a0:a1], off[a0:a1], hdr[a0:a1] = p, o, h
    ptrs = np.where(page >= 0, BASE + page * PS + off + 4, 0).astype(np.uint64)
    return before, st, ptrs, hdr


def test_live_runs_match_the_serial_oracle():
    T = 1 << 16
    tr = pbm.burst_trace(200, 512, 1.0, 16, 65536, PS, 11)
    before, st, ptrs, hdr = _oracle_run(tr, T, 0.3)
    runs = pbm.burst_live_runs(tr, PS)
    # every live run's owner got a header equal to its page count
    for r, (n, members) in runs.items():
        assert hdr[r] == n
        assert (hdr[members[1:]] == 0).all() if len(members) > 1 else True
    end = pbm.check_burst_end_state(tr, ptrs, BASE, PS, before, st.bm)
    assert end["live_pages"] == orc.count_free(before, T) - st.free[0]
    assert end["failed_runs"] == 0 and end["live_runs"] == len(runs)


def test_end_state_checker_catches_planted_violations():
    T = 1 << 16
    tr = pbm.burst_trace(120, 512, 1.0, 16, 65536, PS, 13)
    before, st, ptrs, _ = _oracle_run(tr, T, 0.3)
    runs = sorted(pbm.burst_live_runs(tr, PS))
    # two live runs on the same pages
    bad = ptrs.copy()
    bad[runs[1]] = bad[runs[0]]
    with pytest.raises(AssertionError):
        pbm.check_burst_end_state(tr, bad, BASE, PS, before, st.bm)
    # a page leaked in the map
    leak = st.bm.copy()
    free_word = int(np.nonzero(leak != np.uint64(~np.uint64(0)))[0][0])
    w = int(leak[free_word])
    leak[free_word] = np.uint64(w | ((~w) & -(~w) & 0xFFFFFFFFFFFFFFFF))
    with pytest.raises(AssertionError):
        pbm.check_burst_end_state(tr, ptrs, BASE, PS, before, leak)
    # a run that does not start on a page
    mis = ptrs.copy()
    mis[runs[2]] = mis[runs[2]] + np.uint64(8)
    with pytest.raises(AssertionError):
        pbm.check_burst_end_state(tr, mis, BASE, PS, before, st.bm)


def test_c4_trace_oracle_fails_nothing_and_leaves_38875_pages():
    # the benchmark's trace on 262,144 empty pages: the figures the GPU
    # tests and the benchmark check their concurrent runs against
    T = 1 << 18
    tr = pbm.burst_trace(4096, 2048, 1.0, 16, 65536, PS, 7)
    n_a, n_f = int(tr["alloc_offset"][-1]), int(tr["free_offset"][-1])
    assert (n_a, n_f) == (6779874, 6761328)
    before = orc.cleared(T)
    _, st, ptrs, _ = _oracle_run(tr, T, 0.0)
    assert int((ptrs == 0).sum()) == 0
    end = pbm.check_burst_end_state(tr, ptrs, BASE, PS, before, st.bm)
    assert end["live_pages"] == 38875 == T - st.free[0]
