"""The committed known-answer vectors (tests/golden/golden.json) against
the oracle, and their hand-checkable entries against first principles."""
import json
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
import make_golden  # noqa: E402

GOLDEN = json.loads((HERE / "golden" / "golden.json").read_text())


@pytest.fixture(scope="module")
def fresh():
    return make_golden.compute()


@pytest.mark.parametrize("key", sorted(k for k in GOLDEN if k != "about"))
def test_oracle_reproduces_golden(fresh, key):
    assert fresh[key] == GOLDEN[key]


def test_golden_lcg_is_the_ansi_c_recurrence():
    # x' = (1103515245 x + 12345) mod 2^31 from 42, by integer arithmetic
    x, out = 42, []
    for _ in range(16):
        x = (1103515245 * x + 12345) % (1 << 31)
        out.append(x)
    assert GOLDEN["lcg_from_42"] == out


def test_golden_rank_is_free_pages_in_order():
    g = GOLDEN["rank_20000_0.97"]
    ids = [i for i in g["ids"] if i >= 0]
    assert ids == sorted(ids) and len(set(ids)) == len(ids)
    assert all(s == 1 for s, i in zip(g["steps"], g["ids"]) if i >= 0)


def test_golden_rwbm_ids_are_unique_and_were_free():
    g = GOLDEN["rwbm_4096_0.9"]
    bm = [int(w, 16) for w in g["map_after"]]
    ids = [i for i in g["ids"] if i >= 0]
    assert len(set(ids)) == len(ids)
    assert all((bm[i // 64] >> (i % 64)) & 1 for i in ids)


def test_golden_crw_basic_takes_the_next_page_when_free():
    # after a grant the next request probes the page after the hint first
    # and counts one step when it is free (parallelPage.cu:200-211)
    g = GOLDEN["crw_basic_3000_0.8"]
    ids, steps = g["ids"], g["steps"]
    for k in range(1, len(ids)):
        if ids[k] == (ids[k - 1] + 1) % g["T"]:
            assert steps[k] == 1
    assert g["hint"] == ids[-1]


def test_golden_kmalloc_collab_headers_are_consistent():
    g = GOLDEN["kmalloc_collab_1024_0.5"]
    hdr = dict((a, v) for a, v in g["headers"])
    none = 2 ** 64 - 1
    for o, size in zip(g["offsets"], g["sizes"]):
        if o == none:
            continue
        h = hdr[o - 8]
        s = ((size + 7) & ~7) + 8
        if h >> 63:                                   # a chunk of its own
            assert (h >> 16) & 0xFFFF == s & 0xFFFF and (o - 8) % g["page_size"] == 0
        else:                                         # a member: offset, bytes, chunk
            assert (h >> 32) == s and (o - 8 - (h & 0xFFFFFFFF)) in hdr


def test_golden_ouroboros_is_fifo():
    # after the frees the queue holds the leftover carved pages, then the
    # freed pages in free order
    g = GOLDEN["ouroboros_2048_0.6"]
    freed = g["freed"]
    second = [i for i in g["second"] if i >= 0]
    pos = [second.index(f) for f in freed if f in second]
    assert pos == sorted(pos)
