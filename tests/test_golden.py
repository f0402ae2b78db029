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
