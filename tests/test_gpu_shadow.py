"""The debug owner shadow through the C-ABI: owners follow every grant and
free of the id-returning calls exactly, clean runs report nothing, and
grants or frees recorded twice are reported with the page and both owners."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu


def _held_bits(words, T):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:T].astype(bool)


@pytest.mark.parametrize("kind", ["rank", "walk", "plain"])
def test_owners_follow_a_clean_cycle(gpu, kind):
    T, n = 1 << 20, 262144
    log_case(f"owner shadow, clean cycle, {kind}")
    lnch = {"rank": pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK),
            "walk": pbm.launch(pbm.MODE_BATCHED, waves_per_cu=8, walk_budget=16,
                               cluster_words=256),
            "plain": pbm.launch(mode=pbm.MODE_PLAIN)}[kind]
    with pbm.Pool(T, seed=42) as pool:
        pool.prefill_bernoulli(0.9, 42)
        before = pool.snapshot()
        pool.shadow_enable()
        own = pool.shadow_owners()
        held = _held_bits(before, T)
        assert np.array_equal(own == pbm.SHADOW_HELD, held)
        assert np.array_equal(own == pbm.SHADOW_FREE, ~held)
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        base = 5 << 32
        pool.get_pages(n, ids, None, base, lnch)
        fl = lnch if lnch.mode == pbm.MODE_BATCHED else pbm.launch(mode=pbm.MODE_PLAIN)
        i_np = ids.cpu().numpy()
        g = np.nonzero(i_np >= 0)[0]
        assert g.size == orc.count_free(before, T)
        own = pool.shadow_owners()
        assert np.array_equal(own[i_np[g]], base + g.astype(np.uint64))
        assert int((own == pbm.SHADOW_FREE).sum()) == 0
        rep = pool.shadow_report()
        assert rep["double_grants"] == 0 and rep["grant_page"] == -1
        pool.free_pages(n, ids, fl)
        own = pool.shadow_owners()
        assert np.array_equal(own == pbm.SHADOW_HELD, held)
        assert np.array_equal(own == pbm.SHADOW_FREE, ~held)
        rep = pool.shadow_report()
        assert rep["bad_frees"] == 0 and rep["free_page"] == -1
        assert np.array_equal(pool.snapshot(), before)


def test_a_second_grant_of_held_pages_is_reported_with_both_owners(gpu):
    T, n = 1 << 16, 5000
    log_case("owner shadow, planted double grant")
    with pbm.Pool(T, seed=3) as pool:
        pool.prefill_bernoulli(0.5, 7)
        pool.shadow_enable()
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        pool.get_pages(n, ids, None, 100, pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK))
        assert pool.shadow_report()["double_grants"] == 0
        # the same pages recorded as granted again, to requests 9000..
        pool.shadow_note(n, ids, None, 9000)
        rep = pool.shadow_report()
        i_np = ids.cpu().numpy()
        assert rep["double_grants"] == n
        k = int(np.nonzero(i_np == rep["grant_page"])[0][0])
        assert rep["grant_owner"] == 9000 + k and rep["grant_holder"] == 100 + k
        # a pre-filled page granted: its holder reads as held
        own = pool.shadow_owners()
        held = int(np.nonzero(own == pbm.SHADOW_HELD)[0][0])
        one = gpu.tensor([held], dtype=gpu.int32, device="cuda")
        pool.shadow_enable()          # a clean report, owners from the map
        pool.shadow_note(1, one, None, 77)
        rep = pool.shadow_report()
        assert (rep["double_grants"], rep["grant_page"], rep["grant_owner"],
                rep["grant_holder"]) == (1, held, 77, pbm.SHADOW_HELD)


def test_frees_of_free_pages_are_reported(gpu):
    T, n = 1 << 16, 3000
    log_case("owner shadow, planted double free")
    with pbm.Pool(T, seed=5) as pool:
        pool.shadow_enable()
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        pool.get_pages(n, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN))
        pool.free_pages(n, ids)
        assert pool.shadow_report()["bad_frees"] == 0
        pool.shadow_note(n, ids, None, 0, freeing=True)
        rep = pool.shadow_report()
        assert rep["bad_frees"] == n and rep["free_page"] in set(ids.cpu().numpy().tolist())
        # ids past the pool count too and are never touched
        bad = gpu.tensor([T, T + 5], dtype=gpu.int32, device="cuda")
        pool.shadow_note(2, bad, None, 0, freeing=True)
        assert pool.shadow_report()["bad_frees"] == n + 2


def test_multi_page_runs_and_reset(gpu):
    T, n = 1 << 18, 20000
    log_case("owner shadow, getNPage runs")
    counts = np.random.default_rng(4).integers(0, 17, size=n).astype(np.uint32)
    with pbm.Pool(T, strategy=pbm.RW_MALLOC, seed=8) as pool:
        pool.prefill_bernoulli(0.3, 2)
        before = pool.snapshot()
        pool.shadow_enable()
        d_counts = gpu.from_numpy(counts.view(np.int32)).cuda()
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        pool.get_n_pages(n, d_counts, ids, None, 40,
                         pbm.launch(mode=pbm.MODE_PLAIN, cluster_words=256))
        i_np = ids.cpu().numpy()
        own = pool.shadow_owners()
        expect = np.where(_held_bits(before, T), pbm.SHADOW_HELD, pbm.SHADOW_FREE).astype(np.uint64)
        for r in np.nonzero((i_np >= 0) & (counts > 0))[0]:
            expect[i_np[r]:i_np[r] + counts[r]] = 40 + r
        assert np.array_equal(own, expect)
        assert pool.shadow_report()["double_grants"] == 0
        pool.free_n_pages(n, ids, d_counts)
        assert pool.shadow_report()["bad_frees"] == 0
        assert np.array_equal(pool.snapshot(), before)
        pool.reset()
        assert (pool.shadow_owners() == pbm.SHADOW_FREE).all()
        pool.shadow_enable(False)
        with pytest.raises(pbm.PbmError, match="off"):
            pool.shadow_report()


def test_shadow_needs_a_bitmap_strategy(gpu):
    with pbm.Pool(4096, strategy=pbm.LINKED_LIST) as pool:
        with pytest.raises(pbm.PbmError, match="bitmap"):
            pool.shadow_enable()
