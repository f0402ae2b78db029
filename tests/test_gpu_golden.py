"""The device path, through the C-ABI, against the committed known-answer
vectors (tests/golden/golden.json): serial launches reproduce them bit for
bit, and the rank-batched launch does so at full concurrency."""
import json
from pathlib import Path

import numpy as np
import pytest

import pbm
from conftest import log_case

pytestmark = pytest.mark.gpu

G = json.loads((Path(__file__).resolve().parent / "golden" / "golden.json").read_text())


def _words(hexlist):
    return np.array([int(w, 16) for w in hexlist], dtype=np.uint64)


def _get(torch, pool, n, mode, req_base=0, flags=0):
    ids = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    steps = torch.zeros((n,), dtype=torch.int32, device="cuda")
    pool.get_pages(n, ids, steps, req_base, pbm.launch(mode=mode, flags=flags))
    torch.cuda.synchronize()
    return ids.cpu().numpy().tolist(), steps.cpu().numpy().view(np.uint32).tolist()


def test_golden_prefill(gpu):
    log_case("golden prefill")
    with pbm.Pool(1000, strategy=pbm.RW_BM) as pool:
        pool.prefill_bernoulli(0.5, 3)
        assert np.array_equal(pool.snapshot(), _words(G["prefill_1000_0.5_3"]))
    with pbm.Pool(4096, strategy=pbm.RW_BM) as pool:
        pool.prefill_bernoulli(0.9, 5)
        assert pool.free_count() == G["prefill_4096_0.9_5_free"]


def test_golden_rwbm(gpu):
    g = G["rwbm_4096_0.9"]
    log_case("golden rwbm")
    with pbm.Pool(g["T"], strategy=pbm.RW_BM, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_SERIAL, g["req_base"])
        assert ids == g["ids"] and steps == g["steps"]
        assert np.array_equal(pool.snapshot(), _words(g["map_after"]))


@pytest.mark.parametrize("order", [0, 1])
def test_golden_collab(gpu, order):
    g = G[f"collab_4096_0.9_order{order}"]
    log_case(f"golden collab order={order}")
    with pbm.Pool(g["T"], strategy=pbm.COLLAB_RW_BM, seed=g["seed"], order=order) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_SERIAL, g["req_base"])
        assert ids == g["ids"] and steps == g["steps"]


def test_golden_rank_concurrent(gpu):
    g = G["rank_20000_0.97"]
    log_case("golden rank")
    with pbm.Pool(g["T"], strategy=pbm.RW_BM) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK)
        assert ids == g["ids"] and steps == g["steps"]
        assert pool.free_count() == g["free_after"]


def test_golden_crw(gpu):
    g = G["crw_4096_0.5"]
    log_case("golden crw")
    with pbm.Pool(g["T"], strategy=pbm.CRW_BM, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_SERIAL)
        assert ids == g["ids"] and steps == g["steps"]


def test_golden_clocks(gpu):
    log_case("golden clocks")
    g = G["single_clock_2048_0.7"]
    with pbm.Pool(g["T"], strategy=pbm.SINGLE_CLOCK) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_SERIAL)
        assert ids == g["ids"] and steps == g["steps"]
    g = G["dist_clock_2048_0.7"]
    with pbm.Pool(g["T"], strategy=pbm.DIST_CLOCK, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_SERIAL)
        assert ids == g["ids"] and steps == g["steps"]


def test_golden_list(gpu):
    g = G["list_512_0.8"]
    log_case("golden list")
    with pbm.Pool(g["T"], strategy=pbm.LINKED_LIST) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        a, _ = _get(gpu, pool, len(g["first"]), pbm.MODE_SERIAL)
        assert a == g["first"]
        back = gpu.tensor(g["freed"], dtype=gpu.int32, device="cuda")
        pool.free_pages(len(g["freed"]), back, pbm.launch(mode=pbm.MODE_SERIAL))
        b, _ = _get(gpu, pool, len(g["second"]), pbm.MODE_SERIAL)
        assert b == g["second"]


def test_golden_getnpage(gpu):
    g = G["getnpage_4096_0.5"]
    log_case("golden getNPage")
    n = len(g["counts"])
    with pbm.Pool(g["T"], strategy=pbm.RW_MALLOC, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        counts = gpu.tensor(g["counts"], dtype=gpu.int32, device="cuda")
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.get_n_pages(n, counts, ids, steps, 0, pbm.launch(mode=pbm.MODE_SERIAL))
        gpu.cuda.synchronize()
        assert ids.cpu().numpy().tolist() == g["ids"]
        assert steps.cpu().numpy().view(np.uint32).tolist() == g["steps"]


def test_golden_malloc(gpu):
    g = G["malloc_4096_0.5"]
    log_case("golden mallocRW")
    n, ps = len(g["sizes"]), g["page_size"]
    with pbm.Pool(g["T"], ps, strategy=pbm.RW_MALLOC, backing=True, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        sizes = gpu.tensor(g["sizes"], dtype=gpu.int64, device="cuda")
        ptrs = gpu.zeros((n,), dtype=gpu.int64, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.malloc(n, sizes, ptrs, steps, 0, pbm.launch(mode=pbm.MODE_SERIAL))
        gpu.cuda.synchronize()
        p = ptrs.cpu().numpy().view(np.uint64).astype(np.int64)
        page = np.array(g["page"], dtype=np.int64)
        off = np.array(g["offset"], dtype=np.int64)
        ok = page >= 0
        assert np.array_equal(p > 0, ok)
        assert np.array_equal(p[ok] - 4 - pool.page_address(0), page[ok] * ps + off[ok])
        assert steps.cpu().numpy().view(np.uint32).tolist() == g["steps"]


@pytest.mark.parametrize("key,strategy", [("rwbasic_3000_0.8", pbm.RW_BASIC),
                                          ("crw_basic_3000_0.8", pbm.CRW_BASIC),
                                          ("collab_basic_3000_0.8", pbm.COLLAB_RW_BASIC)])
def test_golden_page_walks(gpu, key, strategy):
    g = G[key]
    log_case(f"golden {key}")
    with pbm.Pool(g["T"], strategy=strategy, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        ids, steps = _get(gpu, pool, g["n"], pbm.MODE_SERIAL)
        assert ids == g["ids"] and steps == g["steps"]


def test_golden_getxpage(gpu):
    g = G["getxpage_4096_0.4"]
    log_case("golden getXPage")
    n = len(g["counts"])
    with pbm.Pool(g["T"], strategy=pbm.RW_BM, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        counts = gpu.tensor(g["counts"], dtype=gpu.int32, device="cuda")
        ids = gpu.full((n,), -7, dtype=gpu.int32, device="cuda")
        steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
        pool.get_x_pages(n, counts, ids, steps, 0, pbm.launch(mode=pbm.MODE_SERIAL))
        gpu.cuda.synchronize()
        assert ids.cpu().numpy().tolist() == g["ids"]
        assert steps.cpu().numpy().view(np.uint32).tolist() == g["steps"]


@pytest.mark.parametrize("key,kind", [("kmalloc_collab_1024_0.5", pbm.KMALLOC_COLLAB),
                                      ("kmalloc_simple_1024_0.5", pbm.KMALLOC_SIMPLE)])
def test_golden_kmalloc_headers(gpu, key, kind):
    g = G[key]
    log_case(f"golden {key}")
    n, ps = len(g["sizes"]), g["page_size"]
    with pbm.Pool(g["T"], ps, strategy=pbm.RW_BM, backing=True, seed=g["seed"]) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        with pbm.KMalloc(pool, kind, max_threads=64) as km:
            sizes = gpu.tensor(g["sizes"], dtype=gpu.int32, device="cuda")
            ptrs = gpu.zeros((n,), dtype=gpu.int64, device="cuda")
            steps = gpu.zeros((n,), dtype=gpu.int32, device="cuda")
            km.malloc(n, sizes, ptrs, steps, 0, pbm.launch(mode=pbm.MODE_SERIAL))
            gpu.cuda.synchronize()
            p = ptrs.cpu().numpy().view(np.uint64)
            base = np.uint64(pool.page_address(0))
            off = np.where(p == 0, np.uint64(2 ** 64 - 1), p - base)
            assert [int(o) for o in off] == g["offsets"]
            assert steps.cpu().numpy().view(np.uint32).tolist() == g["steps"]
            mem = pool.read_pool(0, g["T"] * ps)
            for at, val in g["headers"]:
                assert int.from_bytes(mem[at:at + 8].tobytes(), "little") == val, at
            if kind == pbm.KMALLOC_SIMPLE:
                assert km.current()[0] == g["current"]


def test_golden_ouroboros(gpu):
    g = G["ouroboros_2048_0.6"]
    log_case("golden Ouroboros")
    with pbm.Pool(g["T"], g["page_size"], strategy=pbm.OUROBOROS) as pool:
        pool.prefill_bernoulli(g["fill"], g["prefill_seed"])
        a, sa = _get(gpu, pool, len(g["first"]), pbm.MODE_SERIAL)
        assert a == g["first"] and sa == g["first_steps"]
        back = gpu.tensor(g["freed"], dtype=gpu.int32, device="cuda")
        pool.free_pages(len(g["freed"]), back, pbm.launch(mode=pbm.MODE_SERIAL))
        b, sb = _get(gpu, pool, len(g["second"]), pbm.MODE_SERIAL)
        assert b == g["second"] and sb == g["second_steps"]
