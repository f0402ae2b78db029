"""The serving loop through the C-ABI: its admission/finish schedule equals
the serial oracle's, every sequence ends with its target length and the
checksum of its own payloads (so no page was shared), each result is the
same alone and in the mix, graph-captured steps equal eager ones, and the
pool comes back to its starting state."""
import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu

SEED = 99


def _trace(n, rate, median, sigma, max_pages, seed=7):
    # arrivals per step = rate * step_s with step_s = 1 ms
    return pbm.serve_trace(n, rate, 1e-3, median, sigma, seed, max_pages)


def _run_to_end(server, n_steps, chunk=64):
    done = 0
    while done < n_steps:
        k = min(chunk, n_steps - done)
        server.run(k)
        done += k


def _check_results(res, keys, target, payload=True):
    fin = res["finish_step"] > 0
    assert fin.all()
    assert np.array_equal(res["length"], target)
    if payload:
        exp = np.array([orc.serve_checksum(SEED, int(k), int(t)) for k, t in zip(keys, target)],
                       dtype=np.uint64)
        assert np.array_equal(res["checksum"], exp)


@pytest.mark.parametrize("strategy", [pbm.RW_BM, pbm.COLLAB_RW_BM, pbm.CRW_BM, pbm.RW_MALLOC,
                                      pbm.SINGLE_CLOCK, pbm.DIST_CLOCK, pbm.LINKED_LIST,
                                      pbm.RW_BASIC, pbm.CRW_BASIC, pbm.COLLAB_RW_BASIC,
                                      pbm.OUROBOROS])
def test_serve_schedule_matches_oracle(gpu, strategy):
    T, n = 1 << 16, 2500
    log_case(f"serve schedule strategy={strategy}")
    arr, tgt = _trace(n, 8000, 60, 0.8, 2048)
    with pbm.Pool(T, 64, strategy=strategy, backing=True, seed=5) as pool:
        pool.prefill_bernoulli(0.5, 3)
        A = pool.free_count()
        adm, fin, tot = orc.serve_sim(A, 512, arr, tgt, 100000)
        assert (fin > 0).all()
        steps = int(fin.max()) + 2
        with pbm.Server(pool, 512, 2048, n, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
            srv.submit(arr, tgt)
            _run_to_end(srv, steps)
            res = srv.results()
            st = srv.stats()
        assert np.array_equal(res["admit_step"], adm)
        assert np.array_equal(res["finish_step"], fin)
        _check_results(res, np.arange(n), tgt)
        assert st["alloc_failures"] == 0 and st["payload_errors"] == 0
        assert st["allocs"] == tot[0] and st["frees"] == tot[1]
        assert st["admitted"] == st["finished"] == n and st["active"] == 0
        assert pool.free_count() == A
        assert np.array_equal(pool.snapshot(), orc.prefill(T, 0.5, 3))


def _run_loop_to_end(server, n_steps, chunk=64):
    done = 0
    while done < n_steps:
        k = min(chunk, n_steps - done)
        server.run_loop(k)
        done += k


@pytest.mark.parametrize("strategy", [pbm.RW_BM, pbm.COLLAB_RW_BM, pbm.CRW_BM, pbm.RW_MALLOC,
                                      pbm.SINGLE_CLOCK, pbm.DIST_CLOCK, pbm.LINKED_LIST,
                                      pbm.RW_BASIC, pbm.CRW_BASIC, pbm.COLLAB_RW_BASIC,
                                      pbm.OUROBOROS])
def test_serve_loop_schedule_matches_oracle(gpu, strategy):
    # the device-resident loop (many steps per kernel, phases ordered by
    # device-side counters) gives the oracle's schedule too
    T, n = 1 << 16, 2500
    log_case(f"serve loop schedule strategy={strategy}")
    arr, tgt = _trace(n, 8000, 60, 0.8, 2048)
    with pbm.Pool(T, 64, strategy=strategy, backing=True, seed=5) as pool:
        pool.prefill_bernoulli(0.5, 3)
        A = pool.free_count()
        adm, fin, tot = orc.serve_sim(A, 512, arr, tgt, 100000)
        steps = int(fin.max()) + 2
        with pbm.Server(pool, 512, 2048, n, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
            srv.submit(arr, tgt)
            _run_loop_to_end(srv, steps, chunk=97)
            res = srv.results()
            st = srv.stats()
            assert srv.loop_timeouts() == 0
        assert np.array_equal(res["admit_step"], adm)
        assert np.array_equal(res["finish_step"], fin)
        _check_results(res, np.arange(n), tgt)
        assert st["alloc_failures"] == 0 and st["payload_errors"] == 0
        assert st["allocs"] == tot[0] and st["frees"] == tot[1]
        assert st["admitted"] == st["finished"] == n and st["active"] == 0
        assert pool.free_count() == A
        assert np.array_equal(pool.snapshot(), orc.prefill(T, 0.5, 3))


@pytest.mark.parametrize("loop", [False, True])
def test_serve_many_sequences_finishing_in_one_step(gpu, loop):
    # 500 one-page sequences finish at the grow of step 1 and about 250
    # more at step 2: the release walks a done list far past the first 64
    # entries that each wave loads with the step word
    T = 1 << 16
    log_case(f"serve many finishing at once loop={loop}")
    rng = np.random.default_rng(11)
    arr = np.repeat(np.array([1, 2, 3], dtype=np.uint32), [500, 500, 500])
    tgt = np.concatenate([np.ones(500), rng.integers(1, 3, 500),
                          rng.integers(1, 9, 500)]).astype(np.uint32)
    n = arr.size
    with pbm.Pool(T, 64, strategy=pbm.RW_BM, backing=True, seed=5) as pool:
        pool.prefill_bernoulli(0.5, 3)
        A = pool.free_count()
        adm, fin, tot = orc.serve_sim(A, 1024, arr, tgt, 1000)
        assert (fin > 0).all()
        assert np.bincount(fin)[1] == 500 and np.bincount(fin)[2] > 64
        steps = int(fin.max()) + 2
        with pbm.Server(pool, 1024, 16, n, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
            srv.submit(arr, tgt)
            if loop:
                srv.run_loop(steps)
            else:
                srv.run(steps)
            res = srv.results()
            st = srv.stats()
        assert np.array_equal(res["admit_step"], adm)
        assert np.array_equal(res["finish_step"], fin)
        _check_results(res, np.arange(n), tgt)
        assert st["alloc_failures"] == 0 and st["payload_errors"] == 0
        assert st["allocs"] == tot[0] and st["frees"] == tot[1]
        assert st["admitted"] == st["finished"] == n and st["active"] == 0
        assert pool.free_count() == A
        assert np.array_equal(pool.snapshot(), orc.prefill(T, 0.5, 3))


@pytest.mark.parametrize("slots", [3000, 4096, 32768])
def test_serve_loop_equals_eager_steps(gpu, slots):
    # 3, 4 and 32 loop workgroups (the first with a partial last group of
    # slots, which the slot order deals round them too): loop runs, eager
    # steps and a graph of loop runs give the same schedule, lengths,
    # checksums and counters, with live submission block by block
    T, n, B = 1 << 20, 12000, 50
    log_case(f"serve loop vs eager slots={slots}")
    arr, tgt = _trace(n, 20000, 200, 0.8, 4096, seed=17)
    keys = (np.arange(n) * 13 + 1).astype(np.uint32)
    out = []
    for mode in ("eager", "loop", "loop_graph"):
        with pbm.Pool(T, 64, strategy=pbm.RW_BM, backing=True, seed=9) as pool:
            pool.prefill_bernoulli(0.2, 2)
            with pbm.Server(pool, slots, 4096, n, pbm.SERVE_PAYLOAD | pbm.SERVE_LATENCY,
                            seed=SEED, lat_capacity=1 << 22) as srv:
                sub = 0
                g = None
                for blk in range(60):
                    j = int(np.searchsorted(arr, (blk + 1) * B, side="right"))
                    if j > sub:
                        srv.submit(arr[sub:j], tgt[sub:j], keys[sub:j])
                        sub = j
                    if mode == "eager":
                        srv.run(B)
                    elif mode == "loop":
                        srv.run_loop(B)
                    else:
                        if g is None:
                            g = srv.graph_loop(B)
                        g.launch()
                if g is not None:
                    g.close()
                gpu.cuda.synchronize()
                res = srv.results()
                st = srv.stats()
                assert srv.loop_timeouts() == 0
                assert (st["payload_errors"], st["alloc_failures"]) == (0, 0), mode
                out.append((res, st))
    (a, sa) = out[0]
    for b, sb in out[1:]:
        for k in ("admit_step", "finish_step", "length", "checksum"):
            assert np.array_equal(a[k], b[k]), k
        assert sa == sb
    assert (a["finish_step"] > 0).sum() > 1000


def test_serve_alone_equals_in_mix(gpu):
    T, n = 1 << 16, 3000
    log_case("serve alone vs in mix")
    arr, tgt = _trace(n, 10000, 80, 0.8, 4096, seed=11)
    keys = (np.arange(n) * 7919 + 13).astype(np.uint32)
    with pbm.Pool(T, 64, strategy=pbm.RW_BM, backing=True, seed=8) as pool:
        pool.prefill_bernoulli(0.3, 4)
        with pbm.Server(pool, 1024, 4096, n, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
            srv.submit(arr, tgt, keys)
            _, fin, _ = orc.serve_sim(pool.free_count(), 1024, arr, tgt, 100000)
            _run_to_end(srv, int(fin.max()) + 2)
            mix = srv.results()
            assert srv.stats()["payload_errors"] == 0
        _check_results(mix, keys, tgt)
        pick = np.random.default_rng(0).choice(n, size=6, replace=False)
        for q in pick:
            with pbm.Server(pool, 4, 4096, 1, pbm.SERVE_PAYLOAD, seed=SEED) as one:
                one.submit([1], [tgt[q]], [keys[q]])
                one.run(int(tgt[q]) + 1)
                alone = one.results()
            assert alone["length"][0] == mix["length"][q]
            assert alone["checksum"][0] == mix["checksum"][q]


def test_serve_graph_steps_equal_eager_steps(gpu):
    T, n = 1 << 15, 1500
    log_case("serve graph vs eager")
    arr, tgt = _trace(n, 6000, 40, 0.8, 1024, seed=3)
    out = []
    for use_graph in (False, True):
        with pbm.Pool(T, 64, strategy=pbm.CRW_BM, backing=True, seed=2) as pool:
            pool.prefill_bernoulli(0.4, 6)
            with pbm.Server(pool, 256, 1024, n, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
                srv.submit(arr, tgt)
                if use_graph:
                    g = srv.graph(50)
                    for _ in range(40):          # 2000 steps, eager work between replays
                        g.launch()
                        gpu.cuda.synchronize()
                        srv.stats()
                    g.close()
                else:
                    _run_to_end(srv, 2000)
                res = srv.results()
                out.append((res, srv.stats()))
    (a, sa), (b, sb) = out
    for k in ("admit_step", "finish_step", "length", "checksum"):
        assert np.array_equal(a[k], b[k]), k
    assert sa == sb


def test_serve_live_submission_matches_the_oracle_schedule(gpu):
    # sequences are submitted block by block while earlier steps are still
    # queued or running (no host synchronisation in between): each block
    # is submitted before the steps of its arrivals are enqueued, so the
    # schedule is the one of the whole trace submitted up front
    T, n, B = 1 << 16, 3000, 25
    log_case("serve live submission")
    arr, tgt = _trace(n, 10000, 60, 0.8, 2048, seed=21)
    keys = (np.arange(n) * 31 + 5).astype(np.uint32)
    with pbm.Pool(T, 64, strategy=pbm.RW_BM, backing=True, seed=4) as pool:
        pool.prefill_bernoulli(0.5, 9)
        A = pool.free_count()
        adm, fin, tot = orc.serve_sim(A, 512, arr, tgt, 100000)
        steps = int(fin.max()) + 2
        with pbm.Server(pool, 512, 2048, n, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
            i = done = 0
            while done < steps:
                j = int(np.searchsorted(arr, done + B, side="right"))
                if j > i:
                    srv.submit(arr[i:j], tgt[i:j], keys[i:j])
                    i = j
                srv.run(B)
                done += B
            res = srv.results()
            st = srv.stats()
            waits = srv.wait_ticks()
        assert np.array_equal(res["admit_step"], adm)
        assert np.array_equal(res["finish_step"], fin)
        _check_results(res, keys, tgt)
        assert st["allocs"] == tot[0] and st["frees"] == tot[1] and st["payload_errors"] == 0
        assert pool.free_count() == A
        # admitted at their arrival step: no wait; queued: a positive wait,
        # longer for the ones that queued longer
        assert (waits != 0xFFFFFFFF).all()
        on_time = adm == arr
        assert (waits[on_time] == 0).all()
        late = adm > arr
        assert late.any() and (waits[late] > 0).all()
        d = adm[late] - arr[late]
        assert np.median(waits[late][d >= 10]) > np.median(waits[late][d <= 2])


def test_serve_sequences_submitted_after_their_arrival_step_wait_for_the_next_step(gpu):
    log_case("serve late submission")
    with pbm.Pool(1 << 12, 64, strategy=pbm.RW_BM, backing=True, seed=1) as pool:
        with pbm.Server(pool, 16, 64, 8, pbm.SERVE_PAYLOAD, seed=SEED) as srv:
            srv.submit([1, 2], [3, 3])
            srv.run(10)
            srv.submit([4, 4, 12], [2, 5, 1])   # steps 1..10 are already on the stream
            srv.run(10)
            res = srv.results(0, 5)
            waits = srv.wait_ticks(0, 5)
            hz = srv.clock_hz()
    assert res["admit_step"].tolist() == [1, 2, 11, 11, 12]
    assert res["finish_step"].tolist() == [3, 4, 12, 15, 12]
    # the late ones waited from their submission to step 11's start: a
    # positive span, and well under a second
    assert waits[0] == 0 and waits[1] == 0 and waits[4] == 0
    assert 0 < waits[2] < hz and 0 < waits[3] < hz


def test_serve_latency_samples_and_clock(gpu):
    log_case("serve latency")
    arr, tgt = _trace(400, 5000, 30, 0.5, 256, seed=1)
    with pbm.Pool(1 << 14, strategy=pbm.RW_BM, seed=1) as pool:
        with pbm.Server(pool, 128, 256, 400, pbm.SERVE_LATENCY, lat_capacity=1000) as srv:
            srv.submit(arr, tgt)
            srv.run(400)
            st = srv.stats()
            lat = srv.latency_ticks(5000)
            hz = srv.clock_hz()
    assert st["lat_samples"] == st["allocs"] > 1000
    assert lat.size == 1000 and (lat > 0).all()
    assert 1e6 <= hz <= 1e10


def test_serve_c5_pool_metadata_only(gpu):
    # 67,108,864 pages, no page memory: schedule and counts against the oracle
    T, n = 1 << 26, 20000
    log_case("serve C5-size pool")
    arr, tgt = _trace(n, 20000, 512, 0.8, 16384, seed=42)
    with pbm.Pool(T, 4096, strategy=pbm.RW_BM, seed=42) as pool:
        A = pool.free_count()
        adm, fin, tot = orc.serve_sim(A, 32768, arr, tgt, 1 << 20)
        steps = int(fin.max()) + 2
        with pbm.Server(pool, 32768, 16384, n, 0) as srv:
            srv.submit(arr, tgt)
            _run_to_end(srv, steps, chunk=256)
            res = srv.results()
            st = srv.stats()
        assert np.array_equal(res["admit_step"], adm)
        assert np.array_equal(res["finish_step"], fin)
        assert st["allocs"] == tot[0] == st["frees"] and st["alloc_failures"] == 0
        assert pool.free_count() == A and pool.count_used() == 0


@pytest.mark.parametrize("strategy", [pbm.RW_BM, pbm.COLLAB_RW_BM, pbm.LINKED_LIST, pbm.OUROBOROS])
@pytest.mark.parametrize("loop", [False, True])
def test_serve_destroy_gives_back_what_it_holds(gpu, strategy, loop):
    # a server destroyed with sequences running: their pages and the pages
    # promised to them go back, and the pool is as it was before
    T, n = 1 << 16, 3000
    log_case(f"serve destroy mid-run strategy={strategy} loop={loop}")
    arr, tgt = _trace(n, 8000, 60, 0.8, 2048, seed=23)
    with pbm.Pool(T, 64, strategy=strategy, seed=5) as pool:
        pool.prefill_bernoulli(0.5, 3)
        A = pool.free_count()
        before = pool.snapshot()
        srv = pbm.Server(pool, 512, 2048, n, 0, seed=SEED)
        srv.submit(arr, tgt)
        if loop:
            srv.run_loop(40)
        else:
            srv.run(40)
        st = srv.stats()
        assert st["active"] > 100 and st["reserved_pages"] > 0
        assert pool.free_count() == A - st["reserved_pages"] - (st["allocs"] - st["frees"])
        srv.close()
        assert pool.free_count() == A
        assert np.array_equal(pool.snapshot(), before)


def test_serve_rejects_bad_submissions(gpu):
    log_case("serve bad submissions")
    with pbm.Pool(1024, strategy=pbm.RW_BM) as pool:
        with pbm.Server(pool, 8, 16, 4) as srv:
            with pytest.raises(pbm.PbmError):
                srv.submit([1], [17])             # longer than max_pages
            with pytest.raises(pbm.PbmError):
                srv.submit([3, 2], [1, 1])        # arrivals decrease
            srv.submit([1, 1, 2, 2], [1, 2, 3, 4])
            with pytest.raises(pbm.PbmError):
                srv.submit([5], [1])              # beyond max_seqs
        with pytest.raises(pbm.PbmError):
            pbm.Server(pool, 8, 16, 4, pbm.SERVE_PAYLOAD)   # payload needs backing
