#!/usr/bin/env python3
"""Headline benchmark: page allocs+frees per second at 90% pool fill,
persistent-batched versus plain per-thread getPage, with average walk steps.

Workload (one step = one alloc phase + one free phase, inputs in HBM):
  1,048,576 pages x 4 KiB (page memory allocated), seeded Bernoulli
  pre-occupancy f = 0.9 (seed 42), 262,144 one-page requests, then every
  granted page freed.  About 104,858 pages are free, so about 157,286
  requests fail by admission; allocs+frees counts the grants and the frees.

A step is the alloc kernel then the free kernel, launched back to back on
torch's current stream; the step time is the mean over the timed steps
between two HIP events.  The same step captured as one HIP graph is timed
too (ms_per_step_graph).  Each phase kernel is also timed on its own, with
HIP events the library records at the kernel's own start and end
(hipExtLaunchKernelGGL): the roofline divides by those durations, which are
the ones a kernel trace reports.  Events around each launch are reported
beside them.  The batched launch (rank or
walk homes, window width, walk budget, waves per CU) comes from the
library's autotuner; the best walk-homes launch is timed too.  The plain
variant is the reference's RandomWalkBasic_BM getPage/freePage, one thread
per request.  The same step under every other one-page policy (the other
walks, the clocks, the FIFO list and the Ouroboros page queue) is timed with
plain launches as comparators, and the kernel-malloc managers run the
reference's malloc benchmark.
The CPU baseline is the serial oracle running the reference's
RandomWalkBasic_BM policy on the identical pre-fill and request stream, on
one host core.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "parallel-db-buffer-management_amd"))

from metrics import walk_stats  # noqa: E402

METRIC = "page allocs+frees/sec at 90% fill, persistent-batched vs plain; avg walk steps"
T_PAGES, PAGE_SIZE, FILL, SEED, N_REQ = 1 << 20, 4096, 0.9, 42, 262144
HBM_PEAK_GBS = 8000.0
GRAPH_STEPS = 32  # alloc+free steps captured in one replayed graph


_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """One progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def _rwbm_tas(T: int, A: int, N: int):
    """The reference's RWBM_TAS closed form for 64-bit words (None when it
    does not apply: every free page granted makes its last term infinite)."""
    import models
    try:
        return models.RWBM_TAS(T, A, min(N, A - 1), W=64)
    except (ValueError, ZeroDivisionError, OverflowError):
        return None


def traffic_of(leg: str, kernel: str):
    """HBM bytes per launch of `kernel` in `leg`, from the committed
    rocprofv3 FETCH_SIZE and WRITE_SIZE runs of these workloads (each
    scaled by what the counter reads for a 1 GiB device copy); None when
    that kernel was not profiled."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    for k in json.loads(f.read_text())["kernels"]:
        if k["leg"] == leg and k["kernel"] == kernel:
            return k["hbm_bytes"]
    return None


def time_region(torch, fn, steps: int, warmup: int, stream) -> float:
    """Mean ms of fn() over `steps` calls launched back to back between two
    events on `stream`, after `warmup` untimed calls."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def time_phases(torch, pbm, pool, n, ids, gl, fl, steps: int, stream) -> tuple[dict, dict]:
    """Mean ms of each phase kernel over `steps` eager alloc+free steps
    launched back to back.  First: each kernel's own duration, from HIP
    events the library records at the kernel's start and end
    (pbm_time_next_kernel, hipExtLaunchKernelGGL), the figure a kernel
    trace reports.  Second: events recorded on the stream around every
    launch, which add the dispatch of the launch and of the markers."""
    ev = [[pbm.KernelEvent() for _ in range(4)] for _ in range(steps)]
    for e in ev:
        pool.time_next_kernel(e[0], e[1])
        pool.get_pages(n, ids, None, 0, gl, stream)
        pool.time_next_kernel(e[2], e[3])
        pool.free_pages(n, ids, fl, stream)
    torch.cuda.synchronize()
    own = {"alloc": float(np.mean([e[0].elapsed_ms(e[1]) for e in ev])),
           "free": float(np.mean([e[2].elapsed_ms(e[3]) for e in ev]))}
    for e in ev:
        for x in e:
            x.close()
    br = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for e in br:
        e[0].record(stream)
        pool.get_pages(n, ids, None, 0, gl, stream)
        e[1].record(stream)
        pool.free_pages(n, ids, fl, stream)
        e[2].record(stream)
    torch.cuda.synchronize()
    bracket = {"alloc": float(np.mean([e[0].elapsed_time(e[1]) for e in br])),
               "free": float(np.mean([e[1].elapsed_time(e[2]) for e in br]))}
    return own, bracket


def serving_leg(torch, pbm, stream, warm_steps: int, timed_steps: int, block: int = 50) -> dict:
    """Serving loop on 67,108,864 x 4 KiB pages (page map only): Poisson
    arrivals at 20,000/s binned into 1 ms steps, lognormal lengths (median
    512 pages, sigma 0.8, capped at 16,384), one page per sequence per step,
    admission by free-page count.  Arrivals are live: before each block of
    `block` steps is put on the stream, the sequences arriving in it are
    submitted (stream-ordered, no host synchronisation), while earlier
    blocks still run.  Measured: ops/s over `timed_steps` eager steps and
    over as many steps replayed from a captured `block`-step graph and as
    many run by the device-resident loop (`block` steps per launch), the
    per-allocation latency of the device getPage calls and the
    arrival-to-admission wait of the sequences admitted in the eager region
    and in the loop region, and each step kernel's own duration."""
    T, n_steps = 1 << 26, warm_steps + 3 * timed_steps + 200 + 2 * block
    n = int(20000 * n_steps * 1e-3 * 1.2) + 64
    arr, tgt = pbm.serve_trace(n, 20000.0, 1e-3, 512.0, 0.8, 7, 16384)
    pool = pbm.Pool(T, 4096, strategy=pbm.RW_BM, backing=False, seed=7)
    srv = pbm.Server(pool, 32768, 16384, n, pbm.SERVE_LATENCY, lat_capacity=1 << 26)
    state = {"sub": 0, "done": 0}

    def submit_until(step_hi: int) -> None:
        j = int(np.searchsorted(arr, step_hi, side="right"))
        if j > state["sub"]:
            srv.submit(arr[state["sub"]:j], tgt[state["sub"]:j], stream=stream)
            state["sub"] = j

    def steps(k: int, graph=None, loop=False) -> None:
        for _ in range(k // block):
            submit_until(state["done"] + block)
            if graph is not None:
                graph.launch(stream)
            elif loop:
                srv.run_loop(block, stream)
            else:
                srv.run(block, stream)
            state["done"] += block

    steps(warm_steps)
    torch.cuda.synchronize()
    progress(f"serving: {warm_steps} warm-up steps done")
    s0 = srv.stats(stream)
    c0 = pool.counters(stream)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    steps(timed_steps)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    s1 = srv.stats(stream)
    progress(f"serving: {timed_steps} timed eager steps, {ms:.1f} ms")
    g = srv.graph(block)
    t2 = torch.cuda.Event(enable_timing=True)
    t3 = torch.cuda.Event(enable_timing=True)
    t2.record(stream)
    steps(timed_steps, graph=g)
    t3.record(stream)
    torch.cuda.synchronize()
    ms_graph = t2.elapsed_time(t3)
    g.close()
    s2 = srv.stats(stream)
    progress(f"serving: {timed_steps} steps replayed from {block}-step graphs, {ms_graph:.1f} ms")
    # the device-resident loop: `block` steps per kernel launch
    t4 = torch.cuda.Event(enable_timing=True)
    t5 = torch.cuda.Event(enable_timing=True)
    t4.record(stream)
    steps(timed_steps, loop=True)
    t5.record(stream)
    torch.cuda.synchronize()
    ms_loop = t4.elapsed_time(t5)
    s2l = srv.stats(stream)
    loop_timeouts = srv.loop_timeouts(stream)
    progress(f"serving: {timed_steps} steps in {block}-step device-resident loop launches, "
             f"{ms_loop:.1f} ms")
    # each kernel's own duration over 200 more steps, with device counters
    submit_until(state["done"] + 200)
    pool.clear_counters(stream)
    s3 = srv.stats(stream)
    kms = srv.run_timed(200, stream)
    state["done"] += 200
    s4 = srv.stats(stream)
    c4 = pool.counters(stream)
    lat_all = srv.latency_ticks(1 << 26, stream)
    lat = lat_all[s0["lat_samples"]:s1["lat_samples"]]
    lat_loop = lat_all[s2["lat_samples"]:s2l["lat_samples"]]
    res = srv.results(0, state["sub"], stream)
    waits = srv.wait_ticks(0, state["sub"], stream)
    hz = srv.clock_hz()
    in_eager = (res["admit_step"] > s0["step"]) & (res["admit_step"] <= s1["step"])
    w = waits[in_eager].astype(np.float64)
    in_loop = (res["admit_step"] > s2["step"]) & (res["admit_step"] <= s2l["step"])
    w_loop = waits[in_loop].astype(np.float64)
    w_steps = (res["admit_step"][in_eager].astype(np.int64) - arr[:state["sub"]][in_eager])
    ops = (s1["allocs"] - s0["allocs"]) + (s1["frees"] - s0["frees"])
    ops_graph = (s2["allocs"] - s1["allocs"]) + (s2["frees"] - s1["frees"])
    ops_loop = (s2l["allocs"] - s2["allocs"]) + (s2l["frees"] - s2["frees"])
    # roofline of the grow kernel: every slot's state word is scanned (4 B);
    # a grant reads the slot's length and target (8 B), reads map words
    # (8 B per probe), claims with read-modify-writes (16 B each), writes
    # the page id, the new length and a latency sample (12 B)
    grants = s4["allocs"] - s3["allocs"]
    frees = s4["frees"] - s3["frees"]
    grow_atomics = c4["atomics"] - c4["frees"]
    grow_bytes = (4 * 32768 * 200 + 8 * c4["probes"] + 16 * grow_atomics + 20 * grants) / 200
    achieved = grow_bytes / (kms["grow"] * 1e-3) / 1e9
    out = {
        "workload": "67,108,864 x 4 KiB pages (page map only), Poisson 20,000 seq/s in 1 ms "
                    "steps, lognormal length median 512 sigma 0.8, +1 page per step, RW_BM "
                    f"device getPage per slot; arrivals submitted live, {block} steps ahead",
        "warm_steps": warm_steps, "timed_steps": timed_steps, "ms_per_step": ms / timed_steps,
        "ops_per_s": ops / (ms * 1e-3), "allocs": s1["allocs"] - s0["allocs"],
        "frees": s1["frees"] - s0["frees"],
        "graph": {"steps_per_graph": block, "ms_per_step": ms_graph / timed_steps,
                  "ops_per_s": ops_graph / (ms_graph * 1e-3)},
        "loop": {"steps_per_launch": block, "ms_per_step": ms_loop / timed_steps,
                 "ops_per_s": ops_loop / (ms_loop * 1e-3), "wait_timeouts": loop_timeouts,
                 "p50_alloc_us": float(np.percentile(lat_loop, 50)) * 1e6 / hz if lat_loop.size else None,
                 "p99_alloc_us": float(np.percentile(lat_loop, 99)) * 1e6 / hz if lat_loop.size else None,
                 "latency_samples": int(lat_loop.size), "admitted": int(in_loop.sum()),
                 "p50_wait_us": float(np.percentile(w_loop, 50)) * 1e6 / hz if w_loop.size else None,
                 "p99_wait_us": float(np.percentile(w_loop, 99)) * 1e6 / hz if w_loop.size else None},
        "fastest": min(({"mode": "eager", "ms_per_step": ms / timed_steps},
                        {"mode": "graph", "ms_per_step": ms_graph / timed_steps},
                        {"mode": "loop", "ms_per_step": ms_loop / timed_steps}),
                       key=lambda r: r["ms_per_step"]),
        "kernel_ms": kms,
        "active_at_end": s4["active"], "pages_in_use_at_end": pool.count_used(stream),
        "p50_alloc_us": float(np.percentile(lat, 50)) * 1e6 / hz if lat.size else None,
        "p99_alloc_us": float(np.percentile(lat, 99)) * 1e6 / hz if lat.size else None,
        "latency_samples": int(lat.size), "clock_hz": hz,
        "admitted_timed": int(in_eager.sum()),
        "p50_wait_us": float(np.percentile(w, 50)) * 1e6 / hz if w.size else None,
        "p99_wait_us": float(np.percentile(w, 99)) * 1e6 / hz if w.size else None,
        "mean_wait_steps": float(w_steps.mean()) if w.size else None,
        "alloc_failures": s4["alloc_failures"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_of("serving", "k_serve_grow<0>"),
                     "kernel": "k_serve_grow<0>", "bytes_per_launch": grow_bytes,
                     "launch_ms": kms["grow"], "grants_per_launch": grants / 200,
                     "frees_per_launch": frees / 200},
        "_cpu_sample_map": pool.snapshot(stream), "_cpu_sample_grants": int(grants / 200 * 100),
    }
    srv.close()
    pool.close()
    return out


def bursts_leg(torch, pbm, stream, n_bursts: int, tune_bursts: int, persistent: bool) -> dict:
    """Bursts of mallocRW/freeRW: 262,144 x 4 KiB pages (backed), empty at
    the start (a page-level random pre-fill leaves no runs for multi-page
    requests); n_bursts bursts with lognormal request counts
    (median 2,048, sigma 1.0), each request a free of a live allocation or
    a mallocRW of log-uniform 16 B .. 64 KiB with probability 1/2.  The
    autotuner picks the launch (phase kernels or the persistent kernel,
    waves per CU, walk budget, getNPage window) on the first tune_bursts
    bursts; the whole trace is then timed from the same starting state,
    replayed from HIP graphs, and its end state checked against the
    trace: the live runs disjoint and in use, the map the start plus
    exactly those runs, no wait given up."""
    T, PS = 1 << 18, 4096
    tr = pbm.burst_trace(n_bursts, 2048.0, 1.0, 16, 65536, PS, 7)
    pool = pbm.Pool(T, PS, strategy=pbm.RW_MALLOC, backing=True, seed=7)
    start = pool.snapshot(stream)
    bs = pbm.Bursts(pool, tr)
    progress(f"bursts: trace of {n_bursts} bursts built; autotuning on {tune_bursts}")
    flags = pbm.TUNE_QUICK | (0 if persistent else pbm.TUNE_NO_PERSISTENT)
    tune = bs.autotune(min(tune_bursts, n_bursts), flags, stream)
    progress(f"bursts: tuned {tune['waves_per_cu']} waves/CU, budget {tune['walk_budget']}, "
             f"window {tune['cluster_words']} words, persistent {tune['persistent']}, "
             f"{tune['ms']:.2f} ms")
    best = tune["launch"]
    pool.restore(start, stream)
    bs.reset(stream)
    stride = n_bursts if tune["persistent"] else 256
    graphs = [bs.graph(first, min(stride, n_bursts - first), best)
              for first in range(0, n_bursts, stride)]
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for g in graphs:
        g.launch(stream)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    progress(f"bursts: whole trace {ms:.1f} ms")
    for g in graphs:
        g.close()
    ptrs = bs.ptrs(stream=stream)
    timeouts = bs.wait_timeouts(stream)
    end = pbm.check_burst_end_state(tr, ptrs, pool.page_address(0), PS, start,
                                    pool.snapshot(stream))
    n_malloc, n_free = int(tr["alloc_offset"][-1]), int(tr["free_offset"][-1])
    n_req = n_malloc + n_free
    # the same run once more, eager, with device counters and HIP events at
    # the main kernel's own start and end (the persistent kernel)
    pool.restore(start, stream)
    bs.reset(stream)
    pool.clear_counters(stream)
    counted = pbm.Launch(best.mode, best.block, best.waves_per_cu, best.walk_budget,
                         best.cluster_words, best.flags & ~pbm.LAUNCH_NO_COUNTERS)
    e0, e1 = pbm.KernelEvent(), pbm.KernelEvent()
    pool.time_next_kernel(e0, e1)
    bs.run(0, n_bursts, counted, stream)
    torch.cuda.synchronize()
    kernel_ms = e0.elapsed_ms(e1) if tune["persistent"] else None
    e0.close()
    e1.close()
    cnt = pool.counters(stream)
    end2 = pbm.check_burst_end_state(tr, bs.ptrs(stream=stream), pool.page_address(0), PS, start,
                                     pool.snapshot(stream))
    timeouts += bs.wait_timeouts(stream)
    # algorithmic bytes: a mallocRW reads its map word (8 B), claims with one
    # read-modify-write (16 B), writes its 4-byte header and its 8-byte
    # pointer; a freeRW reads the pointer and the header and releases with
    # one read-modify-write
    alg = n_malloc * (8 + 16 + 4 + 8) + n_free * (8 + 4 + 16)
    dom_ms = kernel_ms if kernel_ms else ms
    achieved = alg / (dom_ms * 1e-3) / 1e9
    traffic = traffic_of("bursts", "k_bursts_persistent") if tune["persistent"] else None
    out = {
        "workload": f"{n_bursts} bursts, lognormal count median 2048 sigma 1.0, 50/50 "
                    "mallocRW (16 B..64 KiB log-uniform) / freeRW, 262,144 x 4 KiB pages "
                    "empty at the start",
        "requests": n_req, "mallocs": n_malloc, "frees": n_free, "ms": ms,
        "us_per_burst": 1e3 * ms / n_bursts, "requests_per_s": n_req / (ms * 1e-3),
        "kernel_ms": kernel_ms,
        "malloc_failures": int((ptrs == 0).sum()),
        "launch": {k: tune[k] for k in ("waves_per_cu", "walk_budget", "cluster_words",
                                        "persistent")},
        "tune_ms_first_bursts": tune["ms"], "tune_ms_longer_run": tune["ms_longer_run"],
        "tune_candidates": tune["candidates"],
        "wait_timeouts": timeouts,
        "live_runs_at_end": end["live_runs"], "pages_live_at_end": end["live_pages"],
        # the serial oracle serves every request of this trace
        "verified": bool(timeouts == 0 and int((ptrs == 0).sum()) == 0
                         and end["failed_runs"] == 0
                         and end2["live_pages"] == end["live_pages"]),
        "counters": cnt,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_bursts_persistent" if kernel_ms else "k_burst_alloc+k_burst_free",
                     "bytes_per_launch": alg, "launch_ms": dom_ms,
                     "per_malloc": 36, "per_free": 28,
                     "atomic": {"achieved_ops": cnt["atomics"] / (dom_ms * 1e-3),
                                "unit": "u64 RMW/s on the page map"}},
    }
    out["_cpu_trace"] = tr
    bs.close()
    pool.close()
    return out


def comparator_leg(torch, pbm, stream, strategy: int, steps: int) -> dict:
    """The alloc+free step of the headline workload under another one-page
    policy, plain launches (one thread per request, wave-collective
    policies collective over each wave): the other bitmap walks, the clocks,
    the FIFO list (parallelPage.cu:342-418) and the Ouroboros page queue
    (Ouroboros.cuh:22-62)."""
    pool = pbm.Pool(T_PAGES, PAGE_SIZE, strategy=strategy, seed=SEED)
    pool.prefill_bernoulli(FILL, SEED, stream)
    before = pool.snapshot(stream)
    granted = min(N_REQ, pool.free_count(stream))
    ids = torch.full((N_REQ,), -1, dtype=torch.int32, device="cuda")
    lp = pbm.launch(mode=pbm.MODE_PLAIN, flags=pbm.LAUNCH_NO_COUNTERS)

    def step():
        pool.get_pages(N_REQ, ids, None, 0, lp, stream)
        pool.free_pages(N_REQ, ids, lp, stream)

    ms = time_region(torch, step, steps, 2, stream)
    ok = bool(np.array_equal(pool.snapshot(stream), before)) and pool.counters(stream)["errors"] == 0
    pool.close()
    progress(f"comparator {pbm.STRATEGY_NAMES[strategy]} measured")
    return {"value": 2 * granted / (ms * 1e-3), "ms_per_step": ms, "steps": steps,
            "verified": ok}


def policy_rows(torch, pbm, stream, n_pages: int, n_req: int, fill: float, steps: int) -> dict:
    """One configuration under plain RandomWalkBasic_BM and wave64
    CollabRW_BM getPage (one thread per request) and the rank-batched
    launch: n_req requests then frees on n_pages x 4 KiB pages with a
    Bernoulli pre-fill f (seed 42).  Per policy: walk steps per grant (TAS,
    WAS over 64- and 32-lane groups, max), failures against N - A,
    allocs+frees/s; verified = unique grants (the map returns exactly) and
    failures = max(0, N - A)."""
    row = {}
    for name, strategy, lnch in (
            ("plain_rw_bm", pbm.RW_BM, pbm.launch(mode=pbm.MODE_PLAIN)),
            ("plain_collab_rw_bm", pbm.COLLAB_RW_BM, pbm.launch(mode=pbm.MODE_PLAIN)),
            ("batched_rank", pbm.RW_BM, pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK))):
        pool = pbm.Pool(n_pages, PAGE_SIZE, strategy=strategy, seed=SEED)
        pool.prefill_bernoulli(fill, SEED, stream)
        before = pool.snapshot(stream)
        A = pool.free_count(stream)
        ids = torch.full((n_req,), -1, dtype=torch.int32, device="cuda")
        stp = torch.zeros((n_req,), dtype=torch.int32, device="cuda")
        fl = lnch if lnch.mode == pbm.MODE_BATCHED else pbm.launch(mode=pbm.MODE_PLAIN)
        pool.get_pages(n_req, ids, stp, 0, lnch, stream)
        torch.cuda.synchronize()
        i_np = ids.cpu().numpy()
        ws = walk_stats(stp.cpu().numpy().view(np.uint32), i_np)
        yields = None
        if strategy == pbm.COLLAB_RW_BM:
            # collectWarpData on a second cycle: free bits each wave's
            # window showed per step, against the pool's share 4096 A / T
            pool.free_pages(n_req, ids, fl, stream)
            waves, K = (n_req + 63) // 64, 8
            y = torch.zeros((waves * K,), dtype=torch.int32, device="cuda")
            ns = torch.zeros((waves,), dtype=torch.int32, device="cuda")
            pool.collect_warp_data(n_req, ids, y, K, ns, 0, lnch, stream)
            torch.cuda.synchronize()
            y_np = y.cpu().numpy().reshape(waves, K)
            ns_np = ns.cpu().numpy()
            ran = ns_np > 0
            yields = {"first_step_mean": float(y_np[ran, 0].mean()) if ran.any() else 0.0,
                      "window_share": 4096.0 * A / n_pages, "waves": int(ran.sum())}
        fails = int((i_np < 0).sum())
        pool.free_pages(n_req, ids, fl, stream)
        quiet = pbm.Launch(lnch.mode, lnch.block, lnch.waves_per_cu, lnch.walk_budget,
                           lnch.cluster_words, lnch.flags | pbm.LAUNCH_NO_COUNTERS)
        qf = pbm.Launch(fl.mode, fl.block, fl.waves_per_cu, fl.walk_budget, fl.cluster_words,
                        fl.flags | pbm.LAUNCH_NO_COUNTERS)

        def step():
            pool.get_pages(n_req, ids, None, 0, quiet, stream)
            pool.free_pages(n_req, ids, qf, stream)

        ms = time_region(torch, step, steps, 1, stream)
        own = None
        if lnch.mode == pbm.MODE_BATCHED:
            # on a small map the two kernels are shorter than the two host
            # launches; their own durations show the device's part
            own, _ = time_phases(torch, pbm, pool, n_req, ids, quiet, qf, steps, stream)
        ok = (bool(np.array_equal(pool.snapshot(stream), before))
              and fails == max(0, n_req - A) and pool.counters(stream)["errors"] == 0)
        row[name] = {"walk_steps": ws, "failures": fails, "expected_failures": max(0, n_req - A),
                     "value": 2 * min(n_req, A) / (ms * 1e-3), "ms_per_step": ms, "verified": ok}
        if own:
            row[name]["phases_ms"] = own
        if yields:
            row[name]["warp_yield"] = yields
        pool.close()
    return row


def fill_sweep_leg(torch, pbm, stream, fills, steps: int) -> dict:
    """The 1,048,576-page, 262,144-request configuration at each pre-fill f."""
    out = {}
    for f in fills:
        out[str(f)] = policy_rows(torch, pbm, stream, T_PAGES, N_REQ, f, steps)
        progress(f"fill sweep: f={f} measured")
    return out


def kmalloc_leg(torch, pbm, stream, n: int, steps: int) -> dict:
    """KernelMallocCollab malloc + free and KernelMallocSimple malloc of the
    reference's malloc benchmark (unitTests/unitTestMalloc1.cu:14-41):
    n requests, request t asking (1103515245 t + 12345) % 8188 + 4 bytes,
    on 262,144 x 4 KiB pages (page memory allocated), empty at the start.
    A wave serves 64 consecutive requests (the reference's 32-thread blocks
    are one warp each)."""
    T = 1 << 18
    t = np.arange(n, dtype=np.int64)
    sizes = ((1103515245 * t + 12345) % 8188 + 4).astype(np.uint32)
    d_sizes = torch.from_numpy(sizes.view(np.int32)).cuda()
    ptrs = torch.zeros((n,), dtype=torch.int64, device="cuda")
    lp = pbm.launch(mode=pbm.MODE_PLAIN, flags=pbm.LAUNCH_NO_COUNTERS)
    out = {"requests": n, "sizes": "(1103515245 t + 12345) % 8188 + 4 bytes", "pages": T}
    with pbm.Pool(T, 4096, strategy=pbm.RW_BM, backing=True, seed=SEED) as pool:
        with pbm.KMalloc(pool, pbm.KMALLOC_COLLAB) as km:
            def cycle():
                km.malloc(n, d_sizes, ptrs, None, 0, lp, stream)
                km.free(n, ptrs, lp, stream)
            ms = time_region(torch, cycle, steps, 2, stream)
            ok = pool.count_used(stream) == 0 and pool.counters(stream)["errors"] == 0
            out["collab"] = {"value": 2 * n / (ms * 1e-3), "unit": "mallocs+frees/s",
                             "ms_per_cycle": ms, "verified": bool(ok)}
        with pbm.KMalloc(pool, pbm.KMALLOC_SIMPLE, max_threads=n) as km:
            total = 0.0
            for k in range(steps + 2):
                km.reset(stream)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                km.malloc(n, d_sizes, ptrs, None, 0, lp, stream)
                e1.record(stream)
                torch.cuda.synchronize()
                if k >= 2:
                    total += e0.elapsed_time(e1)
            p = ptrs.cpu().numpy()
            ms = total / steps
            out["simple"] = {"value": n / (ms * 1e-3), "unit": "mallocs/s", "ms_per_call": ms,
                             "verified": bool((p != 0).all())}
            km.reset(stream)
    progress("kernel malloc measured")
    return out


def cache_leg(torch, pbm, stream, n_refs: int, batch: int) -> dict:
    """Buffer cache: 16,777,216 x 4 KiB frames (page map only) allocated
    with CRW_BM, 4,194,304 logical pages held in 1,048,576 slots (the
    resident cap), CLOCK replacement, a Zipf(1.1) trace of n_refs
    references from seed 7, taken `batch` at a time.  The trace is timed
    whole, then run again on a fresh cache with each kernel's own duration
    recorded (pbm_cache_run_timed)."""
    T, L, C = 1 << 24, 1 << 22, 1 << 20
    refs = torch.empty(n_refs, dtype=torch.int32, device="cuda")
    pbm.zipf_trace(refs, L, 1.1, 7, stream=stream)
    pool = pbm.Pool(T, 4096, strategy=pbm.CRW_BM, backing=False, seed=7)
    cache = pbm.Cache(pool, L, C, batch)
    torch.cuda.synchronize()
    progress(f"cache: {n_refs} Zipf references generated")
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    cache.run(refs, stream=stream)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    st = cache.stats(stream)
    cache.close()
    pool.close()
    progress(f"cache: {ms:.1f} ms")
    pool = pbm.Pool(T, 4096, strategy=pbm.CRW_BM, backing=False, seed=7)
    cache = pbm.Cache(pool, L, C, batch)
    kms = cache.run_timed(refs, stream=stream)
    st2 = cache.stats(stream)
    cache.close()
    pool.close()
    n_batches = (n_refs + batch - 1) // batch
    m = st["misses"] / n_batches       # misses per batch
    v = st["evictions"] / n_batches    # victims per batch
    # algorithmic bytes per launch (one batch): hits reads each reference
    # and its page's slot (8 B) and raises the first-miss tag of each miss
    # (16 B read-modify-write); plan reads each reference and its tag
    # (12 B), writes each miss (4 B) and evicts each victim (owner, page and
    # frame words: 16 B); load takes a frame per miss (8 B probe, 16 B
    # claim) and writes the slot, its page and bit (16 B)
    bytes_k = {"hits": 8 * batch + 16 * m, "plan": 12 * batch + 4 * m + 16 * v,
               "load": 40 * m}
    dom = max(kms, key=kms.get)
    achieved = bytes_k[dom] / (kms[dom] * 1e-3) / 1e9
    out = {
        "workload": f"{n_refs} Zipf(1.1) references (seed 7) over 4,194,304 logical pages, "
                    "1,048,576 slots, CLOCK replacement, frames from CRW_BM on 16,777,216 "
                    f"x 4 KiB pages, batches of {batch}",
        "ms": ms, "refs_per_s": n_refs / (ms * 1e-3), "hits": st["hits"], "misses": st["misses"],
        "hit_rate": st["hits"] / n_refs, "evictions": st["evictions"],
        "resident": st["resident"], "alloc_failures": st["alloc_failures"],
        "scan_timeouts": st["scan_timeouts"] + st2["scan_timeouts"],
        "kernel_ms": kms, "verified": bool(st2["hits"] == st["hits"]
                                           and st2["evictions"] == st["evictions"]),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_of("cache", {"hits": "k_cache_hits", "plan": "k_cache_plan",
                                                     "load": "k_cache_load<2>"}[dom]),
                     "kernel": {"hits": "k_cache_hits", "plan": "k_cache_plan",
                                "load": "k_cache_load<2>"}[dom],
                     "bytes_per_launch": bytes_k[dom], "launch_ms": kms[dom]},
    }
    # larger batches: fewer launches, judged against an older cache state
    out["batch_sweep"] = {}
    for b2 in (1 << 18, 1 << 20):
        pool = pbm.Pool(T, 4096, strategy=pbm.CRW_BM, backing=False, seed=7)
        cache = pbm.Cache(pool, L, C, b2)
        torch.cuda.synchronize()
        t0.record(stream)
        cache.run(refs, stream=stream)
        t1.record(stream)
        torch.cuda.synchronize()
        ms2 = t0.elapsed_time(t1)
        st3 = cache.stats(stream)
        cache.close()
        pool.close()
        out["batch_sweep"][str(b2)] = {"ms": ms2, "refs_per_s": n_refs / (ms2 * 1e-3),
                                       "hits": st3["hits"], "evictions": st3["evictions"],
                                       "hit_rate": st3["hits"] / n_refs,
                                       "scan_timeouts": st3["scan_timeouts"]}
        progress(f"cache: batches of {b2}: {ms2:.1f} ms")
    # the references on the host for the serial CLOCK baseline
    out["_cpu_refs"] = refs.cpu().numpy()
    del refs
    return out


def rank_at_scale_leg(torch, pbm, stream, n_pages: int, fill: float, n_req: int,
                      steps: int) -> dict:
    """The rank allocation at the serving pool's size: n_pages x 4 KiB
    (page map only), Bernoulli pre-fill `fill` (seed 42), n_req one-page
    requests then frees, with each kernel's own duration (the tile scan
    holds four map words per thread on maps of 2^20 words and more)."""
    pool = pbm.Pool(n_pages, PAGE_SIZE, strategy=pbm.RW_BM, backing=False, seed=SEED)
    pool.prefill_bernoulli(fill, SEED, stream)
    before = pool.snapshot(stream)
    A = pool.free_count(stream)
    ids = torch.full((n_req,), -1, dtype=torch.int32, device="cuda")
    # the free's waves per CU from the autotuner (the rank scan has one shape)
    tuned = pbm.autotune(pool, n_req, ids, pbm.TUNE_RANK_ONLY | pbm.TUNE_QUICK, stream)
    gl = pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK)
    fl = pbm.launch(pbm.MODE_BATCHED, waves_per_cu=tuned["free_waves_per_cu"])
    pool.clear_counters(stream)
    pool.get_pages(n_req, ids, None, 0, gl, stream)
    torch.cuda.synchronize()
    ca = pool.counters(stream)
    i_np = ids.cpu().numpy()
    pool.free_pages(n_req, ids, fl, stream)
    torch.cuda.synchronize()
    cc = pool.counters(stream)
    grants = int((i_np >= 0).sum())
    ok = (grants == min(n_req, A) and len(np.unique(i_np[i_np >= 0])) == grants
          and bool(np.array_equal(pool.snapshot(stream), before)))
    q = pbm.Launch(gl.mode, gl.block, gl.waves_per_cu, gl.walk_budget, gl.cluster_words,
                   gl.flags | pbm.LAUNCH_NO_COUNTERS)
    qf = pbm.Launch(fl.mode, fl.block, fl.waves_per_cu, fl.walk_budget, fl.cluster_words,
                    fl.flags | pbm.LAUNCH_NO_COUNTERS)

    def step():
        pool.get_pages(n_req, ids, None, 0, q, stream)
        pool.free_pages(n_req, ids, qf, stream)

    ms = time_region(torch, step, steps, 2, stream)
    own, _ = time_phases(torch, pbm, pool, n_req, ids, q, qf, steps, stream)
    L = n_pages // 64
    alloc_bytes = 8 * L + 16 * ca["atomics"] + 4 * n_req
    free_bytes = 4 * n_req + 16 * (cc["atomics"] - ca["atomics"])
    achieved = alloc_bytes / (own["alloc"] * 1e-3) / 1e9
    # the same step with exclusive claims (plain stores; nothing else runs
    # on the pool meanwhile): one checked cycle, then timed
    qx = pbm.Launch(q.mode, q.block, q.waves_per_cu, q.walk_budget, q.cluster_words,
                    q.flags | pbm.LAUNCH_EXCLUSIVE)
    pool.get_pages(n_req, ids, None, 0, qx, stream)
    torch.cuda.synchronize()
    x_np = ids.cpu().numpy()
    pool.free_pages(n_req, ids, qf, stream)
    torch.cuda.synchronize()
    ok_x = (np.array_equal(x_np, i_np) and bool(np.array_equal(pool.snapshot(stream), before)))

    def step_x():
        pool.get_pages(n_req, ids, None, 0, qx, stream)
        pool.free_pages(n_req, ids, qf, stream)

    ms_x = time_region(torch, step_x, steps, 2, stream)
    own_x, _ = time_phases(torch, pbm, pool, n_req, ids, qx, qf, steps, stream)
    # exclusive claims write each claimed word (8 B) instead of a 16 B RMW
    x_bytes = 8 * L + 8 * ca["atomics"] + 4 * n_req
    x_ach = x_bytes / (own_x["alloc"] * 1e-3) / 1e9
    pool.close()
    return {"workload": f"{n_pages:,} x 4 KiB pages (page map only), Bernoulli pre-fill "
                        f"f={fill} seed 42, {n_req:,} one-page requests then frees, rank "
                        "allocation",
            "free_pages": A, "grants": grants, "ms_per_step": ms,
            "free_waves_per_cu": tuned["free_waves_per_cu"],
            "value": 2 * grants / (ms * 1e-3), "phases_ms": own, "verified": ok,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": traffic_of("rank_at_scale", "k_get_rank<4, false>"),
                         "kernel": "k_get_rank<4, false>", "bytes_per_launch": alloc_bytes,
                         "launch_ms": own["alloc"], "free_bytes_per_launch": free_bytes,
                         "free_gbs": free_bytes / (own["free"] * 1e-3) / 1e9,
                         "free_traffic": traffic_of("rank_at_scale", "k_free_batched")},
            "exclusive": {"ms_per_step": ms_x, "value": 2 * grants / (ms_x * 1e-3),
                          "phases_ms": own_x, "verified": ok_x,
                          "same_ids_as_atomic_claims": bool(np.array_equal(x_np, i_np)),
                          "roofline": {"bound": "hbm", "achieved": x_ach, "peak": HBM_PEAK_GBS,
                                       "unit": "GB/s", "frac": x_ach / HBM_PEAK_GBS,
                                       "traffic": traffic_of("rank_at_scale",
                                                             "k_get_rank<4, true>"),
                                       "kernel": "k_get_rank<4, true>",
                                       "bytes_per_launch": x_bytes,
                                       "launch_ms": own_x["alloc"]}}}


SWEEP_PAGES, SWEEP_PAGE_SIZE = 1 << 24, 512  # the reference's default pool: 8 GiB of 512 B pages
SWEEP_FREE = [round(0.1 - 0.001 * i, 3) for i in range(96)]  # unitTest1.cu:17 (0.1 .. 0.005)
SWEEP_THREADS = [1 << i for i in range(21)]                   # unitTest1.5.cu:16
SWEEP_SIZES = [4, 16, 64, 128] + list(range(192, 257, 64)) + list(range(512, 8193, 256))


def reference_sweeps_leg(torch, pbm, stream, samples: int) -> dict:
    """The reference's three measurement grids on its default pool
    (16,777,216 pages of 512 B, Paging.cuh:4-7), with `samples` samples per
    point where the reference takes 96, 30 and 100:

    free_fraction (unitTests/unitTest1.cu:14-17, 95-125): 5,000 one-page
      requests at each free fraction 0.1, 0.099, .., 0.005, for the three
      bitmap walks.  The pool is filled to each point by the strategy's own
      plain getPage requests, (int)((1 - p) T) of them as prefillBuffer
      counts them; each sample restores that map and walks with new
      request streams.  TAS, WAS (64- and 32-lane groups) and the kernel
      time, beside the closed forms of models.sweep at the same T, A, N.
    free_time (unitTests/unitTest1.5.cu:14-16, 66-87): N one-page grants on
      the empty pool, then the timed free of all N, N = 1 .. 1,048,576; for
      the array walk RandomWalkBasic and the bitmap RandomWalkBasic_BM
      (plain, one thread per page) and the batched bitmap free.
    getnpage_size (unitTests/unitTest2.cu:12-18, 70-92): 10,000 and 100,000
      RWMalloc getNPage requests of ceil(size / 512) pages each on the empty
      pool, sizes 4 B .. 8 KiB; the reference's window (one word per lane)
      and the 256-word window.

    Times are events recorded on the stream around each launch, as the
    reference's KernelTiming brackets its kernels.  Every sample checks
    its grants: min(N, A) of them, no page twice, the map restored."""
    import models
    T = SWEEP_PAGES
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn) -> float:
        ev0.record(stream)
        fn()
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_time(ev1)

    ok = True
    plain = pbm.launch(mode=pbm.MODE_PLAIN)
    plain_q = pbm.launch(mode=pbm.MODE_PLAIN, flags=pbm.LAUNCH_NO_COUNTERS)

    # ---- free fraction grid, 5,000 requests
    n_req = 5000
    ff = {"pages": T, "page_size": SWEEP_PAGE_SIZE, "threads": n_req, "samples": samples,
          "free": SWEEP_FREE}
    ids = torch.full((n_req,), -1, dtype=torch.int32, device="cuda")
    stp = torch.zeros((n_req,), dtype=torch.int32, device="cuda")
    fill_ids = torch.empty((T,), dtype=torch.int32, device="cuda")
    for name, strategy in (("RandomWalkBasic_BM", pbm.RW_BM), ("CollabRW_BM", pbm.COLLAB_RW_BM),
                           ("CRW_BM", pbm.CRW_BM)):
        cols = {"tas": [], "was64": [], "was32": [], "ms": []}
        with pbm.Pool(T, SWEEP_PAGE_SIZE, strategy=strategy, seed=SEED) as pool:
            filled = 0
            for p in SWEEP_FREE:
                target = int(np.float32(1.0 - np.float32(p)) * np.float32(T))
                if target > filled:
                    pool.get_pages(target - filled, fill_ids, None, (0xFFFE << 48) + filled,
                                   plain_q, stream)
                    filled = target
                words = pool.snapshot(stream)
                A = pool.free_count(stream)
                tas, w64, w32, ms = [], [], [], []
                for k in range(samples):
                    ids.fill_(-1)
                    ms.append(timed(lambda: pool.get_pages(n_req, ids, stp, k * n_req, plain,
                                                           stream)))
                    i_np = ids.cpu().numpy()
                    ws = walk_stats(stp.cpu().numpy().view(np.uint32), i_np)
                    g = i_np[i_np >= 0]
                    ok = ok and g.size == min(n_req, A) and np.unique(g).size == g.size
                    tas.append(ws["tas"])
                    w64.append(ws["was64"])
                    w32.append(ws["was32"])
                    pool.restore(words, stream)
                cols["tas"].append(float(np.mean(tas)))
                cols["was64"].append(float(np.mean(w64)))
                cols["was32"].append(float(np.mean(w32)))
                cols["ms"].append(float(np.mean(ms)))
            ok = ok and filled == int(np.float32(1.0 - np.float32(SWEEP_FREE[-1])) * np.float32(T))
        ff[name] = cols
        progress(f"reference sweeps: free fraction grid, {name}")
    del fill_ids
    rows = models.sweep(T, n_req, [1.0 - p for p in SWEEP_FREE])
    ff["model"] = {k: [r[k] for r in rows] for k in
                   ("RW_TAS", "RWBM_TAS", "RWBM_WAS", "CoRWBM_WAS")}

    # ---- timed free over 1 .. 1,048,576 threads
    ft = {"pages": T, "threads": SWEEP_THREADS, "samples": samples}
    big = torch.full((SWEEP_THREADS[-1],), -1, dtype=torch.int32, device="cuda")
    batched = pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_NO_COUNTERS)
    for name, strategy, fl in (("RandomWalkBasic", pbm.RW_BASIC, plain_q),
                               ("RandomWalkBasic_BM", pbm.RW_BM, plain_q),
                               ("batched_bitmap_free", pbm.RW_BM, batched)):
        col = []
        with pbm.Pool(T, SWEEP_PAGE_SIZE, strategy=strategy, seed=SEED) as pool:
            for n in SWEEP_THREADS:
                ms = []
                for k in range(samples):
                    pool.get_pages(n, big, None, k * n, plain_q, stream)
                    ms.append(timed(lambda: pool.free_pages(n, big, fl, stream)))
                col.append(float(np.mean(ms)))
            ok = ok and pool.free_count(stream) == T
        ft[name] = col
        progress(f"reference sweeps: timed free, {name}")
    del big

    # ---- getNPage over request sizes, 10,000 and 100,000 requests
    gs = {"pages": T, "page_size": SWEEP_PAGE_SIZE, "sizes": SWEEP_SIZES,
          "n_pages": [-(-s_ // SWEEP_PAGE_SIZE) for s_ in SWEEP_SIZES], "samples": samples}
    with pbm.Pool(T, SWEEP_PAGE_SIZE, strategy=pbm.RW_MALLOC, seed=SEED) as pool:
        for n_req in (10000, 100000):
            ids = torch.full((n_req,), -1, dtype=torch.int32, device="cuda")
            stp = torch.zeros((n_req,), dtype=torch.int32, device="cuda")
            cnt = torch.empty((n_req,), dtype=torch.int32, device="cuda")
            for name, gl in (("plain", plain),
                             ("plain_wide", pbm.launch(mode=pbm.MODE_PLAIN, cluster_words=256))):
                cols = {"tas": [], "was64": [], "ms": []}
                for npg in gs["n_pages"]:
                    cnt.fill_(npg)
                    tas, w64, ms = [], [], []
                    for k in range(samples):
                        pool.reset(stream)
                        ids.fill_(-1)
                        ms.append(timed(lambda: pool.get_n_pages(n_req, cnt, ids, stp, k * n_req,
                                                                 gl, stream)))
                        i_np = ids.cpu().numpy()
                        ws = walk_stats(stp.cpu().numpy().view(np.uint32), i_np)
                        g = i_np[i_np >= 0]
                        ok = (ok and g.size == n_req
                              and pool.free_count(stream) == T - n_req * npg)
                        tas.append(ws["tas"])
                        w64.append(ws["was64"])
                    cols["tas"].append(float(np.mean(tas)))
                    cols["was64"].append(float(np.mean(w64)))
                    cols["ms"].append(float(np.mean(ms)))
                gs[f"{name}_{n_req}"] = cols
                progress(f"reference sweeps: getNPage sizes, {name}, {n_req} requests")
        pool.reset(stream)
    return {"free_fraction": ff, "free_time": ft, "getnpage_size": gs, "verified": bool(ok)}


def cpu_legs(orc, c1, bursts, cache, serving, seconds: float) -> None:
    """The CPU baseline of each configuration leg: the oracle's serial
    restatement of the reference's policy on one host core, over the same
    input (bounded samples of about `seconds` each, stated in `sample`)."""
    if c1:
        T, n = 1 << 16, 1 << 15
        empty = orc.cleared(T)
        t_end = time.perf_counter() + seconds
        reps, busy = 0, 0.0
        while time.perf_counter() < t_end:
            st = orc.State(empty, T)
            t = time.perf_counter()
            ids, _ = st.rwbm(n, 0, SEED)
            st.free_ids(ids)
            busy += time.perf_counter() - t
            reps += 1
        c1["cpu_baseline"] = {
            "value": 2 * n * reps / busy, "unit": "allocs+frees/s", "cores": 1, "kind": "port",
            "sample": f"{reps} cycles of 32,768 RandomWalkBasic_BM getPage + freePage on the "
                      f"empty 65,536-page map, {busy:.1f} s on one core"}
        progress("cpu baseline: 65,536-page configuration")
    if bursts and "_cpu_trace" in bursts:
        tr, T, PS = bursts["_cpu_trace"], 1 << 18, 4096
        st = orc.State(orc.cleared(T), T)
        n = int(tr["alloc_offset"][-1])
        page = np.full(n, -1, dtype=np.int64)
        hdr = np.zeros(n, dtype=np.int64)
        nb = len(tr["alloc_offset"]) - 1
        t = time.perf_counter()
        for b in range(nb):
            fi = tr["free_index"][tr["free_offset"][b]:tr["free_offset"][b + 1]]
            live = (page[fi] >= 0) & (hdr[fi] > 0)
            if live.any():
                st.free_n(page[fi][live], hdr[fi][live])
            a0, a1 = int(tr["alloc_offset"][b]), int(tr["alloc_offset"][b + 1])
            if a1 > a0:
                p_, _, h_, _ = st.malloc(tr["sizes"][a0:a1], PS, a0, 7)
                page[a0:a1], hdr[a0:a1] = p_, h_
        busy = time.perf_counter() - t
        bursts["cpu_baseline"] = {
            "value": bursts["requests"] / busy, "unit": "requests/s", "cores": 1, "kind": "port",
            "sample": f"the whole trace ({nb} bursts) by the serial mallocRW/freeRW policy, "
                      f"{busy:.1f} s on one core", "malloc_failures": int((page < 0).sum()),
            "pages_live_at_end": int(T - st.free[0])}
        progress("cpu baseline: bursts")
    if cache and "_cpu_refs" in cache:
        refs = cache["_cpu_refs"]
        cc = orc.ClockCache(1 << 22, 1 << 20, 1)
        t = time.perf_counter()
        cc.run_textbook(refs)
        busy = time.perf_counter() - t
        cache["cpu_baseline"] = {
            "value": refs.size / busy, "unit": "references/s", "cores": 1, "kind": "port",
            "sample": f"the whole trace by textbook second-chance CLOCK, one reference at a "
                      f"time, {busy:.2f} s on one core"}
        cache["serial_clock"] = {
            "hits": int(cc.stats[0]), "misses": int(cc.stats[1]), "evictions": int(cc.stats[2]),
            "hit_rate": float(cc.stats[0]) / refs.size,
            "batched_minus_serial_hits": cache["hits"] - int(cc.stats[0])}
        for row in cache.get("batch_sweep", {}).values():
            row["batched_minus_serial_hits"] = row["hits"] - int(cc.stats[0])
        progress("cpu baseline: cache")
    if serving and "_cpu_sample_map" in serving:
        bm, g = serving["_cpu_sample_map"], serving["_cpu_sample_grants"]
        T = bm.size * 64
        t_end = time.perf_counter() + seconds
        reps, busy = 0, 0.0
        while time.perf_counter() < t_end:
            st = orc.State(bm, T)
            t = time.perf_counter()
            ids, _ = st.rwbm(g, 0, SEED)
            st.free_ids(ids)
            busy += time.perf_counter() - t
            reps += 1
        serving["cpu_baseline"] = {
            "value": 2 * g * reps / busy, "unit": "allocs+frees/s", "cores": 1, "kind": "port",
            "sample": f"{reps} rounds of {g:,} RandomWalkBasic_BM getPage + freePage (100 "
                      f"steps' grants) on the serving pool's map at the end of the run, "
                      f"{busy:.1f} s on one core"}
        progress("cpu baseline: serving")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="host time spent on the CPU-baseline sample (0 = skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner; use fixed rank and walk launches "
                         "(for counter-collection runs)")
    ap.add_argument("--serve-steps", type=int, default=1200,
                    help="timed steps of the serving loop (after as many warm-up steps "
                         "as --serve-warm; 0 = skip)")
    ap.add_argument("--serve-warm", type=int, default=800)
    ap.add_argument("--bursts", type=int, default=4096,
                    help="bursts of the mallocRW/freeRW leg (0 = skip)")
    ap.add_argument("--bursts-persistent", type=int, default=1,
                    help="let the bursts autotuner pick the persistent kernel "
                         "(0 = phase kernels only, replayed from HIP graphs)")
    ap.add_argument("--cache-refs", type=int, default=1 << 26,
                    help="references of the buffer-cache leg (0 = skip)")
    ap.add_argument("--cache-batch", type=int, default=1 << 16)
    ap.add_argument("--comparators", type=int, default=1,
                    help="time the headline step under every other one-page policy (0 = skip)")
    ap.add_argument("--fill-sweep", type=int, default=1,
                    help="plain RW_BM / CollabRW_BM / rank at fills 0.5, 0.9, 0.99, and on "
                         "65,536 empty pages (0 = skip)")
    ap.add_argument("--rank-scale", type=int, default=1,
                    help="time the rank allocation on 67,108,864 pages at 90%% fill with "
                         "4,194,304 requests (0 = skip)")
    ap.add_argument("--ref-sweeps", type=int, default=4,
                    help="samples per point of the reference's free-fraction, timed-free and "
                         "getNPage-size grids (0 skips them)")
    ap.add_argument("--kmalloc", type=int, default=30000,
                    help="requests of the kernel-malloc leg (0 = skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the line, copy a device buffer of this many MiB once "
                         "(the known-size reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the allocator would shard by page range, with no cross-GPU exchange)",
              file=sys.stderr)
        return 2

    import torch
    import pbm

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    dev_name = torch.cuda.get_device_name(0)

    pool = pbm.Pool(T_PAGES, PAGE_SIZE, strategy=pbm.RW_BM, backing=True, seed=SEED)
    pool.prefill_bernoulli(FILL, SEED, stream)
    before = pool.snapshot(stream)
    A = pool.free_count(stream)
    granted = min(N_REQ, A)
    ops_per_step = 2 * granted

    ids = torch.full((N_REQ,), -1, dtype=torch.int32, device="cuda")
    stp = torch.zeros((N_REQ,), dtype=torch.int32, device="cuda")

    # ---- launch shapes from the autotuner
    if args.no_tune:
        def fixed(rank, wpc, cw, budget):
            flags = pbm.LAUNCH_NO_COUNTERS | (pbm.LAUNCH_RANK if rank else 0)
            gl = pbm.launch(pbm.MODE_BATCHED, 256, wpc, budget, cw, flags)
            fl = pbm.launch(pbm.MODE_BATCHED, 256, wpc, 0, 64, pbm.LAUNCH_NO_COUNTERS)
            return {"get_launch": gl, "free_launch": fl, "rank": rank, "cluster_words": cw,
                    "waves_per_cu": wpc, "walk_budget": budget, "free_waves_per_cu": wpc,
                    "ms_alloc": 0.0, "ms_free": 0.0, "candidates": 0}
        tune_rank = fixed(True, 16, 64, 16)
        tune_walk = fixed(False, 8, 256, 16)
    else:
        tune_rank = pbm.autotune(pool, N_REQ, ids, pbm.TUNE_RANK_ONLY, stream)
        progress("alloc/free: rank launches tuned")
        tune_walk = pbm.autotune(pool, N_REQ, ids, pbm.TUNE_WALK_ONLY, stream)
        progress("alloc/free: walk launches tuned")
    best = tune_rank if (tune_rank["ms_alloc"] + tune_rank["ms_free"]
                         <= tune_walk["ms_alloc"] + tune_walk["ms_free"]) else tune_walk
    assert np.array_equal(pool.snapshot(stream), before), "autotune changed the pool"

    def counted(gl, fl):
        """One untimed cycle with device counters and step output."""
        gc = pbm.Launch(gl.mode, gl.block, gl.waves_per_cu, gl.walk_budget, gl.cluster_words,
                        gl.flags & ~pbm.LAUNCH_NO_COUNTERS)
        fc = pbm.Launch(fl.mode, fl.block, fl.waves_per_cu, fl.walk_budget, fl.cluster_words,
                        fl.flags & ~pbm.LAUNCH_NO_COUNTERS)
        pool.clear_counters(stream)
        pool.get_pages(N_REQ, ids, stp, 0, gc, stream)
        torch.cuda.synchronize()
        c_alloc = pool.counters(stream)
        i_np = ids.cpu().numpy()
        s_np = stp.cpu().numpy().view(np.uint32)
        pool.free_pages(N_REQ, ids, fc, stream)
        torch.cuda.synchronize()
        c_all = pool.counters(stream)
        grants = int((i_np >= 0).sum())
        ok = (bool(np.array_equal(pool.snapshot(stream), before)) and grants == granted
              and len(np.unique(i_np[i_np >= 0])) == grants)
        return c_alloc, c_all, walk_stats(s_np, i_np), ok

    def measure(gl, fl, steps, warmup):
        """Counted cycle; step time with the two phase kernels launched
        back to back (eager) and as one captured graph; per-phase time."""
        c_alloc, c_all, ws, ok = counted(gl, fl)

        def eager():
            pool.get_pages(N_REQ, ids, None, 0, gl, stream)
            pool.free_pages(N_REQ, ids, fl, stream)

        ms = time_region(torch, eager, steps, warmup, stream)
        g = pbm.Graph(pool, N_REQ, ids, None, 0, gl, fl, pbm.PHASE_ALLOC | pbm.PHASE_FREE)
        ms_graph1 = time_region(torch, lambda: g.launch(stream), steps, warmup, stream)
        g.close()
        # GRAPH_STEPS steps per graph: a replay runs that many steps
        g = pbm.Graph(pool, N_REQ, ids, None, 0, gl, fl, pbm.PHASE_ALLOC | pbm.PHASE_FREE,
                      n_steps=GRAPH_STEPS)
        reps = max(1, steps // GRAPH_STEPS)
        ms_graph = time_region(torch, lambda: g.launch(stream), reps, 1, stream) / GRAPH_STEPS
        g.close()
        phases, bracket = time_phases(torch, pbm, pool, N_REQ, ids, gl, fl, min(steps, 200),
                                      stream)
        ok = ok and bool(np.array_equal(pool.snapshot(stream), before))
        return {"ms_per_step": ms, "ms_per_step_graph": ms_graph,
                "ms_per_step_graph_1": ms_graph1, "phases_ms": phases,
                "phases_ms_bracketed": bracket,
                "counters_alloc": c_alloc, "counters_cycle": c_all, "walk_steps": ws,
                "verified": ok}

    results = {}
    for name, tuned in (("batched", best), ("batched_walk", tune_walk)):
        results[name] = measure(tuned["get_launch"], tuned["free_launch"], args.steps, args.warmup)
        progress(f"alloc/free: {name} measured")
        results[name]["tuned"] = {k: v for k, v in tuned.items()
                                  if k not in ("get_launch", "free_launch")}
    # the rank launch with exclusive claims (plain stores of the claimed
    # words; the step runs nothing else on the pool)
    rg = tune_rank["get_launch"]
    xg = pbm.Launch(rg.mode, rg.block, rg.waves_per_cu, rg.walk_budget, rg.cluster_words,
                    rg.flags | pbm.LAUNCH_EXCLUSIVE)
    results["batched_exclusive"] = measure(xg, tune_rank["free_launch"], args.steps, args.warmup)
    progress("alloc/free: batched_exclusive measured")

    # ---- plain per-thread getPage / freePage (the reference's policy)
    p_get = pbm.launch(mode=pbm.MODE_PLAIN, flags=pbm.LAUNCH_NO_COUNTERS)
    p_free = pbm.launch(mode=pbm.MODE_PLAIN, flags=pbm.LAUNCH_NO_COUNTERS)
    p_steps = max(3, min(args.steps, 10))
    rp = measure(p_get, p_free, p_steps, 2)
    progress("alloc/free: plain measured")
    cp_alloc, cp_all, ws_p, ok_p = (rp["counters_alloc"], rp["counters_cycle"], rp["walk_steps"],
                                    rp["verified"])

    comparators = {}
    if args.comparators:
        for s_ in (pbm.COLLAB_RW_BM, pbm.CRW_BM, pbm.RW_MALLOC, pbm.LINKED_LIST,
                   pbm.SINGLE_CLOCK, pbm.DIST_CLOCK, pbm.RW_BASIC, pbm.CRW_BASIC,
                   pbm.COLLAB_RW_BASIC, pbm.OUROBOROS):
            comparators[pbm.STRATEGY_NAMES[s_]] = comparator_leg(torch, pbm, stream, s_, p_steps)
    kmalloc = kmalloc_leg(torch, pbm, stream, args.kmalloc, 10) if args.kmalloc > 0 else None
    sweep = (fill_sweep_leg(torch, pbm, stream, (0.5, 0.9, 0.99), p_steps)
             if args.fill_sweep else None)
    c1 = None
    if args.fill_sweep:
        c1 = {"workload": "65,536 x 4 KiB pages, empty, 32,768 requests then frees",
              **policy_rows(torch, pbm, stream, 1 << 16, 1 << 15, 0.0, p_steps)}
        # roofline of the rank row's dominant kernel: the alloc scan reads
        # all 1,024 map words, claims the 512 it grants from with one
        # read-modify-write each and writes 32,768 ids; the free reads the
        # ids and releases the 512 words
        ph = c1["batched_rank"]["phases_ms"]
        c1_bytes = {"alloc": 8 * 1024 + 16 * 512 + 4 * 32768, "free": 4 * 32768 + 16 * 512}
        c1_dom = max(ph, key=ph.get)
        c1_ach = c1_bytes[c1_dom] / (ph[c1_dom] * 1e-3) / 1e9
        c1["roofline"] = {"bound": "hbm", "achieved": c1_ach, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": c1_ach / HBM_PEAK_GBS,
                          "traffic": traffic_of("small_pool", "k_get_rank_wave<false>"
                                                if c1_dom == "alloc" else "k_free_batched"),
                          "kernel": "k_get_rank_wave<false>" if c1_dom == "alloc"
                          else "k_free_batched",
                          "bytes_per_launch": c1_bytes[c1_dom], "launch_ms": ph[c1_dom]}
        progress("65,536-page configuration measured")
    rank_scale = None
    if args.rank_scale:
        rank_scale = rank_at_scale_leg(torch, pbm, stream, 1 << 26, 0.9, 1 << 22,
                                       max(3, min(args.steps, 50)))
        progress("rank allocation at 67,108,864 pages measured")
    ref_sweeps = (reference_sweeps_leg(torch, pbm, stream, args.ref_sweeps)
                  if args.ref_sweeps > 0 else None)

    b = results["batched"]
    ms_b = b["ms_per_step"]
    ms_p = rp["ms_per_step"]
    value = ops_per_step / (ms_b * 1e-3)
    value_p = ops_per_step / (ms_p * 1e-3)

    # ---- roofline of the dominant kernel of the batched step
    ca, cc = b["counters_alloc"], b["counters_cycle"]
    L = T_PAGES // 64
    rank = b["tuned"]["rank"]
    # algorithmic bytes per launch of each kernel: every map word read (8 B),
    # every claim or release a read-modify-write (16 B), every id written by
    # the alloc phase and read by the free phase (4 B)
    bytes_k = {
        "alloc": 8 * ca["probes"] + 16 * ca["atomics"] + 4 * N_REQ,
        "free": 4 * N_REQ + 16 * (cc["atomics"] - ca["atomics"]),
    }
    wpt = pool.ctx().scan_wpt
    kernel_names = {"alloc": f"k_get_rank<{wpt}, false>" if rank else
                    f"k_get_batched<{b['tuned']['cluster_words'] // 64}>",
                    "free": "k_free_batched"}
    dom = max(b["phases_ms"], key=b["phases_ms"].get)
    dom_ms = b["phases_ms"][dom]
    achieved = bytes_k[dom] / (dom_ms * 1e-3) / 1e9
    # HBM bytes per launch of that kernel from the committed FETCH_SIZE /
    # WRITE_SIZE counter runs of this workload (null when not profiled)
    traffic = traffic_of("headline", kernel_names[dom])
    atom_peak = pbm.atomic_peak(L, 64)
    atom_rate = (ca["atomics"] / (b["phases_ms"]["alloc"] * 1e-3))

    serving = (serving_leg(torch, pbm, stream, args.serve_warm, args.serve_steps)
               if args.serve_steps > 0 else None)
    bursts = (bursts_leg(torch, pbm, stream, args.bursts, 1024, bool(args.bursts_persistent))
              if args.bursts > 0 else None)
    cache = (cache_leg(torch, pbm, stream, args.cache_refs, args.cache_batch)
             if args.cache_refs > 0 else None)

    # ---- achievable HBM bandwidth: a 1 GiB device-to-device copy (read +
    # write bytes over the copy's time), the ceiling the rooflines' 8 TB/s
    # nominal peak is checked against
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ms_copy = time_region(torch, lambda: dst.copy_(src), 5, 2, stream)
    hbm_copy = {"bytes": 2 * src.numel() * 4, "ms": ms_copy,
                "gbs": 2 * src.numel() * 4 / (ms_copy * 1e-3) / 1e9}
    del src, dst
    progress(f"hbm copy: {hbm_copy['gbs']:.0f} GB/s")

    # ---- CPU baseline: reference policy, serial, same pre-fill and requests
    cpu = None
    if args.cpu_seconds > 0:
        progress(f"cpu baseline: {args.cpu_seconds:.0f} s of the serial policy")
        sys.path.insert(0, str(ROOT / "oracle"))
        import pbm_oracle as orc
        t_end = time.perf_counter() + args.cpu_seconds
        reps, busy = 0, 0.0
        while time.perf_counter() < t_end:
            st = orc.State(before, T_PAGES, A)
            t = time.perf_counter()
            c_ids, _ = st.rwbm(N_REQ, 0, SEED)
            st.free_ids(c_ids)
            busy += time.perf_counter() - t
            reps += 1
        cpu = {"value": ops_per_step * reps / busy, "unit": "allocs+frees/s", "cores": 1,
               "kind": "port",
               "sample": f"{reps} full cycles of the workload (RandomWalkBasic_BM policy, "
                         f"262,144 requests + frees on the same pre-fill), {busy:.1f} s on "
                         f"one of {os.cpu_count()} host cores"}
        # the same policy on the host cores this process may use (at most
        # 16), racing on one map with atomic claims: every grant, then every
        # free, as in the device step
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        progress(f"cpu baseline: {args.cpu_seconds / 2:.0f} s of the policy on {threads} threads")
        t_end = time.perf_counter() + args.cpu_seconds / 2
        reps_t, busy_t = 0, 0.0
        while time.perf_counter() < t_end:
            st = orc.State(before, T_PAGES, A)
            t = time.perf_counter()
            _, _, g_t = st.rwbm_threads(N_REQ, 0, SEED, threads, free_after=True)
            busy_t += time.perf_counter() - t
            reps_t += 1
            if g_t != granted or not np.array_equal(st.bm, before):
                raise RuntimeError("threaded CPU baseline lost or doubled a page")
        cpu["threads"] = {
            "value": ops_per_step * reps_t / busy_t, "unit": "allocs+frees/s",
            "cores": threads, "kind": "port",
            "sample": f"{reps_t} full cycles of the same workload on {threads} threads sharing "
                      f"one map (atomic claims), {busy_t:.1f} s"}
        cpu_legs(orc, c1, bursts, cache, serving, args.cpu_seconds / 4)
    for leg in (bursts, cache, serving):
        if leg:
            for k in [k for k in leg if k.startswith("_cpu")]:
                del leg[k]

    line = {
        "metric": METRIC, "value": value, "unit": "allocs+frees/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_b,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "1,048,576 x 4 KiB pages, Bernoulli pre-fill f=0.9 seed 42, "
                               "262,144 one-page requests then frees; batched launch "
                               + ("(single-pass rank allocation: request r takes the free page "
                                  "of rank r in map order)" if rank else
                                  "(persistent random-walk windows)"),
                   "pages": T_PAGES, "page_size": PAGE_SIZE, "fill": FILL,
                   "requests": N_REQ, "free_pages": A, "grants": granted,
                   "launch": b["tuned"], "device": dev_name},
        "ms_per_step_graph": b["ms_per_step_graph"],
        "ms_per_step_graph_1": b["ms_per_step_graph_1"], "graph_steps": GRAPH_STEPS,
        "phases_ms": b["phases_ms"],
        "phases_ms_bracketed": b["phases_ms_bracketed"],
        "walk_steps": b["walk_steps"],
        "verified": b["verified"],
        "plain": {"value": value_p, "ms_per_step": ms_p, "steps": p_steps,
                  "phases_ms": rp["phases_ms"], "walk_steps": ws_p,
                  "probes": cp_alloc["probes"], "atomics": cp_all["atomics"],
                  "verified": ok_p},
        "speedup_vs_plain": value / value_p,
        # against the reference's collaborative walk on the same step
        "speedup_vs_plain_collab": (value / comparators["CollabRW_BM"]["value"]
                                    if "CollabRW_BM" in comparators else None),
        # the plain walk's mean steps against the reference's closed form
        # (theoretical_models.py RWBM_TAS, 64-bit words, the grants as
        # requests): (unsigned)seed % L on this power-of-two map keeps the
        # LCG's low 14 bits, whose single cycle every walk follows
        "plain_tas_closed_form": _rwbm_tas(T_PAGES, A, granted),
        "comparators": comparators,
        "batched_walk": {"value": ops_per_step / (results["batched_walk"]["ms_per_step"] * 1e-3),
                         **{k: results["batched_walk"][k] for k in
                            ("ms_per_step", "ms_per_step_graph", "phases_ms", "walk_steps",
                             "verified", "tuned")}},
        "batched_exclusive": {
            "value": ops_per_step / (results["batched_exclusive"]["ms_per_step"] * 1e-3),
            **{k: results["batched_exclusive"][k] for k in
               ("ms_per_step", "ms_per_step_graph", "phases_ms", "verified")}},
        "counters": {"alloc": ca, "cycle": cc},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel_names[dom], "phase": dom, "bytes_per_launch": bytes_k[dom],
                     "launch_ms": dom_ms,
                     "atomic": {"achieved_ops": atom_rate, "peak_ops": atom_peak,
                                "frac": atom_rate / atom_peak, "unit": "u64 RMW/s",
                                "kernel": kernel_names["alloc"]}},
        "cpu_baseline": cpu,
        "serving": serving,
        "bursts": bursts,
        "cache": cache,
        "kmalloc": kmalloc,
        "fill_sweep": sweep,
        "small_pool": c1,
        "rank_at_scale": rank_scale,
        "reference_sweeps": ref_sweeps,
        "hbm_copy": hbm_copy,
    }
    print(json.dumps(line), flush=True)
    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
    pool.close()
    return 0 if (b["verified"] and ok_p and results["batched_walk"]["verified"]
                 and results["batched_exclusive"]["verified"]
                 and all(v["verified"] for v in comparators.values())
                 and (bursts is None or bursts["verified"])
                 and (cache is None or cache["verified"])
                 and (rank_scale is None or (rank_scale["verified"]
                                             and rank_scale["exclusive"]["verified"]))
                 and (ref_sweeps is None or ref_sweeps["verified"])
                 and (kmalloc is None or (kmalloc["collab"]["verified"]
                                          and kmalloc["simple"]["verified"]))
                 and (sweep is None or all(v["verified"] for row in sweep.values()
                                           for v in row.values()))
                 and (c1 is None or all(c1[k]["verified"] for k in
                                        ("plain_rw_bm", "plain_collab_rw_bm", "batched_rank")))
                 ) else 1


if __name__ == "__main__":
    sys.exit(main())
