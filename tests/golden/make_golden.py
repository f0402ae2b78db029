#!/usr/bin/env python3
"""Writes golden.json: small known-answer vectors of the serial oracle.

The reference ships no golden vectors, known-answer tests or fixtures for
the allocation path, so these are produced by this repository's own CPU
oracle (oracle/pbm_oracle.c) on seeded inputs.  They pin the oracle
against drift, and the GPU tests compare the device path with them
directly.  Regenerate with `python tests/golden/make_golden.py` after a
deliberate change of semantics.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
import pbm_oracle as orc  # noqa: E402


def hexwords(bm: np.ndarray) -> list[str]:
    return [f"{int(w):016x}" for w in bm]


def compute() -> dict:
    g: dict = {"about": "serial oracle known answers; see make_golden.py"}

    # RNG_LCG (RNG_LCG.cuh:6-12): 16 outputs from seed 42, and one step
    # from edge inputs
    x, seq = 42, []
    for _ in range(16):
        x = orc.lcg(x)
        seq.append(x)
    g["lcg_from_42"] = seq
    edges = [-1, 0, 1, 2**31 - 1, -2**31, 12345, -987654321]
    g["lcg_edges"] = {"in": edges, "out": [orc.lcg(e) for e in edges]}
    g["mix64"] = {"in": [0, 1, 2, 3, 2**64 - 1], "out": [orc.mix64(v) for v in [0, 1, 2, 3, 2**64 - 1]]}
    g["request_seed_42"] = [orc.request_seed(42, r) for r in range(8)]

    # seeded Bernoulli pre-fill (bit i used iff page_hash(seed, i) < fill * 2^64)
    g["prefill_1000_0.5_3"] = hexwords(orc.prefill(1000, 0.5, 3))
    g["prefill_4096_0.9_5_free"] = orc.count_free(orc.prefill(4096, 0.9, 5), 4096)

    def pool(T, fill, seed):
        return orc.State(orc.prefill(T, fill, seed), T)

    # RandomWalkBasic_BM getPage, serial, 300 requests on 4096 pages at 90%
    st = pool(4096, 0.9, 5)
    ids, steps = st.rwbm(300, 0, 2)
    g["rwbm_4096_0.9"] = {"T": 4096, "fill": 0.9, "prefill_seed": 5, "n": 300, "req_base": 0,
                          "seed": 2, "ids": ids.tolist(), "steps": steps.tolist(),
                          "map_after": hexwords(st.bm)}

    # CollabRW_BM, serial waves, both hand-out orders
    for order in (0, 1):
        st = pool(4096, 0.9, 5)
        ids, steps = st.collab(300, 0, 2, order)
        g[f"collab_4096_0.9_order{order}"] = {"T": 4096, "fill": 0.9, "prefill_seed": 5,
                                              "n": 300, "req_base": 0, "seed": 2,
                                              "order": order, "ids": ids.tolist(),
                                              "steps": steps.tolist()}

    # batched by rank: request r takes the r-th free page
    st = pool(20000, 0.97, 12)
    ids, steps = st.batched_rank(700)
    g["rank_20000_0.97"] = {"T": 20000, "fill": 0.97, "prefill_seed": 12, "n": 700,
                            "ids": ids.tolist(), "steps": steps.tolist(),
                            "free_after": int(st.free[0])}

    # CRW_BM: one thread from no hint on a 50% pool
    st = pool(4096, 0.5, 7)
    ids, steps, hint = st.crw(200, 0, 9)
    g["crw_4096_0.5"] = {"T": 4096, "fill": 0.5, "prefill_seed": 7, "n": 200, "seed": 9,
                         "ids": ids.tolist(), "steps": steps.tolist(), "hint": hint}

    # SingleClock and DistClock (one arm) on a 70% pool
    st = pool(2048, 0.7, 4)
    ids, steps, arm = st.single_clock(150, 0)
    g["single_clock_2048_0.7"] = {"T": 2048, "fill": 0.7, "prefill_seed": 4, "n": 150,
                                  "ids": ids.tolist(), "steps": steps.tolist(), "arm": arm}
    st = pool(2048, 0.7, 4)
    arm0 = orc.dist_clock_arm(11, 0, 2048)
    ids, steps, arm = st.dist_clock(150, arm0)
    g["dist_clock_2048_0.7"] = {"T": 2048, "fill": 0.7, "prefill_seed": 4, "n": 150,
                                "seed": 11, "arm0": arm0, "ids": ids.tolist(),
                                "steps": steps.tolist(), "arm": arm}

    # LinkedList FIFO: built from a pre-fill, 40 gets, 10 frees, 40 gets
    ls = orc.ListState(orc.prefill(512, 0.8, 6), 512)
    a = ls.get(40)
    ls.put(a[:10][::-1].copy())
    b = ls.get(40)
    g["list_512_0.8"] = {"T": 512, "fill": 0.8, "prefill_seed": 6, "first": a.tolist(),
                         "freed": a[:10][::-1].tolist(), "second": b.tolist()}

    # getNPage runs of 1..20 pages on a 50% pool
    rng = np.random.default_rng(3)
    counts = rng.integers(1, 21, size=64).astype(np.uint32)
    st = pool(4096, 0.5, 8)
    ids, steps = st.get_n(counts, 0, 5)
    g["getnpage_4096_0.5"] = {"T": 4096, "fill": 0.5, "prefill_seed": 8, "seed": 5,
                              "counts": counts.tolist(), "ids": ids.tolist(),
                              "steps": steps.tolist()}

    # mallocRW sizes 16 B .. 64 KiB on 4 KiB pages
    sizes = np.exp(rng.uniform(np.log(16), np.log(65536), size=64)).astype(np.uint64)
    st = pool(4096, 0.5, 8)
    page, off, hdr, steps = st.malloc(sizes, 4096, 0, 5)
    g["malloc_4096_0.5"] = {"T": 4096, "fill": 0.5, "prefill_seed": 8, "seed": 5,
                            "page_size": 4096, "sizes": sizes.tolist(), "page": page.tolist(),
                            "offset": off.tolist(), "header": hdr.tolist(),
                            "steps": steps.tolist()}

    # RandomWalkBasic and the legacy clustered / collaborative page walks
    st = pool(3000, 0.8, 9)
    ids, steps = st.rwbasic(200, 0, 4)
    g["rwbasic_3000_0.8"] = {"T": 3000, "fill": 0.8, "prefill_seed": 9, "n": 200, "seed": 4,
                             "ids": ids.tolist(), "steps": steps.tolist()}
    st = pool(3000, 0.8, 9)
    ids, steps, hint = st.crw_basic(200, 0, 4)
    g["crw_basic_3000_0.8"] = {"T": 3000, "fill": 0.8, "prefill_seed": 9, "n": 200, "seed": 4,
                               "ids": ids.tolist(), "steps": steps.tolist(), "hint": hint}
    st = pool(3000, 0.8, 9)
    ids, steps = st.collab_basic(200, 0, 4)
    g["collab_basic_3000_0.8"] = {"T": 3000, "fill": 0.8, "prefill_seed": 9, "n": 200,
                                  "seed": 4, "ids": ids.tolist(), "steps": steps.tolist()}

    # per-thread getXPage, runs of 1..12 pages on a 40% pool
    xc = rng.integers(1, 13, size=80).astype(np.uint32)
    st = pool(4096, 0.4, 10)
    ids, steps = st.get_x(xc, 0, 6)
    g["getxpage_4096_0.4"] = {"T": 4096, "fill": 0.4, "prefill_seed": 10, "seed": 6,
                              "counts": xc.tolist(), "ids": ids.tolist(),
                              "steps": steps.tolist()}

    # kernel malloc headers: the bytes written, as (offset, 8-byte value)
    def headers(mem, wr):
        at = np.flatnonzero(wr[::8]) * 8
        return [[int(a), int.from_bytes(mem[a:a + 8].tobytes(), "little")] for a in at]

    T, ps = 1024, 4096
    ksz = np.exp(rng.uniform(0, np.log(9000), size=100)).astype(np.uint32)
    st = pool(T, 0.5, 11)
    mem, wr = np.zeros(T * ps, np.uint8), np.zeros(T * ps, np.uint8)
    off, steps = st.kmalloc_collab(mem, wr, ksz, ps, 0, 7)
    g["kmalloc_collab_1024_0.5"] = {"T": T, "fill": 0.5, "prefill_seed": 11, "seed": 7,
                                    "page_size": ps, "sizes": ksz.tolist(),
                                    "offsets": [int(o) for o in off], "steps": steps.tolist(),
                                    "headers": headers(mem, wr)}
    st = pool(T, 0.5, 11)
    mem, wr = np.zeros(T * ps, np.uint8), np.zeros(T * ps, np.uint8)
    off, steps, cur = st.kmalloc_simple(mem, wr, ksz, ps, 0, 7)
    g["kmalloc_simple_1024_0.5"] = {"T": T, "fill": 0.5, "prefill_seed": 11, "seed": 7,
                                    "page_size": ps, "sizes": ksz.tolist(),
                                    "offsets": [int(o) for o in off], "steps": steps.tolist(),
                                    "current": cur, "headers": headers(mem, wr)}

    # Ouroboros queue at 512-byte pages (16 per chunk): gets, frees, gets
    q = orc.QueueState(orc.prefill(2048, 0.6, 12), 2048, orc.ouroboros_chunk_pages(512))
    a, sa = q.get(120)
    back = a[a >= 0][::3].copy()
    q.put(back)
    b, sb = q.get(120)
    g["ouroboros_2048_0.6"] = {"T": 2048, "page_size": 512, "fill": 0.6, "prefill_seed": 12,
                               "first": a.tolist(), "first_steps": sa.tolist(),
                               "freed": back.tolist(), "second": b.tolist(),
                               "second_steps": sb.tolist()}

    return g


def main() -> None:
    (HERE / "golden.json").write_text(json.dumps(compute(), separators=(",", ":")) + "\n")


if __name__ == "__main__":
    main()
