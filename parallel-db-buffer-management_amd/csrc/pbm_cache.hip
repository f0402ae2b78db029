// Buffer cache over a pool: logical pages are kept in `capacity` slots, each
// slot holding one frame (a pool page) got with the pool strategy's device
// getPage; CLOCK replacement.  References are taken a batch at a time, in
// three kernels:
//   hits   closes the last batch (the arm moves past its last victim) and
//          clears this batch's look-back state; every reference whose page
//          is present sets its slot's bit, and the first reference to an
//          absent page is its miss (an atomicMax per page keeps the first)
//   plan   persistent workgroups take tiles by ticket: first the tiles of
//          the batch, then the tiles of the ring.
//          - a batch tile compacts its misses in reference order (decoupled
//            look-back); the last one plans the batch (misses K, victims
//            V) and publishes V with the batch's tag in one atomic word
//          - a ring tile is the CLOCK sweep from the arm: ring position j is
//            slot (arm + j) mod capacity, and a slot is a candidate when it
//            was occupied at batch start and its bit is clear.  Look-back
//            ranks the candidates; rank r < V is victim r, evicted where it
//            is taken (its page becomes absent, its frame goes back to the
//            pool), and a set bit is cleared when fewer than V candidates
//            precede it.  Once a tile at or past the V-th candidate is
//            known, no later tile is taken
//   load   when the ring held Z < V candidates, every set bit has been
//          cleared and the load kernel first sweeps again, taking the first
//          V - Z now-clear slots as victims Z.. (victims carry bit value 2
//          until loaded, so it passes them).  Then misses load into empty
//          slots first and victims' slots after, each with a new frame and
//          its bit set
// With a batch of one reference this is the textbook second-chance CLOCK.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <cmath>
#include <vector>
#include "pbm_internal.hpp"

using pbm::u64;

namespace {

enum : uint32_t {
    cArm = 0,
    cResident,
    cHits,
    cMisses,
    cEvictions,
    cBatches,
    cK,             // misses of the current batch
    cV,             // victims of the current batch
    cZ,             // candidates in the ring when the first sweep ran short
    cStop,          // ring position of the last victim
    cAllocFail,
    cR0,            // resident at batch start
    cEmpty,         // empty slots at batch start
    cV2,            // victims left for the second sweep
    cTicket,        // tiles of the plan kernel taken (batch tiles, then ring tiles)
    cSweepTicket1,  // tiles of the second sweep taken
    cSweepDone0,    // a ring tile at or past the first sweep's last victim (or kNoStop)
    cSweepDone1,    // the same for the second sweep
    cOpen,          // 1 while the last batch waits to be closed
    cScanTimeouts,  // waits given up (never, while tiles run on)
    cPlan,          // (batch tag << 32) | V, published by the batch's last tile
    cInvalid,       // references to ids >= n_logical, skipped
    cInvalidBatch,  // the same in the current batch (cleared by its plan)
    cPrevTiles,     // ring tiles the last batch's first sweep used (0: not known)
    cCtlWords = 32
};

constexpr u64 kLow = 0xFFFFFFFFull;
constexpr uint32_t kPer = 4;                                // items per thread
constexpr uint32_t kTile = pbm::kScanThreads * kPer;        // items per look-back tile
constexpr uint32_t kVictimBit = 2u;                         // bit value of a chosen victim
constexpr uint32_t kGroups = 64;                            // workgroups of the sweep and load
constexpr uint32_t kNoTile = ~0u;

// Look-back word: [63:62] 0 unpublished, 1 the tile's own count, 2 its
// inclusive prefix; [61:0] the count.  Cleared before every launch.
constexpr u64 kLbOwn = 1ull << 62;
constexpr u64 kLbIncl = 2ull << 62;
constexpr u64 kLbCount = kLbOwn - 1ull;
constexpr u64 kNoStop = ~0ull;
constexpr uint32_t kLbSpin = 1u << 20;

struct CacheDev {
    pbm_ctx c;
    int32_t* slot_of;   // [n_logical] slot of a present page, -1 absent
    u64* first;         // [n_logical] (batch << 32) | ~index of the first miss
    int32_t* owner;     // [capacity] page in the slot
    int32_t* frame;     // [capacity] pool page holding it
    uint32_t* refbit;   // [capacity]
    int32_t* miss;      // [batch]
    int32_t* victim;    // [batch]
    u64* lb_miss;       // [miss_tiles] look-back words of the miss compaction
    u64* lb_sweep;      // [2 * sweep_tiles] look-back words of the two sweeps
    int32_t* log;       // [log_cap] evicted pages in order
    u64 log_cap;
    u64* ctl;           // [cCtlWords]
    uint32_t n_logical, cap, batch;
    uint32_t miss_tiles, sweep_tiles, plan_groups, load_groups;
};

__device__ __forceinline__ u64 ldc(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stc(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 lb_clamp(u64 v) { return v > kLbCount ? kLbCount : v; }

// The batch counters, read all at once and written back all at once (one
// thread; every load is issued before the first store).
struct CtlState {
    u64 arm, resident, evictions, batches, K, V, stop, r0, open;
};

__device__ __forceinline__ CtlState load_ctl(const CacheDev& d) {
    CtlState c;
    c.arm = d.ctl[cArm];
    c.resident = d.ctl[cResident];
    c.evictions = d.ctl[cEvictions];
    c.batches = d.ctl[cBatches];
    c.K = d.ctl[cK];
    c.V = d.ctl[cV];
    c.stop = d.ctl[cStop];
    c.r0 = d.ctl[cR0];
    c.open = d.ctl[cOpen];
    return c;
}

// Closes the last batch: the arm moves past its last victim.
__device__ __forceinline__ void close_batch(const CacheDev& d) {
    CtlState c = load_ctl(d);
    if (!c.open) return;
    if (c.V) c.arm = (c.arm + c.stop + 1u) % d.cap;
    c.resident = c.r0 + c.K < d.cap ? c.r0 + c.K : d.cap;
    d.ctl[cArm] = c.arm;
    d.ctl[cResident] = c.resident;
    d.ctl[cEvictions] = c.evictions + c.V;
    d.ctl[cBatches] = c.batches + 1u;
    d.ctl[cOpen] = 0;
}

// Workgroup scan of per-thread counts: this thread's offset in the tile
// (returned) and the tile's total (agg).  Ends with the workgroup in step.
__device__ __forceinline__ uint32_t tile_scan(uint32_t cnt, uint32_t* s_wsum, uint32_t& agg) {
    const unsigned lane = pbm::lane_id(), wv = threadIdx.x >> 6;
    uint32_t inc = cnt;
#pragma unroll
    for (unsigned d = 1; d < 64u; d <<= 1) {
        const uint32_t v = pbm::shfl32(inc, lane >= d ? lane - d : lane);
        if (lane >= d) inc += v;
    }
    if (lane == 63) s_wsum[wv] = inc;
    __syncthreads();
    uint32_t woff = 0;
    agg = 0;
#pragma unroll
    for (unsigned i = 0; i < pbm::kScanThreads / 64u; ++i) {
        if (i < wv) woff += s_wsum[i];
        agg += s_wsum[i];
    }
    return woff + inc - cnt;
}

// Decoupled look-back, run by wave 0: publishes the tile's own count, sums
// the counts of the preceding tiles back to the nearest one that published
// its inclusive prefix (64 at a time), publishes this tile's inclusive
// prefix and returns the exclusive one.  Every preceding tile is held by a
// workgroup that has started (a workgroup's first tile is its own index,
// and the dispatcher starts each XCD's workgroups in index order; later
// tiles go by ticket, after all of those), so the wait ends, unless that
// workgroup stopped because the sweep's stop (*done) lies before it, and
// then so does this tile's: kNoStop is returned.  A bounded number of polls is
// kept all the same, and giving up is counted.
__device__ u64 lb_prefix(u64* words, uint32_t tile, u64 agg, const u64* done, u64* timeouts) {
    const unsigned lane = pbm::lane_id();
    if (tile == 0) {
        if (lane == 0) stc(&words[0], kLbIncl | lb_clamp(agg));
        return 0;
    }
    if (lane == 0) stc(&words[tile], kLbOwn | lb_clamp(agg));
    int64_t k = (int64_t)tile - 1;
    u64 excl = 0;
    uint32_t spins = 0;
    for (;;) {
        const int64_t idx = k - (int64_t)lane;
        u64 v = 0;
        bool ready = true, incl = false;
        if (idx >= 0) {
            v = ldc(&words[idx]);
            ready = (v >> 62) != 0u;
            incl = (v >> 62) == 2u;
        }
        const u64 P = pbm::ballot(incl);
        const u64 R = pbm::ballot(ready);
        const unsigned fp = P ? (unsigned)__ffsll((long long)P) - 1u : 63u;
        const u64 upto = fp == 63u ? pbm::kFull : ((2ull << fp) - 1ull);
        if ((R & upto) != upto) {
            if (done && ldc(done) < tile) return kNoStop;
            if (++spins > kLbSpin) {
                if (lane == 0) pbm::atom_add_u64(timeouts, 1ull);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        // one inclusive word at most is summed, so the sum cannot wrap
        u64 part = (lane <= fp && idx >= 0) ? (v & kLbCount) : 0ull;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) part += pbm::shfl64(part, lane ^ (unsigned)d);
        excl += part;
        if (P) break;
        k -= 64;
    }
    if (lane == 0) stc(&words[tile], kLbIncl | lb_clamp(excl + agg));
    return excl;
}

// Shared words of a tile-taking workgroup.
struct TileShared {
    uint32_t tile, wsum[4];
    u64 excl, v;
    uint32_t freed;  // frames a first-sweep tile gave back, credited once per tile
};

// tag: the batch's number since the cache was made, from 1.  A reference
// to an id >= n_logical is skipped and counted.
__global__ __launch_bounds__(256) void k_cache_hits(CacheDev d, const int32_t* refs, uint32_t nb,
                                                    u64 tag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t threads = gridDim.x * blockDim.x;
    // the look-back words and tickets of this batch start clear
    for (uint32_t w = i; w < d.miss_tiles; w += threads) d.lb_miss[w] = 0;
    for (uint32_t w = i; w < 2u * d.sweep_tiles; w += threads) d.lb_sweep[w] = 0;
    if (i == 0) {
        close_batch(d);
        const u64 pd = d.ctl[cSweepDone0];
        d.ctl[cPrevTiles] = pd == kNoStop ? 0u : pd + 1u;
        d.ctl[cTicket] = 0;
        d.ctl[cSweepTicket1] = 0;
        d.ctl[cSweepDone0] = kNoStop;
        d.ctl[cSweepDone1] = kNoStop;
        d.ctl[cV2] = 0;
    }
    const u64 tag_hi = tag << 32;
    bool bad = false;
    if (i < nb && (uint32_t)refs[i] >= d.n_logical) {
        bad = true;  // neither a hit nor a miss
    } else if (i < nb) {
        const int32_t l = refs[i];
        const int32_t s = d.slot_of[l];
        if (s >= 0) {
            // hot slots are hit many times a batch: store only a change
            if (d.refbit[s] != 1u) d.refbit[s] = 1u;
        } else {
            const u64 mine = tag_hi | (kLow - i);
            if (ldc(&d.first[l]) < mine) pbm::atom_max_u64(&d.first[l], mine);
        }
    }
    const u64 B = pbm::ballot(bad);
    if (B && pbm::lane_id() == (unsigned)__ffsll((long long)B) - 1u) {
        pbm::atom_add_u64(&d.ctl[cInvalid], (u64)__popcll(B));
        pbm::atom_add_u64(&d.ctl[cInvalidBatch], (u64)__popcll(B));
    }
}

// A batch tile: first references to absent pages, compacted in reference
// order.  The last tile plans the batch; hits are the references that are
// not misses.
__device__ void miss_tile(const CacheDev& d, const int32_t* refs, uint32_t nb, u64 tag,
                          uint32_t tile, TileShared& sh) {
    const u64 tag_hi = tag << 32;
    const uint32_t i0 = tile * kTile + threadIdx.x * kPer;
    // the last tile's plan reads the counters the previous kernels left:
    // loaded now, in flight with the tile's references
    u64 r0 = 0, hits = 0, misses = 0, bad = 0;
    const bool plans = tile == d.miss_tiles - 1u && threadIdx.x == 0;
    if (plans) {
        r0 = d.ctl[cResident];
        hits = d.ctl[cHits];
        misses = d.ctl[cMisses];
        bad = d.ctl[cInvalidBatch];
    }
    int32_t lv[kPer];
    uint32_t fm = 0, cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = i0 + k;
        lv[k] = -1;
        if (i < nb && (uint32_t)refs[i] < d.n_logical) {
            const int32_t l = refs[i];
            // both loads in flight at once; the tag decides only for absent pages
            const int32_t sl = d.slot_of[l];
            const u64 fi = d.first[l];
            if (sl < 0 && fi == (tag_hi | (kLow - i))) {
                lv[k] = l;
                fm |= 1u << k;
                ++cnt;
            }
        }
    }
    uint32_t agg = 0;
    const uint32_t off = tile_scan(cnt, sh.wsum, agg);
    if (threadIdx.x < 64) {
        const u64 e = lb_prefix(d.lb_miss, tile, agg, nullptr, &d.ctl[cScanTimeouts]);
        if (threadIdx.x == 0) sh.excl = e;
    }
    __syncthreads();
    u64 r = sh.excl + off;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k)
        if (fm >> k & 1u) d.miss[r++] = lv[k];
    if (plans) {
        const u64 K = sh.excl + agg;
        const u64 empty = d.cap - r0;
        const u64 V = K > empty ? K - empty : 0u;
        d.ctl[cK] = K;
        d.ctl[cR0] = r0;
        d.ctl[cEmpty] = empty;
        d.ctl[cV] = V;
        d.ctl[cMisses] = misses + K;
        d.ctl[cHits] = hits + (nb - K - bad);
        d.ctl[cInvalidBatch] = 0;
        d.ctl[cOpen] = 1;
        // the ring tiles of this launch read V from here
        stc(&d.ctl[cPlan], ((tag & kLow) << 32) | V);
    }
}

// A ring tile of a CLOCK sweep (pass 0 in the plan kernel, pass 1 in the
// load kernel).  v: the sweep's victims (kNoStop: read the plan word of
// batch `tag` once the tile's candidates are counted).  Victims of pass 1
// are published with a release store after their eviction, for the load
// threads that wait on them.
__device__ void sweep_tile(const CacheDev& d, uint32_t pass, uint32_t tile, u64 v, u64 tag,
                           u64 arm, u64 r0, u64 vbase, u64 ev0, TileShared& sh, pbm::Stats& st) {
    u64* words = d.lb_sweep + (pass ? d.sweep_tiles : 0u);
    u64* done = &d.ctl[pass ? cSweepDone1 : cSweepDone0];
    const uint32_t j0 = tile * kTile + threadIdx.x * kPer;
    uint32_t slot[kPer], bit[kPer];
    int32_t own[kPer], frm[kPer];
    uint32_t occ = 0, cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t j = j0 + k;
        slot[k] = 0;
        bit[k] = 0;
        own[k] = -1;
        frm[k] = -1;
        if (j < d.cap) {
            u64 s = arm + j;
            if (s >= d.cap) s -= d.cap;
            slot[k] = (uint32_t)s;
            if (s < r0) {
                occ |= 1u << k;
                bit[k] = d.refbit[s];
                if (bit[k] == 0u) {
                    ++cnt;
                    // a candidate may become a victim: its page and frame
                    // are read now, off the path that waits for V
                    own[k] = d.owner[s];
                    frm[k] = d.frame[s];
                }
            }
        }
    }
    uint32_t agg = 0;
    const uint32_t off = tile_scan(cnt, sh.wsum, agg);
    // the look-back does not need V, so it overlaps the batch tiles' work
    if (threadIdx.x < 64) {
        const u64 e = lb_prefix(words, tile, agg, done, &d.ctl[cScanTimeouts]);
        if (threadIdx.x == 0) sh.excl = e;
    }
    if (threadIdx.x == 0) {
        u64 vv = v;
        if (vv == kNoStop) {
            // the batch's last tile publishes V; it was taken before this one
            const u64 want = tag & kLow;
            uint32_t spins = 0;
            u64 w = ldc(&d.ctl[cPlan]);
            while ((w >> 32) != want) {
                if (++spins > kLbSpin) {
                    pbm::atom_add_u64(&d.ctl[cScanTimeouts], 1ull);
                    w = want << 32;  // no victims
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                w = ldc(&d.ctl[cPlan]);
            }
            vv = w & kLow;
        }
        sh.v = vv;
        sh.freed = 0;
        const u64 e = sh.excl;
        // a tile that holds the last victim, or lies past it, stops the
        // taking of further tiles (with no victims, every tile does)
        if (vv == 0)
            __hip_atomic_fetch_min(done, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else if (e != kNoStop && e + agg >= vv && ldc(done) > tile)
            __hip_atomic_fetch_min(done, (u64)tile, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pass == 0 && e != kNoStop && tile == d.sweep_tiles - 1u && e + agg < vv) {
            // fewer candidates than victims: every set bit is cleared here,
            // and the second sweep takes the rest
            d.ctl[cZ] = e + agg;
            d.ctl[cV2] = vv - (e + agg);
        }
    }
    __syncthreads();
    v = sh.v;
    if (v == 0) return;
    const u64 excl = sh.excl;  // kNoStop (past the stop) is never below v
    if (pass == 0 && excl != kNoStop && tile == d.sweep_tiles - 1u && excl + agg < v) {
        // the second sweep's victims start unpublished
        for (u64 k = excl + agg + threadIdx.x; k < v; k += blockDim.x) d.victim[k] = -1;
    }
    u64 r = excl + off;  // candidates before this thread's first position
    // first sweep: the victims' frames go back together after the loop, all
    // their atomics in flight at once and one credit for the tile (a free
    // per victim would wait for each atomic in turn).  The second sweep
    // frees each frame before it publishes the victim, for the load
    // threads that reserve a frame once they see it.
    int32_t give[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) give[k] = -1;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        if (excl >= v || !(occ >> k & 1u) || r >= v) continue;
        const uint32_t s = slot[k];
        if (bit[k] == 0u) {
            const u64 vi = vbase + r;
            d.refbit[s] = kVictimBit;
            const int32_t old = own[k];
            d.slot_of[old] = -1;
            if (pass == 0) give[k] = frm[k];
            else if (frm[k] >= 0) pbm::free_page(d.c, frm[k], &st);
            if (ev0 + vi < d.log_cap) d.log[ev0 + vi] = old;
            if (pass)
                __hip_atomic_store(&d.victim[vi], (int32_t)s, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            else
                d.victim[vi] = (int32_t)s;
            if (r == v - 1u) d.ctl[cStop] = j0 + k;
            ++r;
        } else if (bit[k] == 1u) {
            d.refbit[s] = 0u;  // the sweep passes a set bit
        }
    }
    if (pass == 0) {
        u64 mb[kPer], ob[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            mb[k] = 0;
            ob[k] = 0;
            const int32_t f = give[k];
            if (f >= 0 && (uint32_t)f < d.c.n_pages) {
                mb[k] = 1ull << ((uint32_t)f & 63u);
                ob[k] = pbm::atom_and(&d.c.bitmap[(uint32_t)f >> 6], ~mb[k]);
                st.atomics++;
            } else if (f >= 0) {
                st.errors++;
            }
        }
        uint32_t freed = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            if (!mb[k]) continue;
            if (ob[k] & mb[k]) ++freed;
            else st.errors++;
        }
        st.frees += freed;
        // the tile's credit: one atomic, after every wave's frees
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) freed += (uint32_t)__shfl_xor((int)freed, o, 64);
        if (pbm::lane_id() == 0 && freed) atomicAdd(&sh.freed, freed);
        __syncthreads();
        if (threadIdx.x == 0 && sh.freed) pbm::atom_add_ll(d.c.free_count, (long long)sh.freed);
    }
}

// Batch tiles, then ring tiles of the first sweep: workgroup w's first tile
// is tile w (one ticket per workgroup at the start would queue them all at
// the ticket's address), its later ones go by ticket from gridDim.x.  Ring
// tiles start before the batch's victims are known.  A ring tile at or
// beyond `hold` (twice what the last batch's sweep used, and a margin) is
// held, nothing loaded, until the victims are published and the tiles
// before `hold` have counted their candidates; it is dropped when those
// already cover the victims, or the tile that holds the last victim lies
// before it.  (A dropped ticket is safe only past that tile: every tile
// before it must run for the look-back.)
__global__ __launch_bounds__(256) void k_cache_plan(CacheDev d, const int32_t* refs, uint32_t nb,
                                                    u64 tag) {
    __shared__ TileShared sh;
    // the sweep's counters are loaded at the workgroup's first ring tile:
    // a batch tile needs none of them and starts without waiting for them
    u64 arm = 0, r0 = 0, ev0 = 0;
    bool ring_ctl = false;
    u64 hold = ~0ull;  // thread 0's
    pbm::Stats st;
    for (bool first = true;; first = false) {
        if (threadIdx.x == 0) {
            const u64 t = first ? (u64)blockIdx.x
                                : (u64)gridDim.x + pbm::atom_add_u64(&d.ctl[cTicket], 1ull);
            uint32_t tile = kNoTile;
            if (t < d.miss_tiles) {
                tile = (uint32_t)t;
            } else {
                const u64 rt = t - d.miss_tiles;
                bool drop = false;
                if (rt < d.sweep_tiles && hold == ~0ull) {
                    const u64 prev = ldc(&d.ctl[cPrevTiles]);
                    hold = prev ? 2u * prev + 8u : (u64)d.sweep_tiles;
                }
                if (rt < d.sweep_tiles && rt >= hold) {
                    // the batch tile that publishes the victims, and the
                    // ring tiles before `hold`, were taken before this one
                    // and wait on no held tile
                    const u64 want = tag & kLow;
                    uint32_t spins = 0;
                    for (;;) {
                        if (ldc(&d.ctl[cSweepDone0]) < rt) break;
                        const u64 w = ldc(&d.ctl[cPlan]);
                        const u64 c = ldc(&d.lb_sweep[hold - 1u]);
                        if ((w >> 32) == want && (c >> 62) == 2u) {
                            drop = (c & kLbCount) >= (w & kLow);
                            break;
                        }
                        if (++spins > kLbSpin) {
                            pbm::atom_add_u64(&d.ctl[cScanTimeouts], 1ull);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
                if (rt < d.sweep_tiles && !drop && !(ldc(&d.ctl[cSweepDone0]) < rt))
                    tile = d.miss_tiles + (uint32_t)rt;
            }
            sh.tile = tile;
        }
        __syncthreads();
        const uint32_t tile = sh.tile;
        if (tile == kNoTile) break;
        if (tile < d.miss_tiles) {
            miss_tile(d, refs, nb, tag, tile, sh);
        } else {
            if (!ring_ctl) {
                arm = ldc(&d.ctl[cArm]);
                r0 = ldc(&d.ctl[cResident]);
                ev0 = ldc(&d.ctl[cEvictions]);
                ring_ctl = true;
            }
            sweep_tile(d, 0, tile - d.miss_tiles, kNoStop, tag, arm, r0, 0, ev0, sh, st);
        }
        __syncthreads();  // the shared words are taken again
    }
    pbm::flush_stats(d.c, st);
}

// The second sweep when the first ran short, then the loads: misses go
// into empty slots first, then into the victims' slots.
template <int S>
__global__ __launch_bounds__(256) void k_cache_load(CacheDev d) {
    __shared__ TileShared sh;
    const u64 K = ldc(&d.ctl[cK]), empty = ldc(&d.ctl[cEmpty]), r0 = ldc(&d.ctl[cR0]);
    const u64 batch_no = ldc(&d.ctl[cBatches]);
    const u64 v2 = ldc(&d.ctl[cV2]);
    const u64 z = v2 ? ldc(&d.ctl[cZ]) : 0u;
    pbm::Stats st;
    if (v2) {
        const u64 arm = ldc(&d.ctl[cArm]), ev0 = ldc(&d.ctl[cEvictions]);
        for (;;) {
            if (threadIdx.x == 0) {
                const u64 t = pbm::atom_add_u64(&d.ctl[cSweepTicket1], 1ull);
                sh.tile = (t < d.sweep_tiles && !(ldc(&d.ctl[cSweepDone1]) < t)) ? (uint32_t)t
                                                                                 : kNoTile;
            }
            __syncthreads();
            const uint32_t tile = sh.tile;
            if (tile == kNoTile) break;
            sweep_tile(d, 1, tile, v2, 0, arm, r0, z, ev0, sh, st);
            __syncthreads();
        }
    }
    // every ring tile the second sweep needs was taken by a running
    // workgroup before this one left the loop above
    const u64 stride = (u64)gridDim.x * blockDim.x;
    uint32_t fail = 0;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i - pbm::lane_id() < K; i += stride) {
        if (i >= K) continue;
        const int32_t l = d.miss[i];
        u64 s = r0 + i;
        if (i >= empty) {
            const u64 k = i - empty;
            int32_t vs;
            if (v2 && k >= z) {
                uint32_t spins = 0;
                while ((vs = __hip_atomic_load(&d.victim[k], __ATOMIC_ACQUIRE,
                                               __HIP_MEMORY_SCOPE_AGENT)) < 0) {
                    if (++spins > kLbSpin) {
                        pbm::atom_add_u64(&d.ctl[cScanTimeouts], 1ull);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            } else {
                vs = d.victim[k];
            }
            s = (u64)(uint32_t)vs;
        }
        const int32_t seed = pbm::request_seed(d.c.seed, (batch_no << 32) | i);
        const int32_t f = pbm::get_page_of<S>(d.c, (uint32_t)i, seed, nullptr, &st);
        if (s < d.cap) {
            fail += f < 0;
            d.frame[s] = f;
            d.owner[s] = l;
            d.slot_of[l] = (int32_t)s;
            d.refbit[s] = 1u;
        } else {
            ++fail;  // no slot: cannot happen while the plan holds
            if (f >= 0) pbm::free_page(d.c, f, &st);
        }
    }
    if (fail) pbm::atom_add_u64(&d.ctl[cAllocFail], (u64)fail);
    pbm::flush_stats(d.c, st);
}

__global__ void k_cache_close(CacheDev d) { close_batch(d); }

// First-miss tags start again from batch 1 (the tag field is 32 bits).
__global__ void k_cache_clear_first(CacheDev d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.n_logical) d.first[i] = 0;
}

__global__ void k_cache_init(CacheDev d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.n_logical) {
        d.slot_of[i] = -1;
        d.first[i] = 0;
    }
    if (i < d.cap) {
        d.owner[i] = -1;
        d.frame[i] = -1;
        d.refbit[i] = 0;
    }
    if (i < d.miss_tiles) d.lb_miss[i] = 0;
    if (i < 2u * d.sweep_tiles) d.lb_sweep[i] = 0;
    if (i < cCtlWords) d.ctl[i] = 0;
}

// Zipf trace: u from the top 53 bits of a splitmix64 hash of (seed, i),
// page = the first k with cdf[k] > u * cdf[n-1].
__global__ __launch_bounds__(256) void k_zipf(const double* cdf, uint32_t n, u64 seed, u64 first,
                                              u64 count, int32_t* out) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u64 h = pbm::mix64((seed ^ 0xBE5466CF34E90C6Cull) + (first + i + 1ull) * pbm::kGolden);
    const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
    const double t = __dmul_rn(u, cdf[n - 1]);
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (cdf[mid] > t) hi = mid; else lo = mid + 1u;
    }
    out[i] = (int32_t)lo;
}

}  // namespace

struct pbm_cache {
    pbm_handle* h = nullptr;
    pbm_cache_config cfg{};
    CacheDev d{};
    std::vector<void*> mem;
    u64 tag = 0;  // tag of the last batch launched (1 .. 2^32 - 1)
};

extern "C" int pbm_cache_destroy(pbm_cache* k) {
    if (!k) return PBM_OK;
    if (k->h) (void)pbmi::use_device(k->h);
    (void)hipDeviceSynchronize();
    for (void* p : k->mem)
        if (p) (void)hipFree(p);
    delete k;
    return PBM_OK;
}

extern "C" int pbm_cache_create(pbm_handle* h, const pbm_cache_config* cfg, pbm_cache** out) {
    PBM_CHECK_HANDLE(h);
    if (!cfg || !out) return pbmi::fail(PBM_ERR_INVALID, "null cache config or output");
    *out = nullptr;
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "the cache's frames need a bitmap strategy");
    if (cfg->n_logical == 0 || cfg->n_logical > (1u << 30))
        return pbmi::fail(PBM_ERR_RANGE, "n_logical must be in [1, 2^30]");
    if (cfg->capacity == 0 || cfg->capacity > h->ctx.n_pages)
        return pbmi::fail(PBM_ERR_RANGE, "capacity must be in [1, pool pages]");
    if (cfg->batch == 0 || cfg->batch > cfg->capacity)
        return pbmi::fail(PBM_ERR_RANGE, "batch must be in [1, capacity]");
    pbm_cache* k = new pbm_cache();
    k->h = h;
    k->cfg = *cfg;
    CacheDev& d = k->d;
    d.c = h->ctx;
    d.n_logical = cfg->n_logical;
    d.cap = cfg->capacity;
    d.batch = cfg->batch;
    d.log_cap = cfg->log_capacity;
    d.miss_tiles = (cfg->batch + kTile - 1u) / kTile;
    d.sweep_tiles = (cfg->capacity + kTile - 1u) / kTile;
    d.load_groups = std::min((cfg->batch + 255u) / 256u, kGroups);
    d.plan_groups = d.miss_tiles + std::min(d.sweep_tiles, kGroups);
    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8u);
        if (e != hipSuccess) return pbmi::hip_fail(e, "hipMalloc(cache)", __FILE__, __LINE__);
        k->mem.push_back(*p);
        return PBM_OK;
    };
    int rc = PBM_OK;
    void* p = nullptr;
#define PBM_CALLOC(field, type, count)                  \
    if (!rc) {                                          \
        rc = alloc(&p, (size_t)(count) * sizeof(type)); \
        if (!rc) d.field = (type*)p;                    \
    }
    PBM_CALLOC(slot_of, int32_t, cfg->n_logical)
    PBM_CALLOC(first, u64, cfg->n_logical)
    PBM_CALLOC(owner, int32_t, cfg->capacity)
    PBM_CALLOC(frame, int32_t, cfg->capacity)
    PBM_CALLOC(refbit, uint32_t, cfg->capacity)
    PBM_CALLOC(miss, int32_t, cfg->batch)
    PBM_CALLOC(victim, int32_t, cfg->batch)
    PBM_CALLOC(lb_miss, u64, d.miss_tiles)
    PBM_CALLOC(lb_sweep, u64, 2u * d.sweep_tiles)
    PBM_CALLOC(log, int32_t, cfg->log_capacity)
    PBM_CALLOC(ctl, u64, cCtlWords)
#undef PBM_CALLOC
    if (rc) {
        pbm_cache_destroy(k);
        return rc;
    }
    // (miss_tiles <= 2 * sweep_tiles <= capacity + 1)
    const uint32_t n_init =
        std::max(std::max(cfg->n_logical, cfg->capacity + 1u), (uint32_t)cCtlWords);
    hipLaunchKernelGGL(k_cache_init, dim3((n_init + 255u) / 256u), dim3(256), 0, nullptr, d);
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        pbm_cache_destroy(k);
        return pbmi::hip_fail(e, "cache init", __FILE__, __LINE__);
    }
    *out = k;
    return PBM_OK;
}

namespace {

// A launch that records e0 / e1 at the kernel's own start and end when
// they are given.
template <typename K, typename... A>
void launch_ev(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t e0, hipEvent_t e1, A... a) {
    if (e0 || e1) hipExtLaunchKernelGGL(kernel, g, b, 0, st, e0, e1, 0, a...);
    else hipLaunchKernelGGL(kernel, g, b, 0, st, a...);
}

// One batch: hits, plan, load.  With ev (six events), kernel k records
// ev[2k] and ev[2k+1] at its own start and end.
int cache_batch(pbm_cache* k, const int32_t* refs, uint32_t nb, hipStream_t s,
                hipEvent_t* ev = nullptr) {
    const CacheDev& d = k->d;
    const dim3 b(256);
    if (k->tag == kLow) {
        // the next tag would not fit its 32 bits: every first-miss word is
        // cleared and tags start again from 1
        hipLaunchKernelGGL(k_cache_clear_first, dim3((d.n_logical + 255u) / 256u), b, 0, s, d);
        k->tag = 0;
    }
    const u64 tag = ++k->tag;
    hipEvent_t e[6] = {};
    if (ev)
        for (int i = 0; i < 6; ++i) e[i] = ev[i];
    launch_ev(k_cache_hits, dim3((nb + 255u) / 256u), b, s, e[0], e[1], d, refs, nb, tag);
    launch_ev(k_cache_plan, dim3(d.plan_groups), b, s, e[2], e[3], d, refs, nb, tag);
    switch (k->h->cfg.strategy) {
#define PBM_LOAD_CASE(S) \
    case S: launch_ev(k_cache_load<S>, dim3(d.load_groups), b, s, e[4], e[5], d); break;
        PBM_LOAD_CASE(PBM_RW_BM)
        PBM_LOAD_CASE(PBM_COLLAB_RW_BM)
        PBM_LOAD_CASE(PBM_CRW_BM)
        PBM_LOAD_CASE(PBM_RW_MALLOC)
        PBM_LOAD_CASE(PBM_SINGLE_CLOCK)
        PBM_LOAD_CASE(PBM_DIST_CLOCK)
        PBM_LOAD_CASE(PBM_RW_BASIC)
        PBM_LOAD_CASE(PBM_CRW_BASIC)
        PBM_LOAD_CASE(PBM_COLLAB_RW_BASIC)
#undef PBM_LOAD_CASE
        default: return pbmi::fail(PBM_ERR_UNSUPPORTED, "no cache load for this strategy");
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_cache_run(pbm_cache* k, const int32_t* d_refs, uint64_t n, void* stream) {
    if (!k) return pbmi::fail(PBM_ERR_INVALID, "null cache");
    PBM_CHECK_HANDLE(k->h);
    if (n && !d_refs) return pbmi::fail(PBM_ERR_INVALID, "null references");
    hipStream_t s = pbmi::as_stream(stream);
    for (uint64_t b0 = 0; b0 < n; b0 += k->d.batch) {
        const uint32_t nb = (uint32_t)(n - b0 < k->d.batch ? n - b0 : k->d.batch);
        int rc = cache_batch(k, d_refs + b0, nb, s);
        if (rc) return rc;
    }
    if (n) {
        hipLaunchKernelGGL(k_cache_close, dim3(1), dim3(1), 0, s, k->d);
        PBM_HIP(hipGetLastError());
    }
    return PBM_OK;
}

extern "C" int pbm_cache_run_timed(pbm_cache* k, const int32_t* d_refs, uint64_t n,
                                   float* kernel_ms, void* stream) {
    if (!k || !kernel_ms) return pbmi::fail(PBM_ERR_INVALID, "null cache or output");
    PBM_CHECK_HANDLE(k->h);
    if (n && !d_refs) return pbmi::fail(PBM_ERR_INVALID, "null references");
    const uint64_t nbat = (n + k->d.batch - 1u) / k->d.batch;
    if (nbat == 0 || nbat > 4096u)
        return pbmi::fail(PBM_ERR_RANGE, "pbm_cache_run_timed runs 1 .. 4096 batches");
    hipStream_t s = pbmi::as_stream(stream);
    std::vector<hipEvent_t> ev((size_t)nbat * 6u, nullptr);
    int rc = PBM_OK;
    for (hipEvent_t& e : ev) {
        const hipError_t he = hipEventCreate(&e);
        if (he != hipSuccess) {
            rc = pbmi::hip_fail(he, "hipEventCreate", __FILE__, __LINE__);
            break;
        }
    }
    for (uint64_t q = 0; q < nbat && !rc; ++q) {
        const uint64_t b0 = q * k->d.batch;
        const uint32_t nb = (uint32_t)(n - b0 < k->d.batch ? n - b0 : k->d.batch);
        rc = cache_batch(k, d_refs + b0, nb, s, &ev[q * 6u]);
    }
    if (!rc) {
        hipLaunchKernelGGL(k_cache_close, dim3(1), dim3(1), 0, s, k->d);
        hipError_t he = hipGetLastError();
        if (he == hipSuccess) he = hipStreamSynchronize(s);
        if (he != hipSuccess) rc = pbmi::hip_fail(he, "cache timed run", __FILE__, __LINE__);
    }
    double sum[3] = {0, 0, 0};
    for (uint64_t q = 0; q < nbat && !rc; ++q)
        for (int j = 0; j < 3 && !rc; ++j) {
            float ms = 0;
            const hipError_t he = hipEventElapsedTime(&ms, ev[q * 6u + 2u * j], ev[q * 6u + 2u * j + 1u]);
            if (he != hipSuccess) rc = pbmi::hip_fail(he, "hipEventElapsedTime", __FILE__, __LINE__);
            sum[j] += ms;
        }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int j = 0; j < 3; ++j) kernel_ms[j] = (float)(sum[j] / (double)nbat);
    return PBM_OK;
}

extern "C" int pbm_cache_read_stats(pbm_cache* k, pbm_cache_stats* out, void* stream) {
    if (!k || !out) return pbmi::fail(PBM_ERR_INVALID, "null cache or output");
    PBM_CHECK_HANDLE(k->h);
    u64 ctl[cCtlWords];
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(ctl, k->d.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    out->hits = ctl[cHits];
    out->misses = ctl[cMisses];
    out->evictions = ctl[cEvictions];
    out->resident = ctl[cResident];
    out->arm = ctl[cArm];
    out->batches = ctl[cBatches];
    out->alloc_failures = ctl[cAllocFail];
    out->scan_timeouts = ctl[cScanTimeouts];
    out->invalid_refs = ctl[cInvalid];
    return PBM_OK;
}

extern "C" int pbm_cache_read_evictions(pbm_cache* k, uint64_t first, uint32_t n, int32_t* host,
                                        void* stream) {
    if (!k || (n && !host)) return pbmi::fail(PBM_ERR_INVALID, "null cache or output");
    PBM_CHECK_HANDLE(k->h);
    if (first + n > k->d.log_cap) return pbmi::fail(PBM_ERR_RANGE, "beyond the eviction log");
    if (!n) return PBM_OK;
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(host, k->d.log + first, (size_t)n * 4u, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}

extern "C" int pbm_cache_read_slots(pbm_cache* k, int32_t* owner, int32_t* frame, uint32_t* refbit,
                                    void* stream) {
    if (!k) return pbmi::fail(PBM_ERR_INVALID, "null cache");
    PBM_CHECK_HANDLE(k->h);
    hipStream_t s = pbmi::as_stream(stream);
    const size_t bytes = (size_t)k->d.cap * 4u;
    if (owner) PBM_HIP(hipMemcpyAsync(owner, k->d.owner, bytes, hipMemcpyDeviceToHost, s));
    if (frame) PBM_HIP(hipMemcpyAsync(frame, k->d.frame, bytes, hipMemcpyDeviceToHost, s));
    if (refbit) PBM_HIP(hipMemcpyAsync(refbit, k->d.refbit, bytes, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}

// ---- Zipf trace without allocation: the CDF lives in a static scratch
// array, one copy per device (32 MiB: up to 2^22 logical pages), which each
// device keeps until the process ends.  In host code the array's name is
// not a device address: every use goes through the address
// hipGetSymbolAddress returned for the device (zipf_scratch_ptr).
constexpr uint32_t kZipfScratchPages = 1u << 22;
constexpr uint32_t kZipfStage = 1u << 15;  // CDF entries uploaded per copy
constexpr int kZipfMaxDevices = 64;
static __device__ __align__(256) double zipf_scratch[kZipfScratchPages];

namespace {

std::mutex g_zipf_mu;                        // guards the two arrays below
double* g_zipf_ptr[kZipfMaxDevices] = {};    // zipf_scratch of each device
double g_zipf_stage[kZipfStage];             // host staging of the CDF

// The device address of zipf_scratch on `device` (current), looked up once.
int zipf_scratch_ptr(int device, double** out) {
    if (!g_zipf_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(zipf_scratch));
        if (e != hipSuccess || !p) return pbmi::hip_fail(e, "hipGetSymbolAddress(zipf_scratch)",
                                                         __FILE__, __LINE__);
        g_zipf_ptr[device] = (double*)p;
    }
    *out = g_zipf_ptr[device];
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_zipf_trace_noalloc(uint32_t n_logical, double s, uint64_t seed,
                                      uint64_t first, uint64_t count, int32_t* d_out, int device,
                                      void* stream) {
    if (n_logical == 0 || !(s > 0.0)) return pbmi::fail(PBM_ERR_INVALID, "need n_logical > 0, s > 0");
    if (count && !d_out) return pbmi::fail(PBM_ERR_INVALID, "null output");
    if (n_logical > kZipfScratchPages)
        return pbmi::fail(PBM_ERR_RANGE, "pbm_zipf_trace_noalloc holds the CDF of at most "
                                         "4,194,304 logical pages in its static scratch");
    if (device < 0 || device >= kZipfMaxDevices)
        return pbmi::fail(PBM_ERR_INVALID, "device ordinal out of range");
    if (!count) return PBM_OK;
    PBM_HIP(hipSetDevice(device));
    hipStream_t st = pbmi::as_stream(stream);
    std::lock_guard<std::mutex> lock(g_zipf_mu);
    double* cdf = nullptr;
    int rc = zipf_scratch_ptr(device, &cdf);
    if (rc) return rc;
    // the same running sum as pbm_zipf_trace, uploaded a stage at a time;
    // every entry the kernel reads (0 .. n_logical-1) is written here, so
    // what an earlier call left behind is never read
    double acc = 0;
    for (uint32_t k0 = 0; k0 < n_logical; k0 += kZipfStage) {
        const uint32_t m = n_logical - k0 < kZipfStage ? n_logical - k0 : kZipfStage;
        for (uint32_t i = 0; i < m; ++i) {
            acc += std::pow((double)(k0 + i) + 1.0, -s);
            g_zipf_stage[i] = acc;
        }
        PBM_HIP(hipMemcpyAsync(cdf + k0, g_zipf_stage, (size_t)m * 8u, hipMemcpyHostToDevice, st));
        PBM_HIP(hipStreamSynchronize(st));  // the stage is refilled next
    }
    const u64 blocks = (count + 255u) / 256u;
    hipLaunchKernelGGL(k_zipf, dim3((uint32_t)blocks), dim3(256), 0, st, cdf, n_logical, seed,
                       first, count, d_out);
    PBM_HIP(hipGetLastError());
    // the next call on this device rewrites the CDF: it must not overlap
    // this kernel
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_zipf_trace(uint32_t n_logical, double s, uint64_t seed, uint64_t first,
                              uint64_t count, int32_t* d_out, int device, void* stream) {
    if (n_logical == 0 || !(s > 0.0)) return pbmi::fail(PBM_ERR_INVALID, "need n_logical > 0, s > 0");
    if (count && !d_out) return pbmi::fail(PBM_ERR_INVALID, "null output");
    if (!count) return PBM_OK;
    PBM_HIP(hipSetDevice(device));
    std::vector<double> cdf(n_logical);
    double acc = 0;
    for (uint32_t k = 0; k < n_logical; ++k) {
        acc += std::pow((double)k + 1.0, -s);
        cdf[k] = acc;
    }
    double* d_cdf = nullptr;
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMalloc((void**)&d_cdf, (size_t)n_logical * 8u));
    hipError_t e = hipMemcpyAsync(d_cdf, cdf.data(), (size_t)n_logical * 8u, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const u64 blocks = (count + 255u) / 256u;
        hipLaunchKernelGGL(k_zipf, dim3((uint32_t)blocks), dim3(256), 0, st, d_cdf, n_logical, seed,
                           first, count, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_cdf);
    if (e != hipSuccess) return pbmi::hip_fail(e, "zipf trace", __FILE__, __LINE__);
    return PBM_OK;
}
