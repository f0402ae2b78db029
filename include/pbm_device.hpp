/*
 * pbm_device.hpp — header-only device API of the page manager (gfx950, wave64).
 *
 * Include from a HIP kernel and pass the pool's pbm_ctx (pbm_get_ctx) by
 * value.  This replaces the reference's __device__ API, which reads
 * file-static globals (RandomWalkBasic_BM.cuh:14-21, CollabRW_BM.cuh:14-21,
 * CRW_BM.cuh:14-21, RWMalloc.cuh:14-22, parallelPage.cuh:7-75).
 *
 *   pbm::get_page_rwbm(ctx, seed, &steps)      getPage        RandomWalkBasic_BM.cuh:60-89
 *   pbm::get_page_collab(ctx, seed, &steps)    getPage        CollabRW_BM.cuh:114-160 (wave-collective)
 *   pbm::free_page(ctx, id)                    freePage       RandomWalkBasic_BM.cuh:98-102
 *   pbm::page_address(ctx, id)                 pageAddress    Paging.cuh:30-38
 *   pbm::page_id(ctx, ptr)                     getPageID      Paging.cuh:40-52
 *
 * Failure handling differs from the reference on purpose: the reference
 * walks forever when the pool is full (RandomWalkBasic_BM.cuh:67-85) or
 * traps (RandomWalkBasic.cuh:66-69).  Here every grant first takes a unit
 * of the reservation counter ctx.free_count; a request that finds it empty
 * returns -1 at once, and a reserved walk is bounded by walk_budget random
 * jumps followed by a linear sweep, so no call spins without end.
 *
 * Atomics are device (agent) scope and relaxed, as the reference's
 * atomicOr/atomicAnd are.  Code that writes page contents in one thread and
 * frees the page for another thread of the same kernel to reuse must put a
 * __threadfence() before free_page.
 */
#ifndef PBM_DEVICE_HPP
#define PBM_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pbm.h"

namespace pbm {

typedef unsigned long long u64;
constexpr u64 kFull = ~0ull;
constexpr u64 kGolden = 0x9E3779B97F4A7C15ull;
constexpr u64 kPrefillSalt = 0x243F6A8885A308D3ull;
constexpr u64 kWaveSalt = 0x13198A2E03707344ull;
constexpr u64 kArmSalt = 0xA4093822299F31D0ull;
constexpr u64 kPayloadSalt = 0x452821E638D01377ull;
/* Threads per tile of the single-pass rank scan (ctx.scan_threads: one
 * wave on small maps); a tile covers ctx.scan_threads * ctx.scan_wpt map
 * words. */
constexpr uint32_t kScanThreads = 256;

/* ---------------------------------------------------------------------
 * Random numbers
 * ------------------------------------------------------------------- */

/* RNG_LCG (RNG_LCG.cuh:6-12): x' = (1103515245*x + 12345) % 2^31 computed
 * in 64-bit with C++ truncating modulo, so a negative product gives a
 * negative result in (-2^31, 0]. */
__host__ __device__ __forceinline__ int32_t lcg(int32_t x) {
    long long v = 1103515245LL * (long long)x + 12345LL;
    return (int32_t)(v % 2147483648LL);
}

/* splitmix64 output function (Steele, Lea, Flood 2014). */
__host__ __device__ __forceinline__ u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Walk seed of request r: the high half of output r+1 of splitmix64 seeded
 * with `seed`.  Replaces the reference's (tid<<15)+(int16)clock() seeds
 * (RandomWalkBasic_BM.cuh:62-63), which overflow for tid >= 65536 and are
 * not reproducible. */
__host__ __device__ __forceinline__ int32_t request_seed(u64 seed, u64 r) {
    return (int32_t)(uint32_t)(mix64(seed + (r + 1) * kGolden) >> 32);
}

/* Map an LCG state to [0, L) with its high bits.  The reference reduces
 * with (unsigned)state % L (RandomWalkBasic_BM.cuh:71), which for a
 * power-of-two L keeps the low bits: those follow x' = a*x + c mod L, one
 * cycle shared by every walk, so concurrent walks trail each other through
 * the same words.  The 31-bit residue's high bits do not. */
__host__ __device__ __forceinline__ uint32_t lcg_to_range(int32_t state, uint32_t L) {
    return (uint32_t)(((unsigned long long)((uint32_t)state & 0x7FFFFFFFu) * L) >> 31);
}

/* Pre-fill hash of page i: output i+1 of splitmix64 seeded with
 * mix64(seed ^ kPrefillSalt). */
__host__ __device__ __forceinline__ u64 page_hash(u64 seed, u64 i) {
    return mix64(mix64(seed ^ kPrefillSalt) + (i + 1) * kGolden);
}

/* Payload the serving loop writes at the start of page j of sequence seq
 * (and checks when the sequence ends): depends on (seed, seq, j) only, so
 * a sequence's checksum is the same whatever else runs beside it. */
__host__ __device__ __forceinline__ u64 serve_payload(u64 seed, uint32_t seq, uint32_t j) {
    return mix64((seed ^ kPayloadSalt) + ((((u64)seq << 32) | j) + 1ull) * kGolden);
}

/* ---------------------------------------------------------------------
 * Wave64 helpers (all safe when only some lanes of the wave are active:
 * they use ballots and read only lanes known to be active)
 * ------------------------------------------------------------------- */

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

/* popcount of `mask` restricted to lanes below the calling lane. */
__device__ __forceinline__ unsigned rank_in(u64 mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ u64 ballot(bool p) { return __ballot(p ? 1 : 0); }

/* Value of v in lane l (l uniform across the wave, lane l active). */
__device__ __forceinline__ uint32_t readlane(uint32_t v, unsigned l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ u64 readlane64(u64 v, unsigned l) {
    return ((u64)readlane((uint32_t)(v >> 32), l) << 32) | readlane((uint32_t)v, l);
}

/* Value of v in lane src (per-lane src, lane src active). */
__device__ __forceinline__ uint32_t shfl32(uint32_t v, unsigned src) {
    return (uint32_t)__shfl((int)v, (int)src, 64);
}
__device__ __forceinline__ u64 shfl64(u64 v, unsigned src) {
    return ((u64)shfl32((uint32_t)(v >> 32), src) << 32) | shfl32((uint32_t)v, src);
}

/* Exclusive prefix sum and wave total of a 7-bit per-lane value (0..127)
 * over the active lanes, built from seven ballots. */
__device__ __forceinline__ unsigned scan7(unsigned v, unsigned& total) {
    unsigned pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) {
        u64 m = ballot((v >> b) & 1u);
        pre += rank_in(m) << b;
        tot += (unsigned)__popcll(m) << b;
    }
    total = tot;
    return pre;
}

/* Position of the k-th (0-based) set bit of x; requires k < popcount(x). */
__host__ __device__ __forceinline__ unsigned select_bit(u64 x, unsigned k) {
    unsigned pos = 0;
    unsigned c = (unsigned)__builtin_popcount((uint32_t)x);
    if (k >= c) { k -= c; x >>= 32; pos = 32; }
    uint32_t y = (uint32_t)x;
    c = (unsigned)__builtin_popcount(y & 0xFFFFu); if (k >= c) { k -= c; y >>= 16; pos += 16; }
    c = (unsigned)__builtin_popcount(y & 0xFFu);   if (k >= c) { k -= c; y >>= 8;  pos += 8; }
    c = (unsigned)__builtin_popcount(y & 0xFu);    if (k >= c) { k -= c; y >>= 4;  pos += 4; }
    c = (unsigned)__builtin_popcount(y & 0x3u);    if (k >= c) { k -= c; y >>= 2;  pos += 2; }
    c = y & 1u; if (k >= c) pos += 1;
    return pos;
}

/* The k lowest set bits of x (all of x when k >= popcount(x)). */
__host__ __device__ __forceinline__ u64 lowest_bits(u64 x, unsigned k) {
    if (k == 0) return 0;
    if (k >= (unsigned)__builtin_popcountll(x)) return x;
    return x & ((1ull << select_bit(x, k)) - 1ull);
}

/* ---------------------------------------------------------------------
 * Page-map access: relaxed, device scope
 * ------------------------------------------------------------------- */

__device__ __forceinline__ u64 ld_word(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 atom_or(u64* p, u64 m) {
    return __hip_atomic_fetch_or(p, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 atom_and(u64* p, u64 m) {
    return __hip_atomic_fetch_and(p, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long atom_add_ll(long long* p, long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 atom_add_u64(u64* p, u64 v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 atom_max_u64(u64* p, u64 v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Per-thread event tally; batch kernels sum it over the wave and add it to
 * ctx.counters once (flush_stats). */
struct Stats {
    uint32_t allocs = 0, frees = 0, failures = 0, steps = 0;
    uint32_t probes = 0, atomics = 0, fallbacks = 0, errors = 0;
};

/* Adds the wave's Stats to the pool counters: one atomic per counter per
 * wave when the whole wave is active, per lane otherwise. */
__device__ inline void flush_stats(const pbm_ctx& c, const Stats& s) {
    if (!c.counters) return;
    uint32_t v[8] = {s.allocs, s.frees, s.failures, s.steps,
                     s.probes, s.atomics, s.fallbacks, s.errors};
    if (ballot(true) != kFull) {
        u64* base = (u64*)c.counters;
        for (int i = 0; i < 8; ++i)
            if (v[i]) atom_add_u64(base + i, (u64)v[i]);
        return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint32_t x = v[i];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += (uint32_t)__shfl_xor((int)x, d, 64);
        v[i] = x;
    }
    if (lane_id() == 0) {
        u64* base = (u64*)c.counters;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (v[i]) atom_add_u64(base + i, (u64)v[i]);
    }
}

/* ---------------------------------------------------------------------
 * Reservation counter
 * ------------------------------------------------------------------- */

/* Takes one unit of ctx.free_count for every active lane with want=true,
 * one atomic per wave.  Lanes beyond the available count are refused and
 * the excess is given back.  Exactly min(requests, available) grants are
 * made over any sequence of calls that only allocate. */
__device__ inline bool reserve_one(const pbm_ctx& c, bool want) {
    u64 m = ballot(want);
    if (m == 0) return false;
    unsigned cnt = (unsigned)__popcll(m);
    unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
    long long old = 0;
    if (lane_id() == leader) old = atom_add_ll(c.free_count, -(long long)cnt);
    old = (long long)readlane64((u64)old, leader);
    long long granted = old < 0 ? 0 : (old > (long long)cnt ? (long long)cnt : old);
    if (lane_id() == leader && granted < (long long)cnt)
        atom_add_ll(c.free_count, (long long)cnt - granted);
    return want && (long long)rank_in(m) < granted;
}

/* Returns one unit to ctx.free_count for every active lane with give=true,
 * one atomic per wave. */
__device__ inline void release_one(const pbm_ctx& c, bool give) {
    u64 m = ballot(give);
    if (m == 0) return;
    unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
    if (lane_id() == leader) atom_add_ll(c.free_count, (long long)__popcll(m));
}

/* Returns n units to ctx.free_count for every active lane (n may differ by
 * lane, and be 0), one atomic per wave. */
__device__ inline void release_n(const pbm_ctx& c, uint32_t n) {
    const u64 act = ballot(n > 0u);
    if (act == 0) return;
    // the wave's sum, one ballot per bit of n
    u64 total = 0;
    for (unsigned b = 0; b < 32u; ++b) {
        if (!ballot((n >> b) != 0u)) break;
        total += (u64)__popcll(ballot((n >> b) & 1u)) << b;
    }
    if (lane_id() == (unsigned)__ffsll((long long)act) - 1u) atom_add_ll(c.free_count, (long long)total);
}

/* ---------------------------------------------------------------------
 * Paging (Paging.cuh:30-52): one contiguous pool, O(1) both ways
 * ------------------------------------------------------------------- */

__device__ __forceinline__ void* page_address(const pbm_ctx& c, int32_t id) {
    if (!c.pool || id < 0 || (uint32_t)id >= c.n_pages) return nullptr;
    return (void*)(c.pool + (size_t)(uint32_t)id * c.page_size);
}

__device__ __forceinline__ int32_t page_id(const pbm_ctx& c, const void* p) {
    if (!c.pool) return -1;
    const char* q = (const char*)p;
    if (q < c.pool) return -1;
    size_t off = (size_t)(q - c.pool);
    size_t id = off / c.page_size;
    return id < c.n_pages ? (int32_t)id : -1;
}

/* ---------------------------------------------------------------------
 * RandomWalkBasic_BM (RandomWalkBasic_BM.cuh:60-89) on 64-bit words
 * ------------------------------------------------------------------- */

/* Claims the lowest free bit of word g, retrying while the word is not
 * full; returns the page id or -1 when the word is (or becomes) full. */
__device__ __forceinline__ int32_t claim_in_word(const pbm_ctx& c, uint32_t g, u64 w, Stats& st) {
    while (w != kFull) {
        unsigned b = (unsigned)__ffsll((long long)~w) - 1u;
        u64 bit = 1ull << b;
        u64 old = atom_or(&c.bitmap[g], bit);
        st.atomics++;
        if (!(old & bit)) return (int32_t)(g * 64u + b);
        w = old | bit;
    }
    return -1;
}

/* Linear sweep from word `from`+1 over at most `passes` rounds of the map.
 * A reserved walk in a phase without frees finds a page within one pass. */
__device__ inline int32_t sweep_claim(const pbm_ctx& c, uint32_t from, uint32_t& steps,
                                      Stats& st, uint32_t passes) {
    const uint32_t L = c.n_words;
    uint32_t g = from;
    for (uint64_t k = 0; k < (uint64_t)passes * L; ++k) {
        g = (g + 1u == L) ? 0u : g + 1u;
        steps++;
        u64 w = ld_word(&c.bitmap[g]);
        st.probes++;
        if (w == kFull) continue;
        int32_t id = claim_in_word(c, g, w, st);
        if (id >= 0) return id;
    }
    return -1;
}

/* Random walk of one reserved request.  Step = one random jump, as
 * RandomWalkBasic_BM.cuh:68 counts it.  The page id is g*64+b (the
 * reference computes g<<(5+b) by operator precedence, :81). */
__device__ inline int32_t walk_rwbm(const pbm_ctx& c, int32_t& seed, uint32_t budget,
                                    uint32_t& steps, Stats& st) {
    const uint32_t L = c.n_words;
    uint32_t g = 0;
    for (uint32_t j = 0; j < budget; ++j) {
        steps++;
        seed = lcg(seed);
        g = (uint32_t)seed % L;
        u64 w = ld_word(&c.bitmap[g]);
        st.probes++;
        int32_t id = claim_in_word(c, g, w, st);
        if (id >= 0) return id;
    }
    st.fallbacks++;
    return sweep_claim(c, g, steps, st, 4u);
}

/* getPage of RandomWalkBasic_BM: reserve, then walk.  Returns -1 when the
 * pool has no page for this request. */
__device__ inline int32_t get_page_rwbm(const pbm_ctx& c, int32_t seed, uint32_t* steps_out,
                                        Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    uint32_t steps = 0;
    int32_t id = -1;
    if (reserve_one(c, true)) {
        id = walk_rwbm(c, seed, c.walk_budget, steps, st);
        if (id < 0) { atom_add_ll(c.free_count, 1); st.errors++; }
    }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* Page walk of RandomWalkBasic from LCG state `seed`: probe pages
 * (unsigned)lcg^k(seed) % n_pages until one is claimed, a step per probe;
 * after walk_budget probes, sweep from the last probed page's word. */
__device__ inline int32_t walk_pages(const pbm_ctx& c, int32_t seed, uint32_t& steps,
                                     Stats& st) {
    uint32_t p = 0;
    for (uint32_t j = 0; j < c.walk_budget; ++j) {
        seed = lcg(seed);
        p = (uint32_t)seed % c.n_pages;
        steps++;
        st.probes++;
        const u64 bit = 1ull << (p & 63u);
        const u64 old = atom_or(&c.bitmap[p >> 6], bit);
        st.atomics++;
        if (!(old & bit)) return (int32_t)p;
    }
    st.fallbacks++;
    return sweep_claim(c, p >> 6, steps, st, 4u);
}

/* getPage of RandomWalkBasic (RandomWalkBasic.cuh:54-80): a walk over
 * single pages.  Each probe takes the next LCG state, page
 * (unsigned)seed % n_pages, and claims it if its bit was clear; a step is
 * one probe, the first included (:59-65).  The reference keeps one int per
 * page (atomicExch to 1) and traps after n_pages/2 probes; here the pages
 * are bits of the shared map and a walk that spends its budget sweeps. */
__device__ inline int32_t get_page_rwbasic(const pbm_ctx& c, int32_t seed, uint32_t* steps_out,
                                           Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    uint32_t steps = 0;
    int32_t id = -1;
    if (reserve_one(c, true)) {
        id = walk_pages(c, seed, steps, st);
        if (id < 0) { atom_add_ll(c.free_count, 1); st.errors++; }
    }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* freePage (RandomWalkBasic_BM.cuh:98-102): clears the page's bit and
 * returns a reservation unit.  Freeing a free page or an id out of range
 * changes nothing and counts an error. */
__device__ inline void free_page(const pbm_ctx& c, int32_t id, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    bool ok = false;
    if (id >= 0 && (uint32_t)id < c.n_pages) {
        u64 bit = 1ull << ((uint32_t)id & 63u);
        u64 old = atom_and(&c.bitmap[(uint32_t)id >> 6], ~bit);
        st.atomics++;
        ok = (old & bit) != 0;
    }
    if (ok) st.frees++; else st.errors++;
    release_one(c, ok);
}

/* ---------------------------------------------------------------------
 * Collaborative window step (CollabRW_BM.cuh:114-160, wave64 rebuild)
 *
 * Probing lanes each read one word of the window; `need` lanes ask for a
 * page.  The lanes whose words the order reaches before the needy count
 * (lane-major, PBM_ORDER_PACKED, or round-robin over words,
 * PBM_ORDER_REFERENCE, the order of the reference's has/need ballot loop,
 * :140-153) each take all free bits they saw with one atomicOr and own
 * free & ~old, whatever no racing wave took first; the owned bits go to
 * the needy lanes in the same order and the unused ones are given back
 * with one atomicAnd.  Where the reference exchanges every word of the
 * window to all-ones and releases leftovers with atomicExch(&w, ~r)
 * (:126-128, :155-156), which overwrites concurrent frees, only the words
 * needed are touched and the release is an atomicAnd.  A needy lane left
 * without a bit (its bits went to a racing wave) stays needy.
 * ------------------------------------------------------------------- */

struct WindowResult {
    unsigned seen_free;   /* free bits the probing lanes read (uniform)       */
    unsigned designated;  /* bits designated to needy lanes (uniform)         */
};

/* One step over the words held by probing lanes: `word` is the lane's word
 * index and `w` the value it read (kFull for lanes that do not probe).  On
 * return, a needy lane that got a page has need=false and id set, and
 * *w_after (when given) is the lane's word as its own claim left it (w when
 * the lane claimed nothing). */
__device__ inline WindowResult window_claim(const pbm_ctx& c, uint32_t word, u64 w,
                                            bool& need, int32_t& id, int32_t order, Stats& st,
                                            u64* w_after = nullptr) {
    if (w_after) *w_after = w;
    WindowResult res{0u, 0u};
    const u64 fr = ~w;
    const unsigned cnt = (unsigned)__popcll(fr);
    const u64 needy = ballot(need);
    const unsigned n_need = (unsigned)__popcll(needy);
    const unsigned q = rank_in(needy);
    unsigned total = 0;
    const unsigned pre = scan7(cnt, total);
    res.seen_free = total;
    if (n_need == 0 || total == 0) return res;
    res.designated = n_need < total ? n_need : total;

    // Givers: the lanes the order reaches before n_need bits.  Round-robin
    // reaches a lane only if round 0 does (a later round needs every lane
    // of round 0 to have given).
    bool giver;
    if (order == PBM_ORDER_REFERENCE) {
        const u64 B0 = ballot(cnt > 0);
        giver = cnt > 0 && rank_in(B0) < n_need;
    } else {
        giver = cnt > 0 && pre < n_need;
    }
    // Grab every free bit a giver saw; own what no racing wave took first.
    u64 owned = 0, old = 0;
    if (giver) {
        old = atom_or(&c.bitmap[word], fr);
        st.atomics++;
        owned = fr & ~old;
    }
    // Hand the owned bits to the needy lanes in the order.
    const unsigned oc = (unsigned)__popcll(owned);
    unsigned used = 0, src = 0, gbit = 0;
    bool assigned = false;
    if (order == PBM_ORDER_REFERENCE) {
        unsigned H = 0;
        for (unsigned j = 0; j < 64u && H < n_need; ++j) {
            const u64 B = ballot(oc > j);
            const unsigned h = (unsigned)__popcll(B);
            if (h == 0) break;
            const unsigned take = (n_need - H < h) ? n_need - H : h;
            if (oc > j && rank_in(B) < take) used++;
            if (need && q >= H && q < H + take) {
                src = select_bit(B, q - H);
                gbit = j;
                assigned = true;
            }
            H += take;
        }
    } else {
        unsigned o_tot = 0;
        const unsigned opre = scan7(oc, o_tot);
        const unsigned avail = n_need > opre ? n_need - opre : 0u;
        used = oc < avail ? oc : avail;
        u64 G = ballot(used > 0);
        unsigned base = 0;
        while (G) {
            const unsigned g = (unsigned)__ffsll((long long)G) - 1u;
            G &= G - 1ull;
            const unsigned ug = readlane(used, g);
            if (need && q >= base && q < base + ug) {
                src = g;
                gbit = q - base;
                assigned = true;
            }
            base += ug;
        }
    }
    // Give back what the wave does not use.
    const u64 extra = owned & ~lowest_bits(owned, used);
    if (extra) {
        atom_and(&c.bitmap[word], ~extra);
        st.atomics++;
    }
    if (w_after && giver) *w_after = (old | fr) & ~extra;
    const unsigned from = assigned ? src : lane_id();
    const u64 g_owned = shfl64(owned, from);
    const uint32_t g_word = shfl32(word, from);
    if (assigned) {
        id = (int32_t)(g_word * 64u + select_bit(g_owned, gbit));
        need = false;
    }
    return res;
}

/* Loads the lane's word (when probing) and runs window_claim. */
__device__ inline WindowResult window_step(const pbm_ctx& c, bool probe, uint32_t word,
                                           bool& need, int32_t& id, int32_t order, Stats& st) {
    u64 w = kFull;
    if (probe) { w = ld_word(&c.bitmap[word]); st.probes++; }
    return window_claim(c, word, w, need, id, order, st);
}

/* getPage of CollabRW_BM: wave-collective over the active lanes, which all
 * probe and all ask for a page.  The window of step s starts at
 * (uint32)lcg^s(seed) % n_words, lane l probing word start+l (mod n_words);
 * lanes l >= n_words do not probe, so a small map is never read twice in
 * one window.  seed is taken from the lowest active lane.  Step = one
 * window, the same for every lane served in it (CollabRW_BM.cuh:124).
 *
 * collectWarpData (:131-138), which prints the free bits the warp found in
 * each step: with `yields`, the lowest active lane stores the free bits the
 * wave's probing lanes read in step s at yields[s] for s < max_yields, and
 * *n_iters (when given) receives the wave's step count. */
__device__ inline int32_t get_page_collab(const pbm_ctx& c, int32_t seed, uint32_t* steps_out,
                                          Stats* stats = nullptr, uint32_t* yields = nullptr,
                                          uint32_t max_yields = 0, uint32_t* n_iters = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t L = c.n_words;
    seed = (int32_t)__builtin_amdgcn_readfirstlane(seed);
    bool need = reserve_one(c, true);
    int32_t id = -1;
    uint32_t steps = 0;
    const unsigned lane = lane_id();
    uint32_t budget = c.walk_budget;
    uint32_t start = 0;
    uint32_t jumps = 0, sweeps = 0, it = 0;
    const uint32_t max_sweeps = 4u * ((L + 63u) / 64u) + 4u;
    const unsigned first = (unsigned)__ffsll((long long)ballot(true)) - 1u;
    while (ballot(need)) {
        if (jumps < budget) {
            seed = lcg(seed);
            start = (uint32_t)seed % L;
            jumps++;
        } else {
            if (sweeps == 0) st.fallbacks++;
            start = (uint32_t)(((u64)start + 64u) % L);
            if (++sweeps > max_sweeps) break;
        }
        if (need) steps++;
        uint32_t word = (uint32_t)(((u64)start + lane) % L);
        const WindowResult res = window_step(c, lane < L, word, need, id, c.order, st);
        if (yields && it < max_yields && lane == first) yields[it] = res.seen_free;
        it++;
    }
    if (n_iters && lane == first) *n_iters = it;
    if (need) { atom_add_ll(c.free_count, 1); st.errors++; }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}


/* ---------------------------------------------------------------------
 * CRW_BM (CRW_BM.cuh:75-142): clustered walk with a per-thread hint
 * ------------------------------------------------------------------- */

/* Claims the lowest free bit of word g until it is full, counting every
 * atomicOr attempt as a step (CRW_BM.cuh:95,121). */
__device__ __forceinline__ int32_t claim_counting(const pbm_ctx& c, uint32_t g, u64 w,
                                                  uint32_t& steps, Stats& st) {
    while (w != kFull) {
        steps++;
        unsigned b = (unsigned)__ffsll((long long)~w) - 1u;
        u64 bit = 1ull << b;
        u64 old = atom_or(&c.bitmap[g], bit);
        st.atomics++;
        if (!(old & bit)) return (int32_t)(g * 64u + b);
        w = old | bit;
    }
    return -1;
}

/* getPage of CRW_BM for hint slot `slot` (slots >= max_threads have no
 * hint).  First the word of the last page this slot got, moving to the
 * next word when that page was the word's top bit (wrapping at n_words;
 * the reference wraps at n_words-1, CRW_BM.cuh:91); then a random walk as
 * RandomWalkBasic_BM, whose jumps to full words are not counted as steps.
 * The hint is the last granted page id. */
__device__ inline int32_t get_page_crw(const pbm_ctx& c, uint32_t slot, int32_t seed,
                                       uint32_t* steps_out, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t L = c.n_words;
    uint32_t steps = 0;
    int32_t id = -1;
    const bool has_slot = c.hint && slot < c.max_threads;
    if (reserve_one(c, true)) {
        if (has_slot) {
            int32_t h = c.hint[slot];
            if (h >= 0) {
                uint32_t g = (uint32_t)h >> 6;
                if (((uint32_t)h & 63u) == 63u) g = (g + 1u == L) ? 0u : g + 1u;
                u64 w = ld_word(&c.bitmap[g]);
                st.probes++;
                id = claim_counting(c, g, w, steps, st);
            }
        }
        uint32_t g = 0;
        for (uint32_t j = 0; id < 0 && j < c.walk_budget; ++j) {
            seed = lcg(seed);
            g = (uint32_t)seed % L;
            u64 w = ld_word(&c.bitmap[g]);
            st.probes++;
            id = claim_counting(c, g, w, steps, st);
        }
        if (id < 0) {
            st.fallbacks++;
            for (uint64_t k = 0; id < 0 && k < 4ull * L; ++k) {
                g = (g + 1u == L) ? 0u : g + 1u;
                u64 w = ld_word(&c.bitmap[g]);
                st.probes++;
                id = claim_counting(c, g, w, steps, st);
            }
        }
        if (id < 0) { atom_add_ll(c.free_count, 1); st.errors++; }
        else if (has_slot) c.hint[slot] = id;
    }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* ---------------------------------------------------------------------
 * SingleClock (parallelPage.cu:471-513): one global arm
 * ------------------------------------------------------------------- */

/* getPageSingleClock: every try takes the next arm position
 * (atomicAdd(arm,1) % N) and claims that page if it is free.  Step = one
 * arm advance.  The arm is 64-bit and the claim checks the old value (the
 * reference's int arm overflows and its atomicExch result is ignored,
 * parallelPage.cu:498-505).  Tickets are taken for the whole wave with one
 * atomicAdd, in lane order. */
__device__ inline int32_t get_page_single_clock(const pbm_ctx& c, uint32_t* steps_out,
                                                Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t N = c.n_pages;
    bool need = reserve_one(c, true);
    const bool reserved = need;
    uint32_t steps = 0;
    int32_t id = -1;
    const uint64_t max_tries = (uint64_t)c.walk_budget + 2ull * N;
    for (uint64_t it = 0; it < max_tries; ++it) {
        u64 m = ballot(need);
        if (!m) break;
        unsigned cnt = (unsigned)__popcll(m);
        unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
        u64 t0 = 0;
        if (lane_id() == leader) t0 = atom_add_u64(c.clock_arm, (u64)cnt);
        t0 = readlane64(t0, leader);
        if (need) {
            steps++;
            uint32_t p = (uint32_t)((t0 + rank_in(m)) % N);
            u64 bit = 1ull << (p & 63u);
            u64 w = ld_word(&c.bitmap[p >> 6]);
            st.probes++;
            if (!(w & bit)) {
                u64 old = atom_or(&c.bitmap[p >> 6], bit);
                st.atomics++;
                if (!(old & bit)) { id = (int32_t)p; need = false; }
            }
        }
    }
    if (need) { atom_add_ll(c.free_count, 1); st.errors++; }
    if (id >= 0) st.allocs++; else st.failures++;
    (void)reserved;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* ---------------------------------------------------------------------
 * DistClock (parallelPage.cu:537-596): per-thread arms
 * ------------------------------------------------------------------- */

/* getPageDistClock for arm slot `slot` (taken modulo max_threads): scan
 * from the arm, claim the first free page, leave the arm one past the page
 * examined last.  Step = one page examined, as the reference's arm
 * advance; a whole word is examined at once and its pages counted. */
__device__ inline int32_t get_page_dist_clock(const pbm_ctx& c, uint32_t slot,
                                              uint32_t* steps_out, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t N = c.n_pages;
    const uint32_t s = slot % c.max_threads;
    uint32_t steps = 0;
    int32_t id = -1;
    if (reserve_one(c, true)) {
        uint32_t arm = c.arms[s] % N;
        const uint64_t max_pages = (uint64_t)c.walk_budget + 2ull * N;
        uint64_t seen = 0;
        while (seen < max_pages) {
            const uint32_t g = arm >> 6;
            const uint32_t b0 = arm & 63u;
            u64 w = ld_word(&c.bitmap[g]);
            st.probes++;
            // pages b0..63 of this word (the padding bits past N are set)
            u64 fr = ~w & (kFull << b0);
            if (fr) {
                unsigned b = (unsigned)__ffsll((long long)fr) - 1u;
                u64 bit = 1ull << b;
                steps += b - b0 + 1u;
                seen += b - b0 + 1u;
                u64 old = atom_or(&c.bitmap[g], bit);
                st.atomics++;
                arm = g * 64u + b + 1u;
                if (arm >= N) arm = 0;
                if (!(old & bit)) { id = (int32_t)(g * 64u + b); break; }
            } else {
                uint32_t last = g * 64u + 63u < N ? 64u : N - g * 64u;
                steps += last - b0;
                seen += last - b0;
                arm = g * 64u + last;
                if (arm >= N) arm = 0;
            }
        }
        c.arms[s] = arm;
        if (id < 0) { atom_add_ll(c.free_count, 1); st.errors++; }
    }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* ---------------------------------------------------------------------
 * LinkedList pool (parallelPage.cu:36-46, 342-418) as a ring FIFO
 *
 * The reference keeps a Node_t list guarded by sentinel spin locks on the
 * head and tail ids.  Here the free list is a ring of n_pages slots; a
 * slot holds (seq << 32 | page id).  Dequeue ticket t owns slot t % N and
 * waits for seq == t+1 (filled by enqueue ticket t), then writes
 * seq = t+N (free for enqueue ticket t+N); enqueue ticket t waits for
 * seq == t, then writes seq = t+1 with the page.  Pages come back in free
 * order (FIFO), and a page freed after the list ran empty is found again
 * (the reference loses it, parallelPage.cu:357-362,405-414).  Tickets are
 * taken once per wave.  The waits are bounded; a wave must not enqueue and
 * dequeue in the same divergent region.  The page map mirrors the list so
 * occupancy queries and checks see the same state.  Step = clock cycles
 * spent in the call, as the reference counts them (parallelPage.cu:343,367).
 * ------------------------------------------------------------------- */

constexpr uint32_t kRingSpin = 1u << 22;

// A slot word carries both its sequence number and the page id, so relaxed
// 64-bit atomics suffice: nothing else is published through it.
__device__ __forceinline__ u64 ld_slot(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_slot(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int32_t get_page_list(const pbm_ctx& c, uint32_t* steps_out,
                                        Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const long long t_start = clock64();
    const uint32_t N = c.n_pages;
    bool want = reserve_one(c, true);
    int32_t id = -1;
    u64 m = ballot(want);
    if (m) {
        unsigned cnt = (unsigned)__popcll(m);
        unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
        u64 base = 0;
        if (lane_id() == leader) base = atom_add_u64(c.ring_head, (u64)cnt);
        base = readlane64(base, leader);
        if (want) {
            const u64 t = base + rank_in(m);
            u64* slot = &c.ring[t % N];
            const uint32_t expect = (uint32_t)(t + 1u);
            for (uint32_t k = 0; k < kRingSpin; ++k) {
                u64 v = ld_slot(slot);
                if ((uint32_t)(v >> 32) == expect) { id = (int32_t)(uint32_t)v; break; }
            }
            st.probes++;
            if (id >= 0) {
                st_slot(slot, (u64)(uint32_t)(t + N) << 32);
                atom_or(&c.bitmap[(uint32_t)id >> 6], 1ull << ((uint32_t)id & 63u));
                st.atomics++;
            } else {
                st.errors++;
            }
        }
    }
    if (id >= 0) st.allocs++; else st.failures++;
    long long cyc = clock64() - t_start;
    uint32_t steps = cyc > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)cyc;
    st.steps += 1;
    if (steps_out) *steps_out = steps;
    return id;
}

__device__ inline void free_page_list(const pbm_ctx& c, int32_t id, uint32_t* steps_out = nullptr,
                                      Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const long long t_start = clock64();
    const uint32_t N = c.n_pages;
    bool ok = false;
    if (id >= 0 && (uint32_t)id < N) {
        u64 bit = 1ull << ((uint32_t)id & 63u);
        u64 old = atom_and(&c.bitmap[(uint32_t)id >> 6], ~bit);
        st.atomics++;
        ok = (old & bit) != 0;
    }
    if (!ok) st.errors++;
    u64 m = ballot(ok);
    if (m) {
        unsigned cnt = (unsigned)__popcll(m);
        unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
        u64 base = 0;
        if (lane_id() == leader) base = atom_add_u64(c.ring_tail, (u64)cnt);
        base = readlane64(base, leader);
        if (ok) {
            const u64 t = base + rank_in(m);
            u64* slot = &c.ring[t % N];
            const uint32_t expect = (uint32_t)t;
            bool ready = false;
            for (uint32_t k = 0; k < kRingSpin && !ready; ++k)
                ready = (uint32_t)(ld_slot(slot) >> 32) == expect;
            if (!ready) st.errors++;
            st_slot(slot, ((u64)(uint32_t)(t + 1u) << 32) | (uint32_t)id);
            st.frees++;
        }
        // a consumer that reserves this unit before the slot store is
        // visible waits on the slot's sequence number
        release_one(c, ok);
    }
    if (steps_out) {
        long long cyc = clock64() - t_start;
        *steps_out = cyc > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)cyc;
    }
}

/* ---------------------------------------------------------------------
 * Ouroboros page queue (Ouroboros.cuh:22-62 over Ouroboros' page queue,
 * Ouroboros/include/device/queues/page/PageQueueVL_impl.cuh:36-111, and
 * its bulk semaphore, Ouroboros/include/device/BulkSemaphore_impl.cuh:
 * 117-232)
 *
 * A FIFO queue of free pages that fills lazily.  Pages enter it when a
 * chunk of chunk_pages pages is carved, in order, from the pages that were
 * free at the last reset (the bump allocation of chunks,
 * Ouroboros/include/device/Ouroboros.cuh:201-225, whose pages are all
 * enqueued at once, PageQueueVL_impl.cuh:36-48), or when they are freed
 * (enqueued at the back, :104-121).  ctx.q_avail counts queued pages no
 * request has been promised, the semaphore's count.  A wave promises
 * itself its requests with one atomic; when the count falls short it
 * carves enough chunks for the shortfall, enqueues their pages and adds
 * them to the count (the semaphore's AllocateChunk path), then dequeues.
 * Where the semaphore makes later waves reserve and sleep until the
 * carving wave signals, here every short wave carves for itself, so no
 * wave waits on another's promise; a chunk's leftover pages serve the
 * next requests.  The ring, its tickets and slot sequence numbers are the
 * list pool's (above); the page map mirrors the queue.  Reservation
 * (ctx.free_count) bounds grants by the free pages, so a promised page is
 * always queued or carvable.  Step = 1 per dequeue, 2 when the wave carved.
 * ------------------------------------------------------------------- */

/* Enqueues pages src[0..k) at consecutive tail tickets, wave-parallel
 * (k and src uniform).  Returns false if a slot stayed busy. */
__device__ inline bool enqueue_run(const pbm_ctx& c, const unsigned* src, uint32_t k) {
    const uint32_t N = c.n_pages;
    u64 t0 = 0;
    const u64 act = ballot(true);
    const unsigned leader = (unsigned)__ffsll((long long)act) - 1u;
    const unsigned lanes = (unsigned)__popcll(act);
    if (lane_id() == leader) t0 = atom_add_u64(c.ring_tail, (u64)k);
    t0 = readlane64(t0, leader);
    bool ok = true;
    for (uint32_t j = rank_in(act); j < k; j += lanes) {
        const u64 t = t0 + j;
        u64* slot = &c.ring[t % N];
        bool ready = false;
        for (uint32_t s = 0; s < kRingSpin && !ready; ++s)
            ready = (uint32_t)(ld_slot(slot) >> 32) == (uint32_t)t;
        ok &= ready;
        st_slot(slot, ((u64)(uint32_t)(t + 1u) << 32) | src[j]);
    }
    return ballot(!ok) == 0;
}

__device__ inline int32_t get_page_queue(const pbm_ctx& c, uint32_t* steps_out,
                                         Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t N = c.n_pages;
    const bool want = reserve_one(c, true);
    int32_t id = -1;
    uint32_t steps = 0;
    const u64 m = ballot(want);
    if (m) {
        const unsigned cnt = (unsigned)__popcll(m);
        const unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
        // promise; on a shortfall, carve whole chunks for it
        u64 base = 0, take = 0;
        if (lane_id() == leader) {
            const long long old = atom_add_ll(c.q_avail, -(long long)cnt);
            const long long have = old <= 0 ? 0 : (old >= (long long)cnt ? (long long)cnt : old);
            const u64 shortfall = (u64)cnt - (u64)have;
            if (shortfall) {
                const u64 cp = c.chunk_pages;
                const u64 ask = (shortfall + cp - 1u) / cp * cp;
                const u64 n_fresh = ld_word(&c.carve[1]);
                base = atom_add_u64(&c.carve[0], ask);
                take = base >= n_fresh ? 0 : (ask < n_fresh - base ? ask : n_fresh - base);
            }
        }
        base = readlane64(base, leader);
        take = readlane64(take, leader);
        if (take) {
            if (!enqueue_run(c, c.fresh + base, (uint32_t)take)) st.errors++;
            if (lane_id() == leader) atom_add_ll(c.q_avail, (long long)take);
        }
        u64 h0 = 0;
        if (lane_id() == leader) h0 = atom_add_u64(c.ring_head, (u64)cnt);
        h0 = readlane64(h0, leader);
        if (want) {
            const u64 t = h0 + rank_in(m);
            u64* slot = &c.ring[t % N];
            const uint32_t expect = (uint32_t)(t + 1u);
            for (uint32_t k = 0; k < kRingSpin; ++k) {
                const u64 v = ld_slot(slot);
                if ((uint32_t)(v >> 32) == expect) { id = (int32_t)(uint32_t)v; break; }
            }
            st.probes++;
            steps = take ? 2u : 1u;
            if (id >= 0) {
                st_slot(slot, (u64)(uint32_t)(t + N) << 32);
                const u64 bit = 1ull << ((uint32_t)id & 63u);
                if (atom_or(&c.bitmap[(uint32_t)id >> 6], bit) & bit) st.errors++;
                st.atomics++;
            } else {
                st.errors++;
            }
        }
    }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* freePage (Ouroboros.cuh:45-47): a page in use goes to the back of the
 * queue and adds one to the count (PageQueueVL_impl.cuh:104-111).
 * Freeing a free page or an id out of range counts an error. */
__device__ inline void free_page_queue(const pbm_ctx& c, int32_t id, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t N = c.n_pages;
    bool ok = false;
    if (id >= 0 && (uint32_t)id < N) {
        const u64 bit = 1ull << ((uint32_t)id & 63u);
        ok = (atom_and(&c.bitmap[(uint32_t)id >> 6], ~bit) & bit) != 0;
        st.atomics++;
    }
    if (!ok) st.errors++;
    const u64 m = ballot(ok);
    if (!m) return;
    const unsigned cnt = (unsigned)__popcll(m);
    const unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
    u64 t0 = 0;
    if (lane_id() == leader) t0 = atom_add_u64(c.ring_tail, (u64)cnt);
    t0 = readlane64(t0, leader);
    if (ok) {
        const u64 t = t0 + rank_in(m);
        u64* slot = &c.ring[t % N];
        bool ready = false;
        for (uint32_t k = 0; k < kRingSpin && !ready; ++k)
            ready = (uint32_t)(ld_slot(slot) >> 32) == (uint32_t)t;
        if (!ready) st.errors++;
        st_slot(slot, ((u64)(uint32_t)(t + 1u) << 32) | (uint32_t)id);
        st.frees++;
    }
    // signal after the slot stores: a promise made before they are visible
    // waits on the slot's sequence number
    if (lane_id() == leader) atom_add_ll(c.q_avail, (long long)cnt);
    release_one(c, ok);
}

/* ---------------------------------------------------------------------
 * RWMalloc (RWMalloc.cuh:113-334)
 * ------------------------------------------------------------------- */

/* Clears `probe` on every lane whose word equals the word of a lower
 * active lane, so one window never claims from a word image twice. */
__device__ inline void dedupe_words(uint32_t word, bool& probe) {
    u64 pending = ballot(probe);
    while (pending) {
        unsigned l = (unsigned)__ffsll((long long)pending) - 1u;
        uint32_t v = readlane(word, l);
        u64 same = ballot(probe && word == v);
        if (probe && word == v && lane_id() != l) probe = false;
        pending &= ~same;
    }
}

/* getPage of RWMalloc: wave-collective like CollabRW_BM, but every lane
 * walks its own LCG stream (RWMalloc.cuh:114-116,125-126), so a step
 * probes up to 64 scattered words (duplicates probed once).  After the
 * budget the wave sweeps a contiguous window from its lowest lane's last
 * word. */
__device__ inline int32_t get_page_rwmalloc(const pbm_ctx& c, int32_t seed, uint32_t* steps_out,
                                            Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const uint32_t L = c.n_words;
    const unsigned lane = lane_id();
    bool need = reserve_one(c, true);
    int32_t id = -1;
    uint32_t steps = 0, jumps = 0, sweeps = 0, word = 0, start = 0;
    const uint32_t max_sweeps = 4u * ((L + 63u) / 64u) + 4u;
    while (ballot(need)) {
        bool probe;
        if (jumps < c.walk_budget) {
            seed = lcg(seed);
            word = (uint32_t)seed % L;
            probe = true;
            ++jumps;
            if (jumps == c.walk_budget) start = __builtin_amdgcn_readfirstlane(word);
        } else {
            if (sweeps == 0) st.fallbacks++;
            start = (uint32_t)(((u64)start + 64u) % L);
            if (++sweeps > max_sweeps) break;
            word = (uint32_t)(((u64)start + lane) % L);
            probe = lane < L;
        }
        if (need) steps++;
        dedupe_words(word, probe);
        window_step(c, probe, word, need, id, c.order, st);
    }
    if (need) { atom_add_ll(c.free_count, 1); st.errors++; }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* First run of zero bits of x: *start, returns its length (0 if none)
 * and marks the run used in x.  find0Range (RWMalloc.cuh:159-180) means
 * this but finds only runs starting at bit 0. */
__device__ __forceinline__ unsigned first_zero_run(u64& x, unsigned& start) {
    u64 z = ~x;
    if (!z) return 0;
    start = (unsigned)__ffsll((long long)z) - 1u;
    u64 above = x >> start;                  // bit 0 is the run's first zero
    unsigned len = above ? (unsigned)__ffsll((long long)above) - 1u : 64u - start;
    u64 run = (len >= 64u ? kFull : ((1ull << len) - 1ull)) << start;
    x |= run;
    return len;
}

__device__ __forceinline__ u64 run_mask(unsigned start, unsigned n) {
    return (n >= 64u ? kFull : ((1ull << n) - 1ull)) << start;
}

/* Takes `want` units per lane from the reservation counter with one atomic
 * per wave; lanes are served in rank order while the running total fits.
 * With `seen`, the counter's value before the wave's take is stored there
 * (unchanged when no lane wants anything). */
__device__ inline bool reserve_n(const pbm_ctx& c, unsigned want, long long* seen = nullptr) {
    u64 act = ballot(want > 0);
    if (!act) return false;
    // the lanes' running total in lane order, one ballot per bit of want
    unsigned total = 0, pre = 0;
    for (unsigned b = 0; b < 32u; ++b) {
        const u64 mb = ballot((want >> b) & 1u);
        if (!mb && !ballot((want >> b) != 0u)) break;
        pre += rank_in(mb) << b;
        total += (unsigned)__popcll(mb) << b;
    }
    unsigned leader = (unsigned)__ffsll((long long)act) - 1u;
    long long old = 0;
    if (lane_id() == leader) old = atom_add_ll(c.free_count, -(long long)total);
    old = (long long)readlane64((u64)old, leader);
    if (seen) *seen = old;
    bool ok = want > 0 && old >= (long long)(pre + want);
    // give back what the refused lanes asked for
    const u64 refused = act & ~ballot(ok);
    unsigned back = 0;
    if (refused) {
        for (unsigned b = 0; b < 32u; ++b) {
            if (!ballot((want >> b) != 0u)) break;
            back += (unsigned)__popcll(ballot((want >> b) & 1u) & refused) << b;
        }
    }
    if (back && lane_id() == leader) atom_add_ll(c.free_count, (long long)back);
    return ok;
}

constexpr uint32_t kNPageBudget = 64;
constexpr uint32_t kMaxWindowSpans = 4;  /* getNPage windows of up to 4 x 64 words */

/* getNPage (RWMalloc.cuh:198-248): wave-collective; lane asks for n
 * consecutive pages inside one word (n <= 64; n == 0 is a coalesced member
 * that needs nothing and gets 0).  Each step reads a window of consecutive
 * words -- `spans` spans of 64 words, one word per lane per span (1 to 4;
 * the reference's window is one word per lane) -- and the spans are
 * searched in order.  In rounds, every lane offers the next zero-run of
 * its word, and for each offering lane in lane order the lowest
 * still-needing lane whose n fits takes the run's first n bits.  Takes are
 * decided first and then claimed together; a claim that finds a bit taken
 * gives back the bits it set (the reference leaks them, :236-237) and the
 * lane keeps needing.  At most kNPageBudget random windows, then one sweep
 * of the whole map (the reference walks forever).  Returns the first page
 * of the run, -1 on failure. */
/* The search of getNPage for lanes that already hold their reservation
 * (need = true; reserve_n took it): returns the first page of the claimed
 * run, 0 for n == 0, -1 when nothing was found (the reservation is then
 * given back) or need is false. */
/* A window of getNPage's map words read ahead of find_n_pages
 * (wave-collective): lane l holds word start + 64 j + l of span j. */
struct FirstWindow {
    uint32_t start;
    u64 imgs[kMaxWindowSpans];
};

__device__ inline FirstWindow load_window(const pbm_ctx& c, uint32_t start, uint32_t spans) {
    FirstWindow w;
    const uint32_t L = c.n_words;
    w.start = start % L;
#pragma unroll
    for (uint32_t j = 0; j < kMaxWindowSpans; ++j) {
        const uint32_t off = 64u * j + lane_id();
        w.imgs[j] = kFull;
        if (j < spans && off < L) w.imgs[j] = ld_word(&c.bitmap[(uint32_t)(((u64)w.start + off) % L)]);
    }
    return w;
}

/* With `first`, the search starts with that window (a step that takes no
 * random jump), then jumps from the seed as without it. */
__device__ inline int32_t find_n_pages(const pbm_ctx& c, uint32_t n, bool need, int32_t seed,
                                       uint32_t* steps_out, Stats& st, uint32_t spans = 1,
                                       const FirstWindow* first = nullptr) {
    const uint32_t L = c.n_words;
    const unsigned lane = lane_id();
    seed = (int32_t)__builtin_amdgcn_readfirstlane(seed);
    spans = spans < 1u ? 1u : (spans > kMaxWindowSpans ? kMaxWindowSpans : spans);
    int32_t out = (n == 0) ? 0 : -1;
    uint32_t steps = 0, jumps = 0, sweeps = 0, start = 0;
    // Free pages need not form a long enough run, so the search is bounded:
    // a few random windows, then one pass over the map, then failure.
    const uint32_t budget = c.walk_budget < kNPageBudget ? c.walk_budget : kNPageBudget;
    const uint32_t width = 64u * spans;
    const uint32_t max_sweeps = (L + width - 1u) / width;
    bool pre = first != nullptr;  // this step's window was read ahead
    while (ballot(need)) {
        if (pre) {
            start = first->start;
        } else if (jumps < budget) {
            seed = lcg(seed);
            start = (uint32_t)seed % L;
            ++jumps;
        } else {
            if (sweeps == 0) st.fallbacks++;
            start = (uint32_t)(((u64)start + width) % L);
            if (++sweeps > max_sweeps) break;
        }
        if (need) steps++;
        // every span's word is loaded before the first is searched
        uint32_t words[kMaxWindowSpans];
        u64 imgs[kMaxWindowSpans];
#pragma unroll
        for (uint32_t j = 0; j < kMaxWindowSpans; ++j) {
            const uint32_t off = 64u * j + lane;
            words[j] = (uint32_t)(((u64)start + off) % L);
            imgs[j] = kFull;
            // (a lane past the map's end probes nothing)
            if (j < spans && off < L) {
                imgs[j] = pre ? first->imgs[j] : ld_word(&c.bitmap[words[j]]);
                st.probes++;
            }
        }
        pre = false;
#pragma unroll
        for (uint32_t j = 0; j < kMaxWindowSpans; ++j) {
            if (j >= spans || !ballot(need)) break;
            const uint32_t word = words[j];
            u64 img = imgs[j];
            for (;;) {
                unsigned rs = 0;
                const unsigned rl = first_zero_run(img, rs);
                const u64 N = ballot(need);
                if (!N) break;
                // the shortest run a needy lane asks for, bit by bit; an
                // offered run shorter than that is taken by no lane below,
                // so it is passed over at once
                unsigned min_need = 0;
                {
                    u64 cand = N;
                    for (int b = 6; b >= 0; --b) {
                        const u64 z = ballot(!((n >> b) & 1u)) & cand;
                        if (z) cand = z;
                        else min_need |= 1u << b;
                    }
                }
                u64 H = ballot(rl > 0 && rl >= min_need);
                if (!ballot(rl > 0)) break;
                if (!H) continue;
                bool assigned = false;
                unsigned src = 0;
                // shortest offered run, bit by bit over the offering lanes
                unsigned min_len = 0;
                {
                    u64 cand = H;
                    for (int b = 6; b >= 0; --b) {
                        const u64 z = ballot(!((rl >> b) & 1u)) & cand;
                        if (z) cand = z;
                        else min_len |= 1u << b;
                    }
                }
                if (!ballot(need && n > min_len)) {
                    // every needy lane fits every offered run, so the pass
                    // below would give the q-th offering word to the q-th
                    // needy lane: ranks give the same assignment at once
                    const unsigned q = rank_in(N);
                    if (need && q < (unsigned)__popcll(H)) {
                        assigned = true;
                        src = select_bit(H, q);
                    }
                } else {
                    // offering lanes in lane order; each goes to the lowest
                    // still-unassigned needy lane it fits.  The masks live in
                    // scalar registers, and the pass ends once every needy
                    // lane has a run.
                    u64 rem = N;
                    unsigned mine = 64u;
                    while (H && rem) {
                        const unsigned t = (unsigned)__ffsll((long long)H) - 1u;
                        H &= H - 1ull;
                        const unsigned len_t = readlane(rl, t);
                        const u64 fit = ballot(n <= len_t) & rem;
                        if (fit) {
                            const unsigned k = (unsigned)__ffsll((long long)fit) - 1u;
                            rem &= ~(1ull << k);
                            mine = lane == k ? t : mine;
                        }
                    }
                    assigned = mine < 64u;
                    src = assigned ? mine : 0u;
                }
                const unsigned src_start = shfl32(rs, assigned ? src : lane);
                uint32_t src_word = shfl32(word, assigned ? src : lane);
                if (assigned) {
                    u64 mask = run_mask(src_start, n);
                    u64 old = atom_or(&c.bitmap[src_word], mask);
                    st.atomics++;
                    u64 owned = mask & ~old;
                    if (owned == mask) {
                        out = (int32_t)(src_word * 64u + src_start);
                        need = false;
                    } else if (owned) {
                        atom_and(&c.bitmap[src_word], ~owned);
                        st.atomics++;
                    }
                }
            }
        }
    }
    if (need) { atom_add_ll(c.free_count, (long long)n); st.errors++; }
    if (n > 0) {
        if (out >= 0) st.allocs++; else st.failures++;
    }
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return out;
}

__device__ inline int32_t get_n_pages(const pbm_ctx& c, uint32_t n, int32_t seed,
                                      uint32_t* steps_out, Stats* stats = nullptr,
                                      uint32_t spans = 1) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const bool need = reserve_n(c, n > 64u ? 0u : n);
    return find_n_pages(c, n, need, seed, steps_out, st, spans);
}

/* Start of the first run of at least x clear bits of w, or -1. */
__device__ __forceinline__ int first_run_at_least(u64 w, uint32_t x) {
    u64 img = w;
    for (;;) {
        unsigned start = 0;
        const unsigned len = first_zero_run(img, start);
        if (len == 0) return -1;
        if (len >= x) return (int)start;
    }
}

/* getXPage of RandomWalk_BM (parallelPage.cu:1104-1142), one thread: x
 * consecutive pages (1 <= x <= 64) inside one word.  Each step jumps to
 * word (unsigned)seed % n_words and, while the word holds a run of x clear
 * bits, claims the first x bits of its first such run with one atomicOr;
 * bits another thread set first make it give back only the bits it set and
 * look at the word again.  The reference claims x bits from the word's
 * first clear bit whether or not they are clear, which repeats forever on a
 * word whose first run is short, and its revert clears bits it did not set
 * (:1124-1136).  Step = one jump (:1112).  As in get_n_pages, a run may not
 * exist, so the walk is bounded: min(walk_budget, kNPageBudget) jumps, one
 * sweep of the map, then failure (-1).  x == 0 returns 0, x > 64 returns -1. */
__device__ inline int32_t get_x_pages(const pbm_ctx& c, uint32_t x, int32_t seed,
                                      uint32_t* steps_out, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    uint32_t steps = 0;
    int32_t out = -1;
    if (x == 0) {
        out = 0;
    } else if (x <= 64u) {
        const long long old = atom_add_ll(c.free_count, -(long long)x);
        if (old >= (long long)x) {
            const uint32_t L = c.n_words;
            uint32_t g = 0;
            auto try_word = [&](uint32_t w_id) -> bool {
                u64 w = ld_word(&c.bitmap[w_id]);
                st.probes++;
                for (int start = first_run_at_least(w, x); start >= 0;
                     start = first_run_at_least(w, x)) {
                    const u64 mask = run_mask((unsigned)start, x);
                    const u64 prev = atom_or(&c.bitmap[w_id], mask);
                    st.atomics++;
                    const u64 owned = mask & ~prev;
                    if (owned == mask) {
                        out = (int32_t)(w_id * 64u + (uint32_t)start);
                        return true;
                    }
                    if (owned) {
                        atom_and(&c.bitmap[w_id], ~owned);
                        st.atomics++;
                    }
                    w = prev;
                }
                return false;
            };
            bool found = false;
            const uint32_t budget = c.walk_budget < kNPageBudget ? c.walk_budget : kNPageBudget;
            for (uint32_t j = 0; j < budget && !found; ++j) {
                seed = lcg(seed);
                g = (uint32_t)seed % L;
                steps++;
                found = try_word(g);
            }
            if (!found) {
                st.fallbacks++;
                for (uint32_t k = 0; k < L && !found; ++k) {
                    g = (g + 1u == L) ? 0u : g + 1u;
                    steps++;
                    found = try_word(g);
                }
            }
            if (!found) { atom_add_ll(c.free_count, (long long)x); st.errors++; }
        } else {
            atom_add_ll(c.free_count, (long long)x);
        }
        if (out >= 0) st.allocs++; else st.failures++;
    } else {
        st.failures++;
    }
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return out;
}

/* releasePages (RWMalloc.cuh:60-70): returns the n pages starting at id
 * (one word).  Pages that were not in use count as errors. */
__device__ inline void free_n_pages(const pbm_ctx& c, int32_t id, uint32_t n,
                                    Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    uint32_t freed = 0;
    // the run must lie in one word and below n_pages: padding bits past
    // n_pages stay set, since the allocators hand out any clear bit
    if (id >= 0 && n >= 1 && n <= 64 && (uint64_t)(uint32_t)id + n <= c.n_pages &&
        ((uint32_t)id & 63u) + n <= 64u) {
        u64 mask = run_mask((uint32_t)id & 63u, n);
        u64 old = atom_and(&c.bitmap[(uint32_t)id >> 6], ~mask);
        st.atomics++;
        freed = (uint32_t)__popcll(old & mask);
        st.frees += freed;
        st.errors += n - freed;
    } else if (n > 0) {
        st.errors++;
    }
    release_n(c, freed);
}

/* mallocRW (RWMalloc.cuh:299-326): wave-collective over the active lanes.
 * size+4 bytes; requests smaller than a page are packed into one run
 * taken by the lowest such lane (the coalescing leader), at offsets given
 * in lane order (the reference orders them by a shared-memory atomicAdd)
 * and rounded up to 4 bytes so every header is aligned; larger ones take
 * ceil((size+4)/page_size) pages.  An int header before each pointer holds
 * the run's page count (leader and large requests) or 0 (packed members).
 * Needs backing memory; returns nullptr on failure or for runs over 64
 * pages. */
/* What mallocRW asks getNPage for, per lane (wave-collective). */
struct RunPlan {
    uint32_t n;       /* pages of the lane's run: 0 for a packed member, 65 = too big */
    uint64_t off;     /* a packed request's header offset inside the leader's run     */
    bool small;       /* packed behind the coalescing leader                          */
    unsigned leader;  /* lowest packed lane                                           */
    u64 small_mask;
};

__device__ inline RunPlan plan_malloc_rw(const pbm_ctx& c, uint64_t size) {
    RunPlan pl;
    const uint64_t need_bytes = (size + 4u + 3u) & ~(uint64_t)3u;
    pl.small = need_bytes < c.page_size;
    pl.small_mask = ballot(pl.small);
    pl.leader = pl.small_mask ? (unsigned)__ffsll((long long)pl.small_mask) - 1u : 0u;
    // packed offsets: the exclusive sum of the small requests' bytes in
    // lane order, one ballot per bit (need_bytes < page_size, a multiple of 4)
    uint64_t off = 0, total = 0;
    if (pl.small_mask) {
        const unsigned bits = 64u - (unsigned)__builtin_clzll((u64)c.page_size);
        for (unsigned b = 2; b < bits; ++b) {
            const u64 mb = ballot(pl.small && ((need_bytes >> b) & 1u));
            off += (uint64_t)rank_in(mb) << b;
            total += (uint64_t)__popcll(mb) << b;
        }
    }
    uint64_t pages;
    if (pl.small) pages = lane_id() == pl.leader ? (total + c.page_size - 1u) / c.page_size : 0u;
    else pages = (need_bytes + c.page_size - 1u) / c.page_size;
    pl.n = pages > 64u ? 65u : (uint32_t)pages;
    pl.off = pl.small ? off : 0u;
    return pl;
}

/* The pointer and header of a planned request whose run starts at page pid
 * (the leader's pid for packed requests; wave-collective). */
__device__ inline void* place_malloc_rw(const pbm_ctx& c, const RunPlan& pl, int32_t pid) {
    if (pl.small_mask) {
        int32_t lp = (int32_t)readlane((uint32_t)pid, pl.leader);
        if (pl.small) pid = lp;
    }
    if (pid < 0 || !c.pool) return nullptr;
    char* p = c.pool + (size_t)(uint32_t)pid * c.page_size + pl.off;
    // device-scope store: a freeRW in another workgroup of the same kernel
    // reads it once the pointer reaches it
    __hip_atomic_store((int32_t*)p, (int32_t)pl.n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return p + 4;
}

__device__ inline void* malloc_rw(const pbm_ctx& c, uint64_t size, int32_t seed,
                                  uint32_t* steps_out, Stats* stats = nullptr,
                                  uint32_t spans = 1) {
    const RunPlan pl = plan_malloc_rw(c, size);
    const int32_t pid = get_n_pages(c, pl.n, seed, steps_out, stats, spans);
    return place_malloc_rw(c, pl, pid);
}

/* freeRW (RWMalloc.cuh:329-334): header 0 (a packed member) frees nothing;
 * otherwise the run is returned.  Freeing the leader's pointer returns the
 * pages of every member packed with it, as in the reference. */
__device__ inline void free_rw(const pbm_ctx& c, void* ptr, Stats* stats = nullptr) {
    if (!ptr) return;
    // a pointer outside the pool is an error, and its header is not read
    const int32_t id = page_id(c, (char*)ptr - 4);
    if (id < 0) {
        if (stats) stats->errors++;
        return;
    }
    const int32_t n =
        __hip_atomic_load((int32_t*)ptr - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (n == 0) return;
    free_n_pages(c, id, (uint32_t)n, stats);
}

/* ---------------------------------------------------------------------
 * Kernel malloc (malloc.cu): sub-page allocators over runs of pages
 *
 * Both take their chunks (runs of consecutive pages inside one map word)
 * from getNPage, which is the wave-collective getXPage the reference's
 * allocators call (GETXPAGE = getXPageCollabRW_BM, malloc.cu:9-12).
 * Headers are 8 bytes, little-endian, and every space starts 8-byte
 * aligned.  Both return nullptr on a pool without page memory or whose
 * page size is not a multiple of 8.
 * ------------------------------------------------------------------- */

/* KernelMallocCollab headers (malloc.cu:200-211):
 *   chunk  u16 n_pages | u16 occupied bytes | u32 first page (bits 0-30)
 *   space  u32 offset of this header from the chunk start |
 *          u32 space bytes with header (bits 0-30), is_chunk (bit 31)
 * A request with a chunk to itself has no space header of its own: bit 31
 * of its chunk header's last word is is_chunk (malloc.cu:262-265). */
constexpr u64 kIsChunk = 1ull << 63;

/* Largest `occupied` a collaborative group may hold: the u16 field, and
 * at most 64 pages with the chunk header. */
__device__ __forceinline__ uint32_t kmalloc_group_cap(uint32_t page_size) {
    const u64 pages = 64ull * page_size - 8u;
    return pages < 65535ull ? (uint32_t)pages : 65535u;
}

/* KernelMallocCollab::malloc (malloc.cu:231-290), wave-collective over the
 * active lanes.  A request takes s = round8(size) + 8 bytes with its
 * header.  Requests with s < page_size - 8 share chunks: in lane order
 * they form groups whose total fits `occupied` (the whole wave when it
 * fits, as in the reference; the reference's u16 field wraps beyond
 * that); the group's lowest lane gets ceil((total + 8) / page_size) pages,
 * writes the chunk header, and every member writes its space header at
 * 8 + (its suffix sum - s), so the highest lane sits first (:243-251,
 * :280-287).  Larger requests get ceil(s / page_size) pages to themselves
 * and the data follows the chunk header (:276-278).  The reference's
 * suffix sum reads lane targetLaneID instead of targetLaneID-1 (:248). */
__device__ inline void* kmalloc_collab(const pbm_ctx& c, uint32_t size, int32_t seed,
                                       uint32_t* steps_out, Stats* stats = nullptr) {
    if (!c.pool || (c.page_size & 7u)) {
        if (steps_out) *steps_out = 0;
        return nullptr;
    }
    const uint32_t ps = c.page_size;
    const u64 s = (((u64)size + 7u) & ~7ull) + 8u;
    const uint32_t cap = kmalloc_group_cap(ps);
    const bool collab = s + 8u < (u64)ps && s <= cap;
    const u64 cm = ballot(collab);
    unsigned leader = lane_id();
    uint32_t suffix = 0;
    // the collaborative sizes' exclusive sum and total in lane order, one
    // ballot per bit (s is a multiple of 8 below 2^16 for them)
    uint32_t pre = 0, total = 0;
    for (unsigned b = 3; b < 16u; ++b) {
        const u64 mb = ballot(collab && ((s >> b) & 1u));
        pre += rank_in(mb) << b;
        total += (uint32_t)__popcll(mb) << b;
    }
    if (cm && total <= cap) {
        // one group: its lowest lane leads, and a member's suffix sum is
        // the total less what the lanes below it hold
        if (collab) {
            leader = (unsigned)__ffsll((long long)cm) - 1u;
            suffix = total - pre;
        }
    } else {
        // groups in lane order (forward), then suffix sums inside each group
        {
            u64 m = cm;
            uint32_t run = 0;
            unsigned cur = 64;
            while (m) {
                const unsigned l = (unsigned)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                const uint32_t v = readlane((uint32_t)s, l);
                if (cur == 64 || run + v > cap) { cur = l; run = 0; }
                run += v;
                if (lane_id() == l) leader = cur;
            }
        }
        {
            u64 m = cm;
            uint32_t run = 0;
            unsigned cur = 64;
            while (m) {
                const unsigned l = 63u - (unsigned)__clzll((long long)m);
                m &= ~(1ull << l);
                const unsigned g = readlane(leader, l);
                if (g != cur) { cur = g; run = 0; }
                run += readlane((uint32_t)s, l);
                if (lane_id() == l) suffix = run;
            }
        }
    }
    uint32_t n;
    if (collab) n = (leader == lane_id()) ? (uint32_t)((suffix + 8u + ps - 1u) / ps) : 0u;
    else n = (u64)(s + ps - 1u) / ps > 64u ? 65u : (uint32_t)((s + ps - 1u) / ps);
    const int32_t pid = get_n_pages(c, n, seed, steps_out, stats);
    const int32_t gpid = (int32_t)shfl32((uint32_t)pid, collab ? leader : lane_id());
    if (gpid < 0) return nullptr;
    char* chunk = c.pool + (size_t)(uint32_t)gpid * ps;
    if (!collab) {
        *(u64*)chunk = (u64)n | ((u64)(uint16_t)s << 16) | ((u64)(uint32_t)gpid << 32) | kIsChunk;
        return chunk + 8;
    }
    if (leader == lane_id())
        *(u64*)chunk = (u64)n | ((u64)suffix << 16) | ((u64)(uint32_t)gpid << 32);
    const uint32_t root = 8u + suffix - (uint32_t)s;
    *(u64*)(chunk + root) = (u64)root | ((u64)(uint32_t)s << 32);
    return chunk + root + 8;
}

/* KernelMallocCollab::free (malloc.cu:300-320) as its comment describes
 * it: a request with its own chunk returns the chunk's pages; a member
 * takes its bytes off the chunk's `occupied` with one atomic, and the last
 * member returns the pages.  The reference leaves the subtraction
 * commented out (:311-312), so shared chunks are never returned. */
__device__ inline void kfree_collab(const pbm_ctx& c, void* ptr, Stats* stats = nullptr) {
    if (!ptr || !c.pool) return;
    char* hdr = (char*)ptr - 8;
    // a pointer outside the pool, or a space header pointing outside it, is
    // an error, and nothing there is read or written
    if (page_id(c, hdr) < 0) {
        if (stats) stats->errors++;
        return;
    }
    const u64 h = *(const u64*)hdr;
    if (h & kIsChunk) {
        free_n_pages(c, (int32_t)((uint32_t)(h >> 32) & 0x7FFFFFFFu), (uint32_t)(uint16_t)h, stats);
        return;
    }
    const uint32_t root = (uint32_t)h;
    const uint32_t bytes = (uint32_t)(h >> 32) & 0x7FFFFFFFu;
    if ((size_t)root > (size_t)(hdr - c.pool)) {
        if (stats) stats->errors++;
        return;
    }
    char* chunk = hdr - root;
    const uint32_t old = __hip_atomic_fetch_sub((uint32_t*)chunk, bytes << 16, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    if ((old >> 16) == bytes) {
        const uint32_t first = __hip_atomic_load((const uint32_t*)(chunk + 4), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
        free_n_pages(c, (int32_t)(first & 0x7FFFFFFFu), old & 0xFFFFu, stats);
    }
}

/* KernelMallocSimple headers (malloc.cu:45-57), u16 fields:
 *   chunk  allocated (space bytes with headers) | available (n*page_size - 8) |
 *          last (offset of the last space header past the chunk header) | 0
 *   space  prev (bytes of the previous space, 0 first) |
 *          bytes (bits 0-14), is_last (bit 15) | header offset from the chunk start | 0 */
constexpr uint32_t kSimpleMaxSize = 32760;

/* KernelMallocSimple::malloc (malloc.cu:129-175) for the calling thread,
 * whose current chunk is *current (a page id, -1 none): bump allocation
 * inside the chunk; a request that does not fit starts a new chunk of
 * ceil((size + 16) / page_size) pages (chunk and space header) and the old
 * one is left as it is.  Collective over the active lanes because the new
 * chunks come from getNPage.  Sizes are rounded up to 8 and limited to
 * kSimpleMaxSize (a 15-bit field) and chunks to 65,535 available bytes;
 * beyond that it returns nullptr.  As the reference intends: the fit test
 * counts the chunk header once (it subtracts it twice, :136), chunks have
 * room for the space header (:141), the pointer returned is the data after
 * the space header (it returns the header, :174), the last space header
 * is found past the chunk header (:160) and the first space has prev = 0. */
__device__ inline void* kmalloc_simple(const pbm_ctx& c, int32_t* current, uint32_t size,
                                       int32_t seed, uint32_t* steps_out, Stats* stats = nullptr) {
    if (!c.pool || (c.page_size & 7u)) {
        if (steps_out) *steps_out = 0;
        return nullptr;
    }
    const uint32_t ps = c.page_size;
    const uint32_t s = (size + 7u) & ~7u;
    const u64 pages = ((u64)s + 16u + ps - 1u) / ps;
    const bool valid = size <= kSimpleMaxSize && pages <= 64u && pages * ps - 8u <= 65535u;
    int32_t cur = valid ? *current : -1;
    bool fresh = true;
    if (cur >= 0) {
        const u64 m = *(const u64*)(c.pool + (size_t)(uint32_t)cur * ps);
        const uint32_t alloc = (uint32_t)(m & 0xFFFFu), avail = (uint32_t)((m >> 16) & 0xFFFFu);
        fresh = avail - alloc < s + 8u;
    }
    const int32_t pid = get_n_pages(c, valid && fresh ? (uint32_t)pages : 0u, seed, steps_out, stats);
    if (!valid) return nullptr;
    if (fresh) {
        if (pid < 0) return nullptr;
        cur = pid;
        *current = pid;
        *(u64*)(c.pool + (size_t)(uint32_t)pid * ps) = (pages * ps - 8u) << 16;
    }
    char* chunk = c.pool + (size_t)(uint32_t)cur * ps;
    const u64 m = *(const u64*)chunk;
    const uint32_t alloc = (uint32_t)(m & 0xFFFFu);
    const uint32_t last = (uint32_t)((m >> 32) & 0xFFFFu);
    const uint32_t off = 8u + alloc;
    uint32_t prev = 0;
    if (alloc) {
        u64* lh = (u64*)(chunk + 8u + last);
        const u64 l = *lh;
        prev = (uint32_t)((l >> 16) & 0x7FFFu);
        *lh = l & ~(1ull << 31);
    }
    *(u64*)(chunk + off) = (u64)prev | ((u64)(s | 0x8000u) << 16) | ((u64)off << 32);
    *(u64*)chunk = (u64)(alloc + 8u + s) | (m & 0xFFFF0000ull) | ((u64)alloc << 32);
    return chunk + off + 8;
}

/* ---------------------------------------------------------------------
 * Legacy page-map walks (parallelPage.cu:185-224, 642-685)
 *
 * The legacy allocators keep one int per page (parallelPage.cu:117-121,
 * 633-638) and claim with atomicExch(map[p], 1).  Here the same state is
 * one bit per page of the shared map, claimed with atomicOr, so every
 * occupancy query, free and batch entry point works on these pools too.
 * ------------------------------------------------------------------- */

/* getPageClusteredRandomWalk (parallelPage.cu:200-218) for hint slot
 * `slot`: first the page after the slot's last grant (wrapping at
 * n_pages), taken if its bit is clear, which counts one step; otherwise the
 * page walk of RandomWalkBasic, counted as that walk counts (the hint probe
 * is not a step then).  The hint becomes the granted page.  Slots >=
 * max_threads have no hint. */
__device__ inline int32_t get_page_crw_basic(const pbm_ctx& c, uint32_t slot, int32_t seed,
                                             uint32_t* steps_out, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    uint32_t steps = 0;
    int32_t id = -1;
    const bool has_slot = c.hint && slot < c.max_threads;
    if (reserve_one(c, true)) {
        if (has_slot) {
            const int32_t h = c.hint[slot];
            if (h >= 0) {
                const uint32_t p = ((uint32_t)h + 1u >= c.n_pages) ? 0u : (uint32_t)h + 1u;
                const u64 bit = 1ull << (p & 63u);
                st.probes++;
                if (!(ld_word(&c.bitmap[p >> 6]) & bit)) {
                    const u64 old = atom_or(&c.bitmap[p >> 6], bit);
                    st.atomics++;
                    if (!(old & bit)) {
                        id = (int32_t)p;
                        steps = 1;
                    }
                }
            }
        }
        if (id < 0) id = walk_pages(c, seed, steps, st);
        if (id < 0) { atom_add_ll(c.free_count, 1); st.errors++; }
        else if (has_slot) c.hint[slot] = id;
    }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* getPageCollabRW (parallelPage.cu:642-680), wave-collective over the
 * active lanes.  Every active lane walks its own LCG stream over single
 * pages: in each round it probes page (unsigned)seed % n_pages (a page
 * drawn by several lanes is probed by the lowest of them) and claims it if
 * its bit is clear.  The pages won in a round go, in lane order, to the
 * requesting lanes still without one, in lane order; pages beyond what the
 * wave still needs are given back at once (:668-670).  Step = one round,
 * the same for every requesting lane (:657-660).  The reference gathers
 * the pages in a shared array and hands them out with atomicSub
 * (:677-679); here they move by rank through registers.  After
 * walk_budget rounds the requests still open sweep the map one lane at a
 * time. */
__device__ inline int32_t get_page_collab_basic(const pbm_ctx& c, int32_t seed,
                                                uint32_t* steps_out, Stats* stats = nullptr) {
    Stats local;
    Stats& st = stats ? *stats : local;
    const bool req = reserve_one(c, true);
    const u64 reqm = ballot(req);
    const unsigned n_req = (unsigned)__popcll(reqm);
    const unsigned q = rank_in(reqm);
    int32_t id = -1;
    uint32_t rounds = 0, p = 0;
    unsigned filled = 0;
    while (filled < n_req && rounds < c.walk_budget) {
        rounds++;
        seed = lcg(seed);
        p = (uint32_t)seed % c.n_pages;
        bool probe = true;
        dedupe_words(p, probe);
        const u64 bit = 1ull << (p & 63u);
        bool won = false;
        if (probe) {
            st.probes++;
            if (!(ld_word(&c.bitmap[p >> 6]) & bit)) {
                const u64 old = atom_or(&c.bitmap[p >> 6], bit);
                st.atomics++;
                won = !(old & bit);
            }
        }
        const u64 wm = ballot(won);
        const unsigned k = (unsigned)__popcll(wm);
        const unsigned take = k < n_req - filled ? k : n_req - filled;
        if (won && rank_in(wm) >= take) {
            atom_and(&c.bitmap[p >> 6], ~bit);
            st.atomics++;
        }
        const bool mine = req && q >= filled && q < filled + take;
        const unsigned src = mine ? select_bit(wm, q - filled) : lane_id();
        const uint32_t got = shfl32(p, src);
        if (mine) id = (int32_t)got;
        filled += take;
    }
    uint32_t steps = req ? rounds : 0u;
    u64 open = ballot(req && id < 0);
    while (open) {
        const unsigned l = (unsigned)__ffsll((long long)open) - 1u;
        open &= open - 1ull;
        if (lane_id() == l) {
            st.fallbacks++;
            id = sweep_claim(c, p >> 6, steps, st, 4u);
        }
    }
    if (req && id < 0) { atom_add_ll(c.free_count, 1); st.errors++; }
    if (id >= 0) st.allocs++; else st.failures++;
    st.steps += steps;
    if (steps_out) *steps_out = steps;
    return id;
}

/* The strategy's one-page getPage for the calling lane, chosen at compile
 * time: `slot` is the hint or arm slot of per-thread policies, `seed` the
 * request's walk seed.  Wave-collective strategies must be called by whole
 * waves' active lanes together. */
template <int S>
__device__ __forceinline__ int32_t get_page_of(const pbm_ctx& c, uint32_t slot, int32_t seed,
                                               uint32_t* steps_out, Stats* st) {
    if (S == PBM_RW_BM) return get_page_rwbm(c, seed, steps_out, st);
    if (S == PBM_COLLAB_RW_BM) return get_page_collab(c, seed, steps_out, st);
    if (S == PBM_CRW_BM) return get_page_crw(c, slot, seed, steps_out, st);
    if (S == PBM_RW_MALLOC) return get_page_rwmalloc(c, seed, steps_out, st);
    if (S == PBM_SINGLE_CLOCK) return get_page_single_clock(c, steps_out, st);
    if (S == PBM_DIST_CLOCK) return get_page_dist_clock(c, slot, steps_out, st);
    if (S == PBM_RW_BASIC) return get_page_rwbasic(c, seed, steps_out, st);
    if (S == PBM_CRW_BASIC) return get_page_crw_basic(c, slot, seed, steps_out, st);
    if (S == PBM_COLLAB_RW_BASIC) return get_page_collab_basic(c, seed, steps_out, st);
    if (S == PBM_OUROBOROS) return get_page_queue(c, steps_out, st);
    return get_page_list(c, steps_out, st);
}

/* The strategy's one-page freePage, chosen at compile time. */
template <int S>
__device__ __forceinline__ void free_page_of(const pbm_ctx& c, int32_t id, Stats* st) {
    if (S == PBM_LINKED_LIST) free_page_list(c, id, nullptr, st);
    else if (S == PBM_OUROBOROS) free_page_queue(c, id, st);
    else free_page(c, id, st);
}

}  // namespace pbm

#endif  /* PBM_DEVICE_HPP */
