// One-page allocation and free kernels and their batch entry points.
//
//   plain    one thread per request calls the strategy's device getPage,
//            as the reference drivers' get1page_kernel does
//            (unitTests/unitTest1.cu:22-30)
//   serial   the same device code run by one thread (one wave for the
//            wave-collective policies) over the requests in index order
//   batched  a persistent grid; each wave takes 64 requests at a time and
//            serves them from a window of 64*J consecutive map words that it
//            keeps while the window still has free bits.  Admission for the
//            whole call is min(n, free pages at kernel start), given to the
//            lowest request indices: every wave reads the reservation
//            counter when it starts and the last wave to finish subtracts
//            the grant, so a batched call must not overlap other calls on
//            the same pool
#include <hip/hip_runtime.h>
#include "pbm_internal.hpp"

using pbm::u64;

// ------------------------------------------------------------- kernels

// RandomWalkBasic_BM getPage, one request per loop iteration.
__global__ __launch_bounds__(1024) void k_get_rwbm(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                   int32_t* __restrict__ ids,
                                                   uint32_t* __restrict__ steps_out) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        uint32_t steps = 0;
        int32_t id = pbm::get_page_rwbm(c, pbm::request_seed(c.seed, req_base + r), &steps, &st);
        ids[r] = id;
        if (steps_out) steps_out[r] = steps;
    }
    pbm::flush_stats(c, st);
}

// CollabRW_BM getPage: each wave serves 64 consecutive requests per call of
// the wave-collective device function; the walk seed is the seed of the
// chunk's first request.
__global__ __launch_bounds__(1024) void k_get_collab(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                     int32_t* __restrict__ ids,
                                                     uint32_t* __restrict__ steps_out) {
    pbm::Stats st;
    const unsigned lane = pbm::lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x + 63u) >> 6;
    const uint32_t n_chunks = (n + 63u) / 64u;
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        if (r < n) {
            uint32_t steps = 0;
            int32_t id = pbm::get_page_collab(
                c, pbm::request_seed(c.seed, req_base + (uint64_t)ch * 64u), &steps, &st);
            ids[r] = id;
            if (steps_out) steps_out[r] = steps;
        }
    }
    pbm::flush_stats(c, st);
}

// CollabRW_BM getPage with collectWarpData: as k_get_collab, and wave w
// (requests 64w..64w+63) records its per-step yields at
// yields[w * max_steps ..] and its step count at n_steps[w].
__global__ __launch_bounds__(1024) void k_collect_warp_data(pbm_ctx c, uint32_t n,
                                                            uint64_t req_base,
                                                            int32_t* __restrict__ ids,
                                                            uint32_t* __restrict__ yields,
                                                            uint32_t max_steps,
                                                            uint32_t* __restrict__ n_steps) {
    pbm::Stats st;
    const unsigned lane = pbm::lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x + 63u) >> 6;
    const uint32_t n_chunks = (n + 63u) / 64u;
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        if (r < n) {
            int32_t id = pbm::get_page_collab(
                c, pbm::request_seed(c.seed, req_base + (uint64_t)ch * 64u), nullptr, &st,
                yields + (uint64_t)ch * max_steps, max_steps, &n_steps[ch]);
            ids[r] = id;
        }
    }
    pbm::flush_stats(c, st);
}

// ---- batched allocation by rank, one pass over the map
//
// The map is cut into tiles of scan_threads * scan_wpt words, one
// workgroup per tile (each thread holding scan_wpt consecutive words), tile
// = workgroup index.  A tile only ever waits for lower-numbered tiles: the
// dispatcher starts the workgroups of an XCD in index order, so a waited-for
// tile on the same XCD has started, and one on another XCD waits for that
// XCD's slots only, never for slots the waiting tiles hold.  A tile counts its free
// pages, publishes the count, and finds the free pages before it by
// decoupled look-back: wave 0 reads the look-back words of the 256
// preceding tiles at once and sums back to the nearest one that published
// its inclusive prefix.  Tile 0 reserves the call's n requests at the
// reservation counter, keeps G = min(n, free count) as the call's grant
// (the rest goes back) and passes it down the chain inside the same words.  Then the
// free page of rank q < G belongs to request q: each word claims its
// granted bits with one atomicOr and each wave writes ids[q] for its ranks
// in rank order (its own bits, lowest first).  A granted bit that a
// concurrent writer took between the read and the claim sends its request
// to a short random walk.
//
// Every call runs every tile once, and a tile's look-back word carries the
// parity of the call that wrote it: a tile takes this call's parity as the
// flip of its own word's, so no counter is shared between tiles.  (Tiles
// numbered by an atomic ticket at their start would wait in line at one
// address, about 12 ns each: 12 us for the 1,024 tiles of a 2^20-word map.)

// Look-back word: [63] call parity, [62] inclusive prefix (else the tile's
// own count), [61:31] grant G, [30:0] free pages (saturating).
constexpr u64 kScanIncl = 1ull << 62;
constexpr uint32_t kScanSat = 0x7FFFFFFFu;
constexpr uint32_t kScanSpin = 1u << 20;  // look-back polls before giving up
constexpr uint32_t kRankStage = 8192;     // ids staged in LDS per pass (large tiles)
constexpr uint32_t kGroupTiles = 64;      // tiles per group of the two-level look-back
constexpr uint32_t kGroupStride = 32;     // words between group totals (256 B: each in
                                          // a line and channel of its own, since every
                                          // later tile polls them)
constexpr int kLookRounds = 4;            // look-back reads 64 * kLookRounds words at once

__device__ __forceinline__ u64 scan_word(uint32_t parity, bool incl, uint32_t grant, u64 v) {
    const uint32_t sv = v > kScanSat ? kScanSat : (uint32_t)v;
    return ((u64)parity << 63) | (incl ? kScanIncl : 0ull) | ((u64)grant << 31) | sv;
}

__device__ __forceinline__ void st_scan(u64* p, u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A reserved request whose rank page is gone: up to 64 random words, then
// sweeps.  Gives the reservation back when nothing is found.
__device__ inline int32_t rank_fallback(const pbm_ctx& c, uint64_t req, uint32_t budget,
                                        uint32_t& steps, pbm::Stats& st) {
    int32_t seed = pbm::request_seed(c.seed, req);
    int32_t id = pbm::walk_rwbm(c, seed, budget < 64u ? budget : 64u, steps, st);
    if (id < 0) {
        pbm::atom_add_ll(c.free_count, 1);
        st.errors++;
    }
    return id;
}

// One tile of THREADS threads (the whole workgroup).  EXCL: the caller
// keeps every other kernel off the map for the call (PBM_LAUNCH_EXCLUSIVE),
// so a tile, the only writer of its words, stores each claimed word as read
// plus its granted bits.
template <int WPT, int THREADS, bool EXCL>
__device__ __forceinline__ void rank_tile(const pbm_ctx& c, uint32_t tile, uint32_t n,
                                          uint64_t req_base, int32_t* __restrict__ ids,
                                          uint32_t* __restrict__ steps_out, uint32_t budget) {
    using namespace pbm;
    constexpr uint32_t kTileWords = THREADS * WPT;
    constexpr uint32_t kWaves = THREADS / 64;
    __shared__ uint32_t s_parity, s_excl, s_grant;
    __shared__ long long s_free;
    __shared__ uint32_t s_wsum[WPT][kWaves];
    const unsigned lane = lane_id();
    const unsigned wv = threadIdx.x >> 6;
    // thread t holds word k * THREADS + t of the tile, k < WPT: each load
    // and each claim instruction of a wave covers 64 consecutive words
    // (eight words a line, so the memory side takes a claim per line, not
    // per two words), and the tile's map order is k-major
    const uint32_t w0 = tile * kTileWords + threadIdx.x;
    Stats st;
    u64 fm[WPT];
    uint32_t cntk[WPT];
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const uint32_t wi = w0 + (uint32_t)k * THREADS;
        fm[k] = 0;
        if (wi < c.n_words) {
            fm[k] = ~ld_word(&c.bitmap[wi]);  // padding bits past n_pages read as used
            st.probes++;
        }
        cntk[k] = (uint32_t)__popcll(fm[k]);
    }
    // this tile's word from the last call, and tile 0's reservation of all
    // n requests (the counter's value before it fixes the grant), in flight
    // with the map words
    u64 own = 0;
    long long fr = 0;
    if (threadIdx.x == 0) {
        own = ld_word(&c.scan[tile]);
        if (tile == 0) fr = atom_add_ll(c.free_count, -(long long)n);
    }
    // inclusive scans of each word row's free counts over the wave
    uint32_t pre[WPT];
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        uint32_t inc = cntk[k];
#pragma unroll
        for (unsigned d = 1; d < 64u; d <<= 1) {
            const uint32_t v = shfl32(inc, lane >= d ? lane - d : lane);
            if (lane >= d) inc += v;
        }
        pre[k] = inc - cntk[k];
        if (lane == 63) s_wsum[k][wv] = inc;
    }
    if (threadIdx.x == 0) {
        s_parity = (uint32_t)(own >> 63) ^ 1u;
        s_free = fr;
    }
    __syncthreads();
    const uint32_t parity = s_parity;
    // wrank[k]: tile-local rank of the first free page of this thread's
    // word k (the rows before, then the waves and lanes before in the row);
    // agg: the tile's free pages
    uint32_t wrank[WPT];
    uint32_t agg = 0;
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        uint32_t woff = 0, row = 0;
#pragma unroll
        for (unsigned i = 0; i < kWaves; ++i) {
            if (i < wv) woff += s_wsum[k][i];
            row += s_wsum[k][i];
        }
        wrank[k] = agg + woff + pre[k];
        agg += row;
    }

    if (wv == 0) {
        uint32_t grant = 0;
        u64 excl = 0;
        if (tile == 0) {
            // the grant is what the reservation got, min(n, free count),
            // exact even while other kernels reserve; the rest goes back
            const long long f = s_free;
            grant = f <= 0 ? 0u : (f > (long long)n ? n : (uint32_t)f);
            if (lane == 0) {
                st_scan(&c.scan[0], scan_word(parity, true, grant, agg));
                if (grant < n) atom_add_ll(c.free_count, (long long)(n - grant));
            }
        } else if (c.n_tiles > kGroupTiles * 4u) {
            // two-level look-back, for maps of more than 256 tiles: tile
            // counts are never replaced by prefixes; the last tile of each
            // group of 64 publishes the group's total (group 0's with the
            // grant) once its group's counts are out, and a tile sums its
            // group's tiles before it and the groups before its own.  No
            // prefix is passed from tile to tile, so no chain forms.
            if (lane == 0) st_scan(&c.scan[tile], scan_word(parity, false, 0, agg));
            u64* gagg = c.scan + c.n_tiles;
            const uint32_t grp = tile / kGroupTiles, first = grp * kGroupTiles;
            const uint32_t npre = tile - first;
            const uint32_t last = first + kGroupTiles - 1u < c.n_tiles - 1u
                                      ? first + kGroupTiles - 1u : c.n_tiles - 1u;
            uint32_t spins = 0;
            bool gave_up = false;
            u64 v = 0;
            for (;;) {  // this group's tiles before this one
                bool ready = true;
                if (lane < npre) {
                    v = ld_word(&c.scan[first + lane]);
                    ready = (uint32_t)(v >> 63) == parity;
                }
                if (!ballot(!ready)) break;
                if (++spins > kScanSpin) { gave_up = true; break; }
                if (spins < 8u) __builtin_amdgcn_s_sleep(1); else __builtin_amdgcn_s_sleep(2);
            }
            u64 part = lane < npre ? (v & kScanSat) : 0ull;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) part += shfl64(part, lane ^ (unsigned)d);
            excl = part;
            if (grp == 0) grant = (uint32_t)((readlane64(v, 0) >> 31) & 0x7FFFFFFFull);
            if (tile == last && lane == 0)
                st_scan(&gagg[grp * kGroupStride],
                        scan_word(parity, true, grp == 0 ? grant : 0u, part + agg));
            for (uint32_t r0 = 0; r0 < grp && !gave_up; r0 += 64u) {  // the groups before
                const uint32_t gi = r0 + lane;
                for (;;) {
                    bool ready = true;
                    v = 0;
                    if (gi < grp) {
                        v = ld_word(&gagg[gi * kGroupStride]);
                        ready = (uint32_t)(v >> 63) == parity;
                    }
                    if (!ballot(!ready)) break;
                    if (++spins > kScanSpin) { gave_up = true; break; }
                    if (spins < 8u) __builtin_amdgcn_s_sleep(1); else __builtin_amdgcn_s_sleep(2);
                }
                part = gi < grp ? (v & kScanSat) : 0ull;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) part += shfl64(part, lane ^ (unsigned)d);
                excl += part;
                if (r0 == 0) grant = (uint32_t)((readlane64(v, 0) >> 31) & 0x7FFFFFFFull);
            }
            if (gave_up) {
                // cannot happen while every started tile runs on; recorded,
                // and this tile grants nothing
                if (lane == 0) atom_add_u64(&c.scan_ctl[1], 1ull);
                st.errors++;
                grant = 0;
            }
        } else {
            if (lane == 0) st_scan(&c.scan[tile], scan_word(parity, false, 0, agg));
            int64_t k = (int64_t)tile - 1;
            uint32_t spins = 0;
            for (;;) {
                // 64 * kLookRounds predecessors at a time: tile k - (64 j + lane)
                u64 v[kLookRounds];
                u64 P[kLookRounds], R[kLookRounds];
#pragma unroll
                for (int j = 0; j < kLookRounds; ++j) {
                    const int64_t idx = k - (int64_t)(64 * j) - (int64_t)lane;
                    v[j] = 0;
                    bool ready = true, incl = false;
                    if (idx >= 0) {
                        v[j] = ld_word(&c.scan[idx]);
                        ready = (uint32_t)(v[j] >> 63) == parity;
                        incl = ready && (v[j] & kScanIncl);
                    }
                    P[j] = ballot(incl);
                    R[j] = ballot(ready);
                }
                // the nearest inclusive word (jf, fp); every nearer one must
                // have published its count
                int jf = kLookRounds;
                unsigned fp = 63u;
#pragma unroll
                for (int j = kLookRounds - 1; j >= 0; --j)
                    if (P[j]) {
                        jf = j;
                        fp = (unsigned)__ffsll((long long)P[j]) - 1u;
                    }
                bool ok = true;
#pragma unroll
                for (int j = 0; j < kLookRounds; ++j) {
                    const u64 need = j < jf ? kFull
                                     : j == jf ? (fp == 63u ? kFull : ((2ull << fp) - 1ull)) : 0ull;
                    if ((R[j] & need) != need) ok = false;
                }
                if (!ok) {
                    if (++spins > kScanSpin) {
                        // cannot happen while every started tile runs on;
                        // recorded, and the chain goes on with no grant
                        if (lane == 0) atom_add_u64(&c.scan_ctl[1], 1ull);
                        st.errors++;
                        grant = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                u64 part = 0, vf = 0;
#pragma unroll
                for (int j = 0; j < kLookRounds; ++j) {
                    if (j < jf || (j == jf && lane <= fp)) part += v[j] & kScanSat;
                    if (j == jf) vf = v[j];
                }
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) part += shfl64(part, lane ^ (unsigned)d);
                excl += part;
                if (jf < kLookRounds) {
                    grant = (uint32_t)((readlane64(vf, fp) >> 31) & 0x7FFFFFFFull);
                    break;
                }
                k -= 64 * kLookRounds;
            }
            if (lane == 0) st_scan(&c.scan[tile], scan_word(parity, true, grant, excl + agg));
        }
        if (lane == 0) {
            s_excl = excl > kScanSat ? kScanSat : (uint32_t)excl;
            s_grant = grant;
        }
    }
    __syncthreads();

    // claim the bits of rank < G in this thread's words; the results are
    // checked once the ids are out
    const uint32_t grant = s_grant;
    const uint32_t base = s_excl + wrank[0];  // rank of this thread's first free page
    u64 mask[WPT], old[WPT];
    uint32_t mine = 0;  // pages granted to this thread
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const uint32_t rk = s_excl + wrank[k];
        const uint32_t pc = cntk[k];
        uint32_t take = 0;
        if (rk < grant) take = grant - rk < pc ? grant - rk : pc;
        mask[k] = lowest_bits(fm[k], take);
        mine += take;
        old[k] = 0;
        if (mask[k]) {
            const uint32_t wi = w0 + (uint32_t)k * THREADS;
            if constexpr (EXCL) {
                c.bitmap[wi] = ~(fm[k] & ~mask[k]);
            } else {
                old[k] = atom_or(&c.bitmap[wi], mask[k]);
                st.atomics++;
            }
        }
    }
    // ids: the granted bits in rank order
    if constexpr (WPT >= 4) {
        // large tiles (thousands of grants): the tile's ids go through LDS
        // a stage at a time and out in consecutive 4-byte stores, so every
        // wave writes whole lines instead of one scattered id per lane
        __shared__ int32_t s_ids[kRankStage];
        const uint32_t t_lo = s_excl;
        const uint32_t t_hi = grant > s_excl ? (grant - s_excl < agg ? grant : s_excl + agg) : s_excl;
        for (uint32_t lo = t_lo; lo < t_hi; lo += kRankStage) {
            const uint32_t hi = t_hi - lo < kRankStage ? t_hi : lo + kRankStage;
#pragma unroll
            for (int k = 0; k < WPT; ++k) {
                uint32_t q = s_excl + wrank[k];
                u64 m = mask[k];
                if (!m || q >= hi || q + (uint32_t)__popcll(m) <= lo) continue;
                const uint32_t wb = (w0 + (uint32_t)k * THREADS) * 64u;
                while (m && q < hi) {
                    const unsigned bit = (unsigned)__ffsll((long long)m) - 1u;
                    m &= m - 1ull;
                    if (q >= lo) s_ids[q - lo] = (int32_t)(wb + bit);
                    ++q;
                }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < hi - lo; i += THREADS) {
                // streamed past the L2: a large call's ids would otherwise
                // sit dirty in the L2s until the end of the kernel writes
                // them back, after the last wave
                __builtin_nontemporal_store(s_ids[i], &ids[lo + i]);
                if (steps_out) __builtin_nontemporal_store(1u, &steps_out[lo + i]);
            }
            __syncthreads();
        }
    } else if (WPT == 1 && ballot(mask[0] == kFull) == kFull) {
        // every lane of the wave was granted its whole word (an empty map):
        // the wave's 4,096 ranks are its 64 words' pages in order, so it
        // writes them as 64 stores of 64 consecutive ids each (lane l's
        // ranks start 64 l after lane 0's, and its word is lane 0's + l)
        const uint32_t r0 = readlane(base, 0);
        const uint32_t p0 = readlane(w0, 0) * 64u;
        for (uint32_t j = 0; j < 64u; ++j) {
            ids[r0 + j * 64u + lane] = (int32_t)(p0 + j * 64u + lane);
            if (steps_out) steps_out[r0 + j * 64u + lane] = 1u;
        }
    } else {
        // each thread its own
#pragma unroll
        for (int k = 0; k < WPT; ++k) {
            u64 m = mask[k];
            uint32_t r = s_excl + wrank[k];
            const uint32_t wb = (w0 + (uint32_t)k * THREADS) * 64u;
            while (m) {
                const unsigned bit = (unsigned)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                ids[r] = (int32_t)(wb + bit);
                if (steps_out) steps_out[r] = 1u;
                ++r;
            }
        }
    }
    st.allocs += mine;
    st.steps += mine;
    // a concurrent writer took granted bits: their requests walk and
    // overwrite what was written above
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        u64 lost = mask[k] & old[k];
        while (lost) {
            const unsigned bit = (unsigned)__ffsll((long long)lost) - 1u;
            lost &= lost - 1ull;
            const uint32_t q =
                s_excl + wrank[k] + (uint32_t)__popcll(fm[k] & ((1ull << bit) - 1ull));
            uint32_t steps = 1;
            st.fallbacks++;
            st.steps--;
            const int32_t id = rank_fallback(c, req_base + q, budget, steps, st);
            ids[q] = id;
            if (steps_out) steps_out[q] = steps;
            st.steps += steps;
            if (id < 0) {
                st.allocs--;
                st.failures++;
            }
        }
    }

    // requests reserved beyond the free pages the scan saw (a concurrent
    // free moved the count ahead of the map): the last tile walks for them
    if (tile + 1u == c.n_tiles) {
        const u64 seen = (u64)s_excl + agg;
        for (u64 q = seen + threadIdx.x; q < grant; q += blockDim.x) {
            uint32_t steps = 0;
            const int32_t id = rank_fallback(c, req_base + q, budget, steps, st);
            ids[q] = id;
            if (steps_out) steps_out[q] = steps;
            if (id >= 0) st.allocs++; else st.failures++;
            st.steps += steps;
        }
    }
    // requests past the grant fail; every tile writes its slice of them
    {
        const uint32_t rest = n - grant;
        const uint32_t per = (uint32_t)(((u64)rest + c.n_tiles - 1u) / c.n_tiles);
        const u64 lo = (u64)grant + (u64)tile * per;
        const u64 hi = lo + per < (u64)n ? lo + per : (u64)n;
        for (u64 q = lo + threadIdx.x; q < hi; q += blockDim.x) {
            ids[q] = -1;
            if (steps_out) steps_out[q] = 0;
            st.failures++;
        }
    }
    flush_stats(c, st);
}

template <int WPT, bool EXCL>
__global__ __launch_bounds__(256) void k_get_rank(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                  int32_t* __restrict__ ids,
                                                  uint32_t* __restrict__ steps_out,
                                                  uint32_t budget) {
    rank_tile<WPT, 256, EXCL>(c, blockIdx.x, n, req_base, ids, steps_out, budget);
}

// One-wave tiles, for small maps: their grants spread over more CUs.
template <bool EXCL>
__global__ __launch_bounds__(64) void k_get_rank_wave(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                      int32_t* __restrict__ ids,
                                                      uint32_t* __restrict__ steps_out,
                                                      uint32_t budget) {
    rank_tile<1, 64, EXCL>(c, blockIdx.x, n, req_base, ids, steps_out, budget);
}

// The same scan on one workgroup that takes every tile of the call in
// order (PBM_LAUNCH_ONE_WAVE): each tile's look-back finds its predecessor
// finished, the fully serial schedule of the same result.
template <int WPT, int THREADS, bool EXCL>
__global__ __launch_bounds__(THREADS) void k_get_rank_serial(pbm_ctx c, uint32_t n,
                                                             uint64_t req_base,
                                                             int32_t* __restrict__ ids,
                                                             uint32_t* __restrict__ steps_out,
                                                             uint32_t budget) {
    for (uint32_t t = 0; t < c.n_tiles; ++t) {
        rank_tile<WPT, THREADS, EXCL>(c, t, n, req_base, ids, steps_out, budget);
        __syncthreads();
    }
}

// Persistent batched allocation.  Each wave serves chunks of 64 requests.
// Window mode: chunk ch's first window starts at its home word
// ch*64*L/granted (its share of the grants); after that, up to `budget`
// random jumps (LCG steps mapped to a word by their high bits,
// lcg_to_range).  A window is 64*J consecutive words, kept while it still
// shows free bits.
// Sweep mode (after the budget, for the rest of the chunk): the map is cut
// into spans of 64*S words, S = min(kSweepMax, ceil(L/64)); lane l owns the
// S words start + l*S .. and keeps a cursor in them.  Each step every lane
// reads forward from its cursor to its first non-full word and offers it;
// the cursor moves past a word once the lane has seen it full.  When every
// lane has finished its segment the pass is over and the next span starts
// (the same span again when one span covers the map).
constexpr uint32_t kSweepMax = 256;
constexpr uint32_t kSweepRound = 16;

template <int J>
__global__ __launch_bounds__(256) void k_get_batched(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                     int32_t* __restrict__ ids,
                                                     uint32_t* __restrict__ steps_out,
                                                     uint32_t budget) {
    using namespace pbm;
    const unsigned lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t L = c.n_words;
    const uint32_t seg = (L + 63u) / 64u < kSweepMax ? (L + 63u) / 64u : kSweepMax;
    const uint32_t sweep_span = 64u * seg;
    const uint32_t max_dry = 2u * ((L + sweep_span - 1u) / sweep_span) + 2u;
    const uint32_t n_chunks = (n + 63u) / 64u;
    const long long avail =
        __hip_atomic_load(c.free_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long granted = avail < 0 ? 0 : (avail > (long long)n ? (long long)n : avail);
    Stats st;
    int32_t wseed = request_seed(c.seed ^ kWaveSalt, req_base + wave);
    const uint32_t first = lane * seg;
    const uint32_t seg_len = first >= L ? 0u : (L - first < seg ? L - first : seg);
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        const bool valid = r < n;
        bool need = valid && (long long)r < granted;
        int32_t id = -1;
        uint32_t steps = 0, jumps = 0, dry = 0, iters = 0;
        // The chunk's first window starts in its own slice of the map, sized
        // by its share of the call's grants, so concurrent chunks begin on
        // disjoint words.
        const uint32_t home =
            granted > 0 ? (uint32_t)(((u64)ch * 64u * L) / (u64)granted % L) : 0u;
        bool have = false, sweep = false, homed = false, found_pass = false;
        uint32_t start = 0, cur = 0;
        // A kept window or a sweep position is retried only while it showed
        // free bits (lost races); the cap bounds the loop even so.
        const uint32_t max_iters = budget + 4u * max_dry * seg + 65536u;
        while (ballot(need)) {
            if (++iters > max_iters) break;
            if (!have) {
                if (!homed) {
                    start = home;
                    homed = true;
                } else if (jumps < budget) {
                    wseed = lcg(wseed);
                    start = lcg_to_range(wseed, L);
                    ++jumps;
                } else {
                    // the sweep starts at the last window
                    st.fallbacks++;
                    sweep = true;
                    cur = 0;
                    found_pass = false;
                }
                have = true;
            }
            if (need) ++steps;
            if (!sweep) {
                u64 wv[J];
                uint32_t wi[J];
#pragma unroll
                for (int i = 0; i < J; ++i) {
                    const uint32_t off = 64u * (uint32_t)i + lane;
                    wi[i] = (uint32_t)(((u64)start + off) % L);
                    wv[i] = kFull;
                    if (off < L) {
                        wv[i] = ld_word(&c.bitmap[wi[i]]);
                        st.probes++;
                    }
                }
                unsigned seen = 0, desig = 0;
                bool examined_all = true;
#pragma unroll
                for (int i = 0; i < J; ++i) {
                    if (!ballot(need)) {
                        examined_all = false;
                        break;
                    }
                    WindowResult wr = window_claim(c, wi[i], wv[i], need, id, c.order, st);
                    seen += wr.seen_free;
                    desig += wr.designated;
                }
                if (examined_all && seen <= desig) have = false;
            } else {
                uint32_t cand = 0;
                u64 cval = kFull;
                while (cur < seg_len && cval == kFull) {
                    u64 v[kSweepRound];
#pragma unroll
                    for (uint32_t i = 0; i < kSweepRound; ++i) {
                        const uint32_t off = cur + i;
                        uint32_t idx = start + first + off;
                        if (idx >= L) idx -= L;
                        v[i] = kFull;
                        if (off < seg_len) {
                            v[i] = ld_word(&c.bitmap[idx]);
                            st.probes++;
                        }
                    }
                    uint32_t hit = kSweepRound;
#pragma unroll
                    for (uint32_t i = kSweepRound; i-- > 0;)
                        if (v[i] != kFull) hit = i;
                    if (hit < kSweepRound) {
                        cur += hit;
                        uint32_t idx = start + first + cur;
                        cand = idx >= L ? idx - L : idx;
#pragma unroll
                        for (uint32_t i = 0; i < kSweepRound; ++i)
                            if (i == hit) cval = v[i];
                    } else {
                        cur = cur + kSweepRound < seg_len ? cur + kSweepRound : seg_len;
                    }
                }
                u64 after = kFull;
                WindowResult wr = window_claim(c, cand, cval, need, id, c.order, st, &after);
                if (cval != kFull && after == kFull) ++cur;
                if (wr.seen_free) found_pass = true;
                if (!ballot(cur < seg_len)) {
                    // pass over this span finished
                    if (ballot(need)) {
                        if (!found_pass) {
                            if (++dry > max_dry) break;
                        } else {
                            dry = 0;
                        }
                    }
                    found_pass = false;
                    cur = 0;
                    start = (uint32_t)(((u64)start + sweep_span) % L);
                }
            }
        }
        if (need) st.errors++;
        release_one(c, need);
        if (valid) {
            ids[r] = id;
            if (steps_out) steps_out[r] = steps;
            if (id >= 0) st.allocs++; else st.failures++;
        }
        st.steps += steps;
    }
    flush_stats(c, st);
    // The last workgroup to finish applies the reservation.  Every wave used
    // the value it loaded from free_count before its workgroup counts itself
    // done, so the subtraction cannot precede any of those loads; no fence
    // is needed (an acquire/release here would write back and invalidate
    // the whole L2 of the XCD under the workgroups still running).
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 prev = atom_add_u64((u64*)&c.phase[1], 1ull);
        if (prev == (u64)gridDim.x - 1u) {
            atom_add_ll(c.free_count, -granted);
            __hip_atomic_store(&c.phase[1], 0ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// freePage, one id per loop iteration.
__global__ __launch_bounds__(1024) void k_free_plain(pbm_ctx c, uint32_t n,
                                                     const int32_t* __restrict__ ids) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        int32_t id = ids[r];
        if (id >= 0) pbm::free_page(c, id, &st);
    }
    pbm::flush_stats(c, st);
}

// Batched free: adjacent lanes whose pages share a word merge their bits
// into one atomicAnd issued by the first lane of the run.  Each wave frees
// a contiguous range of chunks of 64 ids, kFreeUnroll chunks at a time,
// and issues all their atomics before it reads any result, so their round
// trips to the memory side overlap instead of following one another.  Each
// 1024-thread workgroup returns its pages to the reservation counter with
// one atomic.
constexpr uint32_t kFreeUnroll = 8;

// Reservation credit of the batched free: workgroup w adds its freed pages
// to partial counter w % kCreditSlots, packed with an arrival count
// (arrivals << 40 | pages); the last workgroup to arrive at a counter moves
// its sum to the reservation counter and clears it for the next call.  At
// most a sixteenth of the workgroups meet at one address, where one
// reservation update per workgroup would queue them all at one (about
// 12 ns each: 6 us for the 512 workgroups of a 4,194,304-id call).
// Grids of fewer than kCreditMinGroups workgroups add to the reservation
// counter directly: their short queue costs less than the partial
// counter's extra round trip.
constexpr uint32_t kCreditSlots = 16;
constexpr uint32_t kCreditStride = 32;  // words between partial counters (256 B)
constexpr uint32_t kCreditMinGroups = 128;

__global__ __launch_bounds__(1024) void k_free_batched(pbm_ctx c, uint32_t n,
                                                       const int32_t* __restrict__ ids,
                                                       u64* __restrict__ credit) {
    using namespace pbm;
    const unsigned lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t n_chunks = (n + 63u) / 64u;
    const uint32_t per = (n_chunks + n_waves - 1u) / n_waves;
    const uint64_t c_lo = (uint64_t)wave * per;
    const uint64_t c_hi = c_lo + per < n_chunks ? c_lo + per : n_chunks;
    __shared__ uint32_t s_freed;
    if (threadIdx.x == 0) s_freed = 0;
    __syncthreads();
    Stats st;
    const u64 upto = lane == 63 ? kFull : ((2ull << lane) - 1ull);
    for (uint64_t cb = c_lo; cb < c_hi; cb += kFreeUnroll) {
        int32_t id[kFreeUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kFreeUnroll; ++u) {
            const uint64_t r = (cb + u) * 64u + lane;
            id[u] = (cb + u < c_hi && r < n) ? ids[r] : -1;
        }
        u64 m[kFreeUnroll], old[kFreeUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kFreeUnroll; ++u) {
            m[u] = 0;
            old[u] = 0;
            if (cb + u >= c_hi) continue;  // the same for the whole wave
            const bool valid = id[u] >= 0 && (uint32_t)id[u] < c.n_pages;
            if (id[u] >= 0 && !valid) st.errors++;
            const uint32_t w = valid ? ((uint32_t)id[u] >> 6) : (0x80000000u | lane);
            u64 mu = valid ? (1ull << ((uint32_t)id[u] & 63u)) : 0ull;
            const uint32_t prev = shfl32(w, lane == 0 ? 0u : lane - 1u);
            const bool head = lane == 0 || prev != w;
            const u64 H = ballot(head);
            const unsigned seg = 63u - (unsigned)__clzll((long long)(H & upto));
#pragma unroll
            for (unsigned d = 1; d < 64u; d <<= 1) {
                const unsigned src = lane + d < 64u ? lane + d : lane;
                const u64 o = shfl64(mu, src);
                const uint32_t s2 = shfl32(seg, src);
                if (lane + d < 64u && s2 == seg) mu |= o;
            }
            if (head && valid) {
                m[u] = mu;
                old[u] = atom_and(&c.bitmap[w], ~mu);
                st.atomics++;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kFreeUnroll; ++u) {
            st.frees += (uint32_t)__popcll(old[u] & m[u]);
            st.errors += (uint32_t)__popcll(m[u] & ~old[u]);
        }
    }
    // one reservation update per workgroup for all pages it returned
    uint32_t fr = st.frees;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) fr += (uint32_t)__shfl_xor((int)fr, d, 64);
    if (lane == 0 && fr) atomicAdd(&s_freed, fr);
    flush_stats(c, st);
    __syncthreads();
    if (gridDim.x < kCreditMinGroups) {
        if (threadIdx.x == 0 && s_freed) atom_add_ll(c.free_count, (long long)s_freed);
    } else if (threadIdx.x == 0) {
        const uint32_t slot = blockIdx.x % kCreditSlots;
        const u64 arrivals = (gridDim.x - slot + kCreditSlots - 1u) / kCreditSlots;
        u64* w = credit + (size_t)slot * kCreditStride;
        const u64 old = atom_add_u64(w, (1ull << 40) | (u64)s_freed);
        if ((old >> 40) + 1u == arrivals) {
            // every workgroup of this counter has added: none adds again
            // in this call, and the next call starts after this kernel
            __hip_atomic_store(w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const u64 total = (old & ((1ull << 40) - 1ull)) + s_freed;
            if (total) atom_add_ll(c.free_count, (long long)total);
        }
    }
}

// ---------------------------------------------------------- launches

namespace {

int check_launch(const pbm_launch* in, pbm_launch* out) {
    pbm_launch L{};
    if (in) L = *in;
    if (L.block == 0) L.block = 256;
    if (L.block % 64 != 0 || L.block > 1024)
        return pbmi::fail(PBM_ERR_INVALID, "launch.block must be a multiple of 64, at most 1024");
    if (L.waves_per_cu == 0) L.waves_per_cu = 8;
    if (L.cluster_words == 0) L.cluster_words = 64;
    if (L.cluster_words != 64 && L.cluster_words != 128 && L.cluster_words != 256)
        return pbmi::fail(PBM_ERR_INVALID, "launch.cluster_words must be 64, 128 or 256");
    if (L.mode < PBM_MODE_PLAIN || L.mode > PBM_MODE_SERIAL)
        return pbmi::fail(PBM_ERR_INVALID, "unknown launch mode");
    *out = L;
    return PBM_OK;
}

uint32_t batched_blocks(const pbm_handle* h, const pbm_launch& L, uint32_t n) {
    uint64_t waves = (uint64_t)h->n_cu * L.waves_per_cu;
    uint64_t chunks = (n + 63u) / 64u;
    if (waves > chunks) waves = chunks;
    uint64_t blocks = (waves + 3u) / 4u;
    return (uint32_t)(blocks ? blocks : 1);
}

pbm_ctx launch_ctx(const pbm_handle* h, const pbm_launch& L) {
    pbm_ctx c = h->ctx;
    if (L.walk_budget) c.walk_budget = L.walk_budget;
    if (L.flags & PBM_LAUNCH_NO_COUNTERS) c.counters = nullptr;
    return c;
}

}  // namespace

namespace {

// Clears the events pbm_time_next_kernel armed when a batch call returns,
// whichever path it took.
struct Disarm {
    pbm_handle* h;
    ~Disarm() { h->t_start = h->t_stop = nullptr; }
};

}  // namespace

extern "C" int pbm_time_next_kernel(pbm_handle* h, void* start_event, void* stop_event) {
    if (!h) return pbmi::fail(PBM_ERR_INVALID, "null pbm_handle");
    h->t_start = (hipEvent_t)start_event;
    h->t_stop = (hipEvent_t)stop_event;
    return PBM_OK;
}

extern "C" int pbm_event_create(void** event) {
    if (!event) return pbmi::fail(PBM_ERR_INVALID, "null output");
    hipEvent_t e = nullptr;
    PBM_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return PBM_OK;
}

extern "C" int pbm_event_destroy(void* event) {
    if (event) PBM_HIP(hipEventDestroy((hipEvent_t)event));
    return PBM_OK;
}

extern "C" int pbm_event_elapsed_ms(void* start_event, void* stop_event, float* ms) {
    if (!start_event || !stop_event || !ms) return pbmi::fail(PBM_ERR_INVALID, "null argument");
    PBM_HIP(hipEventSynchronize((hipEvent_t)stop_event));
    PBM_HIP(hipEventElapsedTime(ms, (hipEvent_t)start_event, (hipEvent_t)stop_event));
    return PBM_OK;
}

template <bool EXCL>
static void launch_rank(pbm_handle* h, const pbm_ctx& c, const pbm_launch& L, uint32_t n,
                 uint64_t req_base, int32_t* d_ids, uint32_t* d_steps, hipStream_t s) {
    if (L.flags & PBM_LAUNCH_ONE_WAVE) {
        if (c.scan_wpt == 4)
            pbmi::launch_timed(h, k_get_rank_serial<4, 256, EXCL>, dim3(1), dim3(256), s, c, n,
                               req_base, d_ids, d_steps, c.walk_budget);
        else if (c.scan_threads == 64)
            pbmi::launch_timed(h, k_get_rank_serial<1, 64, EXCL>, dim3(1), dim3(64), s, c, n,
                               req_base, d_ids, d_steps, c.walk_budget);
        else
            pbmi::launch_timed(h, k_get_rank_serial<1, 256, EXCL>, dim3(1), dim3(256), s, c, n,
                               req_base, d_ids, d_steps, c.walk_budget);
    } else if (c.scan_wpt == 4) {
        pbmi::launch_timed(h, k_get_rank<4, EXCL>, dim3(c.n_tiles), dim3(256), s, c, n, req_base,
                           d_ids, d_steps, c.walk_budget);
    } else if (c.scan_threads == 64) {
        pbmi::launch_timed(h, k_get_rank_wave<EXCL>, dim3(c.n_tiles), dim3(64), s, c, n,
                           req_base, d_ids, d_steps, c.walk_budget);
    } else {
        pbmi::launch_timed(h, k_get_rank<1, EXCL>, dim3(c.n_tiles), dim3(256), s, c, n, req_base,
                           d_ids, d_steps, c.walk_budget);
    }
}

static int get_pages(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                     uint32_t* d_steps, const pbm_launch* launch, void* stream) {
    Disarm disarm{h};
    if (n == 0) return PBM_OK;
    if (!d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_ids");
    pbm_launch L;
    int rc = check_launch(launch, &L);
    if (rc) return rc;
    hipStream_t s = pbmi::as_stream(stream);
    pbm_ctx c = launch_ctx(h, L);
    const int strat = h->cfg.strategy;

    if (L.mode == PBM_MODE_BATCHED) {
        if (pbmi::is_queue_strategy(strat))
            return pbmi::fail(PBM_ERR_UNSUPPORTED, "batched mode needs a bitmap strategy");
        if (L.flags & PBM_LAUNCH_RANK) {
            if (L.flags & PBM_LAUNCH_EXCLUSIVE)
                launch_rank<true>(h, c, L, n, req_base, d_ids, d_steps, s);
            else
                launch_rank<false>(h, c, L, n, req_base, d_ids, d_steps, s);
            PBM_HIP(hipGetLastError());
            return PBM_OK;
        }
        const bool one = (L.flags & PBM_LAUNCH_ONE_WAVE) != 0;
        const dim3 grid(one ? 1u : batched_blocks(h, L, n));
        const dim3 blk(one ? 64u : 256u);
        switch (L.cluster_words) {
            case 64:
                pbmi::launch_timed(h, k_get_batched<1>, grid, blk, s, c, n, req_base, d_ids, d_steps,
                                   c.walk_budget);
                break;
            case 128:
                pbmi::launch_timed(h, k_get_batched<2>, grid, blk, s, c, n, req_base, d_ids, d_steps,
                                   c.walk_budget);
                break;
            default:
                pbmi::launch_timed(h, k_get_batched<4>, grid, blk, s, c, n, req_base, d_ids, d_steps,
                                   c.walk_budget);
                break;
        }
        PBM_HIP(hipGetLastError());
        return PBM_OK;
    }

    const bool serial = L.mode == PBM_MODE_SERIAL;
    switch (strat) {
        case PBM_RW_BM: {
            dim3 grid = serial ? dim3(1) : dim3((n + L.block - 1) / L.block);
            dim3 blk = serial ? dim3(1) : dim3(L.block);
            pbmi::launch_timed(h, k_get_rwbm, grid, blk, s, c, n, req_base, d_ids, d_steps);
            break;
        }
        case PBM_COLLAB_RW_BM: {
            uint32_t waves = (n + 63u) / 64u;
            dim3 grid = serial ? dim3(1) : dim3((waves * 64u + L.block - 1) / L.block);
            dim3 blk = serial ? dim3(64) : dim3(L.block);
            pbmi::launch_timed(h, k_get_collab, grid, blk, s, c, n, req_base, d_ids, d_steps);
            break;
        }
        default:
            return pbmi::launch_get_strategy(h, c, L, n, req_base, d_ids, d_steps, s);
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_get_pages(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                             uint32_t* d_steps, const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    int rc = get_pages(h, n, req_base, d_ids, d_steps, launch, stream);
    if (rc) return rc;
    return pbmi::shadow_note(h, n, req_base, d_ids, nullptr, false, pbmi::as_stream(stream));
}

extern "C" int pbm_collect_warp_data(pbm_handle* h, uint32_t n, uint64_t req_base,
                                     int32_t* d_ids, uint32_t* d_yields, uint32_t max_steps,
                                     uint32_t* d_n_steps, const pbm_launch* launch,
                                     void* stream) {
    PBM_CHECK_HANDLE(h);
    Disarm disarm{h};
    if (n == 0) return PBM_OK;
    if (!d_ids || !d_n_steps || (max_steps && !d_yields))
        return pbmi::fail(PBM_ERR_INVALID, "null d_ids, d_yields or d_n_steps");
    if (h->cfg.strategy != PBM_COLLAB_RW_BM)
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "collectWarpData belongs to CollabRW_BM");
    pbm_launch L;
    int rc = check_launch(launch, &L);
    if (rc) return rc;
    if (L.mode == PBM_MODE_BATCHED)
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "collectWarpData runs plain or serial");
    pbm_ctx c = launch_ctx(h, L);
    const bool serial = L.mode == PBM_MODE_SERIAL;
    const uint32_t waves = (n + 63u) / 64u;
    const dim3 grid = serial ? dim3(1) : dim3((waves * 64u + L.block - 1) / L.block);
    const dim3 blk = serial ? dim3(64) : dim3(L.block);
    hipLaunchKernelGGL(k_collect_warp_data, grid, blk, 0, pbmi::as_stream(stream), c, n, req_base,
                       d_ids, d_yields, max_steps, d_n_steps);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_free_pages(pbm_handle* h, uint32_t n, const int32_t* d_ids,
                              const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    Disarm disarm{h};
    if (n == 0) return PBM_OK;
    if (!d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_ids");
    pbm_launch L;
    int rc = check_launch(launch, &L);
    if (rc) return rc;
    hipStream_t s = pbmi::as_stream(stream);
    pbm_ctx c = launch_ctx(h, L);
    // the shadow sees the frees before the kernel makes the pages grantable
    if ((rc = pbmi::shadow_note(h, n, 0, d_ids, nullptr, true, s))) return rc;
    if (pbmi::is_queue_strategy(h->cfg.strategy)) {
        if (L.mode == PBM_MODE_BATCHED)
            return pbmi::fail(PBM_ERR_UNSUPPORTED, "batched mode needs a bitmap strategy");
        return pbmi::launch_free_list(c, h->cfg.strategy, L, n, d_ids, s);
    }
    if (L.mode == PBM_MODE_BATCHED) {
        const uint32_t nb = (batched_blocks(h, L, n) + 3u) / 4u;
        pbmi::launch_timed(h, k_free_batched, dim3(nb), dim3(1024), s, c, n, d_ids, h->d_credit);
    } else if (L.mode == PBM_MODE_SERIAL) {
        pbmi::launch_timed(h, k_free_plain, dim3(1), dim3(1), s, c, n, d_ids);
    } else {
        pbmi::launch_timed(h, k_free_plain, dim3((n + L.block - 1) / L.block), dim3(L.block), s, c,
                           n, d_ids);
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

