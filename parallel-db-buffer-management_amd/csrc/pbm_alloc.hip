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

// ---- rank pre-pass: free-rank of every page at the start of a batched call

// Step 1: one wave per 64-word superblock snapshots the free bits of its
// words, ranks the words inside the superblock, and stores the
// superblock's free total at sb_rank[s+1].
__global__ __launch_bounds__(256) void k_rank_local(pbm_ctx c) {
    using namespace pbm;
    const unsigned lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t sb = wave; sb < c.n_sb; sb += n_waves) {
        const uint32_t w = sb * 64u + lane;
        u64 fm = 0;
        if (w < c.n_words) fm = ~ld_word(&c.bitmap[w]);
        unsigned tot = 0;
        const unsigned pre = scan7((unsigned)__popcll(fm), tot);
        if (w < c.n_words) {
            c.snap[w] = fm;
            c.w_rank[w] = pre;
        }
        if (lane == 0) c.sb_rank[sb + 1u] = tot;
    }
}

// Step 2: one workgroup turns the totals into prefixes in place:
// sb_rank[s] = free pages before superblock s, sb_rank[n_sb] = all free.
__global__ __launch_bounds__(1024) void k_rank_scan(pbm_ctx c) {
    __shared__ uint32_t part[1024];
    const uint32_t n = c.n_sb;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = 0; i < per && b + i < n; ++i) sum += c.sb_rank[b + i + 1u];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024u; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1u] : 0u;
    for (uint32_t i = 0; i < per && b + i < n; ++i) {
        run += c.sb_rank[b + i + 1u];
        c.sb_rank[b + i + 1u] = run;
    }
    if (threadIdx.x == 0) c.sb_rank[0] = 0;
}

// First word holding free rank R (R < sb_rank[n_sb]): a 64-way search over
// the superblock prefixes, then a ballot over the superblock's 64 words.
__device__ inline uint32_t rank_search(const pbm_ctx& c, uint32_t R) {
    using namespace pbm;
    const unsigned lane = lane_id();
    uint32_t lo = 0, hi = c.n_sb;
    while (hi - lo > 1u) {
        const uint32_t stride = (hi - lo + 63u) / 64u;
        const uint32_t idx = lo + lane * stride;
        const uint32_t v = idx < hi ? c.sb_rank[idx] : 0xFFFFFFFFu;
        const u64 m = ballot(idx < hi && v <= R);
        const unsigned t = 63u - (unsigned)__clzll((long long)m);
        lo += t * stride;
        hi = lo + stride < hi ? lo + stride : hi;
    }
    const uint32_t w = lo * 64u + lane;
    bool hit = false;
    if (w < c.n_words) {
        const uint32_t rk = c.sb_rank[lo] + c.w_rank[w];
        const uint32_t pc = (uint32_t)__popcll(c.snap[w]);
        hit = rk <= R && R < rk + pc;
    }
    const u64 m = ballot(hit);
    return lo * 64u + (uint32_t)__ffsll((long long)m) - 1u;
}

// One window of the rank path: lane l holds word w0+l, whose snapshot free
// bits have ranks rk .. rk+pc-1; the bits with ranks in [R, R_end) belong
// to this chunk and are claimed with one atomicOr per word.  Lane q of the
// chunk (request of rank R+q) takes the bit of rank R+q.  *more tells
// whether the chunk's range goes on past this window.
__device__ inline void rank_window(const pbm_ctx& c, uint32_t w0, uint32_t R, uint32_t R_end,
                                   bool& need, int32_t& id, pbm::Stats& st, bool& more) {
    using namespace pbm;
    const unsigned lane = lane_id();
    const uint32_t wd = w0 + lane;
    u64 fm = 0;
    uint32_t rk = 0xFFFFFFFFu, pc = 0;
    if (wd < c.n_words) {
        fm = c.snap[wd];
        rk = c.sb_rank[wd >> 6] + c.w_rank[wd];
        pc = (uint32_t)__popcll(fm);
        st.probes++;
    }
    uint32_t lo = 0, hi = 0;
    if (wd < c.n_words) {
        lo = rk >= R ? 0u : (R - rk < pc ? R - rk : pc);
        hi = rk >= R_end ? 0u : (R_end - rk < pc ? R_end - rk : pc);
        if (hi < lo) hi = lo;
    }
    const u64 mask = lowest_bits(fm, hi) & ~lowest_bits(fm, lo);
    u64 owned = 0;
    if (mask) {
        const u64 old = atom_or(&c.bitmap[wd], mask);
        st.atomics++;
        owned = mask & ~old;
    }
    // the range continues iff the last word of the window ends below R_end
    const u64 valid = ballot(wd < c.n_words);
    more = false;
    if (valid == kFull) {
        const uint32_t end63 = readlane(rk + pc, 63u);
        more = end63 < R_end;
    }
    unsigned src = lane, j = 0;
    bool assigned = false;
    u64 G = ballot(mask != 0);
    while (G) {
        const unsigned g = (unsigned)__ffsll((long long)G) - 1u;
        G &= G - 1ull;
        const uint32_t rk_g = readlane(rk, g), lo_g = readlane(lo, g), hi_g = readlane(hi, g);
        const uint32_t first_q = rk_g + lo_g - R;
        if (need && lane >= first_q && lane < first_q + (hi_g - lo_g)) {
            src = g;
            j = lo_g + (lane - first_q);
            assigned = true;
        }
    }
    const u64 s_owned = shfl64(owned, src);
    const u64 s_fm = shfl64(fm, src);
    const uint32_t s_wd = shfl32(wd, src);
    if (assigned) {
        const unsigned bit = select_bit(s_fm, j);
        if ((s_owned >> bit) & 1ull) {
            id = (int32_t)(s_wd * 64u + bit);
            need = false;
        }
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

template <int J, bool RANK>
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
        if (RANK) {
            // Request r takes the free page of rank r; the walk below only
            // serves what a concurrent writer took away.
            const uint32_t total = c.sb_rank[c.n_sb];
            const uint32_t R = ch * 64u;
            const uint32_t lim = granted < (long long)total ? (uint32_t)granted : total;
            if (R < lim) {
                const uint32_t R_end = R + 64u < lim ? R + 64u : lim;
                const uint32_t w0 = rank_search(c, R);
                bool more = true;
                for (uint32_t k = 0; more && ballot(need) && k <= L / 64u + 1u; ++k) {
                    if (need) ++steps;
                    rank_window(c, w0 + 64u * k, R, R_end, need, id, st, more);
                }
            }
        }
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
// into one atomicAnd issued by the first lane of the run.
__global__ __launch_bounds__(256) void k_free_batched(pbm_ctx c, uint32_t n,
                                                      const int32_t* __restrict__ ids) {
    using namespace pbm;
    const unsigned lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t n_chunks = (n + 63u) / 64u;
    __shared__ uint32_t s_freed;
    if (threadIdx.x == 0) s_freed = 0;
    __syncthreads();
    Stats st;
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        const int32_t id = r < n ? ids[r] : -1;
        const bool valid = id >= 0 && (uint32_t)id < c.n_pages;
        if (id >= 0 && !valid) st.errors++;
        const uint32_t w = valid ? ((uint32_t)id >> 6) : (0x80000000u | lane);
        u64 m = valid ? (1ull << ((uint32_t)id & 63u)) : 0ull;
        const uint32_t prev = shfl32(w, lane == 0 ? 0u : lane - 1u);
        const bool head = lane == 0 || prev != w;
        const u64 H = ballot(head);
        const u64 upto = lane == 63 ? kFull : ((2ull << lane) - 1ull);
        const unsigned seg = 63u - (unsigned)__clzll((long long)(H & upto));
#pragma unroll
        for (unsigned d = 1; d < 64u; d <<= 1) {
            const unsigned src = lane + d < 64u ? lane + d : lane;
            const u64 o = shfl64(m, src);
            const uint32_t s2 = shfl32(seg, src);
            if (lane + d < 64u && s2 == seg) m |= o;
        }
        if (head && valid) {
            const u64 old = atom_and(&c.bitmap[w], ~m);
            st.atomics++;
            st.frees += (uint32_t)__popcll(old & m);
            st.errors += (uint32_t)__popcll(m & ~old);
        }
    }
    // one reservation update per workgroup for all pages it returned
    uint32_t fr = st.frees;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) fr += (uint32_t)__shfl_xor((int)fr, d, 64);
    if (lane == 0 && fr) atomicAdd(&s_freed, fr);
    flush_stats(c, st);
    __syncthreads();
    if (threadIdx.x == 0 && s_freed) atom_add_ll(c.free_count, (long long)s_freed);
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

extern "C" int pbm_get_pages(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                             uint32_t* d_steps, const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_ids");
    pbm_launch L;
    int rc = check_launch(launch, &L);
    if (rc) return rc;
    hipStream_t s = pbmi::as_stream(stream);
    pbm_ctx c = launch_ctx(h, L);
    const int strat = h->cfg.strategy;

    if (L.mode == PBM_MODE_BATCHED) {
        if (strat == PBM_LINKED_LIST)
            return pbmi::fail(PBM_ERR_UNSUPPORTED, "batched mode needs a bitmap strategy");
        const bool one = (L.flags & PBM_LAUNCH_ONE_WAVE) != 0;
        const bool rank = (L.flags & PBM_LAUNCH_RANK) != 0;
        const dim3 grid(one ? 1u : batched_blocks(h, L, n));
        const dim3 blk(one ? 64u : 256u);
        if (rank && !(L.flags & PBM_LAUNCH_RANK_READY)) {
            rc = pbm_rank_prepass(h, s);
            if (rc) return rc;
        }
#define PBM_LAUNCH_BATCHED(JJ)                                                              \
    do {                                                                                    \
        if (rank)                                                                           \
            hipLaunchKernelGGL((k_get_batched<JJ, true>), grid, blk, 0, s, c, n, req_base,  \
                               d_ids, d_steps, c.walk_budget);                              \
        else                                                                                \
            hipLaunchKernelGGL((k_get_batched<JJ, false>), grid, blk, 0, s, c, n, req_base, \
                               d_ids, d_steps, c.walk_budget);                              \
    } while (0)
        switch (L.cluster_words) {
            case 64: PBM_LAUNCH_BATCHED(1); break;
            case 128: PBM_LAUNCH_BATCHED(2); break;
            default: PBM_LAUNCH_BATCHED(4); break;
        }
#undef PBM_LAUNCH_BATCHED
        PBM_HIP(hipGetLastError());
        return PBM_OK;
    }

    const bool serial = L.mode == PBM_MODE_SERIAL;
    switch (strat) {
        case PBM_RW_BM: {
            dim3 grid = serial ? dim3(1) : dim3((n + L.block - 1) / L.block);
            dim3 blk = serial ? dim3(1) : dim3(L.block);
            hipLaunchKernelGGL(k_get_rwbm, grid, blk, 0, s, c, n, req_base, d_ids, d_steps);
            break;
        }
        case PBM_COLLAB_RW_BM: {
            uint32_t waves = (n + 63u) / 64u;
            dim3 grid = serial ? dim3(1) : dim3((waves * 64u + L.block - 1) / L.block);
            dim3 blk = serial ? dim3(64) : dim3(L.block);
            hipLaunchKernelGGL(k_get_collab, grid, blk, 0, s, c, n, req_base, d_ids, d_steps);
            break;
        }
        default:
            return pbmi::launch_get_strategy(h, c, L, n, req_base, d_ids, d_steps, s);
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_rank_prepass(pbm_handle* h, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (h->cfg.strategy == PBM_LINKED_LIST)
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "the rank pre-pass needs a bitmap strategy");
    hipStream_t s = pbmi::as_stream(stream);
    const pbm_ctx& c = h->ctx;
    uint32_t nb = (c.n_sb + 3u) / 4u;
    if (nb > 4096u) nb = 4096u;
    hipLaunchKernelGGL(k_rank_local, dim3(nb), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_rank_scan, dim3(1), dim3(1024), 0, s, c);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_free_pages(pbm_handle* h, uint32_t n, const int32_t* d_ids,
                              const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_ids");
    pbm_launch L;
    int rc = check_launch(launch, &L);
    if (rc) return rc;
    hipStream_t s = pbmi::as_stream(stream);
    pbm_ctx c = launch_ctx(h, L);
    if (h->cfg.strategy == PBM_LINKED_LIST) {
        if (L.mode == PBM_MODE_BATCHED)
            return pbmi::fail(PBM_ERR_UNSUPPORTED, "batched mode needs a bitmap strategy");
        return pbmi::launch_free_list(c, L, n, d_ids, s);
    }
    if (L.mode == PBM_MODE_BATCHED) {
        const uint32_t nb = batched_blocks(h, L, n);
        hipLaunchKernelGGL(k_free_batched, dim3(nb), dim3(256), 0, s, c, n, d_ids);
    } else if (L.mode == PBM_MODE_SERIAL) {
        hipLaunchKernelGGL(k_free_plain, dim3(1), dim3(1), 0, s, c, n, d_ids);
    } else {
        hipLaunchKernelGGL(k_free_plain, dim3((n + L.block - 1) / L.block), dim3(L.block), 0, s, c,
                           n, d_ids);
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

