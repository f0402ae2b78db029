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

// Persistent batched allocation.  Each wave serves chunks of 64 requests.
// Window mode: up to `budget` random jumps per chunk, each to a window of
// 64*J consecutive words kept across chunks while it still shows free bits.
// Sweep mode (after the budget, and for the rest of the wave): every lane
// scans kSweepWords consecutive words and offers its first non-full one, so
// one step covers 64*kSweepWords words of the map; the wave advances when
// no lane has a second non-full word in its segment.
constexpr uint32_t kSweepWords = 16;

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
    const uint32_t sweep_span = 64u * kSweepWords;
    const uint32_t max_sweeps = 2u * ((L + sweep_span - 1u) / sweep_span) + 2u;
    const uint32_t n_chunks = (n + 63u) / 64u;
    const long long avail =
        __hip_atomic_load(c.free_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long granted = avail < 0 ? 0 : (avail > (long long)n ? (long long)n : avail);
    Stats st;
    int32_t wseed = request_seed(c.seed ^ kWaveSalt, req_base + wave);
    bool have = false, sweep = false;
    uint32_t start = 0;
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        const bool valid = r < n;
        bool need = valid && (long long)r < granted;
        int32_t id = -1;
        uint32_t steps = 0, jumps = 0, sweeps = 0, iters = 0;
        // A kept window is retried without a new jump only while it showed
        // more free bits than the wave designated (lost races); the cap
        // bounds the loop even so.
        const uint32_t max_iters = budget + max_sweeps + 65536u;
        while (ballot(need)) {
            if (++iters > max_iters) break;
            if (!have) {
                if (jumps < budget) {
                    wseed = lcg(wseed);
                    start = (uint32_t)wseed % L;
                    ++jumps;
                } else {
                    st.fallbacks++;
                    sweep = true;
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
                u64 v[kSweepWords];
#pragma unroll
                for (uint32_t i = 0; i < kSweepWords; ++i) {
                    const uint32_t off = lane * kSweepWords + i;
                    uint32_t idx = start + off;
                    if (idx >= L) idx -= L;
                    v[i] = kFull;
                    if (off < L) v[i] = ld_word(&c.bitmap[idx]);
                }
                st.probes += L < (lane + 1u) * kSweepWords
                                 ? (L > lane * kSweepWords ? L - lane * kSweepWords : 0u)
                                 : kSweepWords;
                uint32_t cand = 0;
                u64 cval = kFull;
                bool more = false;
#pragma unroll
                for (uint32_t i = 0; i < kSweepWords; ++i) {
                    if (v[i] != kFull) {
                        if (cval == kFull) {
                            uint32_t idx = start + lane * kSweepWords + i;
                            cand = idx >= L ? idx - L : idx;
                            cval = v[i];
                        } else {
                            more = true;
                        }
                    }
                }
                window_claim(c, cand, cval, need, id, c.order, st);
                if (ballot(need) && !ballot(more)) {
                    start = (uint32_t)(((u64)start + sweep_span) % L);
                    if (++sweeps > max_sweeps) break;
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
    // The last workgroup to finish applies the reservation; every wave has
    // read free_count before its workgroup counts itself done.
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 prev = __hip_atomic_fetch_add((u64*)&c.phase[1], 1ull, __ATOMIC_ACQ_REL,
                                          __HIP_MEMORY_SCOPE_AGENT);
        if (prev == (u64)gridDim.x - 1u) {
            atom_add_ll(c.free_count, -granted);
            __hip_atomic_store(&c.phase[1], 0ll, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
        const dim3 grid(one ? 1u : batched_blocks(h, L, n));
        const dim3 blk(one ? 64u : 256u);
        switch (L.cluster_words) {
            case 64:
                hipLaunchKernelGGL(k_get_batched<1>, grid, blk, 0, s, c, n, req_base, d_ids,
                                   d_steps, c.walk_budget);
                break;
            case 128:
                hipLaunchKernelGGL(k_get_batched<2>, grid, blk, 0, s, c, n, req_base, d_ids,
                                   d_steps, c.walk_budget);
                break;
            default:
                hipLaunchKernelGGL(k_get_batched<4>, grid, blk, 0, s, c, n, req_base, d_ids,
                                   d_steps, c.walk_budget);
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

