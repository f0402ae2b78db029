// Bursts of mallocRW / freeRW requests: each burst is a free phase (freeRW
// of earlier bursts' allocations) then an alloc phase (mallocRW of its own
// requests).  Phases run as separate kernels (capturable in a HIP graph)
// or all bursts of a range inside one cooperative persistent kernel whose
// workgroups meet at a grid barrier between phases.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>
#include "pbm_internal.hpp"

using pbm::u64;

namespace {

constexpr uint32_t kBarrierSpin = 1u << 20;  // polls of a grid barrier before giving up

struct BurstDev {
    const uint32_t* alloc_off;  // [n_bursts + 1]
    const uint32_t* free_off;   // [n_bursts + 1]
    const uint64_t* sizes;      // [n_allocs]
    const uint32_t* free_index; // [n_frees]
    uint64_t* table;            // [n_allocs] pointer returned for each request
    u64* bar;                   // [0] arrivals, [1] generation, [2] barrier timeouts
};

// Free phase: thread i frees the allocation free_index[i] points at.
__global__ __launch_bounds__(1024) void k_burst_free(pbm_ctx c, BurstDev t, uint32_t f0,
                                                     uint32_t f1) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = f0 + blockIdx.x * blockDim.x + threadIdx.x; i < f1; i += stride)
        pbm::free_rw(c, (void*)t.table[t.free_index[i]], &st);
    pbm::flush_stats(c, st);
}

// Alloc phase: mallocRW, wave-collective over each 64 consecutive
// requests of the burst; the walk seed of a chunk is the seed of its first
// request's global index (the plain mallocRW launch's rule).
__device__ inline void alloc_chunks(const pbm_ctx& c, const BurstDev& t, uint32_t a0, uint32_t a1,
                                    uint32_t wave, uint32_t n_waves, pbm::Stats& st) {
    const unsigned lane = pbm::lane_id();
    const uint32_t n = a1 - a0;
    const uint32_t n_chunks = (n + 63u) / 64u;
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        if (r < n) {
            void* p = pbm::malloc_rw(c, t.sizes[a0 + r],
                                     pbm::request_seed(c.seed, (u64)a0 + (u64)ch * 64u), nullptr,
                                     &st);
            t.table[a0 + r] = (uint64_t)p;
        }
    }
}

__global__ __launch_bounds__(1024) void k_burst_alloc(pbm_ctx c, BurstDev t, uint32_t a0,
                                                      uint32_t a1) {
    pbm::Stats st;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    alloc_chunks(c, t, a0, a1, wave, n_waves, st);
    pbm::flush_stats(c, st);
}

// Grid barrier of a cooperative launch.  The last workgroup to arrive
// rewinds the arrival count and advances the generation the others poll.
// The agent-scope fences publish each workgroup's plain stores (table
// entries, mallocRW headers) before it arrives and make the others' visible
// after it leaves.  A poll that outlasts kBarrierSpin is counted and the
// workgroup goes on, and once one has, no later barrier of any launch
// waits, so the kernel always ends.
__device__ inline void grid_barrier(u64* bar, uint32_t n_groups) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const u64 gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const u64 arrived = pbm::atom_add_u64(&bar[0], 1ull);
        if (arrived + 1u == n_groups) {
            __hip_atomic_store(&bar[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++spins > kBarrierSpin ||
                    __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    pbm::atom_add_u64(&bar[2], 1ull);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// Bursts first .. last-1 in one cooperative launch.
__global__ __launch_bounds__(256) void k_bursts_coop(pbm_ctx c, BurstDev t, uint32_t first,
                                                     uint32_t last) {
    pbm::Stats st;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_threads = gridDim.x * blockDim.x;
    const uint32_t wave = tid >> 6;
    const uint32_t n_waves = n_threads >> 6;
    for (uint32_t b = first; b < last; ++b) {
        const uint32_t f0 = t.free_off[b], f1 = t.free_off[b + 1];
        for (uint32_t i = f0 + tid; i < f1; i += n_threads)
            pbm::free_rw(c, (void*)t.table[t.free_index[i]], &st);
        grid_barrier(t.bar, gridDim.x);
        alloc_chunks(c, t, t.alloc_off[b], t.alloc_off[b + 1], wave, n_waves, st);
        grid_barrier(t.bar, gridDim.x);
    }
    pbm::flush_stats(c, st);
}

__global__ void k_zero_u64(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace

struct pbm_bursts {
    pbm_handle* h = nullptr;
    uint32_t n_bursts = 0, n_allocs = 0, n_frees = 0;
    std::vector<uint32_t> alloc_off, free_off;
    BurstDev d{};
    std::vector<void*> mem;
    int coop_blocks_per_cu = 0;
};

extern "C" int pbm_bursts_destroy(pbm_bursts* b) {
    if (!b) return PBM_OK;
    if (b->h) (void)pbmi::use_device(b->h);
    (void)hipDeviceSynchronize();
    for (void* p : b->mem)
        if (p) (void)hipFree(p);
    delete b;
    return PBM_OK;
}

extern "C" int pbm_bursts_create(pbm_handle* h, const pbm_burst_trace* tr, pbm_bursts** out) {
    PBM_CHECK_HANDLE(h);
    if (!tr || !out) return pbmi::fail(PBM_ERR_INVALID, "null trace or output");
    *out = nullptr;
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "mallocRW bursts need a backed pool");
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "mallocRW bursts need a bitmap strategy");
    if (tr->n_bursts == 0 || !tr->alloc_offset || !tr->free_offset)
        return pbmi::fail(PBM_ERR_INVALID, "a trace needs bursts and both offset arrays");
    const uint32_t nb = tr->n_bursts;
    if (tr->alloc_offset[0] != 0 || tr->free_offset[0] != 0)
        return pbmi::fail(PBM_ERR_INVALID, "offsets must start at 0");
    for (uint32_t i = 0; i < nb; ++i)
        if (tr->alloc_offset[i + 1] < tr->alloc_offset[i] || tr->free_offset[i + 1] < tr->free_offset[i])
            return pbmi::fail(PBM_ERR_INVALID, "offsets must not decrease");
    const uint32_t na = tr->alloc_offset[nb], nf = tr->free_offset[nb];
    if ((na && !tr->sizes) || (nf && !tr->free_index))
        return pbmi::fail(PBM_ERR_INVALID, "null sizes or free_index");
    for (uint32_t i = 0; i < nb; ++i)
        for (uint32_t k = tr->free_offset[i]; k < tr->free_offset[i + 1]; ++k)
            if (tr->free_index[k] >= tr->alloc_offset[i])
                return pbmi::fail(PBM_ERR_INVALID, "a free must target an earlier burst's request");

    pbm_bursts* b = new pbm_bursts();
    b->h = h;
    b->n_bursts = nb;
    b->n_allocs = na;
    b->n_frees = nf;
    b->alloc_off.assign(tr->alloc_offset, tr->alloc_offset + nb + 1);
    b->free_off.assign(tr->free_offset, tr->free_offset + nb + 1);
    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8u);
        if (e != hipSuccess) return pbmi::hip_fail(e, "hipMalloc(bursts)", __FILE__, __LINE__);
        b->mem.push_back(*p);
        return PBM_OK;
    };
    void *p0 = nullptr, *p1 = nullptr, *p2 = nullptr, *p3 = nullptr, *p4 = nullptr, *p5 = nullptr;
    int rc = alloc(&p0, (size_t)(nb + 1) * 4u);
    if (!rc) rc = alloc(&p1, (size_t)(nb + 1) * 4u);
    if (!rc) rc = alloc(&p2, (size_t)na * 8u);
    if (!rc) rc = alloc(&p3, (size_t)nf * 4u);
    if (!rc) rc = alloc(&p4, (size_t)na * 8u);
    if (!rc) rc = alloc(&p5, 32u);
    if (rc) {
        pbm_bursts_destroy(b);
        return rc;
    }
    b->d.alloc_off = (const uint32_t*)p0;
    b->d.free_off = (const uint32_t*)p1;
    b->d.sizes = (const uint64_t*)p2;
    b->d.free_index = (const uint32_t*)p3;
    b->d.table = (uint64_t*)p4;
    b->d.bar = (u64*)p5;
    hipError_t e;
    if ((e = hipMemcpy(p0, tr->alloc_offset, (size_t)(nb + 1) * 4u, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(p1, tr->free_offset, (size_t)(nb + 1) * 4u, hipMemcpyHostToDevice)) ||
        (na && (e = hipMemcpy(p2, tr->sizes, (size_t)na * 8u, hipMemcpyHostToDevice))) ||
        (nf && (e = hipMemcpy(p3, tr->free_index, (size_t)nf * 4u, hipMemcpyHostToDevice))) ||
        (e = hipMemset(p4, 0, (size_t)na * 8u + 8u * (na == 0))) || (e = hipMemset(p5, 0, 32u))) {
        pbm_bursts_destroy(b);
        return pbmi::hip_fail(e, "bursts upload", __FILE__, __LINE__);
    }
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->coop_blocks_per_cu, k_bursts_coop, 256, 0);
    if (e != hipSuccess) {
        pbm_bursts_destroy(b);
        return pbmi::hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor", __FILE__, __LINE__);
    }
    *out = b;
    return PBM_OK;
}

namespace {

int burst_launch(pbm_bursts* b, uint32_t first, uint32_t count, const pbm_launch* in,
                 hipStream_t s, bool capturing) {
    pbm_handle* h = b->h;
    if ((uint64_t)first + count > b->n_bursts) return pbmi::fail(PBM_ERR_RANGE, "burst range beyond the trace");
    pbm_launch L{};
    if (in) L = *in;
    if (L.block == 0) L.block = 256;
    if (L.block % 64 != 0 || L.block > 1024)
        return pbmi::fail(PBM_ERR_INVALID, "launch.block must be a multiple of 64, at most 1024");
    if (L.waves_per_cu == 0) L.waves_per_cu = 8;
    if (L.mode < PBM_MODE_PLAIN || L.mode > PBM_MODE_SERIAL)
        return pbmi::fail(PBM_ERR_INVALID, "unknown launch mode");
    pbm_ctx c = h->ctx;
    if (L.walk_budget) c.walk_budget = L.walk_budget;
    if (L.flags & PBM_LAUNCH_NO_COUNTERS) c.counters = nullptr;
    const BurstDev& d = b->d;
    if (L.flags & PBM_LAUNCH_COOPERATIVE) {
        if (L.mode != PBM_MODE_BATCHED)
            return pbmi::fail(PBM_ERR_INVALID, "a cooperative launch runs in batched mode");
        if (capturing)
            return pbmi::fail(PBM_ERR_UNSUPPORTED, "a cooperative launch is not captured in a graph");
        if (count == 0) return PBM_OK;
        uint32_t per_cu = (L.waves_per_cu + 3u) / 4u;
        if (per_cu > (uint32_t)b->coop_blocks_per_cu) per_cu = (uint32_t)b->coop_blocks_per_cu;
        if (per_cu == 0) return pbmi::fail(PBM_ERR_UNSUPPORTED, "cooperative kernel does not fit a CU");
        const uint32_t last = first + count;
        void* args[] = {(void*)&c, (void*)&d, (void*)&first, (void*)&last};
        PBM_HIP(hipLaunchCooperativeKernel((const void*)k_bursts_coop,
                                           dim3(per_cu * (uint32_t)h->n_cu), dim3(256), args, 0, s));
        return PBM_OK;
    }
    for (uint32_t k = first; k < first + count; ++k) {
        const uint32_t f0 = b->free_off[k], f1 = b->free_off[k + 1];
        const uint32_t a0 = b->alloc_off[k], a1 = b->alloc_off[k + 1];
        if (f1 > f0) {
            dim3 grid(1), blk(1);
            if (L.mode == PBM_MODE_PLAIN) {
                blk = dim3(L.block);
                grid = dim3((f1 - f0 + L.block - 1) / L.block);
            } else if (L.mode == PBM_MODE_BATCHED) {
                blk = dim3(L.block);
                uint64_t thr = (uint64_t)h->n_cu * L.waves_per_cu * 64u;
                if (thr > f1 - f0) thr = f1 - f0;
                grid = dim3((uint32_t)((thr + L.block - 1) / L.block));
            }
            hipLaunchKernelGGL(k_burst_free, grid, blk, 0, s, c, d, f0, f1);
        }
        if (a1 > a0) {
            const uint32_t chunks = (a1 - a0 + 63u) / 64u;
            dim3 grid(1), blk(64);
            if (L.mode == PBM_MODE_PLAIN) {
                blk = dim3(L.block);
                grid = dim3((chunks * 64u + L.block - 1) / L.block);
            } else if (L.mode == PBM_MODE_BATCHED) {
                blk = dim3(L.block);
                uint64_t waves = (uint64_t)h->n_cu * L.waves_per_cu;
                if (waves > chunks) waves = chunks;
                grid = dim3((uint32_t)((waves * 64u + L.block - 1) / L.block));
            }
            hipLaunchKernelGGL(k_burst_alloc, grid, blk, 0, s, c, d, a0, a1);
        }
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_bursts_run(pbm_bursts* b, uint32_t first, uint32_t count,
                              const pbm_launch* launch, void* stream) {
    if (!b) return pbmi::fail(PBM_ERR_INVALID, "null bursts");
    PBM_CHECK_HANDLE(b->h);
    return burst_launch(b, first, count, launch, pbmi::as_stream(stream), false);
}

extern "C" int pbm_bursts_graph(pbm_bursts* b, uint32_t first, uint32_t count,
                                const pbm_launch* launch, pbm_graph** out) {
    if (!b || !out) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output graph");
    PBM_CHECK_HANDLE(b->h);
    *out = nullptr;
    if (count == 0) return pbmi::fail(PBM_ERR_INVALID, "a graph needs count > 0");
    hipStream_t st = nullptr;
    PBM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    pbm_graph* g = new pbm_graph();
    g->device = b->h->device;
    int rc = PBM_OK;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = pbmi::hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    } else {
        rc = burst_launch(b, first, count, launch, st, true);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(st, &graph);
        g->graph = graph;
        if (!rc && e != hipSuccess) rc = pbmi::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        if (!rc) {
            e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = pbmi::hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
        }
    }
    (void)hipStreamDestroy(st);
    if (rc) {
        pbm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return PBM_OK;
}

extern "C" int pbm_bursts_read_ptrs(pbm_bursts* b, uint32_t first, uint32_t n, uint64_t* host,
                                    void* stream) {
    if (!b || (n && !host)) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output");
    PBM_CHECK_HANDLE(b->h);
    if ((uint64_t)first + n > b->n_allocs) return pbmi::fail(PBM_ERR_RANGE, "beyond the trace's requests");
    if (n == 0) return PBM_OK;
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(host, b->d.table + first, (size_t)n * 8u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_bursts_reset(pbm_bursts* b, void* stream) {
    if (!b) return pbmi::fail(PBM_ERR_INVALID, "null bursts");
    PBM_CHECK_HANDLE(b->h);
    if (b->n_allocs) {
        uint32_t blocks = (b->n_allocs + 255u) / 256u;
        if (blocks > 4096u) blocks = 4096u;
        hipLaunchKernelGGL(k_zero_u64, dim3(blocks), dim3(256), 0, pbmi::as_stream(stream),
                           b->d.table, b->n_allocs);
        PBM_HIP(hipGetLastError());
    }
    return PBM_OK;
}

extern "C" int pbm_bursts_barrier_timeouts(pbm_bursts* b, unsigned long long* out, void* stream) {
    if (!b || !out) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output");
    PBM_CHECK_HANDLE(b->h);
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(out, b->d.bar + 2, 8u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_bursts_autotune(pbm_bursts* b, uint32_t count, uint32_t flags,
                                   pbm_tune_result* out, void* stream) {
    if (!b || !out) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output");
    pbm_handle* h = b->h;
    PBM_CHECK_HANDLE(h);
    if (count == 0 || count > b->n_bursts) return pbmi::fail(PBM_ERR_RANGE, "count must be in [1, n_bursts]");
    hipStream_t s = pbmi::as_stream(stream);
    const uint32_t L = h->ctx.n_words;
    std::vector<uint64_t> map(L);
    int rc = pbm_snapshot_bitmap(h, map.data(), L, stream);
    if (rc) return rc;
    struct Cand { pbm_launch l; };
    std::vector<Cand> space;
    const uint32_t wpcs[] = {1, 2, 4, 8, 16};
    const uint32_t budgets[] = {4, 16, 64};
    for (uint32_t wpc : wpcs)
        for (uint32_t bud : budgets)
            for (int coop = 0; coop < ((flags & PBM_TUNE_NO_COOPERATIVE) ? 1 : 2); ++coop) {
                pbm_launch l{};
                l.mode = PBM_MODE_BATCHED;
                l.waves_per_cu = wpc;
                l.walk_budget = bud;
                l.flags = PBM_LAUNCH_NO_COUNTERS | (coop ? PBM_LAUNCH_COOPERATIVE : 0u);
                space.push_back({l});
            }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PBM_HIP(hipEventCreate(&e0));
    PBM_HIP(hipEventCreate(&e1));
    const int reps = (flags & PBM_TUNE_QUICK) ? 1 : 3;
    float best = 1e30f;
    pbm_tune_result res{};
    for (const Cand& k : space) {
        std::vector<float> t;
        for (int r = 0; r < reps && !rc; ++r) {
            rc = pbm_restore_bitmap(h, map.data(), L, stream);
            if (!rc) rc = pbm_bursts_reset(b, stream);
            if (rc) break;
            (void)hipEventRecord(e0, s);
            rc = pbm_bursts_run(b, 0, count, &k.l, stream);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) rc = pbmi::fail(PBM_ERR_HIP, "bursts autotune run");
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            t.push_back(ms);
        }
        if (rc) break;
        std::sort(t.begin(), t.end());
        const float ms = t[t.size() / 2];
        if (ms < best) {
            best = ms;
            res.get_launch = k.l;
            res.free_launch = k.l;
            res.ms_alloc = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) rc = pbm_restore_bitmap(h, map.data(), L, stream);
    if (!rc) rc = pbm_bursts_reset(b, stream);
    if (rc) return rc;
    res.candidates = (uint32_t)space.size();
    *out = res;
    return PBM_OK;
}
