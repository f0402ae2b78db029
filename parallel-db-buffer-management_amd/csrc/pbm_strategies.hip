// Plain and serial launches of the remaining policies (CRW_BM, RWMalloc,
// SingleClock, DistClock, LinkedList, RandomWalkBasic and the legacy
// clustered and collaborative page walks), the list ring rebuild, and the
// multi-page entry points (getNPage, mallocRW, freeRW).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include "pbm_internal.hpp"

using pbm::u64;

// ------------------------------------------------------------- one page

// Per-thread policies: thread gid serves requests gid, gid+stride, ...;
// its hint/arm slot is gid.  A serial launch (one thread) uses slot 0.
template <int S>
__global__ __launch_bounds__(1024) void k_get_thread(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                     int32_t* __restrict__ ids,
                                                     uint32_t* __restrict__ steps_out) {
    pbm::Stats st;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = gid; r < n; r += stride) {
        uint32_t steps = 0;
        const int32_t seed = pbm::request_seed(c.seed, req_base + r);
        const int32_t id = pbm::get_page_of<S>(c, gid, seed, &steps, &st);
        ids[r] = id;
        if (steps_out) steps_out[r] = steps;
    }
    pbm::flush_stats(c, st);
}

// Wave-collective policies whose lanes each walk their own request's LCG
// stream (RWMalloc getPage, the legacy CollabRW): a wave serves chunks of
// 64 consecutive requests.
template <int S>
__global__ __launch_bounds__(1024) void k_get_wave(pbm_ctx c, uint32_t n, uint64_t req_base,
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
            const int32_t id = pbm::get_page_of<S>(c, 0u, pbm::request_seed(c.seed, req_base + r),
                                                   &steps, &st);
            ids[r] = id;
            if (steps_out) steps_out[r] = steps;
        }
    }
    pbm::flush_stats(c, st);
}

// Free for the ring strategies (LinkedList, Ouroboros).
template <int S>
__global__ __launch_bounds__(1024) void k_free_ring(pbm_ctx c, uint32_t n,
                                                    const int32_t* __restrict__ ids) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        int32_t id = ids[r];
        if (id >= 0) pbm::free_page_of<S>(c, id, &st);
    }
    pbm::flush_stats(c, st);
}

namespace pbmi {

int launch_get_strategy(pbm_handle* h, const pbm_ctx& c, const pbm_launch& L, uint32_t n,
                        uint64_t req_base, int32_t* d_ids, uint32_t* d_steps, hipStream_t s) {
    const bool serial = L.mode == PBM_MODE_SERIAL;
    const dim3 tgrid = serial ? dim3(1) : dim3((n + L.block - 1) / L.block);
    const dim3 tblk = serial ? dim3(1) : dim3(L.block);
    const uint32_t waves = (n + 63u) / 64u;
    const dim3 wgrid = serial ? dim3(1) : dim3((waves * 64u + L.block - 1) / L.block);
    const dim3 wblk = serial ? dim3(64) : dim3(L.block);
    switch (h->cfg.strategy) {
#define PBM_THREAD_CASE(S)                                                                      \
    case S:                                                                                     \
        hipLaunchKernelGGL(k_get_thread<S>, tgrid, tblk, 0, s, c, n, req_base, d_ids, d_steps); \
        break;
        PBM_THREAD_CASE(PBM_CRW_BM)
        PBM_THREAD_CASE(PBM_SINGLE_CLOCK)
        PBM_THREAD_CASE(PBM_DIST_CLOCK)
        PBM_THREAD_CASE(PBM_LINKED_LIST)
        PBM_THREAD_CASE(PBM_RW_BASIC)
        PBM_THREAD_CASE(PBM_CRW_BASIC)
        PBM_THREAD_CASE(PBM_OUROBOROS)
#undef PBM_THREAD_CASE
        case PBM_RW_MALLOC:
            hipLaunchKernelGGL(k_get_wave<PBM_RW_MALLOC>, wgrid, wblk, 0, s, c, n, req_base, d_ids,
                               d_steps);
            break;
        case PBM_COLLAB_RW_BASIC:
            hipLaunchKernelGGL(k_get_wave<PBM_COLLAB_RW_BASIC>, wgrid, wblk, 0, s, c, n, req_base,
                               d_ids, d_steps);
            break;
        default:
            return fail(PBM_ERR_UNSUPPORTED, "strategy has no one-page launch here");
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

int launch_free_list(const pbm_ctx& c, int strategy, const pbm_launch& L, uint32_t n,
                     const int32_t* d_ids, hipStream_t s) {
    const bool serial = L.mode == PBM_MODE_SERIAL;
    const dim3 grid = serial ? dim3(1) : dim3((n + L.block - 1) / L.block);
    const dim3 blk = serial ? dim3(1) : dim3(L.block);
    if (strategy == PBM_OUROBOROS)
        hipLaunchKernelGGL(k_free_ring<PBM_OUROBOROS>, grid, blk, 0, s, c, n, d_ids);
    else
        hipLaunchKernelGGL(k_free_ring<PBM_LINKED_LIST>, grid, blk, 0, s, c, n, d_ids);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace pbmi

// ------------------------------------------------------------- list ring

__global__ void k_word_free_counts(const u64* bm, uint32_t n_words, uint32_t* counts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_words) counts[i] = 64u - (uint32_t)__popcll(bm[i]);
}

// Slot i empty and ready for enqueue ticket i.
__global__ void k_ring_empty(u64* ring, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ring[i] = (u64)i << 32;
}

// The free pages of word w, ascending, at ring positions offs[w].. as if
// enqueued with tickets equal to their positions.
__global__ void k_ring_fill(const u64* bm, uint32_t n_words, const uint32_t* offs, u64* ring,
                            u64* head_tail) {
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    u64 fr = ~bm[w];
    uint32_t pos = offs[w];
    while (fr) {
        unsigned b = (unsigned)__ffsll((long long)fr) - 1u;
        fr &= fr - 1ull;
        ring[pos] = ((u64)(pos + 1u) << 32) | (w * 64u + b);
        ++pos;
    }
    if (w == n_words - 1u) {
        head_tail[0] = 0;
        head_tail[1] = pos;
    }
}

// Ouroboros: the free pages of word w, ascending, at fresh positions
// offs[w]..; the ring is empty and nothing is carved or queued yet.
__global__ void k_fresh_fill(const u64* bm, uint32_t n_words, const uint32_t* offs,
                             unsigned* fresh, u64* carve, u64* head_tail, long long* q_avail) {
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    u64 fr = ~bm[w];
    uint32_t pos = offs[w];
    while (fr) {
        unsigned b = (unsigned)__ffsll((long long)fr) - 1u;
        fr &= fr - 1ull;
        fresh[pos++] = w * 64u + b;
    }
    if (w == n_words - 1u) {
        carve[0] = 0;
        carve[1] = pos;
        head_tail[0] = 0;
        head_tail[1] = 0;
        *q_avail = 0;
    }
}

namespace pbmi {

int ring_rebuild(pbm_handle* h, hipStream_t s) {
    const pbm_ctx& c = h->ctx;
    const uint32_t L = c.n_words;
    uint32_t* counts = nullptr;
    uint32_t* offs = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = PBM_OK;
    hipError_t e;
    do {
        if ((e = hipMalloc((void**)&counts, (size_t)L * 4u)) != hipSuccess ||
            (e = hipMalloc((void**)&offs, (size_t)L * 4u)) != hipSuccess) {
            rc = hip_fail(e, "hipMalloc(ring scan)", __FILE__, __LINE__);
            break;
        }
        hipLaunchKernelGGL(k_word_free_counts, dim3((L + 255u) / 256u), dim3(256), 0, s, h->d_bitmap,
                           L, counts);
        if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, offs, (int)L, s)) !=
                hipSuccess ||
            (e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 8u)) != hipSuccess ||
            (e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, offs, (int)L, s)) !=
                hipSuccess) {
            rc = hip_fail(e, "ring scan", __FILE__, __LINE__);
            break;
        }
        hipLaunchKernelGGL(k_ring_empty, dim3((c.n_pages + 255u) / 256u), dim3(256), 0, s,
                           h->d_ring, c.n_pages);
        if (h->cfg.strategy == PBM_OUROBOROS)
            hipLaunchKernelGGL(k_fresh_fill, dim3((L + 255u) / 256u), dim3(256), 0, s, h->d_bitmap,
                               L, offs, h->d_fresh, h->d_carve, h->d_ring_ht, h->d_qavail);
        else
            hipLaunchKernelGGL(k_ring_fill, dim3((L + 255u) / 256u), dim3(256), 0, s, h->d_bitmap,
                               L, offs, h->d_ring, h->d_ring_ht);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess) {
            rc = hip_fail(e, "ring rebuild", __FILE__, __LINE__);
            break;
        }
    } while (0);
    if (tmp) (void)hipFree(tmp);
    if (offs) (void)hipFree(offs);
    if (counts) (void)hipFree(counts);
    return rc;
}

}  // namespace pbmi

// ------------------------------------------------------------- multi-page

// getNPage: wave-collective over chunks of 64 requests; the chunk's walk
// seed is the seed of its first request.
// SPANS: 64-word spans of the window each search round reads (1 = one
// word per lane, RWMalloc.cuh:198-248).
template <int SPANS>
__global__ __launch_bounds__(1024) void k_get_n_pages(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                      const uint32_t* __restrict__ counts,
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
            int32_t id = pbm::get_n_pages(
                c, counts[r], pbm::request_seed(c.seed, req_base + (uint64_t)ch * 64u), &steps, &st,
                SPANS);
            ids[r] = id;
            if (steps_out) steps_out[r] = steps;
        }
    }
    pbm::flush_stats(c, st);
}

__global__ __launch_bounds__(1024) void k_free_n_pages(pbm_ctx c, uint32_t n,
                                                       const int32_t* __restrict__ ids,
                                                       const uint32_t* __restrict__ counts) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
        if (ids[r] >= 0 && counts[r] > 0) pbm::free_n_pages(c, ids[r], counts[r], &st);
    pbm::flush_stats(c, st);
}

// Per-thread getXPage, one request per loop iteration.
__global__ __launch_bounds__(1024) void k_get_x_pages(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                      const uint32_t* __restrict__ counts,
                                                      int32_t* __restrict__ ids,
                                                      uint32_t* __restrict__ steps_out) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        uint32_t steps = 0;
        const int32_t id = pbm::get_x_pages(c, counts[r], pbm::request_seed(c.seed, req_base + r),
                                            &steps, &st);
        ids[r] = id;
        if (steps_out) steps_out[r] = steps;
    }
    pbm::flush_stats(c, st);
}

template <int SPANS>
__global__ __launch_bounds__(1024) void k_malloc(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                 const uint64_t* __restrict__ sizes,
                                                 uint64_t* __restrict__ ptrs,
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
            void* p = pbm::malloc_rw(c, sizes[r],
                                     pbm::request_seed(c.seed, req_base + (uint64_t)ch * 64u),
                                     &steps, &st, SPANS);
            ptrs[r] = (uint64_t)p;
            if (steps_out) steps_out[r] = steps;
        }
    }
    pbm::flush_stats(c, st);
}

__global__ __launch_bounds__(1024) void k_free_ptrs(pbm_ctx c, uint32_t n,
                                                    const uint64_t* __restrict__ ptrs) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
        pbm::free_rw(c, (void*)ptrs[r], &st);
    pbm::flush_stats(c, st);
}

namespace {

int prep(pbm_handle* h, const pbm_launch* in, pbm_launch* L, pbm_ctx* c) {
    pbm_launch x{};
    if (in) x = *in;
    if (x.block == 0) x.block = 256;
    if (x.block % 64 != 0 || x.block > 1024)
        return pbmi::fail(PBM_ERR_INVALID, "launch.block must be a multiple of 64, at most 1024");
    if (x.mode != PBM_MODE_PLAIN && x.mode != PBM_MODE_SERIAL)
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "multi-page calls run plain or serial");
    if (x.cluster_words == 0) x.cluster_words = 64;
    if (x.cluster_words != 64 && x.cluster_words != 128 && x.cluster_words != 256)
        return pbmi::fail(PBM_ERR_INVALID, "launch.cluster_words must be 64, 128 or 256");
    *L = x;
    *c = h->ctx;
    if (x.walk_budget) c->walk_budget = x.walk_budget;
    if (x.flags & PBM_LAUNCH_NO_COUNTERS) c->counters = nullptr;
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "multi-page calls need a bitmap strategy");
    return PBM_OK;
}

dim3 wave_grid(const pbm_launch& L, uint32_t n) {
    if (L.mode == PBM_MODE_SERIAL) return dim3(1);
    uint32_t waves = (n + 63u) / 64u;
    return dim3((waves * 64u + L.block - 1) / L.block);
}

dim3 wave_block(const pbm_launch& L) { return L.mode == PBM_MODE_SERIAL ? dim3(64) : dim3(L.block); }

dim3 thread_grid(const pbm_launch& L, uint32_t n) {
    return L.mode == PBM_MODE_SERIAL ? dim3(1) : dim3((n + L.block - 1) / L.block);
}

dim3 thread_block(const pbm_launch& L) { return L.mode == PBM_MODE_SERIAL ? dim3(1) : dim3(L.block); }

}  // namespace

extern "C" int pbm_get_n_pages(pbm_handle* h, uint32_t n, uint64_t req_base,
                               const uint32_t* d_counts, int32_t* d_ids, uint32_t* d_steps,
                               const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_counts || !d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_counts or d_ids");
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(h, launch, &L, &c);
    if (rc) return rc;
    auto k = L.cluster_words == 256 ? k_get_n_pages<4>
             : L.cluster_words == 128 ? k_get_n_pages<2> : k_get_n_pages<1>;
    hipLaunchKernelGGL(k, wave_grid(L, n), wave_block(L), 0, pbmi::as_stream(stream), c, n,
                       req_base, d_counts, d_ids, d_steps);
    PBM_HIP(hipGetLastError());
    return pbmi::shadow_note(h, n, req_base, d_ids, d_counts, false, pbmi::as_stream(stream));
}

extern "C" int pbm_get_x_pages(pbm_handle* h, uint32_t n, uint64_t req_base,
                               const uint32_t* d_counts, int32_t* d_ids, uint32_t* d_steps,
                               const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_counts || !d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_counts or d_ids");
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(h, launch, &L, &c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_get_x_pages, thread_grid(L, n), thread_block(L), 0,
                       pbmi::as_stream(stream), c, n, req_base, d_counts, d_ids, d_steps);
    PBM_HIP(hipGetLastError());
    return pbmi::shadow_note(h, n, req_base, d_ids, d_counts, false, pbmi::as_stream(stream));
}

extern "C" int pbm_free_n_pages(pbm_handle* h, uint32_t n, const int32_t* d_ids,
                                const uint32_t* d_counts, const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_counts || !d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_counts or d_ids");
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(h, launch, &L, &c);
    if (rc) return rc;
    if ((rc = pbmi::shadow_note(h, n, 0, d_ids, d_counts, true, pbmi::as_stream(stream))))
        return rc;
    hipLaunchKernelGGL(k_free_n_pages, thread_grid(L, n), thread_block(L), 0,
                       pbmi::as_stream(stream), c, n, d_ids, d_counts);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_malloc(pbm_handle* h, uint32_t n, uint64_t req_base, const uint64_t* d_sizes,
                          uint64_t* d_ptrs, uint32_t* d_steps, const pbm_launch* launch,
                          void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_sizes || !d_ptrs) return pbmi::fail(PBM_ERR_INVALID, "null d_sizes or d_ptrs");
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "mallocRW needs a pool with backing");
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(h, launch, &L, &c);
    if (rc) return rc;
    auto k = L.cluster_words == 256 ? k_malloc<4> : L.cluster_words == 128 ? k_malloc<2> : k_malloc<1>;
    hipLaunchKernelGGL(k, wave_grid(L, n), wave_block(L), 0, pbmi::as_stream(stream), c, n,
                       req_base, d_sizes, d_ptrs, d_steps);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_free_ptrs(pbm_handle* h, uint32_t n, const uint64_t* d_ptrs,
                             const pbm_launch* launch, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (n == 0) return PBM_OK;
    if (!d_ptrs) return pbmi::fail(PBM_ERR_INVALID, "null d_ptrs");
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "freeRW needs a pool with backing");
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(h, launch, &L, &c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_free_ptrs, thread_grid(L, n), thread_block(L), 0, pbmi::as_stream(stream),
                       c, n, d_ptrs);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_page_address(pbm_handle* h, int32_t id, uint64_t* addr) {
    if (!h || !addr) return pbmi::fail(PBM_ERR_INVALID, "null argument");
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "pool has no backing memory");
    if (id < 0 || (uint32_t)id >= h->ctx.n_pages) return pbmi::fail(PBM_ERR_RANGE, "page id out of range");
    *addr = (uint64_t)h->d_pool + (uint64_t)(uint32_t)id * h->ctx.page_size;
    return PBM_OK;
}

extern "C" int pbm_page_id(pbm_handle* h, uint64_t addr, int32_t* id) {
    if (!h || !id) return pbmi::fail(PBM_ERR_INVALID, "null argument");
    *id = -1;
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "pool has no backing memory");
    uint64_t base = (uint64_t)h->d_pool;
    uint64_t span = (uint64_t)h->ctx.n_pages * h->ctx.page_size;
    if (addr >= base && addr < base + span) *id = (int32_t)((addr - base) / h->ctx.page_size);
    return PBM_OK;
}
