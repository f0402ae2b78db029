// Kernel malloc managers (KernelMallocSimple / KernelMallocCollab,
// malloc.cuh:14-37): batch malloc and free over the device functions of
// pbm_device.hpp, and a host read of page memory for checking the
// on-page headers.
#include <hip/hip_runtime.h>
#include "pbm_internal.hpp"

struct pbm_kmalloc {
    pbm_handle* h = nullptr;
    int32_t kind = PBM_KMALLOC_SIMPLE;
    uint32_t max_threads = 0;
    int32_t* d_current = nullptr;  // Simple: current chunk per thread slot, -1 none
};

namespace {

// Simple: thread gid < n_threads serves requests gid, gid+n_threads, ...
// from slot gid.
__global__ __launch_bounds__(1024) void k_kmalloc_simple(pbm_ctx c, int32_t* current, uint32_t n,
                                                         uint32_t n_threads, uint64_t req_base,
                                                         const uint32_t* __restrict__ sizes,
                                                         uint64_t* __restrict__ ptrs,
                                                         uint32_t* __restrict__ steps_out) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n_threads) return;
    pbm::Stats st;
    for (uint32_t r = gid; r < n; r += n_threads) {
        uint32_t steps = 0;
        void* p = pbm::kmalloc_simple(c, &current[gid], sizes[r],
                                      pbm::request_seed(c.seed, req_base + r), &steps, &st);
        ptrs[r] = (uint64_t)p;
        if (steps_out) steps_out[r] = steps;
    }
    pbm::flush_stats(c, st);
}

// Collab: a wave serves chunks of 64 consecutive requests.
__global__ __launch_bounds__(1024) void k_kmalloc_collab(pbm_ctx c, uint32_t n, uint64_t req_base,
                                                         const uint32_t* __restrict__ sizes,
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
            void* p = pbm::kmalloc_collab(c, sizes[r], pbm::request_seed(c.seed, req_base + r),
                                          &steps, &st);
            ptrs[r] = (uint64_t)p;
            if (steps_out) steps_out[r] = steps;
        }
    }
    pbm::flush_stats(c, st);
}

__global__ __launch_bounds__(1024) void k_kfree_collab(pbm_ctx c, uint32_t n,
                                                       const uint64_t* __restrict__ ptrs) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
        pbm::kfree_collab(c, (void*)ptrs[r], &st);
    pbm::flush_stats(c, st);
}

__global__ void k_fill_i32(int32_t* p, uint32_t n, int32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int prep(pbm_kmalloc* k, const pbm_launch* in, pbm_launch* L, pbm_ctx* c) {
    if (!k) return pbmi::fail(PBM_ERR_INVALID, "null pbm_kmalloc");
    int rc = pbmi::use_device(k->h);
    if (rc) return rc;
    pbm_launch x{};
    if (in) x = *in;
    if (x.block == 0) x.block = 256;
    if (x.block % 64 != 0 || x.block > 1024)
        return pbmi::fail(PBM_ERR_INVALID, "launch.block must be a multiple of 64, at most 1024");
    if (x.mode != PBM_MODE_PLAIN && x.mode != PBM_MODE_SERIAL)
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "kernel malloc runs plain or serial");
    *L = x;
    *c = k->h->ctx;
    if (x.walk_budget) c->walk_budget = x.walk_budget;
    if (x.flags & PBM_LAUNCH_NO_COUNTERS) c->counters = nullptr;
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_kmalloc_create(pbm_handle* h, int32_t kind, uint32_t max_threads,
                                  pbm_kmalloc** out) {
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output pointer");
    *out = nullptr;
    PBM_CHECK_HANDLE(h);
    if (kind != PBM_KMALLOC_SIMPLE && kind != PBM_KMALLOC_COLLAB)
        return pbmi::fail(PBM_ERR_INVALID, "unknown kernel malloc kind");
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "kernel malloc needs a pool with backing");
    if (h->cfg.page_size % 8u)
        return pbmi::fail(PBM_ERR_INVALID, "kernel malloc needs a page size that is a multiple of 8");
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "kernel malloc needs a bitmap strategy");
    pbm_kmalloc* k = new pbm_kmalloc();
    k->h = h;
    k->kind = kind;
    if (kind == PBM_KMALLOC_SIMPLE) {
        k->max_threads = max_threads ? max_threads : (1u << 20);
        hipError_t e = hipMalloc((void**)&k->d_current, (size_t)k->max_threads * 4u);
        if (e != hipSuccess) {
            delete k;
            return pbmi::hip_fail(e, "hipMalloc(current chunks)", __FILE__, __LINE__);
        }
        hipLaunchKernelGGL(k_fill_i32, dim3((k->max_threads + 255u) / 256u), dim3(256), 0, 0,
                           k->d_current, k->max_threads, -1);
        e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            hipFree(k->d_current);
            delete k;
            return pbmi::hip_fail(e, "initialise current chunks", __FILE__, __LINE__);
        }
    }
    *out = k;
    return PBM_OK;
}

extern "C" int pbm_kmalloc_destroy(pbm_kmalloc* k) {
    if (!k) return PBM_OK;
    int rc = pbmi::use_device(k->h);
    if (k->d_current) hipFree(k->d_current);
    delete k;
    return rc;
}

extern "C" int pbm_kmalloc_reset(pbm_kmalloc* k, void* stream) {
    if (!k) return pbmi::fail(PBM_ERR_INVALID, "null pbm_kmalloc");
    int rc = pbm_reset(k->h, stream);
    if (rc) return rc;
    if (k->d_current) {
        hipLaunchKernelGGL(k_fill_i32, dim3((k->max_threads + 255u) / 256u), dim3(256), 0,
                           pbmi::as_stream(stream), k->d_current, k->max_threads, -1);
        PBM_HIP(hipGetLastError());
    }
    return PBM_OK;
}

extern "C" int pbm_kmalloc_malloc(pbm_kmalloc* k, uint32_t n, uint64_t req_base,
                                  const uint32_t* d_sizes, uint64_t* d_ptrs, uint32_t* d_steps,
                                  const pbm_launch* launch, void* stream) {
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(k, launch, &L, &c);
    if (rc) return rc;
    if (n == 0) return PBM_OK;
    if (!d_sizes || !d_ptrs) return pbmi::fail(PBM_ERR_INVALID, "null d_sizes or d_ptrs");
    const bool serial = L.mode == PBM_MODE_SERIAL;
    hipStream_t s = pbmi::as_stream(stream);
    if (k->kind == PBM_KMALLOC_SIMPLE) {
        // one slot per thread: no more threads than slots
        const uint32_t threads = serial ? 1u : (n < k->max_threads ? n : k->max_threads);
        const dim3 grid = serial ? dim3(1) : dim3((threads + L.block - 1) / L.block);
        const dim3 blk = serial ? dim3(1) : dim3(L.block);
        hipLaunchKernelGGL(k_kmalloc_simple, grid, blk, 0, s, c, k->d_current, n, threads,
                           req_base, d_sizes, d_ptrs, d_steps);
    } else {
        const uint32_t waves = (n + 63u) / 64u;
        const dim3 grid = serial ? dim3(1) : dim3((waves * 64u + L.block - 1) / L.block);
        const dim3 blk = serial ? dim3(64) : dim3(L.block);
        hipLaunchKernelGGL(k_kmalloc_collab, grid, blk, 0, s, c, n, req_base, d_sizes, d_ptrs,
                           d_steps);
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_kmalloc_free(pbm_kmalloc* k, uint32_t n, const uint64_t* d_ptrs,
                                const pbm_launch* launch, void* stream) {
    pbm_launch L;
    pbm_ctx c;
    int rc = prep(k, launch, &L, &c);
    if (rc) return rc;
    // KernelMallocSimple::free is empty in the reference (malloc.cu:178-181):
    // its chunks come back with reset
    if (n == 0 || k->kind == PBM_KMALLOC_SIMPLE) return PBM_OK;
    if (!d_ptrs) return pbmi::fail(PBM_ERR_INVALID, "null d_ptrs");
    const bool serial = L.mode == PBM_MODE_SERIAL;
    hipLaunchKernelGGL(k_kfree_collab, serial ? dim3(1) : dim3((n + L.block - 1) / L.block),
                       serial ? dim3(1) : dim3(L.block), 0, pbmi::as_stream(stream), c, n, d_ptrs);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_kmalloc_current(pbm_kmalloc* k, int32_t* host_current, uint32_t* max_threads) {
    if (!k) return pbmi::fail(PBM_ERR_INVALID, "null pbm_kmalloc");
    if (max_threads) *max_threads = k->max_threads;
    if (host_current) {
        if (!k->d_current) return pbmi::fail(PBM_ERR_UNSUPPORTED, "only KernelMallocSimple keeps chunks");
        int rc = pbmi::use_device(k->h);
        if (rc) return rc;
        PBM_HIP(hipMemcpy(host_current, k->d_current, (size_t)k->max_threads * 4u,
                          hipMemcpyDeviceToHost));
    }
    return PBM_OK;
}

extern "C" int pbm_read_pool(pbm_handle* h, uint64_t offset, uint64_t bytes, void* host,
                             void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "pool has no backing memory");
    const uint64_t size = (uint64_t)h->cfg.n_pages * h->cfg.page_size;
    if (offset > size || bytes > size - offset)
        return pbmi::fail(PBM_ERR_RANGE, "read beyond the page memory");
    if (bytes == 0) return PBM_OK;
    if (!host) return pbmi::fail(PBM_ERR_INVALID, "null host buffer");
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(host, h->d_pool + offset, bytes, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}
