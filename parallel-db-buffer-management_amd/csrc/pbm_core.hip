// Pool lifecycle, pre-fill, occupancy queries and snapshots (C-ABI).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "pbm_internal.hpp"

namespace pbmi {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    g_last_error = std::string("HIP error '") + hipGetErrorString(e) + "' in " + what + " at " +
                   file + ":" + std::to_string(line);
    // the failure is reported by the return code: take it off the thread's
    // HIP error state, or the caller's next HIP call would report it again
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? PBM_ERR_OOM : PBM_ERR_HIP;
}

int use_device(pbm_handle* h) {
    int cur = -1;
    PBM_HIP(hipGetDevice(&cur));
    if (cur != h->device) PBM_HIP(hipSetDevice(h->device));
    return PBM_OK;
}

}  // namespace pbmi

using pbm::u64;

// ---------------------------------------------------------------- kernels

// Every word free, padding bits past n_pages set.
__global__ void k_clear_map(u64* bm, uint32_t n_words, uint32_t n_pages) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_words) return;
    u64 w = 0;
    uint64_t first = (uint64_t)i * 64u;
    if (first + 64u > n_pages) {
        uint32_t valid = (uint32_t)(n_pages - first);
        w = ~((1ull << valid) - 1ull);
    }
    bm[i] = w;
}

// Page p used iff page_hash(seed,p) < thr (or all_used).  Padding bits set.
__global__ void k_prefill_bernoulli(u64* bm, uint32_t n_words, uint32_t n_pages, u64 thr,
                                    int all_used, u64 seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_words) return;
    u64 w = 0;
    const u64 s = pbm::mix64(seed ^ pbm::kPrefillSalt);
    for (uint32_t b = 0; b < 64u; ++b) {
        uint64_t p = (uint64_t)i * 64u + b;
        bool used;
        if (p >= n_pages) used = true;
        else used = all_used || pbm::mix64(s + (p + 1) * pbm::kGolden) < thr;
        w |= (u64)used << b;
    }
    bm[i] = w;
}

// Free pages (zero bits, padding excluded) summed into *out.
__global__ void k_count_free(const u64* bm, uint32_t n_words, uint32_t n_pages, u64* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v = 0;
    for (; i < n_words; i += gridDim.x * blockDim.x) {
        u64 w = bm[i];
        uint64_t first = (uint64_t)i * 64u;
        if (first + 64u > n_pages) {
            uint32_t valid = (uint32_t)(n_pages - first);
            w |= ~((1ull << valid) - 1ull);
        }
        v += 64u - (uint32_t)__popcll(w);
    }
    for (int d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 64);
    if (pbm::lane_id() == 0 && v) pbm::atom_add_u64(out, v);
}

// Sets the padding bits past n_pages in the last word (a map restored from
// the host may have them clear; the allocators hand out any clear bit).
__global__ void k_set_padding(u64* bm, uint32_t n_words, uint32_t n_pages) {
    if (blockIdx.x != 0 || threadIdx.x != 0 || n_words == 0) return;
    const uint64_t first = (uint64_t)(n_words - 1u) * 64u;
    const uint32_t valid = (uint32_t)(n_pages - first);
    if (valid < 64u) bm[n_words - 1u] |= ~((1ull << valid) - 1ull);
}

__global__ void k_set_u64(u64* p, uint32_t n, u64 v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_free_from_scratch(long long* free_count, const u64* scratch) {
    *free_count = (long long)scratch[0];
}

__global__ void k_fill_i32(int32_t* p, uint32_t n, int32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// DistClock arms: arm[t] = (uint32)lcg(request_seed(seed^salt, t)) % n_pages.
// The reference draws LCG((tid<<15)+clock) % N, which can be negative
// (parallelPage.cu:546-554); the unsigned modulo keeps it in range.
__global__ void k_init_arms(unsigned* arms, uint32_t n, u64 seed, uint32_t n_pages) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t s = pbm::lcg(pbm::request_seed(seed ^ pbm::kArmSalt, i));
    arms[i] = (uint32_t)s % n_pages;
}

static inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

namespace pbmi {

int recount_free(pbm_handle* h, hipStream_t s) {
    const pbm_ctx& c = h->ctx;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, h->d_scratch, 1u, 0ull);
    uint32_t nb = blocks_for(c.n_words, 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_count_free, dim3(nb), dim3(256), 0, s, h->d_bitmap, c.n_words, c.n_pages,
                       h->d_scratch);
    hipLaunchKernelGGL(k_free_from_scratch, dim3(1), dim3(1), 0, s, h->d_free, h->d_scratch);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

int reinit_strategy_state(pbm_handle* h, hipStream_t s) {
    const pbm_ctx& c = h->ctx;
    if (int rc = shadow_sync(h, s)) return rc;
    if (h->d_hint)
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(c.max_threads, 256)), dim3(256), 0, s,
                           h->d_hint, c.max_threads, -1);
    if (h->d_arms)
        hipLaunchKernelGGL(k_init_arms, dim3(blocks_for(c.max_threads, 256)), dim3(256), 0, s,
                           h->d_arms, c.max_threads, c.seed, c.n_pages);
    if (h->d_clock)
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, h->d_clock, 1u, 0ull);
    PBM_HIP(hipGetLastError());
    if (h->d_ring) {
        int rc = ring_rebuild(h, s);
        if (rc) return rc;
    }
    return PBM_OK;
}

}  // namespace pbmi

// ---------------------------------------------------------------- C-ABI

static int validate(const pbm_config* cfg) {
    if (!cfg) return pbmi::fail(PBM_ERR_INVALID, "null pbm_config");
    if (cfg->n_pages == 0 || cfg->n_pages > (1u << 30))
        return pbmi::fail(PBM_ERR_RANGE, "n_pages must be in [1, 2^30]");
    if (cfg->page_size < 8) return pbmi::fail(PBM_ERR_INVALID, "page_size must be >= 8");
    if (cfg->strategy < PBM_RW_BM || cfg->strategy > PBM_OUROBOROS)
        return pbmi::fail(PBM_ERR_INVALID, "unknown strategy");
    if (cfg->order != PBM_ORDER_PACKED && cfg->order != PBM_ORDER_REFERENCE)
        return pbmi::fail(PBM_ERR_INVALID, "unknown collaborative order");
    return PBM_OK;
}

extern "C" int pbm_create(const pbm_config* cfg, pbm_handle** out) {
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output handle pointer");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int n_dev = 0;
    PBM_HIP(hipGetDeviceCount(&n_dev));
    if (cfg->device < 0 || cfg->device >= n_dev)
        return pbmi::fail(PBM_ERR_INVALID, "device ordinal out of range");

    pbm_handle* h = new pbm_handle();
    h->cfg = *cfg;
    h->device = cfg->device;
    if (h->cfg.walk_budget == 0) h->cfg.walk_budget = 1u << 20;
    if (h->cfg.max_threads == 0) h->cfg.max_threads = 1u << 20;
    auto cleanup = [&](int code) {
        pbm_destroy(h);
        return code;
    };
    if ((rc = pbmi::use_device(h))) return cleanup(rc);
    hipError_t e = hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e != hipSuccess) return cleanup(pbmi::hip_fail(e, "hipDeviceGetAttribute", __FILE__, __LINE__));

    const uint32_t n_words = (cfg->n_pages + 63u) / 64u;
#define PBM_ALLOC(ptr, bytes)                                                        \
    do {                                                                             \
        hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));                          \
        if (e_ != hipSuccess) return cleanup(pbmi::hip_fail(e_, "hipMalloc(" #ptr ")", __FILE__, __LINE__)); \
    } while (0)
    PBM_ALLOC(h->d_bitmap, (size_t)n_words * 8u);
    PBM_ALLOC(h->d_free, 8u);
    PBM_ALLOC(h->d_counters, sizeof(pbm_counters));
    PBM_ALLOC(h->d_phase, 64u);
    PBM_ALLOC(h->d_scratch, 64u);
    // four words per thread once a map has 8,192 words: fewer, larger
    // tiles shorten dispatch and the look-back chain
    // rank-scan tiles: 4 words per thread on maps of 2^20 words and more;
    // one wave per tile on small maps, so their grants spread over more CUs
    const uint32_t scan_wpt = n_words >= (1u << 20) ? 4u : 1u;
    const uint32_t scan_threads = n_words <= PBM_SMALL_SCAN_WORDS ? 64u : pbm::kScanThreads;
    const uint32_t tile_words = scan_threads * scan_wpt;
    const uint32_t n_tiles = (n_words + tile_words - 1u) / tile_words;
    // one look-back word per tile, then one per group of 64 tiles, 32 words
    // apart (the two-level look-back of maps with more than 256 tiles)
    const size_t scan_words = (size_t)n_tiles + 32u * ((n_tiles + 63u) / 64u);
    PBM_ALLOC(h->d_scan, scan_words * 8u);
    PBM_ALLOC(h->d_scan_ctl, 32u);
    PBM_ALLOC(h->d_credit, 16u * 32u * 8u);
    if (cfg->backing) PBM_ALLOC(h->d_pool, (size_t)cfg->n_pages * cfg->page_size);
    if (cfg->strategy == PBM_CRW_BM || cfg->strategy == PBM_CRW_BASIC)
        PBM_ALLOC(h->d_hint, (size_t)h->cfg.max_threads * 4u);
    if (cfg->strategy == PBM_DIST_CLOCK) PBM_ALLOC(h->d_arms, (size_t)h->cfg.max_threads * 4u);
    if (cfg->strategy == PBM_SINGLE_CLOCK) PBM_ALLOC(h->d_clock, 8u);
    if (cfg->strategy == PBM_LINKED_LIST || cfg->strategy == PBM_OUROBOROS) {
        PBM_ALLOC(h->d_ring, (size_t)cfg->n_pages * 8u);
        PBM_ALLOC(h->d_ring_ht, 16u);
    }
    if (cfg->strategy == PBM_OUROBOROS) {
        PBM_ALLOC(h->d_fresh, (size_t)cfg->n_pages * 4u);
        PBM_ALLOC(h->d_carve, 16u);
        PBM_ALLOC(h->d_qavail, 8u);
    }
#undef PBM_ALLOC

    pbm_ctx& c = h->ctx;
    std::memset(&c, 0, sizeof(c));
    c.bitmap = h->d_bitmap;
    c.free_count = h->d_free;
    c.counters = h->d_counters;
    c.pool = h->d_pool;
    c.hint = h->d_hint;
    c.arms = h->d_arms;
    c.clock_arm = h->d_clock;
    c.ring = h->d_ring;
    c.ring_head = h->d_ring_ht;
    c.ring_tail = h->d_ring_ht ? h->d_ring_ht + 1 : nullptr;
    c.phase = h->d_phase;
    c.scan = h->d_scan;
    c.scan_ctl = h->d_scan_ctl;
    c.n_tiles = n_tiles;
    c.scan_wpt = scan_wpt;
    c.scan_threads = scan_threads;
    c.n_pages = cfg->n_pages;
    c.n_words = n_words;
    c.page_size = cfg->page_size;
    c.walk_budget = h->cfg.walk_budget;
    c.max_threads = h->cfg.max_threads;
    c.order = cfg->order;
    c.seed = cfg->seed;
    c.fresh = h->d_fresh;
    c.carve = h->d_carve;
    c.q_avail = h->d_qavail;
    c.chunk_pages = pbmi::ouroboros_chunk_pages(cfg->page_size);

    e = hipMemset(h->d_counters, 0, sizeof(pbm_counters));
    if (e != hipSuccess) return cleanup(pbmi::hip_fail(e, "hipMemset", __FILE__, __LINE__));
    e = hipMemset(h->d_phase, 0, 64u);
    if (e != hipSuccess) return cleanup(pbmi::hip_fail(e, "hipMemset", __FILE__, __LINE__));
    // Look-back words start with call parity 0: the first call has parity 1
    // and no zeroed word reads as published by it.
    e = hipMemset(h->d_scan, 0, scan_words * 8u);
    if (e == hipSuccess) e = hipMemset(h->d_credit, 0, 16u * 32u * 8u);
    if (e != hipSuccess) return cleanup(pbmi::hip_fail(e, "hipMemset", __FILE__, __LINE__));
    {
        const unsigned long long ctl[4] = {0, 0, 0, 0};
        e = hipMemcpy(h->d_scan_ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(pbmi::hip_fail(e, "hipMemcpy", __FILE__, __LINE__));
    }
    if ((rc = pbm_reset(h, nullptr))) return cleanup(rc);
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return cleanup(pbmi::hip_fail(e, "hipDeviceSynchronize", __FILE__, __LINE__));
    *out = h;
    return PBM_OK;
}

extern "C" int pbm_destroy(pbm_handle* h) {
    if (!h) return PBM_OK;
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != h->device) (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    void* ptrs[] = {h->d_bitmap, h->d_free, h->d_counters, h->d_pool, h->d_hint, h->d_arms,
                    h->d_clock, h->d_ring, h->d_ring_ht, h->d_phase, h->d_scratch,
                    h->d_scan, h->d_scan_ctl, h->d_credit, h->d_fresh, h->d_carve, h->d_qavail,
                    h->d_shadow, h->d_shadow_rep};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete h;
    return PBM_OK;
}

extern "C" int pbm_reset(pbm_handle* h, void* stream) {
    PBM_CHECK_HANDLE(h);
    hipStream_t s = pbmi::as_stream(stream);
    const pbm_ctx& c = h->ctx;
    hipLaunchKernelGGL(k_clear_map, dim3(blocks_for(c.n_words, 256)), dim3(256), 0, s, h->d_bitmap,
                       c.n_words, c.n_pages);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, (u64*)h->d_free, 1u, (u64)c.n_pages);
    PBM_HIP(hipGetLastError());
    return pbmi::reinit_strategy_state(h, s);
}

// floor(fill * 2^64) for 0 <= fill < 1; fill is a double, so the product is exact.
static u64 fill_threshold(double fill, int* all_used) {
    *all_used = 0;
    if (!(fill > 0.0)) return 0;
    if (fill >= 1.0) {
        *all_used = 1;
        return ~0ull;
    }
    return (u64)std::ldexp(fill, 64);
}

extern "C" int pbm_prefill_bernoulli(pbm_handle* h, double fill, uint64_t seed, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!(fill >= 0.0 && fill <= 1.0)) return pbmi::fail(PBM_ERR_INVALID, "fill must be in [0,1]");
    hipStream_t s = pbmi::as_stream(stream);
    const pbm_ctx& c = h->ctx;
    int all_used = 0;
    u64 thr = fill_threshold(fill, &all_used);
    hipLaunchKernelGGL(k_prefill_bernoulli, dim3(blocks_for(c.n_words, 256)), dim3(256), 0, s,
                       h->d_bitmap, c.n_words, c.n_pages, thr, all_used, (u64)seed);
    PBM_HIP(hipGetLastError());
    int rc = pbmi::recount_free(h, s);
    if (rc) return rc;
    return pbmi::reinit_strategy_state(h, s);
}

extern "C" int pbm_prefill_requests(pbm_handle* h, float free_percentage, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!(free_percentage >= 0.0f && free_percentage <= 1.0f))
        return pbmi::fail(PBM_ERR_INVALID, "freePercentage must be in [0,1]");
    // Same float arithmetic as prefillBuffer (RandomWalkBasic_BM.cuh:154).
    int n_req = (int)((1.0f - free_percentage) * (float)h->ctx.n_pages);
    if (n_req <= 0) return PBM_OK;
    hipStream_t s = pbmi::as_stream(stream);
    int32_t* d_ids = nullptr;
    PBM_HIP(hipMalloc((void**)&d_ids, (size_t)n_req * 4u));
    pbm_launch L{};
    L.mode = PBM_MODE_PLAIN;
    L.flags = PBM_LAUNCH_NO_COUNTERS;
    int rc = pbm_get_pages(h, (uint32_t)n_req, 0xFFFF000000000000ull, d_ids, nullptr, &L, stream);
    hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(d_ids);
    if (rc) return rc;
    if (e != hipSuccess) return pbmi::hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    return PBM_OK;
}

extern "C" int pbm_count_used(pbm_handle* h, uint64_t* used, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!used) return pbmi::fail(PBM_ERR_INVALID, "null output");
    hipStream_t s = pbmi::as_stream(stream);
    const pbm_ctx& c = h->ctx;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, h->d_scratch + 1, 1u, 0ull);
    uint32_t nb = blocks_for(c.n_words, 256);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_count_free, dim3(nb), dim3(256), 0, s, h->d_bitmap, c.n_words, c.n_pages,
                       h->d_scratch + 1);
    PBM_HIP(hipGetLastError());
    u64 free_pages = 0;
    PBM_HIP(hipMemcpyAsync(&free_pages, h->d_scratch + 1, 8u, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    *used = (uint64_t)c.n_pages - free_pages;
    return PBM_OK;
}

extern "C" int pbm_free_page_percentage(pbm_handle* h, float* out, void* stream) {
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output");
    uint64_t used = 0;
    int rc = pbm_count_used(h, &used, stream);
    if (rc) return rc;
    // Same float expression as getFreePagePercentage (RandomWalkBasic_BM.cuh:144).
    *out = (float)((int64_t)h->ctx.n_pages - (int64_t)used) / (float)h->ctx.n_pages;
    return PBM_OK;
}

extern "C" int pbm_snapshot_bitmap(pbm_handle* h, uint64_t* host, size_t words, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!host || words != h->ctx.n_words)
        return pbmi::fail(PBM_ERR_INVALID, "snapshot buffer must hold exactly n_words words");
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(host, h->d_bitmap, words * 8u, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}

extern "C" int pbm_restore_bitmap(pbm_handle* h, const uint64_t* host, size_t words, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!host || words != h->ctx.n_words)
        return pbmi::fail(PBM_ERR_INVALID, "restore buffer must hold exactly n_words words");
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(h->d_bitmap, host, words * 8u, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_set_padding, dim3(1), dim3(64), 0, s, h->d_bitmap, h->ctx.n_words,
                       h->ctx.n_pages);
    PBM_HIP(hipGetLastError());
    PBM_HIP(hipStreamSynchronize(s));
    int rc = pbmi::recount_free(h, s);
    if (rc) return rc;
    rc = pbmi::reinit_strategy_state(h, s);
    if (rc) return rc;
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}

extern "C" int pbm_read_free_count(pbm_handle* h, int64_t* out, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output");
    hipStream_t s = pbmi::as_stream(stream);
    long long v = 0;
    PBM_HIP(hipMemcpyAsync(&v, h->d_free, 8u, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    *out = v;
    return PBM_OK;
}

extern "C" int pbm_get_ctx(pbm_handle* h, pbm_ctx* out) {
    if (!h || !out) return pbmi::fail(PBM_ERR_INVALID, "null argument");
    *out = h->ctx;
    return PBM_OK;
}

extern "C" int pbm_read_counters(pbm_handle* h, pbm_counters* out, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output");
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(out, h->d_counters, sizeof(pbm_counters), hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}

extern "C" int pbm_clear_counters(pbm_handle* h, void* stream) {
    PBM_CHECK_HANDLE(h);
    hipStream_t s = pbmi::as_stream(stream);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, (u64*)h->d_counters,
                       (uint32_t)(sizeof(pbm_counters) / 8u), 0ull);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" const char* pbm_last_error(void) { return pbmi::g_last_error.c_str(); }

extern "C" const char* pbm_version(void) { return "pbm 0.1.0 gfx950"; }
