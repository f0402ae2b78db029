// Buffer cache over a pool: logical pages are kept in `capacity` slots, each
// slot holding one frame (a pool page) got with the pool strategy's device
// getPage; CLOCK replacement.  References are taken a batch at a time:
//   hits        every reference whose page is present sets its slot's bit;
//               the first reference to an absent page is a miss, later ones
//               in the batch are hits (an atomicMax per page keeps the first)
//   misses      compacted in reference order by a scan of first-occurrence
//               flags
//   victims     a distributed CLOCK sweep: one scan over the ring of slots
//               from the arm counts, per position, the clear and the set
//               bits of slots occupied when the batch began; the first
//               victims are the clear slots in ring order, and when a pass
//               has too few, the set slots in ring order after every bit
//               was cleared; bits passed before the last victim are cleared
//   evict/load  victims' pages are dropped and their frames freed, then the
//               misses load into empty slots first and victims' slots after,
//               each with a new frame and its bit set
// With a batch of one reference this is the textbook second-chance CLOCK.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
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
    cK,        // misses of the current batch
    cV,        // victims of the current batch
    cZ,        // clear slots in the ring at batch start
    cStop,     // ring position of the last victim
    cAllocFail,
    cR0,       // resident at batch start
    cEmpty,    // empty slots at batch start
    cCtlWords = 16
};

constexpr u64 kLow = 0xFFFFFFFFull;

struct CacheDev {
    pbm_ctx c;
    int32_t* slot_of;   // [n_logical] slot of a present page, -1 absent
    u64* first;         // [n_logical] (batch << 32) | ~index of the first miss
    int32_t* owner;     // [capacity] page in the slot
    int32_t* frame;     // [capacity] pool page holding it
    uint32_t* refbit;   // [capacity]
    uint32_t* mflag;    // [batch]
    uint32_t* mpos;     // [batch]
    int32_t* miss;      // [batch]
    u64* ring;          // [capacity] clear | set << 32, by ring position
    u64* ring_rank;     // [capacity] exclusive scan of ring
    int32_t* victim;    // [batch]
    int32_t* log;       // [log_cap] evicted pages in order
    u64 log_cap;
    u64* ctl;           // [cCtlWords]
    uint32_t n_logical, cap, batch;
};

__device__ __forceinline__ u64 ldc(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_count(u64* dst, bool x) {
    const u64 m = pbm::ballot(x);
    if (m && pbm::lane_id() == (unsigned)__ffsll((long long)m) - 1u)
        pbm::atom_add_u64(dst, (u64)__popcll(m));
}

__global__ __launch_bounds__(256) void k_cache_hits(CacheDev d, const int32_t* refs, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 tag_hi = (ldc(&d.ctl[cBatches]) + 1u) << 32;
    bool hit = false;
    if (i < nb) {
        const int32_t l = refs[i];
        const int32_t s = d.slot_of[l];
        if (s >= 0) {
            d.refbit[s] = 1u;
            hit = true;
        } else {
            pbm::atom_max_u64(&d.first[l], tag_hi | (kLow - i));
        }
    }
    wave_count(&d.ctl[cHits], hit);
}

__global__ __launch_bounds__(256) void k_cache_flag(CacheDev d, const int32_t* refs, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 tag_hi = (ldc(&d.ctl[cBatches]) + 1u) << 32;
    bool later = false;
    if (i < nb) {
        const int32_t l = refs[i];
        uint32_t f = 0;
        if (d.slot_of[l] < 0) {
            f = d.first[l] == (tag_hi | (kLow - i)) ? 1u : 0u;
            later = !f;
        }
        d.mflag[i] = f;
    }
    wave_count(&d.ctl[cHits], later);
}

// Miss list and the batch plan: K misses, empty slots, victims needed.
__global__ __launch_bounds__(256) void k_cache_plan(CacheDev d, const int32_t* refs, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb && d.mflag[i]) d.miss[d.mpos[i]] = refs[i];
    if (i == 0) {
        const u64 K = (u64)d.mpos[nb - 1] + d.mflag[nb - 1];
        const u64 r0 = d.ctl[cResident];
        const u64 empty = d.cap - r0;
        d.ctl[cK] = K;
        d.ctl[cR0] = r0;
        d.ctl[cEmpty] = empty;
        d.ctl[cV] = K > empty ? K - empty : 0u;
        d.ctl[cMisses] += K;
    }
}

// Ring position j is slot (arm + j) mod capacity.
__global__ __launch_bounds__(256) void k_cache_ring(CacheDev d) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.cap || ldc(&d.ctl[cV]) == 0) return;
    const u64 arm = ldc(&d.ctl[cArm]), r0 = ldc(&d.ctl[cR0]);
    const uint32_t s = (uint32_t)((arm + j) % d.cap);
    u64 f = 0;
    if (s < r0) f = d.refbit[s] ? (1ull << 32) : 1ull;
    d.ring[j] = f;
}

__global__ __launch_bounds__(256) void k_cache_select(CacheDev d) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 v = ldc(&d.ctl[cV]);
    if (j >= d.cap || v == 0) return;
    const u64 arm = ldc(&d.ctl[cArm]);
    const u64 z = (d.ring_rank[d.cap - 1] + d.ring[d.cap - 1]) & kLow;
    const u64 f = d.ring[j], r = d.ring_rank[j];
    const int32_t s = (int32_t)((arm + j) % d.cap);
    const u64 zr = r & kLow, orank = r >> 32;
    if (j == 0) d.ctl[cZ] = z;
    if (v <= z) {
        if ((f & kLow) && zr < v) {
            d.victim[zr] = s;
            if (zr == v - 1) d.ctl[cStop] = j;
        }
    } else {
        if (f & kLow) d.victim[zr] = s;
        if ((f >> 32) && orank < v - z) {
            d.victim[z + orank] = s;
            if (orank == v - z - 1) d.ctl[cStop] = j;
        }
    }
}

// Bits the sweep passed are cleared: before the last victim in one pass,
// all of them in two.
__global__ __launch_bounds__(256) void k_cache_clear(CacheDev d) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 v = ldc(&d.ctl[cV]);
    if (j >= d.cap || v == 0) return;
    const u64 z = ldc(&d.ctl[cZ]), stop = ldc(&d.ctl[cStop]);
    if ((d.ring[j] >> 32) && (v > z || j < stop)) {
        const u64 arm = ldc(&d.ctl[cArm]);
        d.refbit[(arm + j) % d.cap] = 0u;
    }
}

// Victims leave: their pages become absent and their frames go back.
__global__ __launch_bounds__(256) void k_cache_evict(CacheDev d) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 v = ldc(&d.ctl[cV]);
    pbm::Stats st;
    if (k < v) {
        const int32_t s = d.victim[k];
        const int32_t old = d.owner[s];
        d.slot_of[old] = -1;
        if (d.frame[s] >= 0) pbm::free_page(d.c, d.frame[s], &st);
        const u64 at = ldc(&d.ctl[cEvictions]) + k;
        if (at < d.log_cap) d.log[at] = old;
    }
    pbm::flush_stats(d.c, st);
}

// Misses load: into empty slots first, then into the victims' slots.
template <int S>
__global__ __launch_bounds__(256) void k_cache_load(CacheDev d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 K = ldc(&d.ctl[cK]), empty = ldc(&d.ctl[cEmpty]), r0 = ldc(&d.ctl[cR0]);
    const u64 batch_no = ldc(&d.ctl[cBatches]);
    pbm::Stats st;
    bool fail = false;
    if (i < K) {
        const int32_t l = d.miss[i];
        const int32_t s = i < empty ? (int32_t)(r0 + i) : d.victim[i - empty];
        const int32_t seed = pbm::request_seed(d.c.seed, (batch_no << 32) | i);
        const int32_t f = pbm::get_page_of<S>(d.c, i, seed, nullptr, &st);
        fail = f < 0;
        d.frame[s] = f;
        d.owner[s] = l;
        d.slot_of[l] = s;
        d.refbit[s] = 1u;
    }
    wave_count(&d.ctl[cAllocFail], fail);
    pbm::flush_stats(d.c, st);
}

// End of batch: arm past the last victim, residency, counters.
__global__ void k_cache_finish(CacheDev d) {
    const u64 v = d.ctl[cV], K = d.ctl[cK], r0 = d.ctl[cR0];
    if (v) d.ctl[cArm] = (d.ctl[cArm] + d.ctl[cStop] + 1u) % d.cap;
    d.ctl[cResident] = r0 + K < d.cap ? r0 + K : d.cap;
    d.ctl[cEvictions] += v;
    d.ctl[cBatches] += 1u;
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
    void* scan_tmp = nullptr;
    size_t scan_bytes_b = 0, scan_bytes_c = 0;
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
    PBM_CALLOC(mflag, uint32_t, cfg->batch)
    PBM_CALLOC(mpos, uint32_t, cfg->batch)
    PBM_CALLOC(miss, int32_t, cfg->batch)
    PBM_CALLOC(ring, u64, cfg->capacity)
    PBM_CALLOC(ring_rank, u64, cfg->capacity)
    PBM_CALLOC(victim, int32_t, cfg->batch)
    PBM_CALLOC(log, int32_t, cfg->log_capacity)
    PBM_CALLOC(ctl, u64, cCtlWords)
#undef PBM_CALLOC
    if (!rc) {
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, k->scan_bytes_b, d.mflag, d.mpos,
                                                        (int)cfg->batch);
        if (e == hipSuccess)
            e = hipcub::DeviceScan::ExclusiveSum(nullptr, k->scan_bytes_c, d.ring, d.ring_rank,
                                                 (int)cfg->capacity);
        if (e != hipSuccess) rc = pbmi::hip_fail(e, "hipcub scan sizing", __FILE__, __LINE__);
        else rc = alloc(&k->scan_tmp, std::max(k->scan_bytes_b, k->scan_bytes_c));
    }
    if (rc) {
        pbm_cache_destroy(k);
        return rc;
    }
    const uint32_t n_init = std::max(std::max(cfg->n_logical, cfg->capacity), (uint32_t)cCtlWords);
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

int cache_batch(pbm_cache* k, const int32_t* refs, uint32_t nb, hipStream_t s) {
    const CacheDev& d = k->d;
    const dim3 b(256), gb((nb + 255u) / 256u), gc((d.cap + 255u) / 256u);
    hipLaunchKernelGGL(k_cache_hits, gb, b, 0, s, d, refs, nb);
    hipLaunchKernelGGL(k_cache_flag, gb, b, 0, s, d, refs, nb);
    size_t bytes = k->scan_bytes_b;
    PBM_HIP(hipcub::DeviceScan::ExclusiveSum(k->scan_tmp, bytes, d.mflag, d.mpos, (int)nb, s));
    hipLaunchKernelGGL(k_cache_plan, gb, b, 0, s, d, refs, nb);
    hipLaunchKernelGGL(k_cache_ring, gc, b, 0, s, d);
    bytes = k->scan_bytes_c;
    PBM_HIP(hipcub::DeviceScan::ExclusiveSum(k->scan_tmp, bytes, d.ring, d.ring_rank, (int)d.cap, s));
    hipLaunchKernelGGL(k_cache_select, gc, b, 0, s, d);
    hipLaunchKernelGGL(k_cache_clear, gc, b, 0, s, d);
    hipLaunchKernelGGL(k_cache_evict, gb, b, 0, s, d);
    switch (k->h->cfg.strategy) {
#define PBM_LOAD_CASE(S) \
    case S: hipLaunchKernelGGL(k_cache_load<S>, gb, b, 0, s, d); break;
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
    hipLaunchKernelGGL(k_cache_finish, dim3(1), dim3(1), 0, s, d);
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
    out->reserved = 0;
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
