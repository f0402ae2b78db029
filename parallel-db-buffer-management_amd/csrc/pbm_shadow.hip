// Debug shadow of page owners: one 64-bit word per page holding the request
// number of the grant that holds it, kFree, or kHeld (held before the shadow
// last read the map).  After every id-returning call on a handle with the
// shadow on, one kernel on the caller's stream moves each granted page from
// kFree to its request number with a compare-and-swap; a page that was not
// free counts as a double grant, and the first one seen is kept with both
// owners.  Frees move pages back to kFree; a page that already was free
// counts as a bad free.  The map itself is never read or changed here, so a
// grant that two kernels both made (the failure the bitmap claims must
// prevent) is caught at the call that made the second, before the caller can
// use it.
//
// The reference has no such check (its tests compare the final map only,
// unitTests/unitTest1.cu:95-125); it is a debugging aid of this library.

#include "pbm_internal.hpp"

namespace {

using u64 = unsigned long long;

constexpr u64 kFree = ~0ull;
constexpr u64 kHeld = ~1ull;
constexpr u64 kNone = ~0ull;

// report words: [0] double grants, [1] bad frees, [2] first double-granted
// page (kNone = none), [3] its new owner, [4] its holder, [5] first page
// freed while free (kNone = none)
constexpr int kRepWords = 6;

__global__ void k_shadow_init(const u64* __restrict__ bm, u64* __restrict__ sh, uint32_t n_pages) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pages; p += stride)
        sh[p] = ((bm[p >> 6] >> (p & 63u)) & 1ull) ? kHeld : kFree;
}

__global__ void k_shadow_clear_report(u64* rep) {
    if (threadIdx.x < kRepWords) rep[threadIdx.x] = threadIdx.x < 2 ? 0ull : kNone;
}

__device__ void note_double_grant(u64* rep, u64 page, u64 owner, u64 holder) {
    atomicAdd(&rep[0], 1ull);
    if (atomicCAS(&rep[2], kNone, page) == kNone) {
        rep[3] = owner;
        rep[4] = holder;
    }
}

__device__ void note_bad_free(u64* rep, u64 page) {
    atomicAdd(&rep[1], 1ull);
    atomicCAS(&rep[5], kNone, page);
}

// One thread per request; counts null means one page each.  A count of 0 or
// an id < 0 is no grant.  Pages at or past n_pages are reported (as a double
// grant held by kFree, or as a bad free) and never touched.
__global__ void k_shadow_note(u64* __restrict__ sh, u64* __restrict__ rep, uint32_t n_pages,
                              uint32_t n, u64 req_base, const int32_t* __restrict__ ids,
                              const uint32_t* __restrict__ counts, int freeing) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        const int32_t id = ids[r];
        const uint32_t cnt = counts ? counts[r] : 1u;
        if (id < 0 || cnt == 0) continue;
        for (uint32_t k = 0; k < cnt; ++k) {
            const u64 p = (u64)(uint32_t)id + k;
            if (p >= n_pages) {
                if (freeing) note_bad_free(rep, p);
                else note_double_grant(rep, p, req_base + r, kFree);
                continue;
            }
            if (freeing) {
                if (atomicExch(&sh[p], kFree) == kFree) note_bad_free(rep, p);
            } else {
                const u64 old = atomicCAS(&sh[p], kFree, req_base + r);
                if (old != kFree) note_double_grant(rep, p, req_base + r, old);
            }
        }
    }
}

uint32_t grid_for(uint64_t n) {
    uint64_t b = (n + 255u) / 256u;
    return (uint32_t)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

namespace pbmi {

int shadow_sync(pbm_handle* h, hipStream_t s) {
    if (!h->d_shadow) return PBM_OK;
    hipLaunchKernelGGL(k_shadow_init, dim3(grid_for(h->ctx.n_pages)), dim3(256), 0, s,
                       (const u64*)h->d_bitmap, (u64*)h->d_shadow, h->ctx.n_pages);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

int shadow_note(pbm_handle* h, uint32_t n, uint64_t req_base, const int32_t* d_ids,
                const uint32_t* d_counts, bool freeing, hipStream_t s) {
    if (!h->d_shadow || n == 0) return PBM_OK;
    hipLaunchKernelGGL(k_shadow_note, dim3(grid_for(n)), dim3(256), 0, s, (u64*)h->d_shadow,
                       (u64*)h->d_shadow_rep, h->ctx.n_pages, n, (u64)req_base, d_ids, d_counts,
                       freeing ? 1 : 0);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace pbmi

extern "C" int pbm_shadow_enable(pbm_handle* h, int on, void* stream) {
    PBM_CHECK_HANDLE(h);
    hipStream_t s = pbmi::as_stream(stream);
    if (!on) {
        PBM_HIP(hipStreamSynchronize(s));
        if (h->d_shadow) (void)hipFree(h->d_shadow);
        if (h->d_shadow_rep) (void)hipFree(h->d_shadow_rep);
        h->d_shadow = h->d_shadow_rep = nullptr;
        return PBM_OK;
    }
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "the owner shadow needs a bitmap strategy");
    if (!h->d_shadow) {
        PBM_HIP(hipMalloc((void**)&h->d_shadow, (size_t)h->ctx.n_pages * 8u));
        hipError_t e = hipMalloc((void**)&h->d_shadow_rep, kRepWords * 8u);
        if (e != hipSuccess) {
            (void)hipFree(h->d_shadow);
            h->d_shadow = nullptr;
            return pbmi::hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        }
    }
    hipLaunchKernelGGL(k_shadow_clear_report, dim3(1), dim3(64), 0, s, (u64*)h->d_shadow_rep);
    PBM_HIP(hipGetLastError());
    return pbmi::shadow_sync(h, s);
}

extern "C" int pbm_shadow_note(pbm_handle* h, uint32_t n, uint64_t req_base,
                               const int32_t* d_ids, const uint32_t* d_counts, int freeing,
                               void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!h->d_shadow) return pbmi::fail(PBM_ERR_INVALID, "the owner shadow is off");
    if (n && !d_ids) return pbmi::fail(PBM_ERR_INVALID, "null d_ids");
    return pbmi::shadow_note(h, n, req_base, d_ids, d_counts, freeing != 0,
                             pbmi::as_stream(stream));
}

extern "C" int pbm_shadow_read(pbm_handle* h, pbm_shadow_report* out, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output");
    if (!h->d_shadow) return pbmi::fail(PBM_ERR_INVALID, "the owner shadow is off");
    hipStream_t s = pbmi::as_stream(stream);
    u64 w[kRepWords];
    PBM_HIP(hipMemcpyAsync(w, h->d_shadow_rep, sizeof w, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    out->double_grants = w[0];
    out->bad_frees = w[1];
    out->grant_page = w[2] == kNone ? -1 : (int64_t)w[2];
    out->grant_owner = w[3];
    out->grant_holder = w[4];
    out->free_page = w[5] == kNone ? -1 : (int64_t)w[5];
    return PBM_OK;
}

extern "C" int pbm_shadow_owners(pbm_handle* h, uint64_t* host, size_t words, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!h->d_shadow) return pbmi::fail(PBM_ERR_INVALID, "the owner shadow is off");
    if (!host || words != h->ctx.n_pages)
        return pbmi::fail(PBM_ERR_INVALID, "owner buffer must hold exactly n_pages words");
    hipStream_t s = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(host, h->d_shadow, words * 8u, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    return PBM_OK;
}
