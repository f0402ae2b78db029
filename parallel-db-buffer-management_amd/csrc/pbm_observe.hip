// Occupancy observation: free-page clusters, and a sampling monitor that
// runs beside the allocators on its own stream.
#include <hip/hip_runtime.h>
#include "pbm_internal.hpp"

using pbm::u64;

namespace {

// Free bits of word i (1 = free), padding past n_pages cleared.
__device__ __forceinline__ u64 free_bits(const pbm_ctx& c, uint32_t i) {
    u64 f = ~pbm::ld_word(&c.bitmap[i]);
    const uint64_t first = (uint64_t)i * 64u;
    if (first + 64u > c.n_pages) f &= (1ull << (uint32_t)(c.n_pages - first)) - 1ull;
    return f;
}

// Counts free pages whose predecessor (circularly) is in use: the first
// page of every maximal run of free pages.
__global__ __launch_bounds__(256) void k_count_clusters(pbm_ctx c, u64* out) {
    uint32_t v = 0;
    const uint32_t last_page = c.n_pages - 1u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n_words;
         i += gridDim.x * blockDim.x) {
        const u64 f = free_bits(c, i);
        u64 carry;
        if (i == 0) {
            carry = (free_bits(c, last_page >> 6) >> (last_page & 63u)) & 1ull;
        } else {
            carry = free_bits(c, i - 1u) >> 63;
        }
        v += (uint32_t)__popcll(f & ~((f << 1) | carry));
    }
    for (int d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 64);
    if (pbm::lane_id() == 0 && v) pbm::atom_add_u64(out, v);
}

__global__ void k_zero(u64* p) { *p = 0; }

// Monitor: one workgroup.  Sample k (k < count) at wall-clock tick
// t0 + k*gap: the free pages of the map, read with device-scope loads so
// that allocations running beside it are seen.  samples[2k] = tick,
// samples[2k+1] = free pages.
__global__ __launch_bounds__(1024) void k_monitor(pbm_ctx c, u64 gap, uint32_t count,
                                                  u64* samples) {
    __shared__ uint32_t part[16];
    __shared__ u64 s_t0;
    if (threadIdx.x == 0) s_t0 = wall_clock64();
    __syncthreads();
    const u64 t0 = s_t0;
    for (uint32_t k = 0; k < count; ++k) {
        // wait for this sample's tick (a bounded wait: the tick is finite)
        const u64 due = t0 + (u64)k * gap;
        while (wall_clock64() < due) __builtin_amdgcn_s_sleep(8);
        const u64 t = wall_clock64();
        uint32_t v = 0;
        for (uint32_t i = threadIdx.x; i < c.n_words; i += blockDim.x)
            v += (uint32_t)__popcll(free_bits(c, i));
        for (int d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 64);
        if (pbm::lane_id() == 0) part[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            u64 tot = 0;
            for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) tot += part[w];
            samples[2u * k] = t;
            samples[2u * k + 1u] = tot;
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" int pbm_count_clusters(pbm_handle* h, uint64_t* clusters, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!clusters) return pbmi::fail(PBM_ERR_INVALID, "null output");
    hipStream_t s = pbmi::as_stream(stream);
    const pbm_ctx& c = h->ctx;
    u64* out = h->d_scratch + 2;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, s, out);
    uint32_t nb = (c.n_words + 255u) / 256u;
    if (nb > 2048u) nb = 2048u;
    hipLaunchKernelGGL(k_count_clusters, dim3(nb), dim3(256), 0, s, c, out);
    PBM_HIP(hipGetLastError());
    u64 v = 0;
    uint64_t used = 0;
    PBM_HIP(hipMemcpyAsync(&v, out, 8u, hipMemcpyDeviceToHost, s));
    PBM_HIP(hipStreamSynchronize(s));
    // every page free: one cluster (it has no first page by the rule above)
    if (v == 0) {
        int rc = pbm_count_used(h, &used, stream);
        if (rc) return rc;
        if (used == 0) v = 1;
    }
    *clusters = v;
    return PBM_OK;
}

extern "C" int pbm_monitor(pbm_handle* h, uint32_t gap_us, uint32_t count,
                           unsigned long long* d_samples, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (count == 0) return PBM_OK;
    if (!d_samples) return pbmi::fail(PBM_ERR_INVALID, "null sample buffer");
    if ((uint64_t)gap_us * count > 60000000ull)
        return pbmi::fail(PBM_ERR_RANGE, "a monitor run is limited to 60 s");
    int khz = 0;
    PBM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
    const u64 gap = (u64)gap_us * (u64)khz / 1000u;
    hipLaunchKernelGGL(k_monitor, dim3(1), dim3(1024), 0, pbmi::as_stream(stream), h->ctx, gap,
                       count, (u64*)d_samples);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_wall_clock_hz(int device, unsigned long long* hz) {
    if (!hz) return pbmi::fail(PBM_ERR_INVALID, "null output");
    int khz = 0;
    PBM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    *hz = (unsigned long long)khz * 1000ull;
    return PBM_OK;
}
