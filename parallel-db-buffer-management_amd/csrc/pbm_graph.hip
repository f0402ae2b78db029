// HIP-graph capture of allocation/free phases, and the atomic-throughput
// measurement helper.
#include <hip/hip_runtime.h>
#include "pbm_internal.hpp"

using pbm::u64;

extern "C" int pbm_graph_destroy(pbm_graph* g) {
    if (!g) return PBM_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return PBM_OK;
}

extern "C" int pbm_graph_create_steps(pbm_handle* h, uint32_t n, uint64_t req_base,
                                      int32_t* d_ids, uint32_t* d_steps,
                                      const pbm_launch* get_launch, const pbm_launch* free_launch,
                                      uint32_t phases, uint32_t n_steps, pbm_graph** out) {
    PBM_CHECK_HANDLE(h);
    if (!out) return pbmi::fail(PBM_ERR_INVALID, "null output graph pointer");
    *out = nullptr;
    if (phases == 0 || (phases & ~(PBM_PHASE_ALLOC | PBM_PHASE_FREE)))
        return pbmi::fail(PBM_ERR_INVALID, "phases must combine PBM_PHASE_ALLOC/FREE");
    if (n == 0 || !d_ids) return pbmi::fail(PBM_ERR_INVALID, "graph needs n > 0 and d_ids");
    if (n_steps == 0 || n_steps > 1024u) return pbmi::fail(PBM_ERR_RANGE, "n_steps must be in [1, 1024]");
    hipStream_t s = nullptr;
    PBM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    pbm_graph* g = new pbm_graph();
    g->device = h->device;
    int rc = PBM_OK;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = pbmi::hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    } else {
        for (uint32_t k = 0; k < n_steps && !rc; ++k) {
            if (phases & PBM_PHASE_ALLOC)
                rc = pbm_get_pages(h, n, req_base, d_ids, d_steps, get_launch, s);
            if (!rc && (phases & PBM_PHASE_FREE)) rc = pbm_free_pages(h, n, d_ids, free_launch, s);
        }
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(s, &graph);
        g->graph = graph;
        if (!rc && e != hipSuccess) rc = pbmi::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        if (!rc) {
            e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = pbmi::hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
        }
    }
    (void)hipStreamDestroy(s);
    if (rc) {
        pbm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return PBM_OK;
}

extern "C" int pbm_graph_create(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                                uint32_t* d_steps, const pbm_launch* get_launch,
                                const pbm_launch* free_launch, uint32_t phases,
                                pbm_graph** out) {
    return pbm_graph_create_steps(h, n, req_base, d_ids, d_steps, get_launch, free_launch, phases,
                                  1u, out);
}

extern "C" int pbm_graph_launch(pbm_graph* g, void* stream) {
    if (!g || !g->exec) return pbmi::fail(PBM_ERR_INVALID, "null graph");
    int cur = -1;
    PBM_HIP(hipGetDevice(&cur));
    if (cur != g->device) PBM_HIP(hipSetDevice(g->device));
    PBM_HIP(hipGraphLaunch(g->exec, pbmi::as_stream(stream)));
    return PBM_OK;
}

// Every thread issues `rounds` fetch-ORs on pseudo-randomly spread words.
__global__ __launch_bounds__(256) void k_atomic_peak(u64* buf, uint32_t words, uint32_t rounds,
                                                     u64* sink) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    u64 acc = 0;
    for (uint32_t k = 0; k < rounds; ++k) {
        uint32_t idx = (uint32_t)(((u64)gid * 2654435761ull + (u64)k * 40503ull) % words);
        acc ^= pbm::atom_or(&buf[idx], 1ull << (k & 63u));
    }
    if (acc == 0x5eed5eed5eed5eedull) sink[0] = acc;
}

extern "C" int pbm_atomic_peak(uint32_t words, uint32_t rounds, int device, double* ops_per_s) {
    if (!ops_per_s || words == 0 || rounds == 0)
        return pbmi::fail(PBM_ERR_INVALID, "bad atomic-peak arguments");
    PBM_HIP(hipSetDevice(device));
    u64* buf = nullptr;
    PBM_HIP(hipMalloc((void**)&buf, (size_t)words * 8u + 8u));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PBM_OK;
    const uint32_t threads = 1u << 18;
    do {
        hipError_t e;
        if ((e = hipMemset(buf, 0, (size_t)words * 8u + 8u)) != hipSuccess ||
            (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
            rc = pbmi::hip_fail(e, "atomic-peak setup", __FILE__, __LINE__);
            break;
        }
        hipLaunchKernelGGL(k_atomic_peak, dim3(threads / 256), dim3(256), 0, nullptr, buf, words,
                           rounds, buf + words);
        hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(k_atomic_peak, dim3(threads / 256), dim3(256), 0, nullptr, buf, words,
                           rounds, buf + words);
        hipEventRecord(e1, nullptr);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) {
            rc = pbmi::hip_fail(e, "atomic-peak run", __FILE__, __LINE__);
            break;
        }
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        *ops_per_s = (double)threads * rounds / (ms * 1e-3);
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}
