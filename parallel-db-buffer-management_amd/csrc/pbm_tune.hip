// Autotuner of the batched launch: times alloc+free cycles of the caller's
// request count over the launch space and returns the fastest shape.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>
#include "pbm_internal.hpp"

namespace {

// Median kernel durations (ms) of the alloc and the free of `reps` cycles,
// after one warm-up, from the events the library records at each kernel's
// own start and end.
int time_cycle(pbm_handle* h, const pbm_launch& get, const pbm_launch& put, uint32_t n,
               int32_t* d_ids, int reps, hipStream_t s, hipEvent_t* ev, float* ms_get,
               float* ms_put) {
    std::vector<float> a, f;
    for (int i = -1; i < reps; ++i) {
        int rc = pbm_time_next_kernel(h, ev[0], ev[1]);
        if (!rc) rc = pbm_get_pages(h, n, 0, d_ids, nullptr, &get, s);
        if (!rc) rc = pbm_time_next_kernel(h, ev[2], ev[3]);
        if (!rc) rc = pbm_free_pages(h, n, d_ids, &put, s);
        if (rc) return rc;
        PBM_HIP(hipEventSynchronize(ev[3]));
        float x = 0, y = 0;
        PBM_HIP(hipEventElapsedTime(&x, ev[0], ev[1]));
        PBM_HIP(hipEventElapsedTime(&y, ev[2], ev[3]));
        if (i >= 0) {
            a.push_back(x);
            f.push_back(y);
        }
    }
    std::nth_element(a.begin(), a.begin() + a.size() / 2, a.end());
    std::nth_element(f.begin(), f.begin() + f.size() / 2, f.end());
    *ms_get = a[a.size() / 2];
    *ms_put = f[f.size() / 2];
    return PBM_OK;
}

pbm_launch batched(uint32_t wpc, uint32_t flags) {
    pbm_launch L{};
    L.mode = PBM_MODE_BATCHED;
    L.waves_per_cu = wpc;
    L.flags = PBM_LAUNCH_NO_COUNTERS | flags;
    return L;
}

}  // namespace

extern "C" int pbm_autotune(pbm_handle* h, uint32_t n, int32_t* d_ids, uint32_t flags,
                            pbm_tune_result* out, void* stream) {
    PBM_CHECK_HANDLE(h);
    if (!out || !d_ids || n == 0) return pbmi::fail(PBM_ERR_INVALID, "autotune needs n > 0, d_ids and out");
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "batched mode needs a bitmap strategy");
    hipStream_t s = pbmi::as_stream(stream);
    int64_t before = 0;
    int rc = pbm_read_free_count(h, &before, stream);
    if (rc) return rc;

    // alloc launches: the rank scan (its grid is the map's tiles, so one
    // shape) and the walk homes over waves, window width and budget; free
    // launches over waves.  The alloc and the free are separate kernels, so
    // each is chosen on its own: allocs with a fixed free after them, then
    // frees after the chosen alloc.
    std::vector<pbm_launch> gets, puts;
    if (!(flags & PBM_TUNE_WALK_ONLY)) {
        pbm_launch r = batched(8, PBM_LAUNCH_RANK);
        r.cluster_words = 64;
        r.walk_budget = 16;
        gets.push_back(r);
    }
    if (!(flags & PBM_TUNE_RANK_ONLY))
        for (uint32_t wpc : {1u, 2u, 4u, 8u, 16u})
            for (uint32_t cw : {64u, 128u, 256u})
                for (uint32_t b : {4u, 16u, 64u}) {
                    pbm_launch w = batched(wpc, 0);
                    w.cluster_words = cw;
                    w.walk_budget = b;
                    gets.push_back(w);
                }
    if (gets.empty()) return pbmi::fail(PBM_ERR_INVALID, "flags exclude every candidate");
    for (uint32_t wpc : {1u, 2u, 4u, 8u, 16u, 32u}) puts.push_back(batched(wpc, 0));

    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    for (auto& e : ev) PBM_HIP(hipEventCreate(&e));
    const int reps = (flags & PBM_TUNE_QUICK) ? 3 : 9;
    pbm_tune_result res{};
    float best = 1e30f;
    for (const pbm_launch& g : gets) {
        float a = 0, f = 0;
        if ((rc = time_cycle(h, g, puts[4], n, d_ids, reps, s, ev, &a, &f))) break;
        if (a < best) {
            best = a;
            res.get_launch = g;
            res.ms_alloc = a;
        }
    }
    best = 1e30f;
    for (size_t i = 0; !rc && i < puts.size(); ++i) {
        float a = 0, f = 0;
        if ((rc = time_cycle(h, res.get_launch, puts[i], n, d_ids, reps, s, ev, &a, &f))) break;
        if (f < best) {
            best = f;
            res.free_launch = puts[i];
            res.ms_free = f;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    res.candidates = (uint32_t)(gets.size() + puts.size());
    int64_t after = 0;
    rc = pbm_read_free_count(h, &after, stream);
    if (rc) return rc;
    if (after != before)
        return pbmi::fail(PBM_ERR_INVALID,
                          "autotune cycles did not return every page (another call overlapped?)");
    *out = res;
    return PBM_OK;
}
