// Autotuner of the batched launch: times alloc+free cycles of the caller's
// request count over the launch space and returns the fastest shape.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>
#include "pbm_internal.hpp"

namespace {

struct Candidate {
    pbm_launch get;
    pbm_launch put;
};

// Median alloc and free time (ms) of `reps` cycles, after one warm-up.
int time_candidate(pbm_handle* h, const Candidate& k, uint32_t n, int32_t* d_ids, int reps,
                   hipStream_t s, hipEvent_t* ev, float* ms_get, float* ms_put) {
    std::vector<float> a, f;
    for (int i = -1; i < reps; ++i) {
        hipEventRecord(ev[0], s);
        int rc = pbm_get_pages(h, n, 0, d_ids, nullptr, &k.get, s);
        if (rc) return rc;
        hipEventRecord(ev[1], s);
        rc = pbm_free_pages(h, n, d_ids, &k.put, s);
        if (rc) return rc;
        hipEventRecord(ev[2], s);
        PBM_HIP(hipEventSynchronize(ev[2]));
        float x = 0, y = 0;
        hipEventElapsedTime(&x, ev[0], ev[1]);
        hipEventElapsedTime(&y, ev[1], ev[2]);
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

    std::vector<Candidate> space;
    const uint32_t wpcs[] = {1, 2, 4, 8, 16};
    const uint32_t cws[] = {64, 128, 256};
    const uint32_t budgets[] = {4, 16, 64};
    for (uint32_t wpc : wpcs) {
        Candidate k{};
        k.get.mode = k.put.mode = PBM_MODE_BATCHED;
        k.get.waves_per_cu = k.put.waves_per_cu = wpc;
        k.get.flags = k.put.flags = PBM_LAUNCH_NO_COUNTERS;
        if (!(flags & PBM_TUNE_WALK_ONLY)) {
            Candidate r = k;
            r.get.flags |= PBM_LAUNCH_RANK;
            r.get.cluster_words = 64;
            r.get.walk_budget = 16;
            space.push_back(r);
        }
        if (!(flags & PBM_TUNE_RANK_ONLY))
            for (uint32_t cw : cws)
                for (uint32_t b : budgets) {
                    Candidate w = k;
                    w.get.cluster_words = cw;
                    w.get.walk_budget = b;
                    space.push_back(w);
                }
    }
    if (space.empty()) return pbmi::fail(PBM_ERR_INVALID, "flags exclude every candidate");

    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    for (auto& e : ev) PBM_HIP(hipEventCreate(&e));
    const int reps = (flags & PBM_TUNE_QUICK) ? 3 : 9;
    float best = 1e30f;
    pbm_tune_result res{};
    for (const Candidate& k : space) {
        float a = 0, f = 0;
        rc = time_candidate(h, k, n, d_ids, reps, s, ev, &a, &f);
        if (rc) break;
        if (a + f < best) {
            best = a + f;
            res.get_launch = k.get;
            res.free_launch = k.put;
            res.ms_alloc = a;
            res.ms_free = f;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    res.candidates = (uint32_t)space.size();
    int64_t after = 0;
    rc = pbm_read_free_count(h, &after, stream);
    if (rc) return rc;
    if (after != before)
        return pbmi::fail(PBM_ERR_INVALID,
                          "autotune cycles did not return every page (another call overlapped?)");
    *out = res;
    return PBM_OK;
}
