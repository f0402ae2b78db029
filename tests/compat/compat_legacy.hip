// A legacy-style driver on parallelPage.cuh's names (initPagesX, getPageX,
// freePageX, printNumPagesLeftX, resetBufferX, getXPageX,
// monitorBufferRandomWalk) for all nine families in one translation unit,
// as the legacy unit tests use them (unitTests/unitTestRandomWalk2.cu,
// unitTestLinkedList2.cu, unitTestSingleClock1.cu, ...): one getPage per
// thread, uniqueness and occupancy checks, freePage per thread, the pool
// back to full.  Prints one line and exits 0 when every check holds.
#define PBM_LEGACY_TOTAL_N_PAGES 262144u
#define PBM_LEGACY_PAGE_SIZE 264u
#include <algorithm>
#include <cstdio>
#include <vector>
#include "pbm_legacy.hpp"

#define CHECK(cond, msg)                                         \
    do {                                                         \
        if (!(cond)) {                                           \
            std::printf("compat legacy FAIL %s: %s\n", name, msg); \
            return 1;                                            \
        }                                                        \
    } while (0)

#define FAMILY_KERNELS(Name)                                                         \
    __global__ void get_##Name(int n, int* ids, int* steps) {                        \
        int tid = blockIdx.x * blockDim.x + threadIdx.x;                             \
        if (tid < n) ids[tid] = getPage##Name(&steps[tid]);                          \
    }                                                                                \
    __global__ void free_##Name(int n, const int* ids) {                             \
        int tid = blockIdx.x * blockDim.x + threadIdx.x;                             \
        if (tid < n && ids[tid] >= 0) freePage##Name(ids[tid]);                      \
    }

FAMILY_KERNELS(RandomWalk)
FAMILY_KERNELS(ClusteredRandomWalk)
FAMILY_KERNELS(LinkedList)
FAMILY_KERNELS(SingleClock)
FAMILY_KERNELS(DistClock)
FAMILY_KERNELS(CollabRW)
FAMILY_KERNELS(CollabRW_BM)
FAMILY_KERNELS(RandomWalk_BM)
FAMILY_KERNELS(ClusteredRandomWalk_BM)

__global__ void getx_kernel(int n, int which, int* ids) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        int x = 1 + tid % 8;
        if (which == 0) ids[tid] = getXPageRandomWalk_BM(x);
        else if (which == 1) ids[tid] = getXPageClusteredRandomWalk_BM(x);
        else ids[tid] = getXPageCollabRW_BM(x);
    }
}

static int used_pages(pbm_handle* h) {
    uint64_t u = 0;
    pbm_count_used(h, &u, nullptr);
    return (int)u;
}

template <class G, class F>
int run(const char* name, void (*init)(), void (*print)(), void (*reset)(), pbm_handle* (*handle)(),
        G get, F fre, int* d_ids, int* d_steps) {
    const int N = 50000, T = (int)PBM_LEGACY_TOTAL_N_PAGES;
    init();
    pbm_handle* h = handle();
    CHECK(used_pages(h) == 0, "a new pool is all free");
    get<<<(N + 255) / 256, 256>>>(N, d_ids, d_steps);
    std::vector<int> ids(N);
    hipMemcpy(ids.data(), d_ids, N * 4, hipMemcpyDeviceToHost);
    std::vector<int> got(ids);
    std::sort(got.begin(), got.end());
    CHECK(got.front() >= 0 && got.back() < T, "every thread got a page inside the pool");
    CHECK(std::adjacent_find(got.begin(), got.end()) == got.end(), "a page was handed out twice");
    CHECK(used_pages(h) == N, "occupancy after gets");
    print();
    fre<<<(N + 255) / 256, 256>>>(N, d_ids);
    hipDeviceSynchronize();
    CHECK(used_pages(h) == 0, "every page came back");
    get<<<(N + 255) / 256, 256>>>(N, d_ids, d_steps);
    hipDeviceSynchronize();
    reset();
    CHECK(used_pages(h) == 0, "resetBuffer frees every page");
    return 0;
}

#define RUN(Name)                                                                            \
    if (run(#Name, initPages##Name, printNumPagesLeft##Name, resetBuffer##Name,              \
            pbmLegacyHandle##Name, get_##Name, free_##Name, d_ids, d_steps))                 \
        return 1;

int main() {
    int *d_ids, *d_steps;
    hipMalloc(&d_ids, 50000 * 4);
    hipMalloc(&d_steps, 50000 * 4);
    RUN(RandomWalk)
    RUN(ClusteredRandomWalk)
    RUN(LinkedList)
    RUN(SingleClock)
    RUN(DistClock)
    RUN(CollabRW)
    RUN(CollabRW_BM)
    RUN(RandomWalk_BM)
    RUN(ClusteredRandomWalk_BM)
    // getXPage: x = 1..8 consecutive pages inside one word, runs disjoint
    pbm_handle* hs[3] = {pbmLegacyHandleRandomWalk_BM(), pbmLegacyHandleClusteredRandomWalk_BM(),
                         pbmLegacyHandleCollabRW_BM()};
    for (int which = 0; which < 3; ++which) {
        const char* name = which == 0 ? "getXPageRandomWalk_BM"
                           : which == 1 ? "getXPageClusteredRandomWalk_BM" : "getXPageCollabRW_BM";
        const int N = 20000;
        getx_kernel<<<(N + 255) / 256, 256>>>(N, which, d_ids);
        std::vector<int> ids(N);
        hipMemcpy(ids.data(), d_ids, N * 4, hipMemcpyDeviceToHost);
        std::vector<std::pair<int, int>> runs;
        long long pages = 0;
        for (int t = 0; t < N; ++t) {
            const int x = 1 + t % 8;
            CHECK(ids[t] >= 0, "a run was found");
            CHECK(ids[t] % 64 + x <= 64, "a run stays inside one word");
            runs.push_back({ids[t], ids[t] + x});
            pages += x;
        }
        std::sort(runs.begin(), runs.end());
        for (size_t k = 1; k < runs.size(); ++k) CHECK(runs[k - 1].second <= runs[k].first, "runs overlap");
        CHECK(used_pages(hs[which]) == (int)pages, "occupancy after getXPage");
    }
    const char* name = "monitor";
    std::printf("monitor begin\n");
    std::fflush(stdout);
    initPagesRandomWalk();
    monitorBufferRandomWalk(1, 3);
    hipDeviceSynchronize();
    std::fflush(stdout);
    std::printf("monitor end\n");
    CHECK(true, "");
    std::printf("compat legacy ok\n");
    return 0;
}
