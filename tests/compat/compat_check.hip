// A reference-style driver on the reference's own API names
// (initMemoryManagement, prefillBuffer, getPage, freePage, pageAddress,
// getPageID, getFreePagePercentage, mallocRW, freeRW), compiled once per
// policy with -DPBM_COMPAT_STRATEGY=...  It follows the flow of the
// reference's drivers (unitTests/unitTest1.cu:22-44, unitTest3.cu:31-43):
// pre-fill, one getPage per thread, check, freePage per thread, check;
// then partial warps (unitTest5.cu, unitTest7.cu), then mallocRW.
// Prints one line and exits 0 when every check holds.
#include <algorithm>
#include <cstdio>
#include <vector>
#include "pbm_compat.hpp"

__global__ void get1page_kernel(int n, int* ids, int* steps) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) ids[tid] = getPage(&steps[tid]);
}

// Partial-warp callers (unitTests/unitTest5.cu:24-35, unitTest7.cu:17-29):
// only lanes with threadIdx.x % 32 < nalloc ask for a page; the others
// record -2.
__global__ void partial_kernel(int n, int nalloc, int* ids) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        if ((int)(threadIdx.x % 32) < nalloc) ids[tid] = getPage();
        else ids[tid] = -2;
    }
}

// collectPageData (unitTests/collectPageData.cu:19-37): getPage with
// collectWarpData, which prints the pages a warp found in each step.
__global__ void collect_kernel(int n, int* ids) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) ids[tid] = getPage(nullptr, true);
}

__global__ void freePage_kernel(int n, const int* ids) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n && ids[tid] >= 0) freePage(ids[tid]);
}

// page id -> address -> id, and one write at the start of each page
__global__ void roundtrip_kernel(int n, const int* ids, int* bad) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n && ids[tid] >= 0) {
        void* p = pageAddress(ids[tid]);
        *(int*)p = ids[tid];
        if (getPageID(p) != ids[tid] || getPageID((char*)p + 100) != ids[tid]) atomicAdd(bad, 1);
    }
}

__global__ void malloc_kernel(int n, const unsigned* sizes, void** ptrs) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        void* p = mallocRW(sizes[tid]);
        ptrs[tid] = p;
        if (p) for (unsigned k = 0; k < sizes[tid]; k += 256) ((char*)p)[k] = (char)(tid & 0x7f);
    }
}

__global__ void check_kernel(int n, const unsigned* sizes, void* const* ptrs, int* bad) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n && ptrs[tid])
        for (unsigned k = 0; k < sizes[tid]; k += 256)
            if (((const char*)ptrs[tid])[k] != (char)(tid & 0x7f)) { atomicAdd(bad, 1); break; }
}

__global__ void free_kernel(int n, void* const* ptrs) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) freeRW(ptrs[tid]);
}

#define CHECK(cond, msg)                                                      \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::printf("compat FAIL strategy=%d: %s\n", PBM_COMPAT_STRATEGY, msg); \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main() {
    initMemoryManagement(1, 4096);  // 262,144 pages
    const int T = 262144, N = 32768;
    CHECK(getFreePagePercentage() == 1.0f, "a new pool is all free");
    prefillBuffer(0.1f);
    const float before = getFreePagePercentage();
    const int free_pages = (int)((double)before * T + 0.5);
    CHECK(free_pages == T - (int)((1.0f - 0.1f) * T), "prefillBuffer leaves the requested share");

    int *d_ids, *d_steps, *d_bad;
    hipMalloc(&d_ids, N * 4);
    hipMalloc(&d_steps, N * 4);
    hipMalloc(&d_bad, 4);
    hipMemset(d_bad, 0, 4);
    get1page_kernel<<<N / 256, 256>>>(N, d_ids, d_steps);
    roundtrip_kernel<<<N / 256, 256>>>(N, d_ids, d_bad);
    std::vector<int> ids(N);
    int bad = 0;
    hipMemcpy(ids.data(), d_ids, N * 4, hipMemcpyDeviceToHost);
    hipMemcpy(&bad, d_bad, 4, hipMemcpyDeviceToHost);
    std::vector<int> got;
    for (int id : ids) if (id >= 0) got.push_back(id);
    std::sort(got.begin(), got.end());
    CHECK(std::adjacent_find(got.begin(), got.end()) == got.end(), "a page was handed out twice");
    CHECK((int)got.size() == std::min(N, free_pages), "grants = min(requests, free pages)");
    CHECK(got.empty() || (got.front() >= 0 && got.back() < T), "ids inside the pool");
    CHECK(bad == 0, "pageAddress / getPageID round trip");
    CHECK(getFreePagePercentage() == (float)(free_pages - (int)got.size()) / T, "occupancy after gets");
    freePage_kernel<<<N / 256, 256>>>(N, d_ids);
    hipDeviceSynchronize();
    CHECK(getFreePagePercentage() == before, "every page came back");

    // partial warps: 5,000 threads as the reference's drivers use, lane
    // groups of 1, 16, 17, 31 and 32 of every 32
    for (int nalloc : {1, 16, 17, 31, 32}) {
        const int P = 5000;
        partial_kernel<<<(P + 255) / 256, 256>>>(P, nalloc, d_ids);
        std::vector<int> pid(P);
        hipMemcpy(pid.data(), d_ids, P * 4, hipMemcpyDeviceToHost);
        int asked = 0;
        std::vector<int> pg;
        for (int t = 0; t < P; ++t) {
            if (pid[t] == -2) continue;
            ++asked;
            if (pid[t] >= 0) pg.push_back(pid[t]);
            CHECK((t % 256) % 32 < nalloc, "only the asking lanes wrote");
        }
        std::sort(pg.begin(), pg.end());
        CHECK(std::adjacent_find(pg.begin(), pg.end()) == pg.end(), "partial warps: a page twice");
        CHECK((int)pg.size() == asked, "partial warps: every asking lane got a page");
        freePage_kernel<<<(P + 255) / 256, 256>>>(P, d_ids);
        hipDeviceSynchronize();
        CHECK(getFreePagePercentage() == before, "partial warps: every page came back");
    }

    if (PBM_COMPAT_STRATEGY == PBM_COLLAB_RW_BM) {
        std::printf("collect begin\n");
        std::fflush(stdout);
        collect_kernel<<<1, 32>>>(32, d_ids);
        hipDeviceSynchronize();
        std::fflush(stdout);
        std::printf("collect end\n");
        freePage_kernel<<<1, 32>>>(32, d_ids);
        hipDeviceSynchronize();
        CHECK(getFreePagePercentage() == before, "collectWarpData: every page came back");
    }

    int malloc_ok = -1;
    if (PBM_COMPAT_STRATEGY == PBM_RW_MALLOC) {
        resetMemoryManager();
        const int M = 8192;
        std::vector<unsigned> sizes(M);
        unsigned x = 12345;
        for (int i = 0; i < M; ++i) { x = x * 1103515245u + 12345u; sizes[i] = 16u + (x >> 8) % 20000u; }
        unsigned* d_sizes;
        void** d_ptrs;
        hipMalloc(&d_sizes, M * 4);
        hipMalloc(&d_ptrs, M * 8);
        hipMemcpy(d_sizes, sizes.data(), M * 4, hipMemcpyHostToDevice);
        hipMemset(d_bad, 0, 4);
        malloc_kernel<<<M / 256, 256>>>(M, d_sizes, d_ptrs);
        check_kernel<<<M / 256, 256>>>(M, d_sizes, d_ptrs, d_bad);
        std::vector<void*> ptrs(M);
        hipMemcpy(ptrs.data(), d_ptrs, M * 8, hipMemcpyDeviceToHost);
        hipMemcpy(&bad, d_bad, 4, hipMemcpyDeviceToHost);
        malloc_ok = 0;
        for (void* p : ptrs) malloc_ok += p != nullptr;
        CHECK(malloc_ok == M, "every mallocRW succeeded");
        CHECK(bad == 0, "mallocRW regions overlap");
        free_kernel<<<M / 256, 256>>>(M, d_ptrs);
        hipDeviceSynchronize();
        CHECK(getFreePagePercentage() == 1.0f, "freeRW returned every page");
    }
    std::printf("compat ok strategy=%d grants=%d free_before=%d mallocs=%d\n", PBM_COMPAT_STRATEGY,
                (int)got.size(), free_pages, malloc_ok);
    return 0;
}
