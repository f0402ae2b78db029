// A reference-style driver on the Ouroboros wrapper's names
// (Ouroboros.cuh:14-62: initMemoryManagement, prefillBuffer, getPage
// returning a pointer, freePage taking one), following the flow of the
// reference's drivers (unitTests/unitTest1.cu:22-44): pre-fill, one getPage
// per thread, writes into the pages, checks, freePage per thread, checks.
// Prints one line and exits 0 when every check holds.
#define PBM_COMPAT_OUROBOROS 1
#include <cstdio>
#include <vector>
#include <algorithm>
#include "pbm_compat.hpp"

__global__ void get1page_kernel(int n, void** ptrs) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        void* p = getPage();
        ptrs[tid] = p;
        if (p) *(int*)p = tid;
    }
}

__global__ void check_kernel(int n, void* const* ptrs, int* bad) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        void* p = ptrs[tid];
        if (!p || *(const int*)p != tid || pageAddress(getPageID(p)) != p) atomicAdd(bad, 1);
    }
}

__global__ void freePage_kernel(int n, void* const* ptrs) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) freePage(ptrs[tid]);
}

#define CHECK(cond, msg)                                        \
    do {                                                        \
        if (!(cond)) {                                          \
            std::printf("compat ouroboros FAIL: %s\n", msg);    \
            return 1;                                           \
        }                                                       \
    } while (0)

int main() {
    initMemoryManagement(1, 4096);  // 262,144 pages
    const int T = 262144, N = 65536;
    CHECK(getFreePagePercentage() == 1.0f, "a new pool is all free");
    prefillBuffer(0.5f);
    const float before = getFreePagePercentage();
    CHECK(before == 0.5f, "prefillBuffer leaves the requested share");
    void** d_ptrs;
    int* d_bad;
    hipMalloc(&d_ptrs, N * sizeof(void*));
    hipMalloc(&d_bad, 4);
    hipMemset(d_bad, 0, 4);
    get1page_kernel<<<N / 128, 128>>>(N, d_ptrs);
    check_kernel<<<N / 256, 256>>>(N, d_ptrs, d_bad);
    int bad = 0;
    hipMemcpy(&bad, d_bad, 4, hipMemcpyDeviceToHost);
    CHECK(bad == 0, "every thread got its own page");
    std::vector<void*> ptrs(N);
    hipMemcpy(ptrs.data(), d_ptrs, N * sizeof(void*), hipMemcpyDeviceToHost);
    std::sort(ptrs.begin(), ptrs.end());
    CHECK(std::adjacent_find(ptrs.begin(), ptrs.end()) == ptrs.end(), "pages are distinct");
    CHECK(getFreePagePercentage() == before - (float)N / T, "each getPage takes one page");
    freePage_kernel<<<N / 128, 128>>>(N, d_ptrs);
    hipDeviceSynchronize();
    CHECK(getFreePagePercentage() == before, "freePage gives every page back");
    // freed pages come back first: the queue is FIFO
    get1page_kernel<<<N / 128, 128>>>(N, d_ptrs);
    std::vector<void*> again(N);
    hipMemcpy(again.data(), d_ptrs, N * sizeof(void*), hipMemcpyDeviceToHost);
    std::sort(again.begin(), again.end());
    CHECK(again == ptrs, "the freed pages are handed out again");
    std::printf("compat ouroboros ok\n");
    return 0;
}
