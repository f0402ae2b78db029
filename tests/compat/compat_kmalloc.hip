// A reference-style kernel-malloc driver on the reference's names
// (KernelMallocCollab, KernelMallocSimple, malloc, free, reset), passing
// the managers to kernels by value as unitTests/unitTestMalloc1.cu:14-41
// and unitTestMalloc2.cu:24-33 do: sizes from the LCG of
// unitTestMalloc1.cu:10-16, a byte pattern written into every space and
// checked, then free (Collab) or reset (Simple) and the pool must be all
// free again.  Prints one line and exits 0 when every check holds.
#include <cstdio>
#ifdef PBM_KMALLOC_LEGACY
// the legacy build: parallelPage.cuh + malloc.cuh (unitTests/unitTestMalloc1.cu:6-7)
#define PBM_LEGACY_TOTAL_N_PAGES 262144u
#define PBM_LEGACY_PAGE_SIZE 4096u
#include "pbm_legacy.hpp"
static float getFreePagePercentage() {
    float f = 0;
    pbm_free_page_percentage(pbmLegacyHandleCollabRW_BM(), &f, nullptr);
    return f;
}
#else
#include "pbm_compat.hpp"
#endif

__device__ unsigned request_size(int tid) {
    return (unsigned)((1103515245LL * (long long)tid + 12345) % 8188) + 4u;
}

template <class M>
__global__ void allocate_kernel(int n, M mm, void** ptrs) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        unsigned size = request_size(tid);
        char* p = (char*)mm.malloc((int)size);
        ptrs[tid] = p;
        if (p)
            for (unsigned k = 0; k < size; ++k) p[k] = (char)(tid & 0x7f);
    }
}

__global__ void check_kernel(int n, void* const* ptrs, int* bad, int* missing) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) {
        const char* p = (const char*)ptrs[tid];
        if (!p) {
            atomicAdd(missing, 1);
            return;
        }
        unsigned size = request_size(tid);
        for (unsigned k = 0; k < size; ++k)
            if (p[k] != (char)(tid & 0x7f)) { atomicAdd(bad, 1); break; }
    }
}

template <class M>
__global__ void free_kernel(int n, M mm, void* const* ptrs) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) mm.free(ptrs[tid]);
}

#define CHECK(cond, msg)                                      \
    do {                                                      \
        if (!(cond)) {                                        \
            std::printf("compat kmalloc FAIL: %s\n", msg);    \
            return 1;                                         \
        }                                                     \
    } while (0)

template <class M>
int run(const char* name, M& mm, int n, void** d_ptrs, int* d_cnt, bool frees) {
    mm.reset();
    CHECK(getFreePagePercentage() == 1.0f, "reset leaves every page free");
    allocate_kernel<<<(n + 31) / 32, 32>>>(n, mm, d_ptrs);
    hipMemset(d_cnt, 0, 8);
    check_kernel<<<(n + 255) / 256, 256>>>(n, d_ptrs, d_cnt, d_cnt + 1);
    int cnt[2] = {0, 0};
    hipMemcpy(cnt, d_cnt, 8, hipMemcpyDeviceToHost);
    CHECK(cnt[1] == 0, "every request got a space");
    CHECK(cnt[0] == 0, "no space overlaps another");
    CHECK(getFreePagePercentage() < 1.0f, "spaces take pages");
    if (frees) {
        free_kernel<<<(n + 31) / 32, 32>>>(n, mm, d_ptrs);
        hipDeviceSynchronize();
        CHECK(getFreePagePercentage() == 1.0f, "free gives every page back");
    } else {
        mm.reset();
        CHECK(getFreePagePercentage() == 1.0f, "reset gives every page back");
    }
    std::printf("%s %d spaces\n", name, n);
    return 0;
}

int main() {
#ifndef PBM_KMALLOC_LEGACY
    initMemoryManagement(1, 4096);  // 262,144 pages of 4 KiB
#endif
    const int n = 30000;
    void** d_ptrs;
    int* d_cnt;
    hipMalloc(&d_ptrs, n * sizeof(void*));
    hipMalloc(&d_cnt, 8);
    KernelMallocCollab collab;
    if (run("KernelMallocCollab", collab, n, d_ptrs, d_cnt, true)) return 1;
    KernelMallocSimple simple(1000000);
    if (run("KernelMallocSimple", simple, n, d_ptrs, d_cnt, false)) return 1;
#ifdef PBM_KMALLOC_LEGACY
    std::printf("compat kmalloc legacy ok\n");
#else
    std::printf("compat kmalloc ok\n");
#endif
    return 0;
}
