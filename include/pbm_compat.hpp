/*
 * pbm_compat.hpp — the reference's host and device API names over the
 * C-ABI, for code written against RandomWalkBasic_BM.cuh / CollabRW_BM.cuh
 * / CRW_BM.cuh / RWMalloc.cuh / RandomWalkBasic.cuh, the legacy walks of
 * parallelPage.cu, malloc.cuh, and (with PBM_COMPAT_OUROBOROS defined)
 * Ouroboros.cuh, whose getPage returns a pointer and whose freePage takes
 * one.
 *
 * The reference picks its policy by which header a translation unit
 * includes (unitTests/unitTest1.cu:7) and keeps one allocator per
 * translation unit in file-static device globals set with
 * cudaMemcpyToSymbol (Paging.cuh:23-28).  Here PBM_COMPAT_STRATEGY picks
 * the policy before the include (default PBM_RW_BM) and this header keeps
 * one pool per translation unit: a handle on the host and a pbm_ctx copy in
 * a file-static __device__ variable.  Host errors print the library's
 * message and exit with the error code, as gpuErrchk does
 * (ErrorHandler.cuh:6-14).  Walk seeds mix the thread index with the wall
 * clock, as the reference mixes it with clock() (RandomWalkBasic_BM.cuh:63).
 *
 * Build: hipcc --offload-arch=gfx950 -I include your.hip -L<dir> -lpbm
 */
#ifndef PBM_COMPAT_HPP
#define PBM_COMPAT_HPP

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include "pbm.h"
#include "pbm_device.hpp"
#include "pbm_kmalloc_classes.hpp"

#ifndef PBM_COMPAT_STRATEGY
#ifdef PBM_COMPAT_OUROBOROS
#define PBM_COMPAT_STRATEGY PBM_OUROBOROS
#else
#define PBM_COMPAT_STRATEGY PBM_RW_BM
#endif
#endif
#ifndef TOTAL_SIZE_GB_DEFAULT
#define TOTAL_SIZE_GB_DEFAULT 8   /* Paging.cuh:5 */
#endif
#ifndef PAGE_SIZE_DEFAULT
#define PAGE_SIZE_DEFAULT 512     /* Paging.cuh:4 */
#endif

static __device__ pbm_ctx pbm_compat_ctx;
static pbm_handle* pbm_compat_handle = nullptr;

static inline void pbm_compat_check(int rc, const char* what) {
    if (rc != PBM_OK) {
        std::fprintf(stderr, "%s: %s\n", what, pbm_last_error());
        std::exit(rc < 0 ? -rc : rc);
    }
}

static inline void pbm_compat_publish() {
    pbm_ctx c;
    pbm_compat_check(pbm_get_ctx(pbm_compat_handle, &c), "pbm_get_ctx");
    if (hipMemcpyToSymbol(HIP_SYMBOL(pbm_compat_ctx), &c, sizeof(c)) != hipSuccess) {
        std::fprintf(stderr, "hipMemcpyToSymbol(pbm_compat_ctx) failed\n");
        std::exit(2);
    }
}

/* initMemoryManagement (RandomWalkBasic_BM.cuh:42-55): nGB GiB of pages of
 * pageSize bytes, all free (h_total_n_pages = nGB * 2^30 / pageSize,
 * Paging.cuh:76). */
static inline void initMemoryManagement(int nGB = TOTAL_SIZE_GB_DEFAULT,
                                        int pageSize = PAGE_SIZE_DEFAULT) {
    pbm_config cfg = {};
    cfg.n_pages = (uint32_t)(((uint64_t)nGB << 30) / (uint64_t)pageSize);
    cfg.page_size = (uint32_t)pageSize;
    cfg.strategy = PBM_COMPAT_STRATEGY;
    cfg.backing = 1;
    cfg.seed = 42;
    if (pbm_compat_handle) pbm_destroy(pbm_compat_handle);
    pbm_compat_check(pbm_create(&cfg, &pbm_compat_handle), "initMemoryManagement");
    pbm_compat_publish();
}

/* resetMemoryManager (RandomWalkBasic_BM.cuh:104-106): every page free. */
static inline void resetMemoryManager() {
    pbm_compat_check(pbm_reset(pbm_compat_handle, nullptr), "resetMemoryManager");
    pbm_compat_check(hipDeviceSynchronize() == hipSuccess ? PBM_OK : PBM_ERR_HIP, "resetMemoryManager");
}

/* getFreePagePercentage (RandomWalkBasic_BM.cuh:109-145): free / total. */
static inline float getFreePagePercentage() {
    float f = 0;
    pbm_compat_check(pbm_free_page_percentage(pbm_compat_handle, &f, nullptr),
                     "getFreePagePercentage");
    return f;
}

/* prefillBuffer (RandomWalkBasic_BM.cuh:148-164): the allocator itself
 * takes (int)((1 - freePercentage) * N) pages, leaving freePercentage free. */
static inline void prefillBuffer(float freePercentage) {
    pbm_compat_check(pbm_prefill_requests(pbm_compat_handle, freePercentage, nullptr),
                     "prefillBuffer");
    pbm_compat_check(hipDeviceSynchronize() == hipSuccess ? PBM_OK : PBM_ERR_HIP, "prefillBuffer");
}

/* The handle behind the names above, for calls the reference has no name
 * for (batched phases, counters, snapshots). */
static inline pbm_handle* pbmCompatHandle() { return pbm_compat_handle; }

/* ---- device side ------------------------------------------------------ */

static __device__ inline int32_t pbm_compat_seed() {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return pbm::request_seed(pbm_compat_ctx.seed, (gid << 32) ^ wall_clock64());
}

/* pageAddress / getPageID (Paging.cuh:30-52). */
static __device__ inline void* pageAddress(int pageID) {
    return pbm::page_address(pbm_compat_ctx, pageID);
}
static __device__ inline int getPageID(void* ptr) { return pbm::page_id(pbm_compat_ctx, ptr); }

static __device__ inline int pbm_compat_get(int* stepCount) {
    uint32_t steps = 0;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t id = pbm::get_page_of<PBM_COMPAT_STRATEGY>(pbm_compat_ctx, gid, pbm_compat_seed(),
                                                             &steps, nullptr);
    if (stepCount) *stepCount = (int)steps;
    return id;
}

#ifdef PBM_COMPAT_OUROBOROS
/* getPage / freePage of the Ouroboros wrapper (Ouroboros.cuh:16-17,
 * 38-47): a page pointer (nullptr when the pool has no page for the
 * request), and free by pointer. */
static __device__ inline void* getPage() { return pageAddress(pbm_compat_get(nullptr)); }
static __device__ inline void freePage(void* ptr) {
    pbm::free_page_of<PBM_COMPAT_STRATEGY>(pbm_compat_ctx, getPageID(ptr), nullptr);
}
static __device__ inline void* getPagePtr() { return getPage(); }
#else
/* getPage (RandomWalkBasic_BM.cuh:60-89, RandomWalkBasic.cuh:54-80, CollabRW_BM.cuh:114-160,
 * CRW_BM.cuh:75-142, RWMalloc.cuh:113-149, and the legacy
 * getPageClusteredRandomWalk / getPageCollabRW, parallelPage.cu:200-218,
 * 642-680): a page id, or -1 when the pool has none (the reference spins).
 * The collaborative policies and RWMalloc are collective over the calling
 * wave's active lanes.  collectWarpData (CollabRW_BM only) prints the free
 * bits the wave found in each step, one line per step. */
static __device__ inline int getPage(int* stepCount = nullptr, bool collectWarpData = false) {
    if (PBM_COMPAT_STRATEGY == PBM_COLLAB_RW_BM && collectWarpData) {
        // the wave's lowest lane prints the free bits found in each step,
        // as CollabRW_BM.cuh:131-138 prints them (the first 64 steps)
        uint32_t yields[64];
        uint32_t it = 0, steps = 0;
        const unsigned first = (unsigned)__ffsll((long long)pbm::ballot(true)) - 1u;
        const int32_t id = pbm::get_page_collab(pbm_compat_ctx, pbm_compat_seed(), &steps,
                                                nullptr, yields, 64u, &it);
        if (pbm::lane_id() == first)
            for (uint32_t k = 0; k < it && k < 64u; ++k) printf("%u\n", yields[k]);
        if (stepCount) *stepCount = (int)steps;
        return id;
    }
    return pbm_compat_get(stepCount);
}

/* freePage (RandomWalkBasic_BM.cuh:98-102; the list pool's
 * freePageLinkedList, parallelPage.cu:399-418, appends to the ring). */
static __device__ inline void freePage(int pageID) {
    pbm::free_page_of<PBM_COMPAT_STRATEGY>(pbm_compat_ctx, pageID, nullptr);
}

/* getPagePtr (RandomWalkBasic_BM.cuh:92-95). */
static __device__ inline void* getPagePtr() { return pageAddress(getPage()); }
#endif

/* getNPage (RWMalloc.cuh:198-248): n <= 64 consecutive pages in one word;
 * wave-collective. */
static __device__ inline int getNPage(int n, int* stepCount = nullptr) {
    uint32_t steps = 0;
    const int32_t id = pbm::get_n_pages(pbm_compat_ctx, n < 0 ? 65u : (uint32_t)n,
                                        pbm_compat_seed(), &steps);
    if (stepCount) *stepCount = (int)steps;
    return id;
}

/* mallocRW / freeRW (RWMalloc.cuh:299-334); wave-collective. */
static __device__ inline void* mallocRW(size_t size) {
    return pbm::malloc_rw(pbm_compat_ctx, size, pbm_compat_seed(), nullptr);
}
static __device__ inline void freeRW(void* ptr) { pbm::free_rw(pbm_compat_ctx, ptr); }

/* ---- kernel malloc (malloc.cuh:14-37) ----------------------------------
 * KernelMallocSimple / KernelMallocCollab over this translation unit's pool
 * (initMemoryManagement runs with its defaults if it has not run). */

struct PbmCompatPool {
    static pbm_ctx ctx() {
        if (!pbm_compat_handle) initMemoryManagement();
        pbm_ctx c;
        pbm_compat_check(pbm_get_ctx(pbm_compat_handle, &c), "pbm_get_ctx");
        return c;
    }
    static void reset() { resetMemoryManager(); }
};

typedef KernelMallocCollabT<PbmCompatPool> KernelMallocCollab;
typedef KernelMallocSimpleT<PbmCompatPool> KernelMallocSimple;

#endif /* PBM_COMPAT_HPP */
