/*
 * pbm_legacy.hpp — the legacy generation's names (parallelPage.cuh:6-80)
 * over the C-ABI, for drivers written against parallelPage.cu.
 *
 * The legacy header gives every policy its own family of names and its own
 * page map, all usable in one translation unit: initPagesX(), getPageX(int*),
 * freePageX(int[, int*]), printNumPagesLeftX(), resetBufferX(), and
 * getXPageX(int, int*) for the bitmap walks.  Here each family is one pool
 * of PBM_LEGACY_TOTAL_N_PAGES pages of PBM_LEGACY_PAGE_SIZE bytes (the
 * header's TOTAL_N_PAGES and PAGE_SIZE, parallelPage.cuh:3-4, unless defined
 * before the include), created by its initPagesX with page memory, and held
 * in a file-static handle and __device__ pbm_ctx.  The policies:
 *
 *   RandomWalk              PBM_RW_BASIC         ClusteredRandomWalk     PBM_CRW_BASIC
 *   LinkedList              PBM_LINKED_LIST      SingleClock             PBM_SINGLE_CLOCK
 *   DistClock               PBM_DIST_CLOCK       CollabRW                PBM_COLLAB_RW_BASIC
 *   CollabRW_BM             PBM_COLLAB_RW_BM     RandomWalk_BM           PBM_RW_BM
 *   ClusteredRandomWalk_BM  PBM_CRW_BM
 *
 * getPage returns -1 when the pool has no page (the legacy code traps or
 * spins); getXPage takes x <= 64 pages inside one map word (per thread for
 * RandomWalk_BM and ClusteredRandomWalk_BM, wave-collective for
 * CollabRW_BM as getXPageCollabRW_BM is, parallelPage.cu:934-991).
 * pageAddress(id) resolves ids of the RandomWalk_BM pool.  malloc.cuh's
 * KernelMallocSimple / KernelMallocCollab allocate from the CollabRW_BM
 * pool, as malloc.cu does.  pbm_compat.hpp's names cannot share a
 * translation unit with these.
 *
 * Build: hipcc --offload-arch=gfx950 -I include your.hip -L<dir> -lpbm
 */
#ifndef PBM_LEGACY_HPP
#define PBM_LEGACY_HPP

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include "pbm.h"
#include "pbm_device.hpp"
#include "pbm_kmalloc_classes.hpp"

#ifndef PBM_LEGACY_TOTAL_N_PAGES
#define PBM_LEGACY_TOTAL_N_PAGES 41900000u /* parallelPage.cuh:3 */
#endif
#ifndef PBM_LEGACY_PAGE_SIZE
#define PBM_LEGACY_PAGE_SIZE 264u          /* parallelPage.cuh:4 */
#endif

static inline void pbm_legacy_check(int rc, const char* what) {
    if (rc != PBM_OK) {
        std::fprintf(stderr, "%s: %s\n", what, pbm_last_error());
        std::exit(rc < 0 ? -rc : rc);
    }
}

static inline pbm_handle* pbm_legacy_create(int strategy, const void* ctx_symbol) {
    pbm_config cfg = {};
    cfg.n_pages = PBM_LEGACY_TOTAL_N_PAGES;
    cfg.page_size = PBM_LEGACY_PAGE_SIZE;
    cfg.strategy = strategy;
    cfg.backing = 1;
    cfg.seed = 42;
    pbm_handle* h = nullptr;
    pbm_legacy_check(pbm_create(&cfg, &h), "initPages");
    pbm_ctx c;
    pbm_legacy_check(pbm_get_ctx(h, &c), "pbm_get_ctx");
    if (hipMemcpyToSymbol(ctx_symbol, &c, sizeof(c)) != hipSuccess) {
        std::fprintf(stderr, "initPages: hipMemcpyToSymbol failed\n");
        std::exit(-PBM_ERR_HIP);
    }
    return h;
}

/* printNumPagesLeftX: "[TAG info] Number of free pages: N" for every
 * family (the list's own line reports its length, parallelPage.cu:459). */
static inline void pbm_legacy_print_free(pbm_handle* h, const char* tag) {
    uint64_t used = 0;
    pbm_legacy_check(pbm_count_used(h, &used, nullptr), "printNumPagesLeft");
    std::printf("[%s info] Number of free pages: %d \n", tag,
                (int)(PBM_LEGACY_TOTAL_N_PAGES - used));
}

static __device__ inline int32_t pbm_legacy_seed(const pbm_ctx& c) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return pbm::request_seed(c.seed, (gid << 32) ^ wall_clock64());
}

/* One family of names for policy STRAT, printing TAG. */
#define PBM_LEGACY_FAMILY(Name, STRAT, TAG)                                                   \
    static __device__ pbm_ctx pbm_legacy_ctx_##Name;                                           \
    static pbm_handle* pbm_legacy_h_##Name = nullptr;                                          \
    static inline void initPages##Name() {                                                     \
        if (pbm_legacy_h_##Name) pbm_destroy(pbm_legacy_h_##Name);                             \
        pbm_legacy_h_##Name = pbm_legacy_create(STRAT, (const void*)&pbm_legacy_ctx_##Name);   \
    }                                                                                          \
    static __device__ inline int getPage##Name(int* step_count = nullptr) {                    \
        uint32_t steps = 0;                                                                    \
        const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;                            \
        const int32_t id = pbm::get_page_of<STRAT>(pbm_legacy_ctx_##Name, gid,                 \
                                                   pbm_legacy_seed(pbm_legacy_ctx_##Name),     \
                                                   &steps, nullptr);                           \
        if (step_count) *step_count = (int)steps;                                              \
        return id;                                                                             \
    }                                                                                          \
    static __device__ inline void freePage##Name(int pageID, int* step_count = nullptr) {      \
        if (step_count) *step_count = 0;                                                       \
        pbm::free_page_of<STRAT>(pbm_legacy_ctx_##Name, pageID, nullptr);                      \
    }                                                                                          \
    static inline void printNumPagesLeft##Name() {                                             \
        pbm_legacy_print_free(pbm_legacy_h_##Name, TAG);                                       \
    }                                                                                          \
    static inline void resetBuffer##Name() {                                                   \
        pbm_legacy_check(pbm_reset(pbm_legacy_h_##Name, nullptr), "resetBuffer");             \
        pbm_legacy_check(hipDeviceSynchronize() == hipSuccess ? PBM_OK : PBM_ERR_HIP,          \
                         "resetBuffer");                                                       \
    }                                                                                          \
    static inline pbm_handle* pbmLegacyHandle##Name() { return pbm_legacy_h_##Name; }

PBM_LEGACY_FAMILY(RandomWalk, PBM_RW_BASIC, "RW")
PBM_LEGACY_FAMILY(ClusteredRandomWalk, PBM_CRW_BASIC, "CRW")
PBM_LEGACY_FAMILY(LinkedList, PBM_LINKED_LIST, "LL")
PBM_LEGACY_FAMILY(SingleClock, PBM_SINGLE_CLOCK, "SC")
PBM_LEGACY_FAMILY(DistClock, PBM_DIST_CLOCK, "DC")
PBM_LEGACY_FAMILY(CollabRW, PBM_COLLAB_RW_BASIC, "CoRW")
PBM_LEGACY_FAMILY(CollabRW_BM, PBM_COLLAB_RW_BM, "CoRW_BM")
PBM_LEGACY_FAMILY(RandomWalk_BM, PBM_RW_BM, "RW")
PBM_LEGACY_FAMILY(ClusteredRandomWalk_BM, PBM_CRW_BM, "CRW_BM")

#undef PBM_LEGACY_FAMILY

/* getXPage of the bitmap walks (parallelPage.cu:934-991, 1104-1142). */
static __device__ inline int getXPageRandomWalk_BM(int X, int* step_count = nullptr) {
    uint32_t steps = 0;
    const int32_t id = pbm::get_x_pages(pbm_legacy_ctx_RandomWalk_BM, X < 0 ? 65u : (uint32_t)X,
                                        pbm_legacy_seed(pbm_legacy_ctx_RandomWalk_BM), &steps);
    if (step_count) *step_count = (int)steps;
    return id;
}
static __device__ inline int getXPageClusteredRandomWalk_BM(int X, int* step_count = nullptr) {
    uint32_t steps = 0;
    const int32_t id = pbm::get_x_pages(pbm_legacy_ctx_ClusteredRandomWalk_BM,
                                        X < 0 ? 65u : (uint32_t)X,
                                        pbm_legacy_seed(pbm_legacy_ctx_ClusteredRandomWalk_BM),
                                        &steps);
    if (step_count) *step_count = (int)steps;
    return id;
}
static __device__ inline int getXPageCollabRW_BM(int X, int* step_count = nullptr) {
    uint32_t steps = 0;
    const int32_t id = pbm::get_n_pages(pbm_legacy_ctx_CollabRW_BM, X < 0 ? 65u : (uint32_t)X,
                                        pbm_legacy_seed(pbm_legacy_ctx_CollabRW_BM), &steps);
    if (step_count) *step_count = (int)steps;
    return id;
}

/* pageAddress (parallelPage.cuh:7) of the RandomWalk_BM pool. */
static __device__ inline void* pageAddress(int pageID) {
    return pbm::page_address(pbm_legacy_ctx_RandomWalk_BM, pageID);
}

/* malloc.cuh's classes over the CollabRW_BM pool, which malloc.cu
 * allocates from (INITBM / RESETBM / GETXPAGE, malloc.cu:9-12); the pool is
 * created by the first manager built (the reference re-creates it with
 * every manager, malloc.cu:15-19). */
struct PbmLegacyMallocPool {
    static pbm_ctx ctx() {
        if (!pbm_legacy_h_CollabRW_BM) initPagesCollabRW_BM();
        pbm_ctx c;
        pbm_legacy_check(pbm_get_ctx(pbm_legacy_h_CollabRW_BM, &c), "pbm_get_ctx");
        return c;
    }
    static void reset() { resetBufferCollabRW_BM(); }
};

typedef KernelMallocCollabT<PbmLegacyMallocPool> KernelMallocCollab;
typedef KernelMallocSimpleT<PbmLegacyMallocPool> KernelMallocSimple;

/* monitorBufferRandomWalk (parallelPage.cu:209-223): on a stream of its
 * own, every msGap ms, maxCount times, prints the RandomWalk pool's free
 * pages and their share, from one workgroup reading the map. */
static __global__ void pbm_legacy_monitor_kernel(uint64_t gap_ticks, int max_count) {
    __shared__ unsigned long long s_free;
    const pbm_ctx& c = pbm_legacy_ctx_RandomWalk;
    unsigned long long t0 = wall_clock64();
    for (int k = 0; k < max_count; ++k) {
        while (wall_clock64() < t0 + (unsigned long long)k * gap_ticks) __builtin_amdgcn_s_sleep(8);
        if (threadIdx.x == 0) s_free = 0;
        __syncthreads();
        unsigned v = 0;
        for (uint32_t i = threadIdx.x; i < c.n_words; i += blockDim.x) {
            unsigned long long f = ~pbm::ld_word(&c.bitmap[i]);
            if ((uint64_t)i * 64u + 64u > c.n_pages)
                f &= (1ull << (uint32_t)(c.n_pages - (uint64_t)i * 64u)) - 1ull;
            v += (unsigned)__popcll(f);
        }
        atomicAdd(&s_free, (unsigned long long)v);
        __syncthreads();
        if (threadIdx.x == 0)
            printf("[RW info] Number of free pages: %d, percentage free: %2.1f %% \n", (int)s_free,
                   (float)s_free / c.n_pages * 100.0f);
        __syncthreads();
    }
}

static inline void monitorBufferRandomWalk(int msGap = 100, int maxCount = 100) {
    static hipStream_t s = nullptr;
    if (!s && hipStreamCreate(&s) != hipSuccess) {
        std::fprintf(stderr, "monitorBufferRandomWalk: hipStreamCreate failed\n");
        std::exit(-PBM_ERR_HIP);
    }
    unsigned long long hz = 0;
    pbm_legacy_check(pbm_wall_clock_hz(0, &hz), "monitorBufferRandomWalk");
    hipLaunchKernelGGL(pbm_legacy_monitor_kernel, dim3(1), dim3(256), 0, s,
                       (uint64_t)msGap * hz / 1000u, maxCount);
}

#endif /* PBM_LEGACY_HPP */
