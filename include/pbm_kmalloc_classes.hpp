/*
 * pbm_kmalloc_classes.hpp — the MemoryManager classes of malloc.cuh:14-37
 * (KernelMallocSimple, KernelMallocCollab) over a pool that a provider
 * names.  pbm_compat.hpp instantiates them over its pool and
 * pbm_legacy.hpp over its CollabRW_BM pool (malloc.cu's INITBM,
 * malloc.cu:9-12).
 *
 * A provider P has two host functions: pbm_ctx P::ctx(), creating the pool
 * on first use, and void P::reset(), freeing every page.  The objects are
 * built on the host and passed to kernels by value, as the reference's
 * tests do (unitTests/unitTestMalloc1.cu:14-41).  Like the reference's
 * classes they have no destructor that frees anything: the copies made for
 * kernel launches must not release the state.  malloc is collective over
 * the calling wave's active lanes, because its chunks come from getNPage.
 */
#ifndef PBM_KMALLOC_CLASSES_HPP
#define PBM_KMALLOC_CLASSES_HPP

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include "pbm.h"
#include "pbm_device.hpp"

static inline pbm_ctx pbm_kmalloc_checked(pbm_ctx c) {
    if (!c.pool || (c.page_size & 7u)) {
        std::fprintf(stderr,
                     "kernel malloc needs page memory and a page size that is a multiple of 8\n");
        std::exit(-PBM_ERR_UNSUPPORTED);
    }
    return c;
}

static __device__ inline int32_t pbm_kmalloc_seed(const pbm_ctx& c) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return pbm::request_seed(c.seed, (gid << 32) ^ wall_clock64());
}

template <class P>
class KernelMallocCollabT {
public:
    /* KernelMallocCollab() (malloc.cu:214-218) */
    __host__ KernelMallocCollabT() : ctx_(pbm_kmalloc_checked(P::ctx())) {}
    /* reset(): every page free (malloc.cu:222-224). */
    __host__ void reset() { P::reset(); }
    /* malloc / free (malloc.cu:231-320). */
    __device__ void* malloc(int size) {
        return pbm::kmalloc_collab(ctx_, size < 0 ? 0u : (uint32_t)size, pbm_kmalloc_seed(ctx_),
                                   nullptr);
    }
    __device__ void free(void* addr) { pbm::kfree_collab(ctx_, addr); }

private:
    pbm_ctx ctx_;
};

template <class P>
class KernelMallocSimpleT {
public:
    /* KernelMallocSimple(maxNThreads) (malloc.cu:76-84): one current chunk
     * per thread index below maxNThreads; threads beyond get nullptr. */
    __host__ explicit KernelMallocSimpleT(int maxNThreads)
        : ctx_(pbm_kmalloc_checked(P::ctx())), max_threads_(maxNThreads) {
        if (hipMalloc((void**)&current_, (size_t)max_threads_ * 4u) != hipSuccess) {
            std::fprintf(stderr, "KernelMallocSimple: hipMalloc failed\n");
            std::exit(-PBM_ERR_OOM);
        }
        clear_current();
    }
    /* reset(): every page free and no current chunks (malloc.cu:99-104). */
    __host__ void reset() {
        P::reset();
        clear_current();
    }
    /* malloc (malloc.cu:129-175); free does nothing (malloc.cu:178-181). */
    __device__ void* malloc(int size) {
        const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
        int32_t dummy = -1;
        const bool slot = tid < (uint32_t)max_threads_;
        void* p = pbm::kmalloc_simple(ctx_, slot ? &current_[tid] : &dummy,
                                      slot && size >= 0 ? (uint32_t)size : 0xFFFFFFFFu,
                                      pbm_kmalloc_seed(ctx_), nullptr);
        return slot ? p : nullptr;
    }
    __device__ void free(void*) {}

private:
    __host__ void clear_current() {
        if (hipMemset(current_, 0xFF, (size_t)max_threads_ * 4u) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            std::fprintf(stderr, "KernelMallocSimple: clearing the current chunks failed\n");
            std::exit(-PBM_ERR_HIP);
        }
    }
    pbm_ctx ctx_;
    int32_t* current_ = nullptr;
    int max_threads_ = 0;
};

#endif /* PBM_KMALLOC_CLASSES_HPP */
