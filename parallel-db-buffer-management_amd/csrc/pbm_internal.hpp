// Internal declarations shared by the library's translation units.
#ifndef PBM_INTERNAL_HPP
#define PBM_INTERNAL_HPP

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <string>
#include "pbm.h"
#include "pbm_device.hpp"

// Maps of at most this many words scan with one-wave rank tiles.
#ifndef PBM_SMALL_SCAN_WORDS
#define PBM_SMALL_SCAN_WORDS 4096u
#endif

struct pbm_handle {
    pbm_config cfg;
    pbm_ctx ctx;
    int device = 0;
    int n_cu = 0;
    // device allocations owned by the handle
    unsigned long long* d_bitmap = nullptr;
    long long* d_free = nullptr;        // reservation counter
    pbm_counters* d_counters = nullptr;
    char* d_pool = nullptr;
    int32_t* d_hint = nullptr;
    unsigned* d_arms = nullptr;
    unsigned long long* d_clock = nullptr;
    unsigned long long* d_ring = nullptr;
    unsigned long long* d_ring_ht = nullptr;  // [0] head, [1] tail
    long long* d_phase = nullptr;
    unsigned long long* d_scratch = nullptr;  // 8 words for reductions
    unsigned long long* d_scan = nullptr;     // rank scan: one look-back word per tile
    unsigned long long* d_scan_ctl = nullptr; // rank scan: [1] look-back give-ups
    unsigned long long* d_credit = nullptr;   // batched free: 16 partial reservation credits
    unsigned* d_fresh = nullptr;              // Ouroboros: free pages at the last reset
    unsigned long long* d_carve = nullptr;    // Ouroboros: [0] carved, [1] fresh pages
    long long* d_qavail = nullptr;            // Ouroboros: unpromised queued pages
    // debug shadow of page owners (pbm_shadow_enable): one word per page and
    // the report words; null while off
    unsigned long long* d_shadow = nullptr;
    unsigned long long* d_shadow_rep = nullptr;
    // events the next batch call's main kernel records at its own start and
    // end (pbm_time_next_kernel); cleared by that call
    hipEvent_t t_start = nullptr;
    hipEvent_t t_stop = nullptr;
};

// An instantiated HIP graph and the device it runs on.
struct pbm_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace pbmi {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

// Pages per chunk of the Ouroboros comparator: requests of page_size bytes
// fall in the power-of-two size class of at least 16 bytes
// (Ouroboros/include/device/Helper.cuh:42-60); a class up to 8 KiB is
// carved from 8 KiB chunks, up to 32 KiB from 32 KiB chunks
// (Ouroboros/include/Parameters.h:9-18, InstanceDefinitions.cuh:26-28);
// larger pages are one per chunk.
inline uint32_t ouroboros_chunk_pages(uint32_t page_size) {
    uint64_t cls = 16;
    while (cls < page_size) cls <<= 1;
    if (cls <= 8192u) return (uint32_t)(8192u / cls);
    if (cls <= 32768u) return (uint32_t)(32768u / cls);
    return 1u;
}

// Strategies whose free pages live in a FIFO ring (LinkedList, Ouroboros):
// no batched, multi-page, cache or burst calls on them.
inline bool is_queue_strategy(int s) { return s == PBM_LINKED_LIST || s == PBM_OUROBOROS; }

// Launches `kernel`; when pbm_time_next_kernel armed the handle, with the
// armed events recorded at the kernel's own start and end.
template <typename F, typename... Args>
inline void launch_timed(pbm_handle* h, F kernel, dim3 grid, dim3 block, hipStream_t s,
                         Args... args) {
    if (h->t_start || h->t_stop) {
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, h->t_start, h->t_stop, 0, args...);
        h->t_start = h->t_stop = nullptr;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
    }
}

// Makes the handle's device current for the calling thread.
int use_device(pbm_handle* h);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Re-initialises strategy state after the page map changed wholesale
// (reset, prefill, restore): hints, arms, clock, free list.
int reinit_strategy_state(pbm_handle* h, hipStream_t s);

// Shadow owners: set every page's owner from the map (held or free), and
// record the grants or frees of a finished id-returning call.  No-ops while
// the shadow is off.
int shadow_sync(pbm_handle* h, hipStream_t s);
int shadow_note(pbm_handle* h, uint32_t n, uint64_t req_base, const int32_t* d_ids,
                const uint32_t* d_counts, bool freeing, hipStream_t s);

// Recomputes the reservation counter from the page map.
int recount_free(pbm_handle* h, hipStream_t s);

// LinkedList: refills the FIFO ring with the free pages of the page map in
// ascending order and resets head/tail tickets.  Ouroboros: empties the
// ring and lists the free pages, ascending, as the pages chunks are carved
// from.
int ring_rebuild(pbm_handle* h, hipStream_t s);

// One-page get launches of CRW_BM, RWMalloc, SingleClock, DistClock and
// LinkedList (plain or serial), and the list's free launch.
int launch_get_strategy(pbm_handle* h, const pbm_ctx& c, const pbm_launch& L, uint32_t n,
                        uint64_t req_base, int32_t* d_ids, uint32_t* d_steps, hipStream_t s);
int launch_free_list(const pbm_ctx& c, int strategy, const pbm_launch& L, uint32_t n,
                     const int32_t* d_ids, hipStream_t s);

}  // namespace pbmi

#define PBM_HIP(expr)                                                            \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) return pbmi::hip_fail(e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define PBM_CHECK_HANDLE(h)                                                      \
    do {                                                                         \
        if (!(h)) return pbmi::fail(PBM_ERR_INVALID, "null pbm_handle");         \
        int rc_ = pbmi::use_device(h);                                           \
        if (rc_) return rc_;                                                     \
    } while (0)

#endif
