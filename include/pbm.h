/*
 * pbm.h — C-ABI of the MI355X (gfx950) parallel page/buffer manager.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns 0 (PBM_OK) or a negative error code; `pbm_last_error()` gives the
 * message of the last failure on the calling thread.  Device pointers are
 * plain `void*`/typed pointers into HIP device memory; `stream` is a
 * `hipStream_t` passed as `void*` (NULL = the legacy default stream).
 *
 * The reference keeps one allocator per translation unit in file-static
 * device globals set by cudaMemcpyToSymbol (Paging.cuh:23-28,
 * RandomWalkBasic_BM.cuh:28-32).  Here a handle owns all device state, so
 * several pools can coexist and `pbm_destroy` frees everything.
 *
 * Bitmap convention (all strategies): one bit per page in 64-bit words,
 * bit = 1 means used, page id = word * 64 + bit.  Padding bits past
 * n_pages in the last word are kept set so they are never handed out.
 */
#ifndef PBM_H
#define PBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define PBM_OK               0
#define PBM_ERR_INVALID     -1   /* bad argument or handle */
#define PBM_ERR_HIP         -2   /* a HIP runtime call failed */
#define PBM_ERR_OOM         -3   /* device allocation failed */
#define PBM_ERR_UNSUPPORTED -4   /* operation not defined for the strategy */
#define PBM_ERR_RANGE       -5   /* size beyond a documented limit */

/* ---- allocation policies ---------------------------------------------
 * The reference selects one policy per translation unit by swapping the
 * included header (unitTests/unitTest1.cu:4-7); here it is a field of the
 * pool configuration. */
typedef enum pbm_strategy {
    PBM_RW_BM           = 0, /* RandomWalkBasic_BM.cuh:60-102  plain bitmap random walk       */
    PBM_COLLAB_RW_BM    = 1, /* CollabRW_BM.cuh:114-160        wave-collaborative window walk */
    PBM_CRW_BM          = 2, /* CRW_BM.cuh:75-156              clustered walk with hints      */
    PBM_RW_MALLOC       = 3, /* RWMalloc.cuh:113-334           collaborative walk + getNPage  */
    PBM_LINKED_LIST     = 4, /* parallelPage.cu:36-46,300-466  FIFO free list                 */
    PBM_SINGLE_CLOCK    = 5, /* parallelPage.cu:470-535        one global clock arm           */
    PBM_DIST_CLOCK      = 6, /* parallelPage.cu:537-618        per-thread clock arms          */
    PBM_RW_BASIC        = 7, /* RandomWalkBasic.cuh:54-84      random walk over single pages  */
    PBM_CRW_BASIC       = 8, /* parallelPage.cu:200-224        clustered page walk, page hint */
    PBM_COLLAB_RW_BASIC = 9, /* parallelPage.cu:642-685        wave-collaborative page walk   */
    PBM_OUROBOROS       = 10 /* Ouroboros.cuh:22-62            queue of free pages, lazy chunks */
} pbm_strategy;

/* Order in which a collaborative step hands the free bits of its window to
 * the requesting lanes. */
typedef enum pbm_order {
    PBM_ORDER_PACKED    = 0, /* lane-major: all free bits of the first word, then the next */
    PBM_ORDER_REFERENCE = 1  /* round-robin over words, as CollabRW_BM.cuh:140-153 ranks them */
} pbm_order;

/* How a batch entry point schedules its requests. */
typedef enum pbm_mode {
    PBM_MODE_PLAIN   = 0, /* one thread per request calling the strategy's get/free       */
    PBM_MODE_BATCHED = 1, /* persistent grid, wave-collective windows, phase reservation  */
    PBM_MODE_SERIAL  = 2  /* one thread (one wave for collective policies) runs the
                             requests in index order: the schedule of the serial oracle  */
} pbm_mode;

/* ---- configuration ----------------------------------------------------
 * Replaces the (nGB, pageSize) arguments of initMemoryManagement
 * (RandomWalkBasic_BM.cuh:42) and the compile-time TOTAL_N_PAGES/PAGE_SIZE
 * of the legacy header (parallelPage.cuh:3-4). */
typedef struct pbm_config {
    uint32_t n_pages;      /* pages in the pool, 1 .. 2^30                                */
    uint32_t page_size;    /* bytes per page (any value >= 8)                             */
    int32_t  strategy;     /* pbm_strategy                                                */
    int32_t  backing;      /* 1: allocate n_pages*page_size bytes of page memory;
                              0: metadata-only pool (page_address returns 0)              */
    uint64_t seed;         /* seeds request walks, clock arms and hints                   */
    uint32_t walk_budget;  /* random jumps before a walk falls back to a linear sweep;
                              0 = default (1<<20)                                         */
    uint32_t max_threads;  /* per-thread state slots (CRW hints, DistClock arms);
                              0 = default (1<<20)                                         */
    int32_t  order;        /* pbm_order for collaborative strategies                      */
    int32_t  device;       /* HIP device ordinal                                          */
} pbm_config;

/* Device-side event counters of a pool (read with pbm_read_counters). */
typedef struct pbm_counters {
    unsigned long long allocs;      /* successful page grants                         */
    unsigned long long frees;       /* pages returned                                 */
    unsigned long long failures;    /* requests that got no page                      */
    unsigned long long steps;       /* walk steps summed over requests                */
    unsigned long long probes;      /* bitmap words read by walks                     */
    unsigned long long atomics;     /* read-modify-write atomics on the page map      */
    unsigned long long fallbacks;   /* walks that exhausted the budget and swept      */
    unsigned long long errors;      /* double frees, ids out of range, ring overruns  */
} pbm_counters;

/* Plain-old-data view of a pool that device code receives by value.
 * Obtained with pbm_get_ctx; valid until pbm_destroy. */
typedef struct pbm_ctx {
    unsigned long long* bitmap;      /* n_words page-map words                          */
    long long*          free_count;  /* reservation counter: pages not yet granted      */
    pbm_counters*       counters;    /* device counters                                 */
    char*               pool;        /* page memory base (NULL without backing)         */
    int32_t*            hint;        /* CRW: last granted page per slot, -1 = none      */
    unsigned*           arms;        /* DistClock: per-slot arm                         */
    unsigned long long* clock_arm;   /* SingleClock: global arm (64-bit, never wraps)   */
    unsigned long long* ring;        /* LinkedList, Ouroboros: n_pages slots (lap<<32 | page id) */
    unsigned long long* ring_head;   /* LinkedList, Ouroboros: next ticket to dequeue   */
    unsigned long long* ring_tail;   /* LinkedList, Ouroboros: next ticket to enqueue   */
    long long*          phase;       /* batched calls: [1] wave completion counter      */
    unsigned long long* scan;        /* n_tiles look-back words of the single-pass rank
                                        scan (call parity, state, grant, prefix)        */
    unsigned long long* scan_ctl;    /* rank scan: [1] look-back give-ups (stays 0)  */
    uint32_t n_pages;
    uint32_t n_words;
    uint32_t page_size;
    uint32_t walk_budget;
    uint32_t max_threads;
    int32_t  order;
    uint64_t seed;
    uint32_t n_tiles;                /* rank-scan tiles: ceil(n_words / (scan_threads*scan_wpt)) */
    uint32_t scan_wpt;               /* rank scan: map words per thread (1 or 4)        */
    unsigned*           fresh;       /* Ouroboros: pages free at the last reset, ascending */
    unsigned long long* carve;       /* Ouroboros: [0] next fresh index, [1] fresh pages */
    long long*          q_avail;     /* Ouroboros: queued pages no wave has been promised */
    uint32_t chunk_pages;            /* Ouroboros: pages carved per chunk                */
    uint32_t scan_threads;           /* rank scan: threads per tile (64 or 256)          */
} pbm_ctx;

/* Launch shape of a batch call. */
typedef struct pbm_launch {
    int32_t  mode;          /* pbm_mode                                                  */
    uint32_t block;         /* threads per workgroup, multiple of 64 (0 = 256)           */
    uint32_t waves_per_cu;  /* batched: persistent waves per CU (0 = 8)                  */
    uint32_t walk_budget;   /* 0 = the pool's walk_budget                                */
    uint32_t cluster_words; /* batched: words per window, 64 or 128 or 256 (0 = 64);
                               getNPage, mallocRW and their bursts: the words one
                               search round reads (64 = the reference's one word per
                               lane)                                                    */
    uint32_t flags;         /* PBM_LAUNCH_* bits                                         */
} pbm_launch;

#define PBM_LAUNCH_NO_COUNTERS 1u  /* skip the device counter updates                  */
#define PBM_LAUNCH_ONE_WAVE    2u  /* batched mode on a single wave: the schedule of the
                                      serial oracle (orc_batched_serial); with
                                      PBM_LAUNCH_RANK, one workgroup scans every tile in
                                      order                                              */
#define PBM_LAUNCH_RANK        4u  /* batched mode by rank, one pass over the map: request r
                                      of the call takes the free page of rank r (map order);
                                      each tile (256 or 1024 words) learns the free pages before it by
                                      decoupled look-back and claims its granted bits with
                                      one atomicOr per word.                               */
#define PBM_LAUNCH_PERSISTENT  16u /* mallocRW bursts: every burst of the range in one
                                      persistent kernel (batched mode): waves take work
                                      items in order from a ticket and wait on the
                                      items they depend on                             */
#define PBM_LAUNCH_EXCLUSIVE   32u /* with PBM_LAUNCH_RANK: the caller guarantees that no
                                      other kernel reads or writes the pool's map while
                                      the call runs (calls on one stream, or ordered by
                                      events).  Each tile then claims its words with
                                      plain stores of the word it read plus its granted
                                      bits, instead of device-scope atomics, which run
                                      at the memory side on MI355X.  Same ids and map as
                                      without it; no reference counterpart.  Ignored
                                      without PBM_LAUNCH_RANK                          */
#define PBM_LAUNCH_ORDERED     64u /* with PBM_LAUNCH_PERSISTENT: every alloc item waits
                                      for all free items of its burst before it reserves
                                      (the order of the phase kernels).  Without it an
                                      alloc item runs ahead of its burst's frees while at
                                      least a quarter of the pool's pages are unreserved,
                                      and one that then finds no run waits for them and
                                      tries once more                                  */

typedef struct pbm_handle pbm_handle;

/* ---- lifecycle -------------------------------------------------------- */

/* Creates a pool: page memory (optional), page map cleared to all-free,
 * reservation counter = n_pages, strategy state initialised.
 * Replaces initMemoryManagement (RandomWalkBasic_BM.cuh:42-55,
 * CollabRW_BM.cuh:43-56, CRW_BM.cuh:46-70, RWMalloc.cuh:42-55),
 * initPages (Paging.cuh:59-85) and the legacy initPages<Strategy>()
 * family (parallelPage.cuh:10,20,26,33,40,47,54,62,70). */
int pbm_create(const pbm_config* cfg, pbm_handle** out);

/* Frees every device allocation of the pool.  The reference has no
 * counterpart (its pools live until process exit). */
int pbm_destroy(pbm_handle* h);

/* Marks every page free and restores the strategy's initial state
 * (hints cleared, arms re-seeded, list refilled 0..n-1).
 * Replaces resetMemoryManager (RandomWalkBasic_BM.cuh:104-106) and
 * resetBuffer<Strategy>() (parallelPage.cuh:14,30,37,44,51,59,67,75). */
int pbm_reset(pbm_handle* h, void* stream);

/* Deterministic pre-occupancy: page i is marked used iff
 * mix(seed, i) < fill * 2^64 (identical on host and device).  Replaces the
 * self-filling prefillBuffer (RandomWalkBasic_BM.cuh:148-164), whose
 * pattern depends on the schedule.  The reservation counter is set to
 * the number of free pages. */
int pbm_prefill_bernoulli(pbm_handle* h, double fill, uint64_t seed, void* stream);

/* Reference-style pre-fill: (int)((1-freePercentage)*n_pages) requests
 * are served by the pool's own allocator and never freed, as
 * prefillBuffer does (RandomWalkBasic_BM.cuh:148-164). */
int pbm_prefill_requests(pbm_handle* h, float free_percentage, void* stream);

/* Number of used pages (popcount of the page map, padding excluded).
 * Replaces popc_kernel + cub::DeviceReduce::Sum (RandomWalkBasic_BM.cuh:109-145). */
int pbm_count_used(pbm_handle* h, uint64_t* used, void* stream);

/* Free pages / n_pages.  Replaces getFreePagePercentage
 * (RandomWalkBasic_BM.cuh:130-145). */
int pbm_free_page_percentage(pbm_handle* h, float* out, void* stream);

/* Copies the page map to / from host memory (`words` must equal n_words).
 * restore also resets the reservation counter to the free pages of the
 * restored map, and sets the padding bits past n_pages in the last word
 * whatever the host copy holds there.  No reference counterpart
 * (checkpoint/resume). */
int pbm_snapshot_bitmap(pbm_handle* h, uint64_t* host, size_t words, void* stream);
int pbm_restore_bitmap(pbm_handle* h, const uint64_t* host, size_t words, void* stream);

/* Reads the reservation counter (pages not yet granted). */
int pbm_read_free_count(pbm_handle* h, int64_t* out, void* stream);

/* Device view for kernels that include pbm_device.hpp.  Replaces the
 * cudaMemcpyToSymbol publication of d_PageMapRandomWalk_BM / Bitmap_length_d
 * (RandomWalkBasic_BM.cuh:48-54) and d_page_groups (Paging.cuh:73-81). */
int pbm_get_ctx(pbm_handle* h, pbm_ctx* out);

int pbm_read_counters(pbm_handle* h, pbm_counters* out, void* stream);
int pbm_clear_counters(pbm_handle* h, void* stream);

/* ---- batch entry points (device-resident inputs and outputs) ----------- */

/* Serves n one-page requests with indices req_base .. req_base+n-1.
 * d_ids[i] = page id or -1; d_steps[i] (optional, may be NULL) = walk
 * steps of request i.  The request index, not the thread, seeds the walk.
 * Batched mode: a call's workgroups share the pool's completion and
 * look-back words, so two batched calls of one pool must not run at once.
 * The batched walk reads the reservation counter when it starts and takes
 * its grant when it ends, so a plain call that reserves from the pool
 * while it runs can leave both short (the excess requests fail after
 * their walks and count as errors).  The rank launch (PBM_LAUNCH_RANK)
 * takes its grant with one reserving atomic and may run beside plain
 * calls.
 * Replaces the drivers' get1page_kernel / getPage_kernel
 * (unitTests/unitTest1.cu:22-30, unitTest1.5.cu:32-36) calling
 * getPage (RandomWalkBasic_BM.cuh:60-89, CollabRW_BM.cuh:114-160,
 * CRW_BM.cuh:75-142, RWMalloc.cuh:113-149, parallelPage.cu:342,491,571). */
int pbm_get_pages(pbm_handle* h, uint32_t n, uint64_t req_base,
                  int32_t* d_ids, uint32_t* d_steps,
                  const pbm_launch* launch, void* stream);

/* Returns the n pages in d_ids (entries < 0 are skipped).  In batched
 * mode the workgroups' reservation credits meet in per-pool partial
 * counters, so two batched frees of one pool must not run at once (as for
 * every batched call).
 * Replaces freePage_kernel (unitTests/unitTest1.5.cu:38-44) calling
 * freePage (RandomWalkBasic_BM.cuh:98-102, parallelPage.cu:399,511,594). */
int pbm_free_pages(pbm_handle* h, uint32_t n, const int32_t* d_ids,
                   const pbm_launch* launch, void* stream);

/* ---- multi-page requests and device malloc (RWMalloc) ------------------ */

/* Serves n requests for d_counts[i] consecutive pages (0 <= count <= 64,
 * one 64-page word); d_ids[i] = first page or -1 (0 for a count of 0, a
 * coalesced member).  Wave-collective per 64 requests.  Replaces the
 * drivers' getNPage kernels (unitTests/unitTest2.cu) calling getNPage
 * (RWMalloc.cuh:198-248) and the legacy getXPage<Strategy>
 * (parallelPage.cuh:56,64,72). */
int pbm_get_n_pages(pbm_handle* h, uint32_t n, uint64_t req_base, const uint32_t* d_counts,
                    int32_t* d_ids, uint32_t* d_steps, const pbm_launch* launch, void* stream);

/* The per-thread form: request i (one thread each) takes d_counts[i]
 * consecutive pages inside one word (0 <= count <= 64) from a random walk
 * over words.  Replaces the legacy getXPageRandomWalk_BM
 * (parallelPage.cu:1104-1142) behind getXPage<Strategy>
 * (parallelPage.cuh:56,64,72).  Plain or serial; free with
 * pbm_free_n_pages. */
int pbm_get_x_pages(pbm_handle* h, uint32_t n, uint64_t req_base, const uint32_t* d_counts,
                    int32_t* d_ids, uint32_t* d_steps, const pbm_launch* launch, void* stream);

/* Returns the runs (d_ids[i], d_counts[i]); ids < 0 or counts 0 skipped.
 * Replaces releasePages (RWMalloc.cuh:60-70). */
int pbm_free_n_pages(pbm_handle* h, uint32_t n, const int32_t* d_ids, const uint32_t* d_counts,
                     const pbm_launch* launch, void* stream);

/* mallocRW for n requests of d_sizes[i] bytes; d_ptrs[i] = device pointer
 * or 0.  Needs a pool created with backing.  Replaces malloc_kernel
 * (unitTests/unitTest3.cu:31-35) calling mallocRW (RWMalloc.cuh:299-326). */
int pbm_malloc(pbm_handle* h, uint32_t n, uint64_t req_base, const uint64_t* d_sizes,
               uint64_t* d_ptrs, uint32_t* d_steps, const pbm_launch* launch, void* stream);

/* freeRW for n pointers (0 skipped).  Replaces free_kernel
 * (unitTests/unitTest4.cu:37-43) calling freeRW (RWMalloc.cuh:329-334). */
int pbm_free_ptrs(pbm_handle* h, uint32_t n, const uint64_t* d_ptrs, const pbm_launch* launch,
                  void* stream);

/* Host-side pageAddress / getPageID (Paging.cuh:30-52) over the pool's
 * single contiguous allocation (the reference scans 1 GiB groups). */
int pbm_page_address(pbm_handle* h, int32_t id, uint64_t* addr);
int pbm_page_id(pbm_handle* h, uint64_t addr, int32_t* id);

/* Copies `bytes` of page memory from byte `offset` of the pool to host
 * memory (synchronous on `stream`): for reading what device code wrote
 * into pages, such as the kernel-malloc headers. */
int pbm_read_pool(pbm_handle* h, uint64_t offset, uint64_t bytes, void* host, void* stream);

/* ---- kernel malloc ---------------------------------------------------
 * The MemoryManager classes of malloc.cuh:14-37 over a pool with backing
 * and a page size that is a multiple of 8: sub-page spaces carved from
 * runs of consecutive pages, with 8-byte chunk and space headers on the
 * pages (malloc.cu:45-57, 200-211; the device API documents the layouts).
 * The manager keeps a reference to the pool, which must outlive it. */
#define PBM_KMALLOC_SIMPLE 0  /* KernelMallocSimple: per-thread bump chunks, free is a no-op */
#define PBM_KMALLOC_COLLAB 1  /* KernelMallocCollab: a wave's small requests share a chunk  */

typedef struct pbm_kmalloc pbm_kmalloc;

/* Constructor (malloc.cu:76-84, 214-218).  max_threads: per-thread chunk
 * slots of KernelMallocSimple (0 = 1<<20; its constructor argument),
 * ignored for KernelMallocCollab. */
int pbm_kmalloc_create(pbm_handle* h, int32_t kind, uint32_t max_threads, pbm_kmalloc** out);
int pbm_kmalloc_destroy(pbm_kmalloc* k);

/* reset (malloc.cu:99-104, 222-224): every page of the pool free and, for
 * KernelMallocSimple, no current chunks. */
int pbm_kmalloc_reset(pbm_kmalloc* k, void* stream);

/* malloc(size) for n requests (d_sizes[r] bytes; d_ptrs[r] = address or
 * 0; d_steps[r] = getNPage steps).  KernelMallocSimple: thread t serves
 * requests t, t+T, ... from chunk slot t (T = min(n, max_threads); one
 * thread and slot 0 in serial mode).  KernelMallocCollab: a wave serves
 * each 64 consecutive requests together.  Replaces allocate_kernel
 * (unitTests/unitTestMalloc1.cu:27-34) calling memManager.malloc. */
int pbm_kmalloc_malloc(pbm_kmalloc* k, uint32_t n, uint64_t req_base, const uint32_t* d_sizes,
                       uint64_t* d_ptrs, uint32_t* d_steps, const pbm_launch* launch,
                       void* stream);

/* free(ptr) for n pointers (0 skipped).  Replaces free_kernel
 * (unitTests/unitTestMalloc1.cu:36-41).  KernelMallocSimple's free does
 * nothing, as in the reference (malloc.cu:178-181). */
int pbm_kmalloc_free(pbm_kmalloc* k, uint32_t n, const uint64_t* d_ptrs,
                     const pbm_launch* launch, void* stream);

/* KernelMallocSimple's current chunk (first page, -1 none) per slot, into
 * host_current[max_threads] when not NULL; *max_threads when not NULL. */
int pbm_kmalloc_current(pbm_kmalloc* k, int32_t* host_current, uint32_t* max_threads);

/* collectWarpData (CollabRW_BM.cuh:131-138) over a batch, as the
 * collectPageData driver gathers it (unitTests/collectPageData.cu:19-58):
 * CollabRW_BM getPage for n requests, one wave per 64 consecutive requests
 * (serial mode: one wave, in index order).  Wave w = request / 64 records
 * the free bits its probing lanes found in each step s < max_steps at
 * d_yields[w * max_steps + s] (later entries are left as they were) and
 * its step count at d_n_steps[w].  Needs a PBM_COLLAB_RW_BM pool; plain or
 * serial launches.  Granted pages stay allocated, as with pbm_get_pages. */
int pbm_collect_warp_data(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                          uint32_t* d_yields, uint32_t max_steps, uint32_t* d_n_steps,
                          const pbm_launch* launch, void* stream);

/* ---- kernel timing ----------------------------------------------------
 * For measuring the allocation kernels without launch overhead: the next
 * pbm_get_pages or pbm_free_pages call on h launches its main kernel with
 * hipExtLaunchKernelGGL, which records start_event and stop_event (hipEvent_t
 * handles, created with timing) at that kernel's own start and end, as a
 * kernel trace reports them.  Rank-batched, batched-walk, plain
 * RandomWalkBasic_BM / CollabRW_BM gets and bitmap frees honour it; every
 * call clears it, recorded or not.  Not for calls inside a graph capture.
 * No counterpart in the reference, which times launches with cudaEvents
 * around them (unitTests/metrics.h:52-70). */
int pbm_time_next_kernel(pbm_handle* h, void* start_event, void* stop_event);
/* hipEventCreate / hipEventDestroy / elapsed ms (waits for stop_event),
 * for callers without a HIP binding of their own. */
int pbm_event_create(void** event);
int pbm_event_destroy(void* event);
int pbm_event_elapsed_ms(void* start_event, void* stop_event, float* ms);

/* ---- observation ------------------------------------------------------ */

/* Free-page clusters: maximal runs of free pages, the page after the last
 * following the first (a fully free pool is one cluster).  Replaces the
 * cluster count of printNumPagesLeftClusteredRandomWalk_kernel
 * (parallelPage.cu:269-291), which reports 0 for a fully free pool. */
int pbm_count_clusters(pbm_handle* h, uint64_t* clusters, void* stream);

/* Samples the pool's free pages `count` times, gap_us microseconds apart,
 * from one workgroup on `stream` (run it on a stream of its own to watch
 * allocations on others).  d_samples (device, 2*count words) receives
 * (wall-clock tick, free pages) pairs.  Replaces monitorBufferRandomWalk
 * (parallelPage.cu:209-223), which prints from the device.  At most 60 s
 * per call. */
int pbm_monitor(pbm_handle* h, uint32_t gap_us, uint32_t count, unsigned long long* d_samples,
                void* stream);

/* Frequency of the device wall clock (wall_clock64), Hz. */
int pbm_wall_clock_hz(int device, unsigned long long* hz);

/* ---- graph-captured phases --------------------------------------------- */

#define PBM_PHASE_ALLOC 1u
#define PBM_PHASE_FREE  2u

typedef struct pbm_graph pbm_graph;

/* Captures a HIP graph of one allocation phase (pbm_get_pages with
 * get_launch), one free phase (pbm_free_pages of d_ids with free_launch),
 * or both in that order (phases = PBM_PHASE_ALLOC | PBM_PHASE_FREE).  The
 * arguments are frozen into the graph: every launch replays the same
 * request indices, so a replay of both phases returns the map to its state
 * before the replay.  No reference counterpart (its drivers launch kernels
 * one by one, unitTests/unitTest1.cu:64). */
int pbm_graph_create(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                     uint32_t* d_steps, const pbm_launch* get_launch,
                     const pbm_launch* free_launch, uint32_t phases, pbm_graph** out);

/* The same with n_steps (1 .. 1024) copies of the captured phases in one
 * graph, one after another: a launch runs n_steps steps.  No reference
 * counterpart. */
int pbm_graph_create_steps(pbm_handle* h, uint32_t n, uint64_t req_base, int32_t* d_ids,
                           uint32_t* d_steps, const pbm_launch* get_launch,
                           const pbm_launch* free_launch, uint32_t phases, uint32_t n_steps,
                           pbm_graph** out);
int pbm_graph_launch(pbm_graph* g, void* stream);
int pbm_graph_destroy(pbm_graph* g);

/* ---- serving loop (continuous batching) -------------------------------
 *
 * A device-resident request loop over one pool.  Sequences are submitted
 * with an arrival step and a target length in pages; every step
 *   1. admits waiting sequences in arrival order while a slot is free and
 *      the pool's free-page count, less the pages still owed to admitted
 *      sequences, covers the sequence's whole target (so an admitted
 *      sequence never fails to grow);
 *   2. grows every admitted sequence by one page with the pool strategy's
 *      device getPage, called from one thread per slot (wave-collective
 *      strategies serve the active lanes of each wave together), timing
 *      each call with wall_clock64;
 *   3. frees the page lists of the sequences that reached their target.
 * Arrivals that come while others run join at the next step; all admitted
 * sequences grow in the same kernel.  With a backed pool and
 * PBM_SERVE_PAYLOAD, page j of a sequence holds a 64-bit payload that
 * depends on (seed, the sequence's key, j) only; at the end its pages are read
 * back, checked and summed, so a sequence's (length, checksum) is the same
 * alone and in any mix.  No reference counterpart: the reference has no
 * serving loop; its device getPage/freePage are what each step calls. */

#define PBM_SERVE_PAYLOAD 1u  /* write, check and sum page payloads (backed pool)  */
#define PBM_SERVE_LATENCY 2u  /* keep per-allocation latencies (first lat_capacity) */

typedef struct pbm_serve_config {
    uint32_t max_slots;     /* sequences admitted at once, 1 .. 2^22                  */
    uint32_t max_pages;     /* page-list capacity of a slot; longer targets refused   */
    uint32_t max_seqs;      /* sequences that can be submitted over the server's life */
    uint32_t flags;         /* PBM_SERVE_*                                            */
    uint32_t lat_capacity;  /* latency samples kept (PBM_SERVE_LATENCY)               */
    uint32_t reserved;
    uint64_t seed;          /* payload seed                                           */
} pbm_serve_config;

typedef struct pbm_serve_stats {
    unsigned long long step;            /* steps run                                 */
    unsigned long long submitted;       /* sequences submitted                       */
    unsigned long long admitted;        /* sequences admitted                        */
    unsigned long long finished;        /* sequences that reached their target       */
    unsigned long long active;          /* admitted and not finished                 */
    unsigned long long reserved_pages;  /* pages still owed to active sequences      */
    unsigned long long allocs;          /* pages granted                             */
    unsigned long long frees;           /* pages returned                            */
    unsigned long long alloc_failures;  /* getPage calls that returned -1 (stays 0)  */
    unsigned long long payload_errors;  /* payloads found changed at the end         */
    unsigned long long lat_samples;     /* latencies recorded (may exceed capacity)  */
} pbm_serve_stats;

typedef struct pbm_server pbm_server;

/* Creates a server on pool h (any strategy).  The pool must outlive it. */
int pbm_serve_create(pbm_handle* h, const pbm_serve_config* cfg, pbm_server** out);
/* Waits for the device, then gives the pool back what the server holds:
 * the pages of its running sequences and the pages promised to them at
 * admission and not yet granted.  The pool's map and reservation counter
 * are then as if the server had never run. */
int pbm_serve_destroy(pbm_server* s);

/* Appends n sequences (host arrays): arrival_step[i] (non-decreasing, also
 * across calls), target[i] in [1, max_pages] and key[i], the request's
 * own id that its payloads derive from (NULL: the sequence id).  Sequence
 * ids count up from 0 in submission order.  Stream-ordered and without
 * host synchronisation: the arrays are copied into the server's pinned
 * mirror, which one kernel on the stream reads from the device and makes
 * visible to admission after the work already on the stream, so sequences
 * can be submitted while steps run; a sequence
 * is admissible from its arrival step or from the first step after its
 * submission, whichever is later.  The host arrays may be reused when the
 * call returns. */
int pbm_serve_submit(pbm_server* s, uint32_t n, const uint32_t* arrival_step,
                     const uint32_t* target, const uint32_t* key, void* stream);

/* Per sequence first .. first+n-1: wall-clock ticks (pbm_serve_clock_hz)
 * from its arrival -- the later of its submission reaching the device and
 * the start of its arrival step -- to the start of the step that admitted
 * it; 0xFFFFFFFF for a sequence not admitted (or whose arrival step lies
 * more than 65,536 steps before its admission). */
int pbm_serve_read_waits(pbm_server* s, uint32_t first, uint32_t n, uint32_t* ticks,
                         void* stream);

/* Runs n_steps steps (three kernels each) on the stream.  Admission
 * reserves an admitted sequence's whole target in the pool's reservation
 * counter, which then reads free pages less the pages owed to running
 * sequences; the grows take pages already reserved. */
int pbm_serve_run(pbm_server* s, uint32_t n_steps, void* stream);

/* Runs n_steps steps in one kernel: the device-resident loop.  One
 * 1,024-thread workgroup per 1,024 slots and one that admits; each step
 * the admitting workgroup admits while every slot workgroup grows the
 * slots it already holds and frees the pages of the sequences that
 * finishes, then, once the admission is visible, grows the slots admitted
 * now and frees what that finishes; the phases are ordered by device-side
 * counters (no kernel boundary between steps).  The same schedule and
 * results as pbm_serve_run.  Admission waits for every slot workgroup, so
 * all must be resident at once: PBM_ERR_UNSUPPORTED when the device cannot hold them,
 * and a wait that gives up (2 s, another kernel holding the device) is
 * counted by pbm_serve_loop_timeouts and leaves the results invalid.
 * Capturable in a HIP graph.  Sequences submitted before the call are seen
 * by its steps; later ones by the next call.  Like every call on a server,
 * it must be ordered with the server's other calls (one stream, or
 * synchronised): two runs of one server at once race on its slot table. */
int pbm_serve_run_loop(pbm_server* s, uint32_t n_steps, void* stream);

/* Captures one pbm_serve_run_loop call of n_steps steps as a HIP graph
 * (every launch runs the next n_steps steps). */
int pbm_serve_graph_loop(pbm_server* s, uint32_t n_steps, pbm_graph** out);

/* Waits of device-resident loop runs that gave up (stays 0). */
int pbm_serve_loop_timeouts(pbm_server* s, unsigned long long* out, void* stream);

/* Runs n_steps steps (1 .. 4096) as pbm_serve_run does, each kernel
 * recording events at its own start and end, waits for them and writes the
 * mean duration of each kernel in ms: kernel_ms[0] admit, [1] grow, [2]
 * release.  A measurement helper; no reference counterpart. */
int pbm_serve_run_timed(pbm_server* s, uint32_t n_steps, float* kernel_ms, void* stream);

/* Captures n_steps steps as one HIP graph (launch with pbm_graph_launch;
 * every launch runs the next n_steps steps). */
int pbm_serve_graph(pbm_server* s, uint32_t n_steps, pbm_graph** out);

int pbm_serve_read_stats(pbm_server* s, pbm_serve_stats* out, void* stream);

/* Per-sequence results for ids first .. first+n-1 (host arrays, any may be
 * NULL): step admitted, step finished (0 = not yet), pages held at the end
 * and the sum of their payloads (0 without PBM_SERVE_PAYLOAD). */
int pbm_serve_read_results(pbm_server* s, uint32_t first, uint32_t n, uint32_t* admit_step,
                           uint32_t* finish_step, uint32_t* length,
                           unsigned long long* checksum, void* stream);

/* Copies up to n recorded latencies (wall-clock ticks) to host; *got =
 * how many were copied. */
int pbm_serve_read_latency(pbm_server* s, uint32_t* ticks, uint32_t n, uint32_t* got,
                           void* stream);

/* Frequency of the wall clock the latencies are counted in, Hz. */
int pbm_serve_clock_hz(pbm_server* s, unsigned long long* hz);

/* ---- buffer cache with CLOCK replacement ------------------------------
 *
 * A page cache of n_logical pages in `capacity` slots over pool h: a slot
 * holds one frame (a pool page got with the pool strategy's device
 * getPage, CRW_BM's clustered walk for a CRW_BM pool) and a reference bit.
 * References are taken `batch` at a time: the pages a batch hits get their
 * bit set; the first reference to an absent page is a miss and later ones
 * in the batch are hits; misses fill empty slots first and then replace
 * victims chosen by a CLOCK sweep from the arm over the slots occupied when
 * the batch began (set bits are cleared and passed; if one pass has too few
 * clear slots, the next pass takes slots in ring order); the sweep stops
 * at the last victim; loaded pages start with the bit set.  A batch of 1
 * is the textbook second-chance CLOCK.  A cache's calls must be ordered
 * (one stream, or synchronized): each batch's kernels hand tickets and
 * look-back words to the next.  No reference counterpart: the reference's
 * clocks are allocators (parallelPage.cu:470-618) and it has no
 * replacement. */

typedef struct pbm_cache_config {
    uint32_t n_logical;       /* logical pages, ids 0 .. n_logical-1               */
    uint32_t capacity;        /* slots, 1 .. pool pages                           */
    uint32_t batch;           /* references per batch, 1 .. capacity              */
    uint32_t reserved;
    uint64_t log_capacity;    /* evicted pages recorded in order (0 = none)       */
} pbm_cache_config;

typedef struct pbm_cache_stats {
    unsigned long long hits;
    unsigned long long misses;
    unsigned long long evictions;
    unsigned long long resident;        /* pages held now                         */
    unsigned long long arm;             /* ring position of the clock arm         */
    unsigned long long batches;
    unsigned long long alloc_failures;  /* frames the pool could not give (0)     */
    unsigned long long scan_timeouts;   /* look-back waits given up (0)           */
    unsigned long long invalid_refs;    /* references to ids >= n_logical, skipped */
} pbm_cache_stats;

typedef struct pbm_cache pbm_cache;

int pbm_cache_create(pbm_handle* h, const pbm_cache_config* cfg, pbm_cache** out);
int pbm_cache_destroy(pbm_cache* k);

/* Runs n references (device array of logical page ids) through the cache:
 * three kernels per batch, and one that closes the last batch.  A
 * reference to an id outside 0 .. n_logical-1 is neither a hit nor a miss;
 * it is skipped and counted (invalid_refs). */
int pbm_cache_run(pbm_cache* k, const int32_t* d_refs, uint64_t n, void* stream);

/* pbm_cache_run of at most 4,096 batches with each kernel recording events
 * at its own start and end; waits for them and writes the mean duration in
 * ms of kernel_ms[0] hits, [1] plan, [2] load.  A measurement helper; no
 * reference counterpart. */
int pbm_cache_run_timed(pbm_cache* k, const int32_t* d_refs, uint64_t n, float* kernel_ms,
                        void* stream);

int pbm_cache_read_stats(pbm_cache* k, pbm_cache_stats* out, void* stream);

/* Evicted pages first .. first+n-1 of the log, in eviction order. */
int pbm_cache_read_evictions(pbm_cache* k, uint64_t first, uint32_t n, int32_t* host,
                             void* stream);

/* Per slot (capacity entries each, any may be NULL): page held (-1 none),
 * its frame, its reference bit. */
int pbm_cache_read_slots(pbm_cache* k, int32_t* owner, int32_t* frame, uint32_t* refbit,
                         void* stream);

/* count references of a Zipf(s) trace over n_logical pages, starting at
 * index first, into device memory: reference i is the first k with
 * cdf[k] > u_i * cdf[n-1], cdf[k] = sum_{j<=k} (j+1)^-s summed in double in
 * index order, u_i = (splitmix64 hash of (seed, i)) >> 11, over 2^53. */
int pbm_zipf_trace(uint32_t n_logical, double s, uint64_t seed, uint64_t first, uint64_t count,
                   int32_t* d_out, int device, void* stream);

/* pbm_zipf_trace with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The CDF goes
 * to a static device array of 4,194,304 doubles (32 MiB, one per device,
 * held until the process ends) instead of a hipMalloc'd table: n_logical
 * above 4,194,304 is rejected with PBM_ERR_RANGE.  The output is the same
 * as pbm_zipf_trace's.  Calls that overlap on one device race on the array
 * (calls from host threads are serialized; the call returns after its
 * kernel). */
int pbm_zipf_trace_noalloc(uint32_t n_logical, double s, uint64_t seed, uint64_t first,
                           uint64_t count, int32_t* d_out, int device, void* stream);

/* ---- autotuning ------------------------------------------------------- */

#define PBM_TUNE_QUICK     1u  /* 3 timed cycles per candidate instead of 9            */
#define PBM_TUNE_WALK_ONLY 2u  /* random-walk homes only (no PBM_LAUNCH_RANK)          */
#define PBM_TUNE_RANK_ONLY 4u  /* rank mode only                                       */
#define PBM_TUNE_NO_PERSISTENT 8u  /* bursts: phase kernels only                         */

typedef struct pbm_tune_result {
    pbm_launch get_launch;   /* fastest batched launch for pbm_get_pages              */
    pbm_launch free_launch;  /* fastest batched launch for pbm_free_pages after it    */
    float      ms_alloc;     /* median alloc kernel of get_launch, milliseconds       */
    float      ms_free;      /* median free kernel of free_launch, milliseconds       */
    uint32_t   candidates;   /* launch shapes timed                                   */
    uint32_t   reserved;
} pbm_tune_result;

/* Times alloc+free cycles of n requests (device buffer d_ids of n ids)
 * over the batched launch space and returns the fastest alloc launch --
 * the rank scan, or walk homes over window width (64/128/256 words), walk
 * budget (4/16/64 jumps) and waves per CU (1/2/4/8/16) -- and, timed after
 * it, the fastest free launch (waves per CU 1/2/4/8/16/32).  Times are the
 * kernels' own (events at each kernel's start and end), medians of 9
 * cycles (3 with PBM_TUNE_QUICK).  Every cycle returns its pages, so the
 * pool ends as it started; no other call may overlap.  No reference
 * counterpart (its sweeps are hard-coded, unitTests/unitTest1.cu:14-18). */
int pbm_autotune(pbm_handle* h, uint32_t n, int32_t* d_ids, uint32_t flags,
                 pbm_tune_result* out, void* stream);

/* ---- bursts of mallocRW / freeRW --------------------------------------
 *
 * A trace of bursts with uneven request counts and sizes.  Burst b frees
 * free_index[free_offset[b] .. free_offset[b+1]) (requests of earlier
 * bursts, by global request index) and then mallocRWs requests
 * alloc_offset[b] .. alloc_offset[b+1] of sizes[], wave-collective over
 * each 64 consecutive requests of the burst.
 *
 * Two ways to run a range of bursts, both capturable in a HIP graph:
 *  - phase kernels: one free kernel and one alloc kernel per burst (plain,
 *    serial or a batched grid);
 *  - PBM_LAUNCH_PERSISTENT: one kernel for the whole range.  Each burst is
 *    cut into free items (at most 64 frees, whole packed runs) and alloc
 *    items (one 64-request chunk); waves take items in order from a ticket.
 *    A free waits until the pointer it frees is published.  An alloc item
 *    of burst b first waits until an earlier burst has finished (every
 *    item): the earliest one after which the bursts up to b ask for at most
 *    half the pool's pages, counted from the trace when it is created.  It
 *    then runs ahead of its own burst's frees while a quarter of the pool
 *    is unreserved; otherwise, or with PBM_LAUNCH_ORDERED, it waits until
 *    every free item of its burst is done, and one that ran ahead and found
 *    no run for a request waits for them and tries again.  Waits are bounded:
 *    one that gives up is counted (pbm_bursts_wait_timeouts) and the run's
 *    results are then not valid.
 *
 * Requests packed together by mallocRW are the small requests (size + 4,
 * rounded up to 4, below the page size) of one 64-request chunk of a
 * burst; freeing the leader returns the whole run (RWMalloc.cuh:329-334).
 * pbm_bursts_create therefore rejects a trace that frees a request twice or
 * frees one packed run's requests in more than one burst.  No reference
 * counterpart: the reference drivers launch one malloc_kernel / free_kernel
 * per test (unitTests/unitTest3.cu:31-35, unitTest4.cu:37-43). */

typedef struct pbm_burst_trace {
    uint32_t n_bursts;
    uint32_t reserved;
    const uint32_t* alloc_offset;  /* n_bursts + 1 entries, from 0                   */
    const uint32_t* free_offset;   /* n_bursts + 1 entries, from 0                   */
    const uint64_t* sizes;         /* alloc_offset[n_bursts] request sizes in bytes  */
    const uint32_t* free_index;    /* free_offset[n_bursts] request indices          */
} pbm_burst_trace;

typedef struct pbm_bursts pbm_bursts;

/* Uploads a trace (host arrays) for pool h (backed, bitmap strategy). */
int pbm_bursts_create(pbm_handle* h, const pbm_burst_trace* trace, pbm_bursts** out);
int pbm_bursts_destroy(pbm_bursts* b);

/* Runs bursts first .. first+count-1. */
int pbm_bursts_run(pbm_bursts* b, uint32_t first, uint32_t count, const pbm_launch* launch,
                   void* stream);

/* Captures bursts first .. first+count-1 as one HIP graph. */
int pbm_bursts_graph(pbm_bursts* b, uint32_t first, uint32_t count, const pbm_launch* launch,
                     pbm_graph** out);

/* Pointers mallocRW returned for requests first .. first+n-1 (0 = failed
 * or not yet run; 1 = not yet run inside a persistent run whose waits gave
 * up). */
int pbm_bursts_read_ptrs(pbm_bursts* b, uint32_t first, uint32_t n, uint64_t* host,
                         void* stream);

/* Forgets every returned pointer and the wait time-out count (the pool is
 * not touched). */
int pbm_bursts_reset(pbm_bursts* b, void* stream);

/* Dependency waits of persistent runs that gave up, since the last reset
 * (stays 0). */
int pbm_bursts_wait_timeouts(pbm_bursts* b, unsigned long long* out, void* stream);

/* Times the first `count` bursts from the pool's current state over the
 * batched launch space (waves per CU 1/2/4/8/16, walk budget 4/16/64,
 * getNPage window 64/128/256 words, phase kernels or one persistent
 * kernel; phase kernels only with PBM_TUNE_NO_PERSISTENT), times the three
 * fastest again over the first min(n_bursts, 4 * count) bursts when that
 * is more, and returns the fastest of those in get_launch (ms_alloc = its
 * time over the first count bursts, ms_free = its time over the longer run,
 * 0 when there was none).  The pool's map and the pointer table are put
 * back after every candidate and at the end, so the pool must hold only
 * what the trace's first burst expects. */
int pbm_bursts_autotune(pbm_bursts* b, uint32_t count, uint32_t flags, pbm_tune_result* out,
                        void* stream);


/* ---- debug: owner shadow ---------------------------------------------- */

/* What the owner shadow has seen since pbm_shadow_enable. */
typedef struct pbm_shadow_report {
    uint64_t double_grants; /* grants of a page the shadow did not hold as free   */
    uint64_t bad_frees;     /* frees of a page it held as free, or past n_pages   */
    int64_t  grant_page;    /* one double-granted page (-1 = none)                */
    uint64_t grant_owner;   /* the request number that was granted it            */
    uint64_t grant_holder;  /* the request number holding it then (PBM_SHADOW_HELD:
                               held before the shadow last read the map;
                               PBM_SHADOW_FREE: a page past n_pages)            */
    int64_t  free_page;     /* one page of a bad free (-1 = none)                 */
} pbm_shadow_report;

#define PBM_SHADOW_FREE 0xFFFFFFFFFFFFFFFFull  /* owner word of a free page        */
#define PBM_SHADOW_HELD 0xFFFFFFFFFFFFFFFEull  /* owner word of a page held before
                                                  the shadow last read the map     */

/* Debug aid with no reference counterpart (the reference checks only the
 * final map; a double grant shows there only as a lost page).  on != 0:
 * keeps one owner word per page (8 bytes x n_pages of device memory), set
 * from the map, and clears the report; on = 0 frees it.  While on, every
 * successful pbm_get_pages, pbm_free_pages, pbm_get_n_pages,
 * pbm_get_x_pages and pbm_free_n_pages launches one more kernel on the same
 * stream after its own: each granted page moves from free to its request
 * number (req_base + i) by compare-and-swap, so a page granted twice is
 * reported at the call that granted it the second time; each freed page
 * moves back to free.  pbm_reset, pbm_prefill_bernoulli and
 * pbm_restore_bitmap set the owners from the new map.  Other calls that
 * grant or free pages (mallocRW, bursts, the cache, the server) do not
 * update it: call pbm_shadow_enable(h, 1) again after them, or record
 * their pages with pbm_shadow_note.  Calls on several streams at once see
 * the owners in no defined order.  Bitmap strategies only. */
int pbm_shadow_enable(pbm_handle* h, int on, void* stream);

/* Records the grants (freeing = 0) or frees (freeing = 1) of n requests
 * (d_ids[i], d_counts[i] pages; d_counts null = one page each; ids < 0 or
 * counts 0 skipped) made outside the entry points above, for callers with
 * kernels of their own. */
int pbm_shadow_note(pbm_handle* h, uint32_t n, uint64_t req_base, const int32_t* d_ids,
                    const uint32_t* d_counts, int freeing, void* stream);

/* Reads the report (synchronizes the stream). */
int pbm_shadow_read(pbm_handle* h, pbm_shadow_report* out, void* stream);

/* Copies the owner words (`words` must equal n_pages) to host memory. */
int pbm_shadow_owners(pbm_handle* h, uint64_t* host, size_t words, void* stream);

/* ---- measurement ------------------------------------------------------ */

/* Device-scope 64-bit atomicOr throughput on `words` distinct words
 * (uncontended within a launch, every word hit `rounds` times), the
 * ceiling the allocators' claim/free atomics run against.  Writes
 * atomics/s to *ops_per_s.  Measurement helper; no reference counterpart. */
int pbm_atomic_peak(uint32_t words, uint32_t rounds, int device, double* ops_per_s);

/* Human-readable message of the last error on this thread. */
const char* pbm_last_error(void);

/* Library version string. */
const char* pbm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PBM_H */
