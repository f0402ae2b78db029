/*
 * pbm_oracle.h — serial CPU restatement of the reference's page-allocation
 * policies.  TEST INFRASTRUCTURE ONLY: linked by tests/, by the smoke check
 * and by the benchmark's CPU-baseline leg as the checker.  The product
 * library never links, loads or calls it.
 *
 * Every walk runs one request at a time in request-index order (the
 * schedule of one thread, or one wave for the wave-collective policies);
 * the device library reproduces that schedule in PBM_MODE_SERIAL, where its
 * results must equal these bit for bit.
 */
#ifndef PBM_ORACLE_H
#define PBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* random numbers */
int32_t  orc_lcg(int32_t x);
uint64_t orc_mix64(uint64_t z);
int32_t  orc_request_seed(uint64_t seed, uint64_t r);
uint64_t orc_page_hash(uint64_t seed, uint64_t i);

/* page map */
uint32_t orc_n_words(uint32_t n_pages);
void     orc_clear(uint64_t* bm, uint32_t n_pages);
void     orc_prefill_bernoulli(uint64_t* bm, uint32_t n_pages, double fill, uint64_t seed);
int64_t  orc_count_free(const uint64_t* bm, uint32_t n_pages);

/* Plain RandomWalkBasic_BM, serial.  free_count is the reservation counter
 * (in/out).  Returns the number of grants. */
uint32_t orc_rwbm_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                         uint64_t req_base, uint64_t seed, uint32_t budget,
                         int32_t* ids, uint32_t* steps);

/* The same policy on `threads` host threads sharing the map and the
 * reservation counter (atomic claims and counts): blocks of 64 consecutive
 * requests are dealt round-robin, each reserving with one atomic; with
 * free_after, once every thread is done, each frees its grants.  With one
 * thread the allocation equals orc_rwbm_serial.  Returns the number of
 * grants.
 * A CPU baseline for timing, not a schedule the device reproduces. */
uint32_t orc_rwbm_threads(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                          uint64_t req_base, uint64_t seed, uint32_t budget, uint32_t threads,
                          int free_after, int32_t* ids, uint32_t* steps);

/* CollabRW_BM device getPage called by one wave over chunks of 64
 * requests (lanes of the last chunk beyond n inactive). order: 0 packed,
 * 1 reference round-robin. */
uint32_t orc_collab_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                           uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                           int32_t* ids, uint32_t* steps);
/* The same with collectWarpData (CollabRW_BM.cuh:131-138): per wave, the
 * free bits its probing lanes read in each step. */
uint32_t orc_collab_yields(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                           uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                           int32_t* ids, uint32_t* yields, uint32_t max_steps,
                           uint32_t* n_steps);

/* Batched persistent allocation with one wave: phase reservation of
 * min(n, free) for the first requests by index, windows of 64*J words kept
 * across chunks while they have free bits. */
uint32_t orc_batched_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                            uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                            uint32_t cluster_words, int32_t* ids, uint32_t* steps);

/* Batched allocation by rank: request r takes the r-th free page. */
uint32_t orc_batched_rank(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                          int32_t* ids, uint32_t* steps);

/* CRW_BM for one thread (hint in/out, -1 = none). */
uint32_t orc_crw_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                        uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* hint,
                        int32_t* ids, uint32_t* steps);

/* SingleClock with a 64-bit arm (in/out). */
uint32_t orc_single_clock_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                                 uint64_t* arm, uint32_t budget, int32_t* ids, uint32_t* steps);

/* DistClock: initial arm of a slot, and one arm's serial requests. */
uint32_t orc_dist_clock_arm(uint64_t seed, uint32_t slot, uint32_t n_pages);
uint32_t orc_dist_clock_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                               uint32_t* arm, uint32_t budget, int32_t* ids, uint32_t* steps);

/* LinkedList FIFO over a ring of n_pages slots. */
uint32_t orc_list_build(const uint64_t* bm, uint32_t n_pages, int32_t* q);
void orc_list_get(uint64_t* bm, uint32_t n_pages, int32_t* q, uint64_t* head, uint64_t* tail,
                  uint32_t n, int32_t* ids);
void orc_list_put(uint64_t* bm, uint32_t n_pages, int32_t* q, uint64_t* head, uint64_t* tail,
                  uint32_t n, const int32_t* ids);

/* RWMalloc getPage, getNPage, releasePages and mallocRW on one wave. */
uint32_t orc_rwmalloc_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                             uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                             int32_t* ids, uint32_t* steps);
uint32_t orc_getnpage_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                             const uint32_t* counts, uint64_t req_base, uint64_t seed,
                             uint32_t budget, int32_t* ids, uint32_t* steps);
/* getNPage with windows of `spans` x 64 words (1..4); _serial is spans = 1. */
uint32_t orc_getnpage_spans(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                            const uint32_t* counts, uint64_t req_base, uint64_t seed,
                            uint32_t budget, uint32_t spans, int32_t* ids, uint32_t* steps);
uint32_t orc_free_n(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                    const int32_t* ids, const uint32_t* counts);
/* Ouroboros comparator (Ouroboros.cuh:22-62): pages per chunk, queue
 * reset, getPage and freePage, one request at a time; st = {carved, fresh
 * pages, head, tail, unpromised queued pages}. */
uint32_t orc_ouroboros_chunk_pages(uint32_t page_size);
uint32_t orc_queue_reset(const uint64_t* bm, uint32_t n_pages, uint32_t* fresh, int64_t* st);
uint32_t orc_queue_get(uint64_t* bm, uint32_t n_pages, int64_t* free_count,
                       const uint32_t* fresh, int32_t* ring, int64_t* st, uint32_t chunk_pages,
                       uint32_t n, int32_t* ids, uint32_t* steps);
uint32_t orc_queue_put(uint64_t* bm, uint32_t n_pages, int64_t* free_count, int32_t* ring,
                       int64_t* st, uint32_t n, const int32_t* ids);
/* KernelMallocCollab malloc / free (malloc.cu:231-320) and
 * KernelMallocSimple malloc (malloc.cu:129-175) on a byte image `mem` of
 * the pool; `written` (optional) marks the bytes written; off[r] = pool
 * offset of the returned pointer or ~0. */
uint32_t orc_kmalloc_collab(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                            int64_t* free_count, uint8_t* mem, uint8_t* written, uint32_t n,
                            const uint32_t* sizes, uint64_t req_base, uint64_t seed,
                            uint32_t budget, uint64_t* off, uint32_t* steps);
uint32_t orc_kfree_collab(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint8_t* mem,
                          uint32_t n, const uint64_t* off);
uint32_t orc_kmalloc_simple(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                            int64_t* free_count, uint8_t* mem, uint8_t* written, uint32_t n,
                            const uint32_t* sizes, uint64_t req_base, uint64_t seed,
                            uint32_t budget, int32_t* current, uint64_t* off, uint32_t* steps);
uint32_t orc_malloc_serial(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                           int64_t* free_count, uint32_t n, const uint64_t* sizes,
                           uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* page,
                           uint64_t* offset, int32_t* header, uint32_t* steps);
uint32_t orc_malloc_spans(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                          int64_t* free_count, uint32_t n, const uint64_t* sizes,
                          uint64_t req_base, uint64_t seed, uint32_t budget, uint32_t spans,
                          int32_t* page, uint64_t* offset, int32_t* header, uint32_t* steps);

/* freePage for every id >= 0; returns pages actually returned (bits that
 * were set).  free_count grows by that number. */
uint32_t orc_free(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                  const int32_t* ids);

/* Checks of one alloc phase against the page map before it (`before`) and
 * after it (`after`): every granted id unique, in range, free before and
 * used after, and after == before | granted bits.  Returns 0 when all hold,
 * else a positive code naming the first violated property. */
int orc_check_alloc(const uint64_t* before, const uint64_t* after, uint32_t n_pages,
                    uint32_t n, const int32_t* ids, uint32_t* n_granted);

/* Expected mean probes of n serial single-page grants on a pool of T pages
 * with A free, words of W bits (theoretical_models.py:34-41 restated). */
double orc_rwbm_expected_steps(double T, double A, double n, int W);


/* Serving loop: payload of page j of sequence seq, a sequence's checksum,
 * and the admission / finish schedule. */
uint64_t orc_serve_payload(uint64_t seed, uint32_t seq, uint32_t j);
uint64_t orc_serve_checksum(uint64_t seed, uint32_t seq, uint32_t len);
void orc_serve_sim(int64_t free_pages, uint32_t max_slots, uint32_t n_seqs,
                   const uint32_t* arrival, const uint32_t* target, uint32_t n_steps,
                   uint32_t* admit_step, uint32_t* finish_step, uint64_t* totals);


/* Buffer cache: Zipf trace and the batched CLOCK schedule. */
void orc_zipf_cdf(uint32_t n, double s, double* cdf);
int32_t orc_zipf_sample(const double* cdf, uint32_t n, uint64_t seed, uint64_t i);
void orc_zipf_trace(uint32_t n, double s, uint64_t seed, uint64_t first, uint64_t count,
                    int32_t* out);
/* A reference outside 0 .. n_logical-1 is skipped (neither hit nor miss). */
void orc_clock_run(uint32_t n_logical, uint32_t capacity, uint32_t batch, const int32_t* refs,
                   uint64_t n_refs, int32_t* slot_of, int32_t* owner, uint8_t* refbit,
                   uint32_t* arm_res, int32_t* evict_log, uint64_t log_cap, uint64_t* stats);
/* Textbook second-chance CLOCK, one reference at a time, in time linear in
 * the references (the batch rule of orc_clock_run at batch 1). */
void orc_clock_textbook(uint32_t n_logical, uint32_t capacity, const int32_t* refs,
                        uint64_t n_refs, int32_t* slot_of, int32_t* owner, uint8_t* refbit,
                        uint32_t* arm_res, int32_t* evict_log, uint64_t log_cap,
                        uint64_t* stats);

/* RandomWalkBasic (single-page walk) serial, and its closed form RW_TAS. */
uint32_t orc_rwbasic_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                            uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* ids,
                            uint32_t* steps);
/* getPageClusteredRandomWalk (parallelPage.cu:200-218), one thread with
 * hint *hint. */
uint32_t orc_crw_basic_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                              uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* hint,
                              int32_t* ids, uint32_t* steps);
/* getPageCollabRW (parallelPage.cu:642-680), one wave per chunk of 64
 * requests. */
uint32_t orc_collab_basic_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                                 uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* ids,
                                 uint32_t* steps);
double orc_rw_tas(double T, double A, double n);

uint32_t orc_getx_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                         const uint32_t* counts, uint64_t req_base, uint64_t seed,
                         uint32_t budget, int32_t* ids, uint32_t* steps);

uint64_t orc_count_clusters(const uint64_t* bm, uint32_t n_pages);

#ifdef __cplusplus
}
#endif

#endif
