/*
 * pbm_oracle.c — serial CPU restatement of the reference's page-allocation
 * policies.  TEST INFRASTRUCTURE ONLY (see pbm_oracle.h): the product
 * library never links, loads or calls this file.
 *
 * Pinning status of each function (how tests/ check it independently):
 *   orc_lcg               glibc random_r TYPE_0 state update (non-negative
 *                         seeds) and numpy int64 fmod (negative seeds)
 *   orc_mix64             published splitmix64 constants; bijectivity via
 *                         its inverse; avalanche statistics
 *   orc_prefill_bernoulli popcount within binomial bounds of f*T
 *   orc_rwbm_serial       hand-computed known answers on tiny pools,
 *                         invariants, mean steps vs the closed form
 *   orc_rwbm_threads      one thread equals orc_rwbm_serial; several keep
 *                         grants unique and counted, and frees restore the map
 *   orc_collab_serial     hand-built window known answers (both orders),
 *                         invariants, brute-force set property
 *   orc_batched_serial    same as collab plus phase admission known answer
 *   orc_check_alloc       planted violations (duplicate, not-free, lost bit)
 */
#include "pbm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
#define FULL (~(u64)0)
#define GOLDEN 0x9E3779B97F4A7C15ull
#define PREFILL_SALT 0x243F6A8885A308D3ull
#define WAVE_SALT 0x13198A2E03707344ull

/* ------------------------------------------------------------ random */

/* RNG_LCG.cuh:6-12 — (1103515245*x + 12345) % (1<<31) in long long with
 * C truncating modulo (LCG_M expands to 1<<31 without parentheses; the
 * sign of a C remainder follows the dividend, so the divisor's sign does
 * not matter). */
int32_t orc_lcg(int32_t x) {
    long long v = 1103515245LL * (long long)x + 12345LL;
    return (int32_t)(v % 2147483648LL);
}

/* splitmix64 output function (Steele, Lea & Flood, OOPSLA 2014). */
uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Seed of request r: high 32 bits of splitmix64 output r+1 from state
 * `seed`.  Stands in for (tid<<15)+(int16)clock()
 * (RandomWalkBasic_BM.cuh:62-63), which is not reproducible. */
int32_t orc_request_seed(uint64_t seed, uint64_t r) {
    return (int32_t)(uint32_t)(orc_mix64(seed + (r + 1) * GOLDEN) >> 32);
}

/* High-bit reduction of an LCG state to [0, L) used by the batched
 * allocator's window jumps (the reference's (unsigned)state % L keeps the
 * low bits, which cycle with period L for a power-of-two L). */
static uint32_t lcg_to_range(int32_t state, uint32_t L) {
    return (uint32_t)(((u64)((uint32_t)state & 0x7FFFFFFFu) * L) >> 31);
}

uint64_t orc_page_hash(uint64_t seed, uint64_t i) {
    return orc_mix64(orc_mix64(seed ^ PREFILL_SALT) + (i + 1) * GOLDEN);
}

/* ------------------------------------------------------------ page map */

uint32_t orc_n_words(uint32_t n_pages) { return (n_pages + 63u) / 64u; }

static u64 pad_mask(uint32_t n_pages, uint32_t word) {
    u64 first = (u64)word * 64u;
    if (first + 64u <= n_pages) return 0;
    uint32_t valid = (uint32_t)(n_pages - first);
    return ~(((u64)1 << valid) - 1u);
}

/* resetMemoryManager (RandomWalkBasic_BM.cuh:104-106): all free.  Padding
 * bits of the last word stay set. */
void orc_clear(uint64_t* bm, uint32_t n_pages) {
    uint32_t L = orc_n_words(n_pages);
    for (uint32_t i = 0; i < L; ++i) bm[i] = pad_mask(n_pages, i);
}

void orc_prefill_bernoulli(uint64_t* bm, uint32_t n_pages, double fill, uint64_t seed) {
    uint32_t L = orc_n_words(n_pages);
    int all = fill >= 1.0;
    u64 thr = (fill > 0.0 && fill < 1.0) ? (u64)ldexp(fill, 64) : 0;
    for (uint32_t i = 0; i < L; ++i) {
        u64 w = 0;
        for (uint32_t b = 0; b < 64; ++b) {
            u64 p = (u64)i * 64u + b;
            int used = p >= n_pages ? 1 : (all || orc_page_hash(seed, p) < thr);
            w |= (u64)used << b;
        }
        bm[i] = w;
    }
}

/* getFreePagePercentage numerator (RandomWalkBasic_BM.cuh:109-145). */
int64_t orc_count_free(const uint64_t* bm, uint32_t n_pages) {
    uint32_t L = orc_n_words(n_pages);
    int64_t f = 0;
    for (uint32_t i = 0; i < L; ++i) f += 64 - __builtin_popcountll(bm[i] | pad_mask(n_pages, i));
    return f;
}

/* ------------------------------------------------------------ RW_BM */

/* Inner loop of RandomWalkBasic_BM getPage (:72-85): while the word is
 * not full, set its lowest 0-bit; success when that bit was 0. */
static int32_t claim_lowest(uint64_t* bm, uint32_t g) {
    while (bm[g] != FULL) {
        u64 v = bm[g];
        int b = __builtin_ctzll(~v);
        u64 bit = (u64)1 << b;
        u64 old = bm[g];
        bm[g] |= bit;
        if (!(old & bit)) return (int32_t)(g * 64u + (uint32_t)b);
    }
    return -1;
}

/* One reserved walk: `budget` random jumps (RandomWalkBasic_BM.cuh:67-71,
 * step = one jump, :68), page id g*64+b, then a linear sweep of at most
 * four passes starting after the last jumped word. */
static int32_t rwbm_walk(uint64_t* bm, uint32_t L, int32_t* seed, uint32_t budget,
                         uint32_t* steps) {
    uint32_t g = 0;
    for (uint32_t j = 0; j < budget; ++j) {
        (*steps)++;
        *seed = orc_lcg(*seed);
        g = (uint32_t)*seed % L;
        int32_t id = claim_lowest(bm, g);
        if (id >= 0) return id;
    }
    for (u64 k = 0; k < 4ull * L; ++k) {
        g = (g + 1u == L) ? 0u : g + 1u;
        (*steps)++;
        int32_t id = claim_lowest(bm, g);
        if (id >= 0) return id;
    }
    return -1;
}

uint32_t orc_rwbm_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                         uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* ids,
                         uint32_t* steps) {
    uint32_t L = orc_n_words(n_pages), granted = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t s = 0;
        int32_t id = -1;
        if (*free_count > 0) {
            (*free_count)--;
            int32_t sd = orc_request_seed(seed, req_base + r);
            id = rwbm_walk(bm, L, &sd, budget, &s);
            if (id < 0) (*free_count)++;
        }
        if (id >= 0) granted++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return granted;
}

/* ------------------------------------------- RW_BM on host threads */

/* claim_lowest with atomic reads and claims (the map is shared). */
static int32_t claim_lowest_atomic(uint64_t* bm, uint32_t g) {
    for (;;) {
        const u64 v = __atomic_load_n(&bm[g], __ATOMIC_RELAXED);
        if (v == FULL) return -1;
        const u64 bit = (u64)1 << __builtin_ctzll(~v);
        const u64 old = __atomic_fetch_or(&bm[g], bit, __ATOMIC_RELAXED);
        if (!(old & bit)) return (int32_t)(g * 64u + (uint32_t)__builtin_ctzll(bit));
    }
}

/* rwbm_walk with atomic claims. */
static int32_t rwbm_walk_atomic(uint64_t* bm, uint32_t L, int32_t* seed, uint32_t budget,
                                uint32_t* steps) {
    uint32_t g = 0;
    for (uint32_t j = 0; j < budget; ++j) {
        (*steps)++;
        *seed = orc_lcg(*seed);
        g = (uint32_t)*seed % L;
        int32_t id = claim_lowest_atomic(bm, g);
        if (id >= 0) return id;
    }
    for (u64 k = 0; k < 4ull * L; ++k) {
        g = (g + 1u == L) ? 0u : g + 1u;
        (*steps)++;
        int32_t id = claim_lowest_atomic(bm, g);
        if (id >= 0) return id;
    }
    return -1;
}

typedef struct {
    uint64_t* bm;
    uint32_t n_pages, n, threads, t, budget;
    int64_t* free_count;
    uint64_t req_base, seed;
    int phase;  /* 0: allocate, 1: free this thread's grants */
    int32_t* ids;
    uint32_t* steps;
    uint32_t granted;
} rwbm_job;

static void* rwbm_thread(void* arg) {
    rwbm_job* j = (rwbm_job*)arg;
    const uint32_t L = orc_n_words(j->n_pages);
    if (j->phase == 1) {
        int64_t freed = 0;
        for (uint64_t b0 = (uint64_t)j->t * 64u; b0 < j->n; b0 += (uint64_t)j->threads * 64u) {
            for (uint32_t k = 0; k < 64u && b0 + k < j->n; ++k) {
                const int32_t id = j->ids[b0 + k];
                if (id < 0) continue;
                const u64 bit = (u64)1 << ((uint32_t)id & 63u);
                const u64 old =
                    __atomic_fetch_and(&j->bm[(uint32_t)id >> 6], ~bit, __ATOMIC_RELAXED);
                freed += (old & bit) != 0;
            }
        }
        __atomic_fetch_add(j->free_count, freed, __ATOMIC_RELAXED);
        return NULL;
    }
    /* blocks of 64 consecutive requests, dealt round-robin; each block
     * reserves with one atomic, and its first grant-many requests walk (one
     * thread thus serves the requests in order, as the serial policy) */
    for (uint64_t b0 = (uint64_t)j->t * 64u; b0 < j->n; b0 += (uint64_t)j->threads * 64u) {
        const uint32_t want = (uint32_t)(j->n - b0 < 64u ? j->n - b0 : 64u);
        const int64_t old = __atomic_fetch_sub(j->free_count, (int64_t)want, __ATOMIC_RELAXED);
        const uint32_t got = old <= 0 ? 0u : (old >= (int64_t)want ? want : (uint32_t)old);
        if (got < want) __atomic_fetch_add(j->free_count, (int64_t)(want - got), __ATOMIC_RELAXED);
        for (uint32_t k = 0; k < want; ++k) {
            const uint32_t r = (uint32_t)b0 + k;
            uint32_t s = 0;
            int32_t id = -1;
            if (k < got) {
                int32_t sd = orc_request_seed(j->seed, j->req_base + r);
                id = rwbm_walk_atomic(j->bm, L, &sd, j->budget, &s);
                if (id < 0) __atomic_fetch_add(j->free_count, 1, __ATOMIC_RELAXED);
            }
            if (id >= 0) j->granted++;
            j->ids[r] = id;
            if (j->steps) j->steps[r] = s;
        }
    }
    return NULL;
}

/* Runs one phase on every job: jobs[0] on the calling thread, the rest on
 * threads of their own (a job whose thread cannot start runs here too). */
static void rwbm_phase(rwbm_job* jobs, uint32_t threads) {
    pthread_t tid[256];
    int up[256];
    for (uint32_t t = 1; t < threads; ++t)
        up[t] = pthread_create(&tid[t], NULL, rwbm_thread, &jobs[t]) == 0;
    rwbm_thread(&jobs[0]);
    for (uint32_t t = 1; t < threads; ++t) {
        if (up[t]) pthread_join(tid[t], NULL);
        else rwbm_thread(&jobs[t]);
    }
}

uint32_t orc_rwbm_threads(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                          uint64_t req_base, uint64_t seed, uint32_t budget, uint32_t threads,
                          int free_after, int32_t* ids, uint32_t* steps) {
    if (threads == 0) threads = 1;
    if (threads > 256) threads = 256;
    rwbm_job jobs[256];
    uint32_t granted = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        rwbm_job* j = &jobs[t];
        j->bm = bm;
        j->n_pages = n_pages;
        j->n = n;
        j->threads = threads;
        j->t = t;
        j->budget = budget;
        j->free_count = free_count;
        j->req_base = req_base;
        j->seed = seed;
        j->phase = 0;
        j->ids = ids;
        j->steps = steps;
        j->granted = 0;
    }
    rwbm_phase(jobs, threads);
    for (uint32_t t = 0; t < threads; ++t) granted += jobs[t].granted;
    if (free_after) {
        /* every grant is made before the first free */
        for (uint32_t t = 0; t < threads; ++t) jobs[t].phase = 1;
        rwbm_phase(jobs, threads);
    }
    return granted;
}

/* ------------------------------------------------------ RandomWalkBasic */

/* Page walk of RandomWalkBasic (RandomWalkBasic.cuh:54-80) from LCG state
 * sd: each probe takes the next LCG state and page (unsigned)sd % n_pages
 * and claims it if free; a step is one probe, the first included.  After
 * `budget` probes the sweep of rwbm_walk starts after the last probed
 * page's word. */
static int32_t page_walk(uint64_t* bm, uint32_t n_pages, int32_t sd, uint32_t budget,
                         uint32_t* s) {
    const uint32_t L = orc_n_words(n_pages);
    uint32_t p = 0;
    for (uint32_t j = 0; j < budget; ++j) {
        sd = orc_lcg(sd);
        p = (uint32_t)sd % n_pages;
        (*s)++;
        const u64 bit = (u64)1 << (p & 63u);
        if (!(bm[p >> 6] & bit)) {
            bm[p >> 6] |= bit;
            return (int32_t)p;
        }
    }
    uint32_t g = p >> 6;
    for (u64 k = 0; k < 4ull * L; ++k) {
        g = (g + 1u == L) ? 0u : g + 1u;
        (*s)++;
        int32_t id = claim_lowest(bm, g);
        if (id >= 0) return id;
    }
    return -1;
}

/* getPage of RandomWalkBasic for requests in index order. */
uint32_t orc_rwbasic_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                            uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* ids,
                            uint32_t* steps) {
    uint32_t granted = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t s = 0;
        int32_t id = -1;
        if (*free_count > 0) {
            (*free_count)--;
            id = page_walk(bm, n_pages, orc_request_seed(seed, req_base + r), budget, &s);
            if (id < 0) (*free_count)++;
        }
        if (id >= 0) granted++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return granted;
}

/* getPageClusteredRandomWalk (parallelPage.cu:200-218) for one thread
 * whose hint is *hint (last granted page, -1 none): the page after the
 * hint (wrapping at n_pages) if free, one step; otherwise the page walk,
 * counted as that walk counts.  The hint becomes the granted page. */
uint32_t orc_crw_basic_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                              uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* hint,
                              int32_t* ids, uint32_t* steps) {
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t s = 0;
        int32_t id = -1;
        if (*free_count > 0) {
            (*free_count)--;
            if (*hint >= 0) {
                const uint32_t p = ((uint32_t)*hint + 1u >= n_pages) ? 0u : (uint32_t)*hint + 1u;
                const u64 bit = (u64)1 << (p & 63u);
                if (!(bm[p >> 6] & bit)) {
                    bm[p >> 6] |= bit;
                    id = (int32_t)p;
                    s = 1;
                }
            }
            if (id < 0) id = page_walk(bm, n_pages, orc_request_seed(seed, req_base + r), budget, &s);
            if (id < 0) (*free_count)++;
            else *hint = id;
        }
        if (id >= 0) got++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return got;
}

/* getPageCollabRW (parallelPage.cu:642-680) on one wave per chunk of 64
 * requests.  The wave reserves min(active, free) pages for its lowest
 * lanes.  Each round every active lane takes the next state of its own LCG
 * stream and probes page (unsigned)state % n_pages (a page drawn by a lower
 * lane in the same round is not probed again); free pages are claimed in
 * lane order.  The j-th page won in the round goes to the reserved lane
 * filled+j if that lane exists and is given back otherwise.  Rounds stop
 * when every reserved lane has a page or after `budget` rounds; step = the
 * rounds, for every reserved lane.  Reserved lanes still without a page
 * then sweep, in lane order, from their last probed page's word. */
uint32_t orc_collab_basic_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                                 uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* ids,
                                 uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    uint32_t got = 0;
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        const uint32_t n_act = (n - ch * 64u) < 64u ? (n - ch * 64u) : 64u;
        int32_t sd[64], cid[64];
        uint32_t p[64], cst[64];
        int probe[64];
        for (uint32_t l = 0; l < n_act; ++l) {
            sd[l] = orc_request_seed(seed, req_base + ch * 64u + l);
            cid[l] = -1;
            cst[l] = 0;
            p[l] = 0;
        }
        int64_t old = *free_count;
        const uint32_t g = (uint32_t)(old < 0 ? 0 : (old > (int64_t)n_act ? (int64_t)n_act : old));
        *free_count -= g;
        uint32_t filled = 0, rounds = 0;
        while (filled < g && rounds < budget) {
            rounds++;
            for (uint32_t l = 0; l < n_act; ++l) {
                sd[l] = orc_lcg(sd[l]);
                p[l] = (uint32_t)sd[l] % n_pages;
            }
            for (uint32_t l = 0; l < n_act; ++l) {
                probe[l] = 1;
                for (uint32_t k = 0; k < l; ++k)
                    if (p[k] == p[l]) { probe[l] = 0; break; }
            }
            uint32_t j = 0;
            for (uint32_t l = 0; l < n_act; ++l) {
                if (!probe[l]) continue;
                const u64 bit = (u64)1 << (p[l] & 63u);
                if (bm[p[l] >> 6] & bit) continue;
                bm[p[l] >> 6] |= bit;
                if (filled + j < g) cid[filled + j] = (int32_t)p[l];
                else bm[p[l] >> 6] &= ~bit;
                j++;
            }
            filled += (j < g - filled) ? j : g - filled;
        }
        for (uint32_t l = 0; l < g; ++l) cst[l] = rounds;
        for (uint32_t l = filled; l < g; ++l) {
            uint32_t w = p[l] >> 6;
            for (u64 k = 0; k < 4ull * L && cid[l] < 0; ++k) {
                w = (w + 1u == L) ? 0u : w + 1u;
                cst[l]++;
                cid[l] = claim_lowest(bm, w);
            }
            if (cid[l] < 0) (*free_count)++;
        }
        for (uint32_t l = 0; l < n_act; ++l) {
            ids[ch * 64u + l] = cid[l];
            if (steps) steps[ch * 64u + l] = cst[l];
            if (cid[l] >= 0) got++;
        }
    }
    return got;
}

/* Start of the first run of at least x clear bits of w, or -1. */
static int first_run_at_least(u64 w, uint32_t x) {
    uint32_t b = 0;
    while (b < 64u) {
        if ((w >> b) & 1u) { b++; continue; }
        uint32_t e = b;
        while (e < 64u && !((w >> e) & 1u)) e++;
        if (e - b >= x) return (int)b;
        b = e;
    }
    return -1;
}

/* Per-thread getXPage (parallelPage.cu:1104-1142, intended semantics):
 * random words, the first x bits of the first run of >= x clear bits,
 * step = one jump; after min(budget, 64) jumps one sweep of the map. */
uint32_t orc_getx_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                         const uint32_t* counts, uint64_t req_base, uint64_t seed,
                         uint32_t budget, int32_t* ids, uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t x = counts[r];
        uint32_t s = 0;
        int32_t id = -1;
        if (x == 0) {
            id = 0;
        } else if (x <= 64u && *free_count >= (int64_t)x) {
            *free_count -= x;
            int32_t sd = orc_request_seed(seed, req_base + r);
            uint32_t g = 0;
            const uint32_t jumps = budget < 64u ? budget : 64u;
            for (uint32_t j = 0; j < jumps && id < 0; ++j) {
                sd = orc_lcg(sd);
                g = (uint32_t)sd % L;
                s++;
                int st = first_run_at_least(bm[g], x);
                if (st >= 0) {
                    bm[g] |= (x >= 64u ? FULL : (((u64)1 << x) - 1u)) << st;
                    id = (int32_t)(g * 64u + (uint32_t)st);
                }
            }
            for (uint32_t k = 0; k < L && id < 0; ++k) {
                g = (g + 1u == L) ? 0u : g + 1u;
                s++;
                int st = first_run_at_least(bm[g], x);
                if (st >= 0) {
                    bm[g] |= (x >= 64u ? FULL : (((u64)1 << x) - 1u)) << st;
                    id = (int32_t)(g * 64u + (uint32_t)st);
                }
            }
            if (id < 0) *free_count += x;
        }
        if (id >= 0 && x > 0) got++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return got;
}

/* RW_TAS (theoretical_models.py:20-21), the expected mean probes of n
 * serial page-granular walks on T pages with A free: T/n * ln(A/(A-n)). */
double orc_rw_tas(double T, double A, double n) {
    return T / n * log(A / (A - n));
}

/* ------------------------------------------------------------ CollabRW */

/* One collaborative step over a window.  probe[l]: lane l read word
 * word[l]; need[l]: lane l asks for a page.  Designated bits are set in
 * the map (serial: nothing races), served lanes get ids and need=0.
 * Returns the bits designated; *seen = free bits read.
 *
 * order 1 follows the reference's has/need loop literally
 * (CollabRW_BM.cuh:129-153): the s-th needy lane takes the lowest free bit
 * of the (s+1)-th has-lane (fks, :81-94); every has-lane ranked below the
 * number of needy lanes drops its lowest bit (:148-149); masks are
 * recomputed until nobody needs or nobody has; leftovers return to the map
 * (the atomicExch(&w, ~r) of :155-156 restores exactly them when nothing
 * races).  order 0 hands out all free bits of lane 0's word first, then
 * lane 1's, and so on. */
static uint32_t collab_claim(uint64_t* bm, const int* probe, const uint32_t* word, int* need,
                             int32_t* ids, int order, uint32_t* seen) {
    u64 r[64];
    uint32_t total = 0, given = 0;
    for (int l = 0; l < 64; ++l) {
        r[l] = probe[l] ? ~bm[word[l]] : 0;
        total += (uint32_t)__builtin_popcountll(r[l]);
    }
    *seen = total;
    if (order == 1) {
        for (;;) {
            u64 has = 0, needm = 0;
            for (int l = 0; l < 64; ++l) {
                if (r[l]) has |= (u64)1 << l;
                if (need[l]) needm |= (u64)1 << l;
            }
            if (!has || !needm) break;
            int n_need = __builtin_popcountll(needm);
            int32_t give_id[64];
            int give_src[64];
            for (int l = 0; l < 64; ++l) {
                give_src[l] = -1;
                if (!need[l]) continue;
                int s = __builtin_popcountll(needm & (((u64)1 << l) - 1u));
                /* lane of the (s+1)-th set bit of has, if any */
                int cnt = 0;
                for (int j = 0; j < 64; ++j)
                    if ((has >> j) & 1u) {
                        if (cnt == s) { give_src[l] = j; break; }
                        cnt++;
                    }
                if (give_src[l] >= 0) {
                    int j = give_src[l];
                    give_id[l] = (int32_t)(word[j] * 64u + (uint32_t)__builtin_ctzll(r[j]));
                }
            }
            for (int l = 0; l < 64; ++l)
                if (need[l] && give_src[l] >= 0) {
                    ids[l] = give_id[l];
                    bm[(uint32_t)give_id[l] >> 6] |= (u64)1 << ((uint32_t)give_id[l] & 63u);
                    need[l] = 0;
                    given++;
                }
            for (int l = 0; l < 64; ++l)
                if (r[l] && __builtin_popcountll(has & (((u64)1 << l) - 1u)) < n_need)
                    r[l] &= r[l] - 1u;
        }
    } else {
        int l = 0;
        for (int q = 0; q < 64; ++q) {
            if (!need[q]) continue;
            while (l < 64 && r[l] == 0) l++;
            if (l == 64) break;
            int b = __builtin_ctzll(r[l]);
            r[l] &= r[l] - 1u;
            ids[q] = (int32_t)(word[l] * 64u + (uint32_t)b);
            bm[word[l]] |= (u64)1 << b;
            need[q] = 0;
            given++;
        }
    }
    return given;
}

static int any(const int* v) {
    for (int l = 0; l < 64; ++l)
        if (v[l]) return 1;
    return 0;
}

/* CollabRW_BM getPage on one wave per chunk of 64 requests.  With yields,
 * also collectWarpData (CollabRW_BM.cuh:131-138): yields[ch * max_steps + s]
 * = free bits the wave's probing lanes read in its step s (s < max_steps),
 * and n_steps[ch] = the wave's steps. */
static uint32_t collab_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                              uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                              int32_t* ids, uint32_t* steps, uint32_t* yields,
                              uint32_t max_steps, uint32_t* n_steps);

uint32_t orc_collab_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                           uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                           int32_t* ids, uint32_t* steps) {
    return collab_serial(bm, n_pages, free_count, n, req_base, seed, budget, order, ids, steps,
                         NULL, 0, NULL);
}

uint32_t orc_collab_yields(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                           uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                           int32_t* ids, uint32_t* yields, uint32_t max_steps,
                           uint32_t* n_steps) {
    return collab_serial(bm, n_pages, free_count, n, req_base, seed, budget, order, ids, NULL,
                         yields, max_steps, n_steps);
}

static uint32_t collab_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                              uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                              int32_t* ids, uint32_t* steps, uint32_t* yields,
                              uint32_t max_steps, uint32_t* n_steps) {
    const uint32_t L = orc_n_words(n_pages);
    const uint32_t max_sweeps = 4u * ((L + 63u) / 64u) + 4u;
    uint32_t granted_total = 0;
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        int active[64], need[64], probe[64];
        int32_t cid[64];
        uint32_t cst[64], word[64];
        uint32_t n_act = 0;
        for (int l = 0; l < 64; ++l) {
            active[l] = ch * 64u + (uint32_t)l < n;
            n_act += (uint32_t)active[l];
            cid[l] = -1;
            cst[l] = 0;
        }
        /* wave-aggregated reservation: the lowest `granted` active lanes */
        int64_t old = *free_count;
        int64_t g = old < 0 ? 0 : (old > (int64_t)n_act ? (int64_t)n_act : old);
        *free_count -= g;
        for (int l = 0; l < 64; ++l) need[l] = active[l] && l < g;
        int32_t sd = orc_request_seed(seed, req_base + (u64)ch * 64u);
        uint32_t jumps = 0, sweeps = 0, start = 0, it = 0;
        while (any(need)) {
            if (jumps < budget) {
                sd = orc_lcg(sd);
                start = (uint32_t)sd % L;
                jumps++;
            } else {
                start = (uint32_t)(((u64)start + 64u) % L);
                if (++sweeps > max_sweeps) break;
            }
            for (int l = 0; l < 64; ++l) {
                if (need[l]) cst[l]++;
                probe[l] = active[l] && (uint32_t)l < L;
                word[l] = (uint32_t)(((u64)start + (uint32_t)l) % L);
            }
            uint32_t seen;
            collab_claim(bm, probe, word, need, cid, order, &seen);
            if (yields && it < max_steps) yields[(u64)ch * max_steps + it] = seen;
            it++;
        }
        if (n_steps) n_steps[ch] = it;
        for (int l = 0; l < 64; ++l) {
            if (need[l]) (*free_count)++;
            if (!active[l]) continue;
            ids[ch * 64u + l] = cid[l];
            if (steps) steps[ch * 64u + l] = cst[l];
            if (cid[l] >= 0) granted_total++;
        }
    }
    return granted_total;
}

#define SWEEP_MAX 256u

/* Batched persistent allocation on one wave (pbm_alloc.hip k_get_batched):
 * chunk ch's first window starts at its home word ch*64*L/granted, then
 * window mode makes up to `budget` random jumps to windows of 64*J words,
 * kept while they show free bits; after the budget the chunk sweeps: with
 * S = min(256, ceil(L/64)), lane l
 * owns words start+S*l .. +S-1 with a cursor, offers the first non-full word
 * from its cursor each step and moves past a word once it is full; when
 * every lane has finished its segment the next span of 64*S words starts
 * (a wave that never jumped first draws a random start). */
uint32_t orc_batched_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                            uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                            uint32_t cluster_words, int32_t* ids, uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    const uint32_t J = cluster_words / 64u;
    const uint32_t seg = (L + 63u) / 64u < SWEEP_MAX ? (L + 63u) / 64u : SWEEP_MAX;
    const uint32_t sweep_span = 64u * seg;
    const uint32_t max_dry = 2u * ((L + sweep_span - 1u) / sweep_span) + 2u;
    /* phase admission: the first min(n, free) requests by index */
    int64_t old = *free_count;
    int64_t granted = old < 0 ? 0 : (old > (int64_t)n ? (int64_t)n : old);
    *free_count -= granted;
    int32_t wseed = orc_request_seed(seed ^ WAVE_SALT, req_base + 0);
    uint32_t got = 0;
    uint32_t cur[64], seg_len[64];
    for (int l = 0; l < 64; ++l) {
        uint32_t first = (uint32_t)l * seg;
        seg_len[l] = first >= L ? 0u : (L - first < seg ? L - first : seg);
    }
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        /* the chunk's home word: its share of the grants */
        const uint32_t home = granted > 0 ? (uint32_t)(((u64)ch * 64u * L) / (u64)granted % L) : 0u;
        int have = 0, sweep = 0, homed = 0, found_pass = 0;
        uint32_t start = 0;
        for (int l = 0; l < 64; ++l) cur[l] = 0;
        int need[64], valid[64];
        int32_t cid[64];
        uint32_t cst[64];
        for (int l = 0; l < 64; ++l) {
            uint32_t r = ch * 64u + (uint32_t)l;
            valid[l] = r < n;
            need[l] = valid[l] && (int64_t)r < granted;
            cid[l] = -1;
            cst[l] = 0;
        }
        uint32_t jumps = 0, dry = 0, iters = 0;
        const uint32_t max_iters = budget + 4u * max_dry * seg + 65536u;
        while (any(need)) {
            if (++iters > max_iters) break;
            if (!have) {
                if (!homed) {
                    start = home;
                    homed = 1;
                } else if (jumps < budget) {
                    wseed = orc_lcg(wseed);
                    start = lcg_to_range(wseed, L);
                    jumps++;
                } else {
                    sweep = 1;
                    for (int l = 0; l < 64; ++l) cur[l] = 0;
                    found_pass = 0;
                }
                have = 1;
            }
            for (int l = 0; l < 64; ++l)
                if (need[l]) cst[l]++;
            if (!sweep) {
                uint32_t seen_all = 0, desig_all = 0;
                int examined_all = 1;
                for (uint32_t i = 0; i < J; ++i) {
                    if (!any(need)) {
                        examined_all = 0;
                        break;
                    }
                    int probe[64];
                    uint32_t word[64];
                    for (int l = 0; l < 64; ++l) {
                        uint32_t off = 64u * i + (uint32_t)l;
                        probe[l] = off < L;
                        word[l] = (uint32_t)(((u64)start + off) % L);
                    }
                    uint32_t seen;
                    desig_all += collab_claim(bm, probe, word, need, cid, order, &seen);
                    seen_all += seen;
                }
                if (examined_all && seen_all <= desig_all) have = 0;
            } else {
                int probe[64];
                uint32_t word[64];
                for (int l = 0; l < 64; ++l) {
                    probe[l] = 0;
                    word[l] = 0;
                    while (cur[l] < seg_len[l]) {
                        uint32_t idx = start + (uint32_t)l * seg + cur[l];
                        if (idx >= L) idx -= L;
                        if (bm[idx] != FULL) {
                            probe[l] = 1;
                            word[l] = idx;
                            break;
                        }
                        cur[l]++;
                    }
                }
                uint32_t seen;
                collab_claim(bm, probe, word, need, cid, order, &seen);
                for (int l = 0; l < 64; ++l)
                    if (probe[l] && bm[word[l]] == FULL) cur[l]++;
                if (seen) found_pass = 1;
                int open_seg = 0;
                for (int l = 0; l < 64; ++l) open_seg |= cur[l] < seg_len[l];
                if (!open_seg) {
                    if (any(need)) {
                        if (!found_pass) {
                            if (++dry > max_dry) break;
                        } else {
                            dry = 0;
                        }
                    }
                    found_pass = 0;
                    for (int l = 0; l < 64; ++l) cur[l] = 0;
                    start = (uint32_t)(((u64)start + sweep_span) % L);
                }
            }
        }
        for (int l = 0; l < 64; ++l) {
            if (need[l]) (*free_count)++;
            if (!valid[l]) continue;
            ids[ch * 64u + l] = cid[l];
            if (steps) steps[ch * 64u + l] = cst[l];
            if (cid[l] >= 0) got++;
        }
    }
    return got;
}

/* Batched allocation by rank (PBM_LAUNCH_RANK, pbm_alloc.hip k_get_rank):
 * the call admits min(n, free count) requests and request r takes the free
 * page of rank r, i.e. the r-th free page in map order; each grant costs
 * one probe of its word (steps 1), a refusal none.  The outcome does not
 * depend on the schedule. */
uint32_t orc_batched_rank(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                          int32_t* ids, uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    int64_t old = *free_count;
    int64_t granted = old < 0 ? 0 : (old > (int64_t)n ? (int64_t)n : old);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (granted > 0 ? (size_t)granted : 1u));
    if (!order) return 0;
    int64_t k = 0;
    for (uint32_t w = 0; w < L && k < granted; ++w) {
        u64 fr = ~bm[w];
        while (fr && k < granted) {
            int b = __builtin_ctzll(fr);
            fr &= fr - 1u;
            order[k++] = (int32_t)(w * 64u + (uint32_t)b);
        }
    }
    *free_count -= k;
    for (uint32_t r = 0; r < n; ++r) {
        ids[r] = -1;
        if (steps) steps[r] = 0;
    }
    for (int64_t r = 0; r < k; ++r) {
        ids[r] = order[r];
        if (steps) steps[r] = 1u;
    }
    for (int64_t r = 0; r < k; ++r)
        bm[(uint32_t)order[r] >> 6] |= (u64)1 << ((uint32_t)order[r] & 63u);
    free(order);
    return (uint32_t)k;
}

/* ------------------------------------------------------------ CRW_BM */

/* Claim loop of CRW_BM (:84-107, :114-127): every atomicOr attempt is a
 * step. */
static int32_t claim_counting(uint64_t* bm, uint32_t g, uint32_t* steps) {
    while (bm[g] != FULL) {
        (*steps)++;
        int b = __builtin_ctzll(~bm[g]);
        u64 bit = (u64)1 << b;
        u64 old = bm[g];
        bm[g] |= bit;
        if (!(old & bit)) return (int32_t)(g * 64u + (uint32_t)b);
    }
    return -1;
}

/* CRW_BM getPage (CRW_BM.cuh:75-142) for one thread whose hint is *hint
 * (last granted page, -1 none): the hinted word (the next word when the
 * hint was its top bit, wrapping at L), then the random walk, whose jumps
 * to full words are not steps, then a sweep. */
uint32_t orc_crw_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                        uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* hint,
                        int32_t* ids, uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t s = 0;
        int32_t id = -1;
        if (*free_count > 0) {
            (*free_count)--;
            if (*hint >= 0) {
                uint32_t g = (uint32_t)*hint >> 6;
                if (((uint32_t)*hint & 63u) == 63u) g = (g + 1u == L) ? 0u : g + 1u;
                id = claim_counting(bm, g, &s);
            }
            int32_t sd = orc_request_seed(seed, req_base + r);
            uint32_t g = 0;
            for (uint32_t j = 0; id < 0 && j < budget; ++j) {
                sd = orc_lcg(sd);
                g = (uint32_t)sd % L;
                id = claim_counting(bm, g, &s);
            }
            for (u64 k = 0; id < 0 && k < 4ull * L; ++k) {
                g = (g + 1u == L) ? 0u : g + 1u;
                id = claim_counting(bm, g, &s);
            }
            if (id < 0) (*free_count)++;
            else *hint = id;
        }
        if (id >= 0) got++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return got;
}

/* ------------------------------------------------------------ clocks */

/* getPageSingleClock (parallelPage.cu:491-509): take arm++ % N until the
 * page there is free; step = arm advance. */
uint32_t orc_single_clock_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                                 uint64_t* arm, uint32_t budget, int32_t* ids, uint32_t* steps) {
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t s = 0;
        int32_t id = -1;
        if (*free_count > 0) {
            (*free_count)--;
            u64 max_tries = (u64)budget + 2ull * n_pages;
            for (u64 it = 0; it < max_tries; ++it) {
                s++;
                uint32_t p = (uint32_t)((*arm)++ % n_pages);
                u64 bit = (u64)1 << (p & 63u);
                if (!(bm[p >> 6] & bit)) {
                    bm[p >> 6] |= bit;
                    id = (int32_t)p;
                    break;
                }
            }
            if (id < 0) (*free_count)++;
        }
        if (id >= 0) got++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return got;
}

/* Initial DistClock arm of slot t. */
uint32_t orc_dist_clock_arm(uint64_t seed, uint32_t slot, uint32_t n_pages) {
    return (uint32_t)orc_lcg(orc_request_seed(seed ^ 0xA4093822299F31D0ull, slot)) % n_pages;
}

/* getPageDistClock (parallelPage.cu:571-592) for one arm: examine pages
 * from the arm on, take the first free one, leave the arm one past it;
 * step = page examined. */
uint32_t orc_dist_clock_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                               uint32_t* arm, uint32_t budget, int32_t* ids, uint32_t* steps) {
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t s = 0;
        int32_t id = -1;
        if (*free_count > 0) {
            (*free_count)--;
            u64 max_pages = (u64)budget + 2ull * n_pages;
            uint32_t a = *arm % n_pages;
            for (u64 k = 0; k < max_pages; ++k) {
                s++;
                uint32_t p = a;
                a = (p + 1u == n_pages) ? 0u : p + 1u;
                u64 bit = (u64)1 << (p & 63u);
                if (!(bm[p >> 6] & bit)) {
                    bm[p >> 6] |= bit;
                    id = (int32_t)p;
                    break;
                }
            }
            *arm = a;
            if (id < 0) (*free_count)++;
        }
        if (id >= 0) got++;
        ids[r] = id;
        if (steps) steps[r] = s;
    }
    return got;
}

/* ------------------------------------------------------------ LinkedList */

/* initPagesLinkedList (parallelPage.cu:303-337) on a pre-filled map: the
 * free pages in ascending order form the list; returns how many. */
uint32_t orc_list_build(const uint64_t* bm, uint32_t n_pages, int32_t* q) {
    uint32_t k = 0;
    for (uint32_t p = 0; p < n_pages; ++p)
        if (!((bm[p >> 6] >> (p & 63u)) & 1u)) q[k++] = (int32_t)p;
    return k;
}

/* getPageLinkedList (:342-369): pop the head; -1 when empty (the reference
 * traps).  q is a ring of n_pages slots, head/tail are tickets. */
void orc_list_get(uint64_t* bm, uint32_t n_pages, int32_t* q, uint64_t* head, uint64_t* tail,
                  uint32_t n, int32_t* ids) {
    for (uint32_t r = 0; r < n; ++r) {
        if (*head == *tail) {
            ids[r] = -1;
            continue;
        }
        int32_t id = q[*head % n_pages];
        (*head)++;
        bm[(uint32_t)id >> 6] |= (u64)1 << ((uint32_t)id & 63u);
        ids[r] = id;
    }
}

/* freePageLinkedList (:399-418): append at the tail (ids < 0 and pages
 * not in use are skipped). */
void orc_list_put(uint64_t* bm, uint32_t n_pages, int32_t* q, uint64_t* head, uint64_t* tail,
                  uint32_t n, const int32_t* ids) {
    (void)head;
    for (uint32_t r = 0; r < n; ++r) {
        int32_t id = ids[r];
        if (id < 0 || (uint32_t)id >= n_pages) continue;
        u64 bit = (u64)1 << ((uint32_t)id & 63u);
        if (!(bm[(uint32_t)id >> 6] & bit)) continue;
        bm[(uint32_t)id >> 6] &= ~bit;
        q[*tail % n_pages] = id;
        (*tail)++;
    }
}

/* ------------------------------------------------------ Ouroboros queue */

/* Pages per chunk for requests of page_size bytes: the power-of-two size
 * class of at least 16 bytes (Ouroboros/include/device/Helper.cuh:42-60)
 * carved from 8 KiB chunks, 32 KiB chunks above 8 KiB
 * (Ouroboros/include/InstanceDefinitions.cuh:26-28), one page per chunk
 * above 32 KiB. */
uint32_t orc_ouroboros_chunk_pages(uint32_t page_size) {
    u64 cls = 16;
    while (cls < page_size) cls <<= 1;
    if (cls <= 8192u) return (uint32_t)(8192u / cls);
    if (cls <= 32768u) return (uint32_t)(32768u / cls);
    return 1u;
}

/* Queue state after a reset of the page map: fresh = the free pages,
 * ascending; st = {carved, fresh pages, head, tail, q_avail}. */
uint32_t orc_queue_reset(const uint64_t* bm, uint32_t n_pages, uint32_t* fresh, int64_t* st) {
    uint32_t k = 0;
    for (uint32_t p = 0; p < n_pages; ++p)
        if (!((bm[p >> 6] >> (p & 63u)) & 1u)) fresh[k++] = p;
    st[0] = 0;
    st[1] = k;
    st[2] = 0;
    st[3] = 0;
    st[4] = 0;
    return k;
}

/* getPage of the Ouroboros comparator, one request at a time (a wave of
 * one lane).  A reserved request promises itself one queued page; with
 * none unpromised it carves the next chunk_pages fresh pages (fewer at the
 * end) onto the back of the queue first.  It takes the front of the queue.
 * Step = 1, 2 when it carved. */
uint32_t orc_queue_get(uint64_t* bm, uint32_t n_pages, int64_t* free_count,
                       const uint32_t* fresh, int32_t* ring, int64_t* st, uint32_t chunk_pages,
                       uint32_t n, int32_t* ids, uint32_t* steps) {
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        ids[r] = -1;
        if (steps) steps[r] = 0;
        if (*free_count <= 0) continue;
        (*free_count)--;
        const int64_t old = st[4]--;
        uint32_t s = 1;
        if (old <= 0) {
            const int64_t base = st[0];
            st[0] += chunk_pages;
            const int64_t take = base >= st[1] ? 0
                                 : ((int64_t)chunk_pages < st[1] - base ? (int64_t)chunk_pages
                                                                          : st[1] - base);
            for (int64_t k = 0; k < take; ++k) ring[(st[3]++) % n_pages] = (int32_t)fresh[base + k];
            st[4] += take;
            if (take) s = 2;
        }
        const int32_t id = ring[(st[2]++) % n_pages];
        bm[(uint32_t)id >> 6] |= (u64)1 << ((uint32_t)id & 63u);
        ids[r] = id;
        if (steps) steps[r] = s;
        got++;
    }
    return got;
}

/* freePage of the comparator: a page in use goes to the back of the queue;
 * other ids are skipped.  Returns the pages freed. */
uint32_t orc_queue_put(uint64_t* bm, uint32_t n_pages, int64_t* free_count, int32_t* ring,
                       int64_t* st, uint32_t n, const int32_t* ids) {
    uint32_t freed = 0;
    for (uint32_t r = 0; r < n; ++r) {
        const int32_t id = ids[r];
        if (id < 0 || (uint32_t)id >= n_pages) continue;
        const u64 bit = (u64)1 << ((uint32_t)id & 63u);
        if (!(bm[(uint32_t)id >> 6] & bit)) continue;
        bm[(uint32_t)id >> 6] &= ~bit;
        ring[(st[3]++) % n_pages] = id;
        st[4]++;
        (*free_count)++;
        freed++;
    }
    return freed;
}

/* ------------------------------------------------------------ RWMalloc */

/* RWMalloc getPage (RWMalloc.cuh:113-149) on one wave: every active lane
 * walks its own LCG stream; a word probed by several lanes is probed by
 * the lowest of them only; after the budget the wave sweeps from its
 * lowest lane's last word. */
uint32_t orc_rwmalloc_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                             uint64_t req_base, uint64_t seed, uint32_t budget, int order,
                             int32_t* ids, uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    const uint32_t max_sweeps = 4u * ((L + 63u) / 64u) + 4u;
    uint32_t got = 0;
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        int active[64], need[64], probe[64];
        int32_t cid[64], sd[64];
        uint32_t cst[64], word[64], n_act = 0;
        for (int l = 0; l < 64; ++l) {
            uint32_t r = ch * 64u + (uint32_t)l;
            active[l] = r < n;
            n_act += (uint32_t)active[l];
            cid[l] = -1;
            cst[l] = 0;
            sd[l] = active[l] ? orc_request_seed(seed, req_base + r) : 0;
            word[l] = 0;
        }
        int64_t old = *free_count;
        int64_t g = old < 0 ? 0 : (old > (int64_t)n_act ? (int64_t)n_act : old);
        *free_count -= g;
        for (int l = 0; l < 64; ++l) need[l] = active[l] && l < g;
        uint32_t jumps = 0, sweeps = 0, start = 0;
        while (any(need)) {
            if (jumps < budget) {
                for (int l = 0; l < 64; ++l) {
                    if (!active[l]) continue;
                    sd[l] = orc_lcg(sd[l]);
                    word[l] = (uint32_t)sd[l] % L;
                    probe[l] = 1;
                }
                jumps++;
                if (jumps == budget) start = word[0];
            } else {
                start = (uint32_t)(((u64)start + 64u) % L);
                if (++sweeps > max_sweeps) break;
                for (int l = 0; l < 64; ++l) {
                    word[l] = (uint32_t)(((u64)start + (uint32_t)l) % L);
                    probe[l] = active[l] && (uint32_t)l < L;
                }
            }
            for (int l = 0; l < 64; ++l)
                if (need[l]) cst[l]++;
            for (int l = 0; l < 64; ++l) {
                if (!probe[l] || !active[l]) { probe[l] = 0; continue; }
                for (int k = 0; k < l; ++k)
                    if (probe[k] && active[k] && word[k] == word[l]) { probe[l] = 0; break; }
            }
            uint32_t seen;
            collab_claim(bm, probe, word, need, cid, order, &seen);
        }
        for (int l = 0; l < 64; ++l) {
            if (need[l]) (*free_count)++;
            if (!active[l]) continue;
            ids[ch * 64u + l] = cid[l];
            if (steps) steps[ch * 64u + l] = cst[l];
            if (cid[l] >= 0) got++;
        }
    }
    return got;
}

/* First zero run of *x: sets *start, returns its length (0 if none) and
 * marks it used in *x (find0Range, RWMalloc.cuh:159-180, as intended). */
static unsigned first_zero_run(u64* x, unsigned* start) {
    u64 z = ~*x;
    if (!z) return 0;
    unsigned s = (unsigned)__builtin_ctzll(z);
    u64 above = *x >> s;
    unsigned len = above ? (unsigned)__builtin_ctzll(above) : 64u - s;
    u64 run = (len >= 64u ? FULL : (((u64)1 << len) - 1u)) << s;
    *x |= run;
    *start = s;
    return len;
}

static u64 run_mask(unsigned start, unsigned n) {
    return (n >= 64u ? FULL : (((u64)1 << n) - 1u)) << start;
}

/* getNPage (RWMalloc.cuh:198-248) on one wave over chunks of 64 requests
 * for counts[r] pages.  Reservation: lanes in order while the running sum
 * of their counts fits the free pages.  Per step a window of `spans` x 64
 * consecutive words (the reference's window is one word per lane, spans =
 * 1), searched span by span; per round every lane offers the next zero run
 * of its word in the span and, for each offering lane in order, the lowest
 * lane that still needs and whose count fits takes the run's first pages.
 * At most 64 random windows, then one sweep of the map, then failure. */
uint32_t orc_getnpage_spans(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                            const uint32_t* counts, uint64_t req_base, uint64_t seed,
                            uint32_t budget, uint32_t spans, int32_t* ids, uint32_t* steps) {
    const uint32_t L = orc_n_words(n_pages);
    if (spans < 1u) spans = 1u;
    if (spans > 4u) spans = 4u;
    const uint32_t width = 64u * spans;
    /* bounded search: at most 64 random windows, then one pass of the map */
    const uint32_t max_sweeps = (L + width - 1u) / width;
    if (budget > 64u) budget = 64u;
    uint32_t got = 0;
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        int active[64], need[64];
        int32_t out[64];
        uint32_t cst[64], cnt[64];
        u64 pre = 0, total = 0;
        for (int l = 0; l < 64; ++l) {
            uint32_t r = ch * 64u + (uint32_t)l;
            active[l] = r < n;
            cnt[l] = active[l] ? counts[r] : 0;
            out[l] = cnt[l] == 0 ? 0 : -1;
            cst[l] = 0;
            need[l] = 0;
            if (active[l] && cnt[l] >= 1 && cnt[l] <= 64) total += cnt[l];
        }
        int64_t old = *free_count;
        *free_count -= (int64_t)total;
        u64 back = 0;
        for (int l = 0; l < 64; ++l) {
            if (!(active[l] && cnt[l] >= 1 && cnt[l] <= 64)) continue;
            if (old >= (int64_t)(pre + cnt[l])) need[l] = 1;
            else back += cnt[l];
            pre += cnt[l];
        }
        *free_count += (int64_t)back;
        int32_t sd = orc_request_seed(seed, req_base + (u64)ch * 64u);
        uint32_t jumps = 0, sweeps = 0, start = 0;
        while (any(need)) {
            if (jumps < budget) {
                sd = orc_lcg(sd);
                start = (uint32_t)sd % L;
                jumps++;
            } else {
                start = (uint32_t)(((u64)start + width) % L);
                if (++sweeps > max_sweeps) break;
            }
            for (int l = 0; l < 64; ++l)
                if (need[l]) cst[l]++;
            /* the spans hold distinct words, so searching them in order on
             * a serial schedule is the same as reading them all first */
            for (uint32_t j = 0; j < spans && any(need); ++j) {
                u64 img[64];
                uint32_t word[64];
                for (int l = 0; l < 64; ++l) {
                    const uint32_t off = 64u * j + (uint32_t)l;
                    word[l] = (uint32_t)(((u64)start + off) % L);
                    img[l] = (active[l] && off < L) ? bm[word[l]] : FULL;
                }
                for (;;) {
                    unsigned rs[64], rl[64];
                    int has = 0;
                    for (int l = 0; l < 64; ++l) {
                        rs[l] = 0;
                        rl[l] = active[l] ? first_zero_run(&img[l], &rs[l]) : 0;
                        has |= rl[l] > 0;
                    }
                    if (!has || !any(need)) break;
                    for (int t = 0; t < 64; ++t) {
                        if (!rl[t]) continue;
                        for (int k = 0; k < 64; ++k) {
                            if (!need[k] || cnt[k] > rl[t]) continue;
                            bm[word[t]] |= run_mask(rs[t], cnt[k]);
                            out[k] = (int32_t)(word[t] * 64u + rs[t]);
                            need[k] = 0;
                            break;
                        }
                    }
                }
            }
        }
        for (int l = 0; l < 64; ++l) {
            if (need[l]) (*free_count) += cnt[l];
            if (!active[l]) continue;
            ids[ch * 64u + l] = out[l];
            if (steps) steps[ch * 64u + l] = cst[l];
            if (cnt[l] > 0 && out[l] >= 0) got++;
        }
    }
    return got;
}

uint32_t orc_getnpage_serial(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                             const uint32_t* counts, uint64_t req_base, uint64_t seed,
                             uint32_t budget, int32_t* ids, uint32_t* steps) {
    return orc_getnpage_spans(bm, n_pages, free_count, n, counts, req_base, seed, budget, 1u, ids,
                              steps);
}

/* releasePages (RWMalloc.cuh:60-70). */
uint32_t orc_free_n(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                    const int32_t* ids, const uint32_t* counts) {
    uint32_t freed = 0;
    for (uint32_t r = 0; r < n; ++r) {
        int32_t id = ids[r];
        uint32_t c = counts[r];
        /* a run past n_pages would clear padding bits, which must stay set */
        if (id < 0 || c == 0 || c > 64 || (uint64_t)(uint32_t)id + c > n_pages ||
            ((uint32_t)id & 63u) + c > 64u)
            continue;
        u64 m = run_mask((uint32_t)id & 63u, c);
        freed += (uint32_t)__builtin_popcountll(bm[(uint32_t)id >> 6] & m);
        bm[(uint32_t)id >> 6] &= ~m;
    }
    *free_count += freed;
    return freed;
}

/* ------------------------------------------------------- kernel malloc */

#define OFF_NONE (~(u64)0)

static void put64(uint8_t* mem, uint8_t* written, u64 at, u64 v) {
    memcpy(mem + at, &v, 8);
    if (written) memset(written + at, 1, 8);
}

static u64 get64(const uint8_t* mem, u64 at) {
    u64 v;
    memcpy(&v, mem + at, 8);
    return v;
}

/* KernelMallocCollab::malloc (malloc.cu:231-290) on one wave per chunk of
 * 64 requests, writing the headers into `mem` (the pool's bytes; `written`
 * marks every byte written when not NULL).  s = round8(size) + 8; requests
 * with s + 8 < page_size (and s within the group cap) form groups in lane
 * order whose totals fit min(65535, 64 page_size - 8); a group's lowest
 * lane asks getNPage for ceil((total + 8) / page_size) pages and writes the
 * chunk header (n | total << 16 | page << 32); each member writes its space
 * header (root | s << 32) at root = 8 + suffix - s, suffix being the sum
 * of s over the group's lanes from itself up.  Other requests ask for
 * ceil(s / page_size) pages and write (n | (s & 0xFFFF) << 16 | page << 32
 * | 1 << 63).  off[r]: pool offset of the returned pointer, or ~0. */
uint32_t orc_kmalloc_collab(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                            int64_t* free_count, uint8_t* mem, uint8_t* written, uint32_t n,
                            const uint32_t* sizes, uint64_t req_base, uint64_t seed,
                            uint32_t budget, uint64_t* off, uint32_t* steps) {
    const u64 ps = page_size;
    const u64 cap = 64u * ps - 8u < 65535u ? 64u * ps - 8u : 65535u;
    uint32_t got = 0;
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        const uint32_t n_act = (n - ch * 64u) < 64u ? (n - ch * 64u) : 64u;
        u64 sz[64], suffix[64];
        int collab[64], leader[64];
        uint32_t cnt[64], st[64];
        int32_t ids[64];
        int cur = -1;
        u64 run = 0;
        for (uint32_t l = 0; l < n_act; ++l) {
            const uint32_t r = ch * 64u + l;
            sz[l] = (((u64)sizes[r] + 7u) & ~(u64)7) + 8u;
            collab[l] = sz[l] + 8u < ps && sz[l] <= cap;
            leader[l] = (int)l;
            suffix[l] = 0;
            if (!collab[l]) continue;
            if (cur < 0 || run + sz[l] > cap) { cur = (int)l; run = 0; }
            run += sz[l];
            leader[l] = cur;
        }
        cur = -1;
        run = 0;
        for (int l = (int)n_act - 1; l >= 0; --l) {
            if (!collab[l]) continue;
            if (leader[l] != cur) { cur = leader[l]; run = 0; }
            run += sz[l];
            suffix[l] = run;
        }
        for (uint32_t l = 0; l < n_act; ++l) {
            if (collab[l]) {
                cnt[l] = leader[l] == (int)l ? (uint32_t)((suffix[l] + 8u + ps - 1u) / ps) : 0u;
            } else {
                const u64 p = (sz[l] + ps - 1u) / ps;
                cnt[l] = p > 64u ? 65u : (uint32_t)p;
            }
        }
        orc_getnpage_serial(bm, n_pages, free_count, n_act, cnt, req_base + (u64)ch * 64u, seed,
                            budget, ids, st);
        for (uint32_t l = 0; l < n_act; ++l) {
            const uint32_t r = ch * 64u + l;
            if (steps) steps[r] = st[l];
            const int32_t gp = collab[l] ? ids[leader[l]] : ids[l];
            if (gp < 0) {
                off[r] = OFF_NONE;
                continue;
            }
            got++;
            const u64 chunk = (u64)(uint32_t)gp * ps;
            if (!collab[l]) {
                put64(mem, written, chunk,
                      (u64)cnt[l] | ((u64)(uint16_t)sz[l] << 16) | ((u64)(uint32_t)gp << 32) |
                          ((u64)1 << 63));
                off[r] = chunk + 8u;
                continue;
            }
            if (leader[l] == (int)l)
                put64(mem, written, chunk,
                      (u64)cnt[l] | (suffix[l] << 16) | ((u64)(uint32_t)gp << 32));
            const u64 root = 8u + suffix[l] - sz[l];
            put64(mem, written, chunk + root, root | (sz[l] << 32));
            off[r] = chunk + root + 8u;
        }
    }
    return got;
}

/* KernelMallocCollab::free as intended (malloc.cu:300-320), in index
 * order: a header with bit 63 returns its chunk's pages; a space header
 * takes its bytes off the chunk's occupied field and the last one returns
 * the chunk's pages. */
uint32_t orc_kfree_collab(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint8_t* mem,
                          uint32_t n, const uint64_t* off) {
    uint32_t freed = 0;
    for (uint32_t r = 0; r < n; ++r) {
        if (off[r] == OFF_NONE) continue;
        const u64 hdr = off[r] - 8u;
        const u64 h = get64(mem, hdr);
        int32_t id;
        uint32_t cnt;
        if (h >> 63) {
            id = (int32_t)((uint32_t)(h >> 32) & 0x7FFFFFFFu);
            cnt = (uint32_t)(h & 0xFFFFu);
        } else {
            const u64 chunk = hdr - (uint32_t)h;
            const uint32_t bytes = (uint32_t)(h >> 32) & 0x7FFFFFFFu;
            uint32_t w;
            memcpy(&w, mem + chunk, 4);
            const uint32_t occ = w >> 16;
            w -= bytes << 16;
            memcpy(mem + chunk, &w, 4);
            if (occ != bytes) continue;
            uint32_t first;
            memcpy(&first, mem + chunk + 4, 4);
            id = (int32_t)(first & 0x7FFFFFFFu);
            cnt = w & 0xFFFFu;
        }
        freed += orc_free_n(bm, n_pages, free_count, 1, &id, &cnt);
    }
    return freed;
}

/* KernelMallocSimple::malloc (malloc.cu:129-175) for one thread whose
 * current chunk is *current, requests in index order.  s = round8(size);
 * valid when size <= 32760, the chunk of ceil((s + 16) / page_size)
 * pages has at most 64 pages and 65,535 available bytes.  A request that
 * does not fit the current chunk (available - allocated < s + 8) asks
 * getNPage for a new chunk, whose header becomes (available << 16).  The
 * space header (prev | (s | 0x8000) << 16 | offset << 32) goes at offset
 * 8 + allocated; the previous last space loses bit 15 of its second field
 * and gives its size as prev; the chunk header becomes (allocated + 8 + s
 * | available << 16 | old allocated << 32). */
uint32_t orc_kmalloc_simple(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                            int64_t* free_count, uint8_t* mem, uint8_t* written, uint32_t n,
                            const uint32_t* sizes, uint64_t req_base, uint64_t seed,
                            uint32_t budget, int32_t* current, uint64_t* off, uint32_t* steps) {
    const u64 ps = page_size;
    uint32_t got = 0;
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t s = (sizes[r] + 7u) & ~7u;
        const u64 pages = ((u64)s + 16u + ps - 1u) / ps;
        const int valid = sizes[r] <= 32760u && pages <= 64u && pages * ps - 8u <= 65535u;
        int32_t cur = valid ? *current : -1;
        int fresh = 1;
        if (cur >= 0) {
            const u64 m = get64(mem, (u64)(uint32_t)cur * ps);
            fresh = (uint32_t)((m >> 16) & 0xFFFFu) - (uint32_t)(m & 0xFFFFu) < s + 8u;
        }
        uint32_t st = 0;
        int32_t pid = 0;
        if (valid && fresh) {
            const uint32_t c = (uint32_t)pages;
            orc_getnpage_serial(bm, n_pages, free_count, 1, &c, req_base + r, seed, budget, &pid,
                                &st);
        }
        if (steps) steps[r] = st;
        off[r] = OFF_NONE;
        if (!valid) continue;
        if (fresh) {
            if (pid < 0) continue;
            cur = pid;
            *current = pid;
            put64(mem, written, (u64)(uint32_t)pid * ps, (pages * ps - 8u) << 16);
        }
        const u64 chunk = (u64)(uint32_t)cur * ps;
        const u64 m = get64(mem, chunk);
        const uint32_t alloc = (uint32_t)(m & 0xFFFFu);
        const uint32_t last = (uint32_t)((m >> 32) & 0xFFFFu);
        const uint32_t at = 8u + alloc;
        uint32_t prev = 0;
        if (alloc) {
            const u64 l = get64(mem, chunk + 8u + last);
            prev = (uint32_t)((l >> 16) & 0x7FFFu);
            put64(mem, written, chunk + 8u + last, l & ~((u64)1 << 31));
        }
        put64(mem, written, chunk + at, (u64)prev | ((u64)(s | 0x8000u) << 16) | ((u64)at << 32));
        put64(mem, written, chunk,
              (u64)(alloc + 8u + s) | (m & 0xFFFF0000ull) | ((u64)alloc << 32));
        off[r] = chunk + at + 8u;
        got++;
    }
    return got;
}

/* mallocRW (RWMalloc.cuh:299-326) on one wave: the plan of every request
 * (page count asked, offset inside the run, header) and the getNPage
 * grants.  Outputs per request: page (first page of its run, -1 on
 * failure), offset (bytes from the page start to the header) and header. */
uint32_t orc_malloc_spans(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                          int64_t* free_count, uint32_t n, const uint64_t* sizes,
                          uint64_t req_base, uint64_t seed, uint32_t budget, uint32_t spans,
                          int32_t* page, uint64_t* offset, int32_t* header, uint32_t* steps) {
    uint32_t got = 0;
    for (uint32_t ch = 0; ch * 64u < n; ++ch) {
        uint32_t m = n - ch * 64u < 64u ? n - ch * 64u : 64u;
        uint32_t cnt[64];
        u64 off[64], need_b[64], total = 0;
        int small[64], leader = -1;
        for (uint32_t l = 0; l < m; ++l) {
            u64 nb = (sizes[ch * 64u + l] + 4u + 3u) & ~(u64)3u;
            need_b[l] = nb;
            small[l] = nb < page_size;
            off[l] = 0;
            if (small[l]) {
                if (leader < 0) leader = (int)l;
                off[l] = total;
                total += nb;
            }
        }
        for (uint32_t l = 0; l < m; ++l) {
            u64 pg;
            if (small[l]) pg = (int)l == leader ? (total + page_size - 1u) / page_size : 0;
            else pg = (need_b[l] + page_size - 1u) / page_size;
            cnt[l] = pg > 64u ? 65u : (uint32_t)pg;
        }
        int32_t ids[64];
        uint32_t st[64];
        orc_getnpage_spans(bm, n_pages, free_count, m, cnt, req_base + (u64)ch * 64u, seed,
                           budget, spans, ids, st);
        for (uint32_t l = 0; l < m; ++l) {
            int32_t pid = small[l] ? ids[leader] : ids[l];
            page[ch * 64u + l] = pid;
            offset[ch * 64u + l] = small[l] ? off[l] : 0;
            header[ch * 64u + l] = (int32_t)cnt[l];
            if (steps) steps[ch * 64u + l] = st[l];
            if (pid >= 0) got++;
        }
    }
    return got;
}

uint32_t orc_malloc_serial(uint64_t* bm, uint32_t n_pages, uint32_t page_size,
                           int64_t* free_count, uint32_t n, const uint64_t* sizes,
                           uint64_t req_base, uint64_t seed, uint32_t budget, int32_t* page,
                           uint64_t* offset, int32_t* header, uint32_t* steps) {
    return orc_malloc_spans(bm, n_pages, page_size, free_count, n, sizes, req_base, seed, budget,
                            1u, page, offset, header, steps);
}

/* ------------------------------------------------------------ free */

/* freePage (RandomWalkBasic_BM.cuh:98-102): clear the page's bit. */
uint32_t orc_free(uint64_t* bm, uint32_t n_pages, int64_t* free_count, uint32_t n,
                  const int32_t* ids) {
    uint32_t freed = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t id = ids[i];
        if (id < 0 || (uint32_t)id >= n_pages) continue;
        u64 bit = (u64)1 << ((uint32_t)id & 63u);
        if (bm[(uint32_t)id >> 6] & bit) {
            bm[(uint32_t)id >> 6] &= ~bit;
            freed++;
        }
    }
    *free_count += freed;
    return freed;
}

/* ------------------------------------------------------------ checks */

int orc_check_alloc(const uint64_t* before, const uint64_t* after, uint32_t n_pages, uint32_t n,
                    const int32_t* ids, uint32_t* n_granted) {
    uint32_t L = orc_n_words(n_pages);
    u64* seen = (u64*)calloc(L ? L : 1, sizeof(u64));
    int rc = 0;
    uint32_t g = 0;
    if (!seen) return 99;
    for (uint32_t i = 0; i < n && !rc; ++i) {
        int32_t id = ids[i];
        if (id < 0) continue;
        if ((uint32_t)id >= n_pages) { rc = 1; break; }
        uint32_t w = (uint32_t)id >> 6;
        u64 bit = (u64)1 << ((uint32_t)id & 63u);
        if (before[w] & bit) { rc = 2; break; }
        if (seen[w] & bit) { rc = 3; break; }
        if (!(after[w] & bit)) { rc = 4; break; }
        seen[w] |= bit;
        g++;
    }
    for (uint32_t w = 0; w < L && !rc; ++w)
        if (after[w] != (before[w] | seen[w])) rc = 5;
    free(seen);
    if (n_granted) *n_granted = g;
    return rc;
}

/* RWBM_TAS (theoretical_models.py:34-41) for words of W bits: the k-th
 * grant sees T-A+k used pages and needs 1/(1-((T-A+k)/T)^W) probes on
 * average.  The reference sums k < N-1 and divides by N; this is the mean
 * over all n grants. */
double orc_rwbm_expected_steps(double T, double A, double n, int W) {
    double s = 0;
    for (double k = 0; k < n; k += 1.0) s += 1.0 / (1.0 - pow((T - A + k) / T, (double)W));
    return n > 0 ? s / n : 0.0;
}

/* ------------------------------------------------------ serving loop */

/* Payload of page j of sequence seq (pbm_device.hpp serve_payload). */
uint64_t orc_serve_payload(uint64_t seed, uint32_t seq, uint32_t j) {
    return orc_mix64((seed ^ 0x452821E638D01377ull) +
                     (((((u64)seq) << 32) | j) + 1ull) * 0x9E3779B97F4A7C15ull);
}

/* Checksum a sequence of `len` pages ends with: the wrapping sum of its
 * payloads. */
uint64_t orc_serve_checksum(uint64_t seed, uint32_t seq, uint32_t len) {
    u64 s = 0;
    for (uint32_t j = 0; j < len; ++j) s += orc_serve_payload(seed, seq, j);
    return s;
}

/* The serving loop's schedule, counts only (pbm_serve.hip k_serve_admit /
 * k_serve_grow / k_serve_release), which fix it: at step k = 1, 2, ...
 * waiting sequences are admitted in arrival order, at most 1024 examined
 * per step, while arrival <= k, a slot is free and free pages - pages owed
 * to admitted sequences >= the target; then every admitted sequence gets
 * one page, and those that reach their target finish at k and return all
 * their pages before step k+1.  Writes admit/finish steps (0 = never) and
 * totals[0..3] = allocs, frees, admitted, finished. */
void orc_serve_sim(int64_t free_pages, uint32_t max_slots, uint32_t n_seqs,
                   const uint32_t* arrival, const uint32_t* target, uint32_t n_steps,
                   uint32_t* admit_step, uint32_t* finish_step, uint64_t* totals) {
    uint32_t* act = (uint32_t*)malloc(sizeof(uint32_t) * (max_slots ? max_slots : 1u));
    uint32_t* len = (uint32_t*)calloc(n_seqs ? n_seqs : 1u, sizeof(uint32_t));
    if (!act || !len) { free(act); free(len); return; }
    for (uint32_t i = 0; i < n_seqs; ++i) admit_step[i] = finish_step[i] = 0;
    uint32_t n_act = 0, head = 0;
    int64_t owed = 0;
    u64 allocs = 0, frees = 0, admitted = 0, finished = 0;
    for (uint32_t k = 1; k <= n_steps; ++k) {
        uint32_t win = n_seqs - head < 1024u ? n_seqs - head : 1024u;
        int64_t budget = free_pages - owed;
        uint32_t m = 0;
        while (m < win && arrival[head + m] <= k && n_act < max_slots &&
               budget >= (int64_t)target[head + m]) {
            budget -= target[head + m];
            owed += target[head + m];
            admit_step[head + m] = k;
            act[n_act++] = head + m;
            ++m;
        }
        head += m;
        admitted += m;
        /* grow: one page each */
        free_pages -= n_act;
        owed -= n_act;
        allocs += n_act;
        uint32_t keep = 0;
        for (uint32_t i = 0; i < n_act; ++i) {
            uint32_t q = act[i];
            if (++len[q] == target[q]) {
                finish_step[q] = k;
                free_pages += len[q];
                frees += len[q];
                finished++;
            } else {
                act[keep++] = q;
            }
        }
        n_act = keep;
    }
    if (totals) { totals[0] = allocs; totals[1] = frees; totals[2] = admitted; totals[3] = finished; }
    free(act);
    free(len);
}

/* ------------------------------------------- buffer cache (CLOCK) */

/* Zipf(s) over n logical pages: cdf[k] = sum_{j<=k} (j+1)^-s, summed in
 * index order in double.  pbm_cache.hip builds the same table the same
 * way. */
void orc_zipf_cdf(uint32_t n, double s, double* cdf) {
    double acc = 0;
    for (uint32_t k = 0; k < n; ++k) {
        acc += pow((double)k + 1.0, -s);
        cdf[k] = acc;
    }
}

/* Reference i of the trace: u = the top 53 bits of mix64 of (seed, i) as a
 * double in [0,1), page = the first k with cdf[k] > u * cdf[n-1]. */
int32_t orc_zipf_sample(const double* cdf, uint32_t n, uint64_t seed, uint64_t i) {
    const u64 h = orc_mix64((seed ^ 0xBE5466CF34E90C6Cull) + (i + 1ull) * 0x9E3779B97F4A7C15ull);
    const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
    const double t = u * cdf[n - 1];
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2u;
        if (cdf[mid] > t) hi = mid; else lo = mid + 1u;
    }
    return (int32_t)lo;
}

void orc_zipf_trace(uint32_t n, double s, uint64_t seed, uint64_t first, uint64_t count,
                    int32_t* out) {
    double* cdf = (double*)malloc(sizeof(double) * n);
    if (!cdf) return;
    orc_zipf_cdf(n, s, cdf);
    for (uint64_t i = 0; i < count; ++i) out[i] = orc_zipf_sample(cdf, n, seed, first + i);
    free(cdf);
}

/* The buffer cache of pbm_cache.hip: `capacity` slots, CLOCK replacement,
 * references taken `batch` at a time.  A batch (i) sets the reference bit
 * of every slot it hits, as the cache stood when the batch began; (ii)
 * counts the first reference to each absent page as a miss and later ones
 * in the batch as hits; (iii) fills empty slots in order with the first
 * misses, and for the rest takes victims by CLOCK over the slots occupied
 * when the batch began, in ring order from the arm: a slot with its bit
 * set has it cleared, one without is a victim; if one pass finds too few,
 * every bit is clear and the next pass takes slots in ring order again;
 * the arm moves past the last victim; (iv) loads the misses, in order,
 * into the empty slots and then the victims' slots, with the bit set.
 * With batch = 1 this is the textbook second-chance CLOCK.
 * state: slot_of[n_logical] (-1 absent), owner[capacity], refbit[capacity],
 * arm_res[2] = arm, resident.  stats += hits, misses, evictions.
 * Evicted pages go to evict_log in order while it has room. */
void orc_clock_run(uint32_t n_logical, uint32_t capacity, uint32_t batch, const int32_t* refs,
                   uint64_t n_refs, int32_t* slot_of, int32_t* owner, uint8_t* refbit,
                   uint32_t* arm_res, int32_t* evict_log, uint64_t log_cap, uint64_t* stats) {
    const uint32_t C = capacity;
    uint32_t arm = arm_res[0], resident = arm_res[1];
    uint32_t* stamp = (uint32_t*)calloc(n_logical ? n_logical : 1u, sizeof(uint32_t));
    int32_t* miss = (int32_t*)malloc(sizeof(int32_t) * (batch ? batch : 1u));
    int32_t* victim = (int32_t*)malloc(sizeof(int32_t) * (batch ? batch : 1u));
    int32_t* ones = (int32_t*)malloc(sizeof(int32_t) * (C ? C : 1u));
    uint32_t* ones_pos = (uint32_t*)malloc(sizeof(uint32_t) * (C ? C : 1u));
    if (!stamp || !miss || !victim || !ones || !ones_pos) goto out;
    uint32_t bno = 0;
    for (uint64_t b0 = 0; b0 < n_refs; b0 += batch, ++bno) {
        const uint32_t nb = n_refs - b0 < batch ? (uint32_t)(n_refs - b0) : batch;
        uint32_t K = 0;
        for (uint32_t i = 0; i < nb; ++i) {
            const int32_t l = refs[b0 + i];
            if ((uint32_t)l >= n_logical) continue;  /* skipped: neither hit nor miss */
            const int32_t s = slot_of[l];
            if (s >= 0) {
                refbit[s] = 1;
                stats[0]++;
            } else if (stamp[l] != bno + 1u) {
                stamp[l] = bno + 1u;
                miss[K++] = l;
                stats[1]++;
            } else {
                stats[0]++;
            }
        }
        const uint32_t r0 = resident, empty = C - r0;
        const uint32_t v = K > empty ? K - empty : 0u;
        if (v > 0) {
            uint32_t z = 0, n1 = 0, stop = 0;
            for (uint32_t j = 0; j < C; ++j) {
                const uint32_t s = (arm + j) % C;
                if (s < r0 && !refbit[s]) z++;
            }
            uint32_t nv = 0;
            if (v <= z) {
                for (uint32_t j = 0; j < C && nv < v; ++j) {
                    const uint32_t s = (arm + j) % C;
                    if (s >= r0) continue;
                    if (refbit[s]) {
                        refbit[s] = 0;
                    } else {
                        victim[nv++] = (int32_t)s;
                        stop = j;
                    }
                }
            } else {
                for (uint32_t j = 0; j < C; ++j) {
                    const uint32_t s = (arm + j) % C;
                    if (s >= r0) continue;
                    if (refbit[s]) {
                        refbit[s] = 0;
                        ones[n1] = (int32_t)s;
                        ones_pos[n1++] = j;
                    } else {
                        victim[nv++] = (int32_t)s;
                    }
                }
                for (uint32_t k = 0; k < v - z; ++k) {
                    victim[nv++] = ones[k];
                    stop = ones_pos[k];
                }
            }
            arm = (arm + stop + 1u) % C;
        }
        for (uint32_t i = 0; i < K; ++i) {
            uint32_t s;
            if (i < empty) {
                s = r0 + i;
            } else {
                s = (uint32_t)victim[i - empty];
                const int32_t old = owner[s];
                slot_of[old] = -1;
                if (evict_log && stats[2] < log_cap) evict_log[stats[2]] = old;
                stats[2]++;
            }
            owner[s] = miss[i];
            slot_of[miss[i]] = (int32_t)s;
            refbit[s] = 1;
        }
        resident = r0 + K < C ? r0 + K : C;
    }
    arm_res[0] = arm;
    arm_res[1] = resident;
out:
    free(stamp);
    free(miss);
    free(victim);
    free(ones);
    free(ones_pos);
}

/* Textbook second-chance CLOCK, one reference at a time (the batch rule
 * at batch 1, in time linear in the references rather than in the frames
 * per miss): a hit sets the frame's bit; a miss fills the next empty frame,
 * else the hand clears set bits until it finds a clear one, replaces that
 * page and moves past it.  Loaded pages start with the bit set.  Every
 * bit the hand clears was set by a reference, so the sweeps total at most
 * the references.  References outside 0 .. n_logical-1 are skipped.
 * stats = {hits, misses, evictions}; the page evicted k-th goes to
 * evict_log[k] while k < log_cap. */
void orc_clock_textbook(uint32_t n_logical, uint32_t capacity, const int32_t* refs,
                        uint64_t n_refs, int32_t* slot_of, int32_t* owner, uint8_t* refbit,
                        uint32_t* arm_res, int32_t* evict_log, uint64_t log_cap,
                        uint64_t* stats) {
    const uint32_t C = capacity;
    uint32_t hand = arm_res[0], resident = arm_res[1];
    for (uint64_t i = 0; i < n_refs; ++i) {
        const int32_t l = refs[i];
        if ((uint32_t)l >= n_logical) continue;
        const int32_t s0 = slot_of[l];
        if (s0 >= 0) {
            refbit[s0] = 1;
            stats[0]++;
            continue;
        }
        stats[1]++;
        uint32_t s;
        if (resident < C) {
            s = resident++;
        } else {
            while (refbit[hand]) {
                refbit[hand] = 0;
                hand = hand + 1u == C ? 0u : hand + 1u;
            }
            s = hand;
            const int32_t old = owner[s];
            slot_of[old] = -1;
            if (evict_log && stats[2] < log_cap) evict_log[stats[2]] = old;
            stats[2]++;
            hand = hand + 1u == C ? 0u : hand + 1u;
        }
        owner[s] = l;
        slot_of[l] = (int32_t)s;
        refbit[s] = 1;
    }
    arm_res[0] = hand;
    arm_res[1] = resident;
}

/* --------------------------------------------------- clusters */

/* Maximal runs of free pages, circular; a fully free pool is one run. */
uint64_t orc_count_clusters(const uint64_t* bm, uint32_t n_pages) {
    uint64_t c = 0, free_pages = 0;
    for (uint32_t i = 0; i < n_pages; ++i) {
        const uint32_t prev = i ? i - 1u : n_pages - 1u;
        const int f = !((bm[i >> 6] >> (i & 63u)) & 1u);
        const int fp = !((bm[prev >> 6] >> (prev & 63u)) & 1u);
        free_pages += (uint64_t)f;
        if (f && !fp) c++;
    }
    if (c == 0 && free_pages == n_pages) c = 1;
    return c;
}
