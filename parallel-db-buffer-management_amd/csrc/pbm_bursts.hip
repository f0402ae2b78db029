// Bursts of mallocRW / freeRW requests: each burst is a free phase (freeRW
// of earlier bursts' allocations) then an alloc phase (mallocRW of its own
// requests).  Two ways to run a range of bursts:
//
//   phase kernels  one free kernel and one alloc kernel per burst (plain,
//                  serial or a batched grid), capturable in a HIP graph
//   persistent     PBM_LAUNCH_PERSISTENT: the whole range in one kernel.
//                  The host cuts every burst into work items -- free items
//                  of at most 64 frees, then alloc items of one 64-request
//                  mallocRW chunk each -- and waves take the items in order
//                  from an atomic ticket.  Dependencies are waited on where
//                  they are, not with grid barriers: a free waits until the
//                  pointer it frees has been published.  An alloc item of
//                  burst b waits until burst behind[b] has finished, the
//                  earliest burst after which the bursts up to b ask for at
//                  most half the pool (from the trace's sizes), so the
//                  allocations running ahead of the frees that later give
//                  their pages back stay bounded.  It runs ahead of its own
//                  burst's frees while a quarter of the pool is unreserved,
//                  and otherwise (or with PBM_LAUNCH_ORDERED) searches only
//                  once every free item of its burst is done; one that ran
//                  ahead and found no run waits for them and tries again.
//                  So bursts overlap wherever the trace allows it, and a
//                  failure is reported only after the burst's frees have
//                  returned their pages.
//
// Progress: a wave waits only on items with lower tickets.  A wave takes a
// ticket only while running, and takes its next one only once its current
// item has nothing left to wait for, so every item taken is being worked on
// by a running wave: the lowest unfinished item always is, and its own
// dependencies are done.  (Fetching the next ticket ahead, while the
// current item may still wait, leaves items that others wait on held by a
// wave that is itself waiting; such chains grew with the wave count.)
// Every wait is bounded anyway (2 s of wall clock); a wait that gives up
// is counted, and the run's results are then not to be trusted.
//
// Memory ordering across workgroups (and XCDs, whose L2 caches are not
// coherent with each other): a pointer is published with a device-scope
// release store after its header, and a free item loads it and then
// acquires before reading the header, so the header it reads is the one
// written for that pointer, never an earlier owner's still held by its
// XCD's L2.  A free item publishes nothing but device-scope atomic
// read-modify-writes (map bits, reservation counter); it waits for them to
// complete before it counts itself done, and the alloc items read those
// words only with device-scope atomics, so no fence is needed that way.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <vector>
#include "pbm_internal.hpp"

using pbm::u64;

namespace {

constexpr u64 kWaitTicks = 200000000ull;  // 2 s of the 100 MHz wall clock: a wait gives up
// A table entry whose mallocRW has not run yet.  mallocRW returns 4-byte
// aligned pointers (or 0 on failure), so 1 is never a result.
constexpr u64 kPending = 1ull;
constexpr uint32_t kMaxBursts = 1u << 24;  // burst numbers fit the item word

struct BurstDev {
    const uint32_t* alloc_off;   // [n_bursts + 1]
    const uint32_t* free_off;    // [n_bursts + 1]
    const uint64_t* sizes;       // [n_allocs]
    const uint32_t* free_index;  // [n_frees] in trace order (phase kernels)
    const uint32_t* free_order;  // [n_frees] grouped by packed run (persistent items)
    const uint2* items;          // persistent work items, burst by burst
    const uint32_t* n_fitems;    // [n_bursts] free items of each burst
    uint32_t* fdone;             // [n_bursts] free items finished
    uint32_t* adone;             // [n_bursts] alloc items finished
    const uint32_t* behind;      // [n_bursts] burst b's alloc items wait until burst
                                 // behind[b] has finished (behind[b] == b: no wait)
    uint64_t* table;             // [n_allocs] pointer returned for each request
    u64* ctl;                    // [0] ticket, [1] waits that gave up
};

// Work item: x = first request (alloc) or first free_order entry (free);
// y = burst << 8 | count << 1 | is_alloc, count in 1..64.
__host__ __device__ inline uint32_t item_word(uint32_t burst, uint32_t count, bool alloc) {
    return (burst << 8) | (count << 1) | (alloc ? 1u : 0u);
}

__device__ __forceinline__ uint32_t ld_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_u64(const uint64_t* p) {
    return __hip_atomic_load((const u64*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_u64(uint64_t* p, u64 v) {
    __hip_atomic_store((u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Free phase: thread i frees the allocation free_index[i] points at.
__global__ __launch_bounds__(1024) void k_burst_free(pbm_ctx c, BurstDev t, uint32_t f0,
                                                     uint32_t f1) {
    pbm::Stats st;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = f0 + blockIdx.x * blockDim.x + threadIdx.x; i < f1; i += stride)
        pbm::free_rw(c, (void*)t.table[t.free_index[i]], &st);
    pbm::flush_stats(c, st);
}

// Alloc phase: mallocRW, wave-collective over each 64 consecutive
// requests of the burst; the walk seed of a chunk is the seed of its first
// request's global index (the plain mallocRW launch's rule).
__device__ inline void alloc_chunks(const pbm_ctx& c, const BurstDev& t, uint32_t a0, uint32_t a1,
                                    uint32_t wave, uint32_t n_waves, uint32_t spans,
                                    pbm::Stats& st) {
    const unsigned lane = pbm::lane_id();
    const uint32_t n = a1 - a0;
    const uint32_t n_chunks = (n + 63u) / 64u;
    for (uint32_t ch = wave; ch < n_chunks; ch += n_waves) {
        const uint32_t r = ch * 64u + lane;
        if (r < n) {
            void* p = pbm::malloc_rw(c, t.sizes[a0 + r],
                                     pbm::request_seed(c.seed, (u64)a0 + (u64)ch * 64u), nullptr,
                                     &st, spans);
            t.table[a0 + r] = (uint64_t)p;
        }
    }
}

__global__ __launch_bounds__(1024) void k_burst_alloc(pbm_ctx c, BurstDev t, uint32_t a0,
                                                      uint32_t a1, uint32_t spans) {
    pbm::Stats st;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    alloc_chunks(c, t, a0, a1, wave, n_waves, spans, st);
    pbm::flush_stats(c, st);
}

// Before a persistent run of bursts b0 .. b1-1: their requests' table
// entries become pending, their free-item counts zero, the ticket zero.
__global__ void k_bursts_prepare(BurstDev t, uint32_t a0, uint32_t a1, uint32_t b0,
                                 uint32_t b1) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = a0 + tid; i < a1; i += stride) t.table[i] = kPending;
    for (uint32_t b = b0 + tid; b < b1; b += stride) t.fdone[b] = t.adone[b] = 0u;
    if (tid == 0) t.ctl[0] = 0ull;
}

// Polls `ready` until it holds, sleeping longer the longer it waits (every
// poll is a device-scope load performed at the memory side, and hundreds of
// waiting waves polling at full rate slow the waves they wait for).  Gives
// up (returns false) after kWaitTicks of the 100 MHz wall clock.
template <typename Ready>
__device__ inline bool wait_until(Ready ready) {
    if (ready()) return true;
    const u64 t0 = wall_clock64();
    for (uint32_t spins = 1;; ++spins) {
        if (spins < 8u) __builtin_amdgcn_s_sleep(2);
        else if (spins < 32u) __builtin_amdgcn_s_sleep(8);
        else __builtin_amdgcn_s_sleep(32);
        if (ready()) return true;
        if ((spins & 15u) == 0u && wall_clock64() - t0 > kWaitTicks) return false;
    }
}

// Work items i0 .. i1-1, one wave per workgroup, items taken by ticket.
__global__ __launch_bounds__(64) void k_bursts_persistent(pbm_ctx c, BurstDev t, uint32_t i0,
                                                          uint32_t i1, uint32_t b0,
                                                          uint32_t spans, long long ahead_min) {
    using namespace pbm;
    Stats st;
    const unsigned lane = lane_id();
    uint32_t gave_up = 0;  // wave-uniform
    u64 tk = 0;
    if (lane == 0) tk = atom_add_u64(&t.ctl[0], 1ull);
    long long room = __hip_atomic_load(c.free_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // burst seen_done-1 is known to have finished (the window's burst only
    // moves forward for a wave: its tickets, and so its bursts, increase,
    // and behind[] does not decrease); a burst before b0 finished with the
    // launches before this one
    uint32_t seen_done = b0;
    u64 it = (u64)i0 + readlane64(tk, 0);
    while (it < (u64)i1) {
        // the next ticket is taken once this item has nothing left to wait
        // for, and is in flight while the item finishes
        u64 nt = 0;
        bool took = false;
        auto take_next = [&] {
            if (took) return;
            if (lane == 0) nt = atom_add_u64(&t.ctl[0], 1ull);
            took = true;
        };
        const uint2 item = t.items[it];
        const uint32_t start = item.x;
        const uint32_t burst = item.y >> 8;
        const uint32_t cnt = (item.y >> 1) & 127u;
        if (item.y & 1u) {
            // alloc item: one mallocRW chunk.  While at least ahead_min
            // pages are unreserved it runs ahead of its burst's frees; below
            // that it waits for them before it reserves (so the items of
            // bursts far ahead cannot hold pages the bursts before them
            // need).  An item that ran ahead and found no run for a lane
            // waits for its burst's frees and tries that lane once more, so
            // a failure is reported only once every page the burst's frees
            // return is back.  Its first window is the item's home, item i
            // at word L * frac(i * phi), so the chunks running together
            // start apart; random jumps from the chunk's seed follow.  The
            // home window is read while the reservation is in flight.
            // (room: the reservation counter as this wave's last reserve
            // saw it, so the decision costs no extra round trip)
            RunPlan pl{};
            if (lane < cnt) pl = plan_malloc_rw(c, t.sizes[start + lane]);
            // window: the bursts after behind[burst] ask for at most half
            // the pool's pages, and it must have finished (all its items)
            // before this item allocates, so the allocations running ahead
            // of the frees the trace gives back after them stay bounded
            const uint32_t wb = t.behind[burst];
            if (wb < burst && wb >= seen_done) {
                const uint32_t nf = t.n_fitems[wb];
                const uint32_t na = (t.alloc_off[wb + 1] - t.alloc_off[wb] + 63u) / 64u;
                if (!wait_until([&] {
                        return ld_u32(&t.fdone[wb]) >= nf && ld_u32(&t.adone[wb]) >= na;
                    }))
                    gave_up++;
                seen_done = wb + 1u;
            }
            bool done = false;
            if (room < ahead_min) {
                const uint32_t nf = t.n_fitems[burst];
                if (!wait_until([&] { return ld_u32(&t.fdone[burst]) >= nf; })) gave_up++;
                done = true;
                take_next();
            }
            const uint32_t want = (lane < cnt && pl.n <= 64u) ? pl.n : 0u;
            int32_t pid = -1;
            bool held = false;
            const int32_t seed = request_seed(c.seed, start);
            if (lane < cnt) {
                const uint32_t home =
                    (uint32_t)(((u64)(uint32_t)((uint32_t)it * 2654435769u) * c.n_words) >> 32);
                const FirstWindow win = load_window(c, home, spans);
                held = reserve_n(c, want, &room);
                pid = find_n_pages(c, pl.n, held, seed, nullptr, st, spans, &win);
            }
            room = (long long)readlane64((u64)room, 0);  // lane 0 is always in the item
            if (!done && ballot(lane < cnt && want > 0u && pid < 0)) {
                const uint32_t nf = t.n_fitems[burst];
                if (!wait_until([&] { return ld_u32(&t.fdone[burst]) >= nf; })) gave_up++;
                const bool again = lane < cnt && want > 0u && pid < 0;
                if (again) {
                    // the first try counted a failure, and an error when it
                    // held a reservation it could not place
                    st.failures--;
                    if (held) st.errors--;
                }
                if (lane < cnt) {
                    const bool held2 = reserve_n(c, again ? want : 0u, &room);
                    const int32_t pid2 =
                        find_n_pages(c, again ? want : 0u, held2, seed, nullptr, st, spans);
                    if (again) pid = pid2;
                }
                room = (long long)readlane64((u64)room, 0);
            }
            take_next();
            if (lane < cnt) {
                void* p = place_malloc_rw(c, pl, pid);
                // publish: the header written before the pointer is visible
                // with it (release; the free's load is followed by an acquire)
                __hip_atomic_store((u64*)&t.table[start + lane], (u64)p, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            if (lane == 0) __hip_atomic_fetch_add(&t.adone[burst], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        } else {
            // free item: whole packed runs, so every member reads its header
            // (one load for the wave) before any run of the item is released
            u64 v = kPending;
            uint32_t r = 0;
            if (lane < cnt) {
                r = t.free_order[start + lane];
                v = ld_u64(&t.table[r]);
            }
            if (ballot(v == kPending && lane < cnt)) {
                // one pending pointer is polled per round; once it is
                // published (its chunk's pointers are published together)
                // every pending lane reads its own again
                const bool ok = wait_until([&] {
                    const u64 pend = ballot(lane < cnt && v == kPending);
                    if (!pend) return true;
                    const unsigned q = (unsigned)__ffsll((long long)pend) - 1u;
                    if (lane == q) v = ld_u64(&t.table[r]);
                    if (readlane64(v, q) == kPending) return false;
                    if (lane < cnt && v == kPending) v = ld_u64(&t.table[r]);
                    return !ballot(lane < cnt && v == kPending);
                });
                if (!ok) gave_up++;
            }
            take_next();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (lane < cnt && v != kPending) free_rw(c, (void*)v, &st);
            // every release has landed before the item counts as done.  The
            // releases are device-scope atomics on the map and the
            // reservation counter, performed where every workgroup sees
            // them, and they are all the item publishes to the alloc items:
            // waiting for them to complete orders them before the count.
            __builtin_amdgcn_s_waitcnt(0);
            if (lane == 0) __hip_atomic_fetch_add(&t.fdone[burst], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        }
        take_next();
        it = (u64)i0 + readlane64(nt, 0);
    }
    if (lane == 0 && gave_up) atom_add_u64(&t.ctl[1], (u64)gave_up);
    flush_stats(c, st);
}

__global__ void k_zero_u64(uint64_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace

struct pbm_bursts {
    pbm_handle* h = nullptr;
    uint32_t n_bursts = 0, n_allocs = 0, n_frees = 0;
    std::vector<uint32_t> alloc_off, free_off, item_off;
    BurstDev d{};
    std::vector<void*> mem;
};

extern "C" int pbm_bursts_destroy(pbm_bursts* b) {
    if (!b) return PBM_OK;
    if (b->h) (void)pbmi::use_device(b->h);
    (void)hipDeviceSynchronize();
    for (void* p : b->mem)
        if (p) (void)hipFree(p);
    delete b;
    return PBM_OK;
}

namespace {

// Work items of every burst, and the frees regrouped so that the requests
// mallocRW packed into one run (the small requests of one 64-request chunk
// of a burst) are freed by one item.  Rejects a trace that frees a request
// twice or frees one packed run over several bursts: freeing the leader
// returns the whole run, so a member freed later would read a header that
// the run's next owner may have overwritten.
int plan_items(const pbm_burst_trace* tr, uint32_t page_size, uint32_t n_pages,
               std::vector<uint2>& items, std::vector<uint32_t>& item_off,
               std::vector<uint32_t>& n_fitems, std::vector<uint32_t>& free_order,
               std::vector<uint32_t>& behind) {
    const uint32_t nb = tr->n_bursts;
    const uint32_t na = tr->alloc_offset[nb];
    // chunk start of each request, and whether mallocRW packs it
    std::vector<uint32_t> chunk_of(na);
    std::vector<uint8_t> small(na), freed(na, 0);
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t a0 = tr->alloc_offset[b], a1 = tr->alloc_offset[b + 1];
        for (uint32_t r = a0; r < a1; ++r) {
            chunk_of[r] = a0 + (r - a0) / 64u * 64u;
            const uint64_t need = (tr->sizes[r] + 4u + 3u) & ~(uint64_t)3u;
            small[r] = need < page_size;
        }
    }
    std::vector<uint32_t> run_burst(na, UINT32_MAX);  // by chunk start: burst freeing its run
    items.clear();
    item_off.assign(nb + 1, 0);
    n_fitems.assign(nb, 0);
    free_order.clear();
    free_order.reserve(tr->free_offset[nb]);
    std::vector<std::pair<u64, uint32_t>> keyed;
    for (uint32_t b = 0; b < nb; ++b) {
        item_off[b] = (uint32_t)items.size();
        keyed.clear();
        for (uint32_t k = tr->free_offset[b]; k < tr->free_offset[b + 1]; ++k) {
            const uint32_t r = tr->free_index[k];
            if (freed[r]) return pbmi::fail(PBM_ERR_INVALID, "a request is freed twice");
            freed[r] = 1;
            u64 key = r;
            if (small[r]) {
                const uint32_t cs = chunk_of[r];
                if (run_burst[cs] != UINT32_MAX && run_burst[cs] != b)
                    return pbmi::fail(PBM_ERR_INVALID,
                                      "the requests of one packed mallocRW run must be freed in "
                                      "one burst");
                run_burst[cs] = b;
                key = (1ull << 63) | cs;
            }
            keyed.emplace_back(key, r);
        }
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const std::pair<u64, uint32_t>& x, const std::pair<u64, uint32_t>& y) {
                             return x.first < y.first;
                         });
        uint32_t item_start = (uint32_t)free_order.size(), item_n = 0;
        for (size_t i = 0; i < keyed.size();) {
            size_t j = i;
            while (j < keyed.size() && keyed[j].first == keyed[i].first) ++j;
            const uint32_t g = (uint32_t)(j - i);  // at most 64: one chunk's requests
            if (item_n + g > 64u) {
                items.push_back(make_uint2(item_start, item_word(b, item_n, false)));
                n_fitems[b]++;
                item_start = (uint32_t)free_order.size();
                item_n = 0;
            }
            for (size_t q = i; q < j; ++q) free_order.push_back(keyed[q].second);
            item_n += g;
            i = j;
        }
        if (item_n) {
            items.push_back(make_uint2(item_start, item_word(b, item_n, false)));
            n_fitems[b]++;
        }
        const uint32_t a0 = tr->alloc_offset[b], a1 = tr->alloc_offset[b + 1];
        for (uint32_t s = a0; s < a1; s += 64u) {
            const uint32_t cnt = a1 - s < 64u ? a1 - s : 64u;
            items.push_back(make_uint2(s, item_word(b, cnt, true)));
        }
    }
    item_off[nb] = (uint32_t)items.size();
    // pages each burst's mallocRW chunks ask for: a large request its own
    // run, the small ones of a chunk one packed run (runs over 64 pages
    // fail and take nothing)
    std::vector<uint64_t> apages(nb, 0);
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t a0 = tr->alloc_offset[b], a1 = tr->alloc_offset[b + 1];
        for (uint32_t s = a0; s < a1; s += 64u) {
            uint64_t packed = 0;
            for (uint32_t r = s; r < a1 && r < s + 64u; ++r) {
                const uint64_t need = (tr->sizes[r] + 4u + 3u) & ~(uint64_t)3u;
                if (small[r]) {
                    packed += need;
                } else {
                    const uint64_t n = (need + page_size - 1u) / page_size;
                    if (n <= 64u) apages[b] += n;
                }
            }
            const uint64_t n = (packed + page_size - 1u) / page_size;
            if (n <= 64u) apages[b] += n;
        }
    }
    // behind[b]: the earliest burst l < b such that bursts l+1 .. b ask for
    // at most half the pool (b - 1 when b alone asks for more); b itself,
    // meaning no wait, while bursts 0 .. b together ask for at most that
    behind.assign(nb, 0);
    const uint64_t budget = n_pages / 2u;
    std::vector<uint64_t> pre(nb + 1, 0);  // pre[i] = apages[0 .. i-1]
    for (uint32_t b = 0; b < nb; ++b) pre[b + 1] = pre[b] + apages[b];
    for (uint32_t b = 0; b < nb; ++b) {
        if (b == 0 || pre[b + 1] <= budget) {
            behind[b] = b;
            continue;
        }
        const uint64_t target = pre[b + 1] - budget;  // need pre[l + 1] >= target
        const uint32_t j = (uint32_t)(std::lower_bound(pre.begin() + 1, pre.begin() + b + 1,
                                                       target) - pre.begin());
        behind[b] = j <= b ? j - 1u : b - 1u;
    }
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_bursts_create(pbm_handle* h, const pbm_burst_trace* tr, pbm_bursts** out) {
    PBM_CHECK_HANDLE(h);
    if (!tr || !out) return pbmi::fail(PBM_ERR_INVALID, "null trace or output");
    *out = nullptr;
    if (!h->d_pool) return pbmi::fail(PBM_ERR_UNSUPPORTED, "mallocRW bursts need a backed pool");
    if (pbmi::is_queue_strategy(h->cfg.strategy))
        return pbmi::fail(PBM_ERR_UNSUPPORTED, "mallocRW bursts need a bitmap strategy");
    if (tr->n_bursts == 0 || !tr->alloc_offset || !tr->free_offset)
        return pbmi::fail(PBM_ERR_INVALID, "a trace needs bursts and both offset arrays");
    const uint32_t nb = tr->n_bursts;
    if (nb > kMaxBursts) return pbmi::fail(PBM_ERR_RANGE, "at most 2^24 bursts");
    if (tr->alloc_offset[0] != 0 || tr->free_offset[0] != 0)
        return pbmi::fail(PBM_ERR_INVALID, "offsets must start at 0");
    for (uint32_t i = 0; i < nb; ++i)
        if (tr->alloc_offset[i + 1] < tr->alloc_offset[i] || tr->free_offset[i + 1] < tr->free_offset[i])
            return pbmi::fail(PBM_ERR_INVALID, "offsets must not decrease");
    const uint32_t na = tr->alloc_offset[nb], nf = tr->free_offset[nb];
    if ((na && !tr->sizes) || (nf && !tr->free_index))
        return pbmi::fail(PBM_ERR_INVALID, "null sizes or free_index");
    for (uint32_t i = 0; i < nb; ++i)
        for (uint32_t k = tr->free_offset[i]; k < tr->free_offset[i + 1]; ++k)
            if (tr->free_index[k] >= tr->alloc_offset[i])
                return pbmi::fail(PBM_ERR_INVALID, "a free must target an earlier burst's request");
    std::vector<uint2> items;
    std::vector<uint32_t> item_off, n_fitems, free_order, behind;
    int rc = plan_items(tr, h->ctx.page_size, h->ctx.n_pages, items, item_off, n_fitems,
                        free_order, behind);
    if (rc) return rc;

    pbm_bursts* b = new pbm_bursts();
    b->h = h;
    b->n_bursts = nb;
    b->n_allocs = na;
    b->n_frees = nf;
    b->alloc_off.assign(tr->alloc_offset, tr->alloc_offset + nb + 1);
    b->free_off.assign(tr->free_offset, tr->free_offset + nb + 1);
    b->item_off = item_off;
    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8u);
        if (e != hipSuccess) return pbmi::hip_fail(e, "hipMalloc(bursts)", __FILE__, __LINE__);
        b->mem.push_back(*p);
        return PBM_OK;
    };
    void* p[12] = {};
    const size_t bytes[12] = {(size_t)(nb + 1) * 4u, (size_t)(nb + 1) * 4u, (size_t)na * 8u,
                              (size_t)nf * 4u,       (size_t)nf * 4u,       items.size() * 8u,
                              (size_t)nb * 4u,       (size_t)nb * 4u,       (size_t)na * 8u,
                              64u,                   (size_t)nb * 4u,       (size_t)nb * 4u};
    for (int i = 0; i < 12 && !rc; ++i) rc = alloc(&p[i], bytes[i]);
    if (rc) {
        pbm_bursts_destroy(b);
        return rc;
    }
    b->d.alloc_off = (const uint32_t*)p[0];
    b->d.free_off = (const uint32_t*)p[1];
    b->d.sizes = (const uint64_t*)p[2];
    b->d.free_index = (const uint32_t*)p[3];
    b->d.free_order = (const uint32_t*)p[4];
    b->d.items = (const uint2*)p[5];
    b->d.n_fitems = (const uint32_t*)p[6];
    b->d.fdone = (uint32_t*)p[7];
    b->d.table = (uint64_t*)p[8];
    b->d.ctl = (u64*)p[9];
    b->d.adone = (uint32_t*)p[10];
    b->d.behind = (const uint32_t*)p[11];
    hipError_t e = hipSuccess;
    const void* src[8] = {tr->alloc_offset, tr->free_offset, tr->sizes, tr->free_index,
                          free_order.data(), items.data(), n_fitems.data(), nullptr};
    for (int i = 0; i < 7 && e == hipSuccess; ++i)
        if (bytes[i] && src[i] && !(i == 2 && !na) && !((i == 3 || i == 4) && !nf))
            e = hipMemcpy(p[i], src[i], bytes[i], hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p[7], 0, bytes[7]);
    if (e == hipSuccess) e = hipMemset(p[10], 0, bytes[10]);
    if (e == hipSuccess) e = hipMemcpy(p[11], behind.data(), bytes[11], hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p[8], 0, bytes[8] ? bytes[8] : 8u);
    if (e == hipSuccess) e = hipMemset(p[9], 0, 64u);
    if (e != hipSuccess) {
        pbm_bursts_destroy(b);
        return pbmi::hip_fail(e, "bursts upload", __FILE__, __LINE__);
    }
    *out = b;
    return PBM_OK;
}

namespace {

int burst_launch(pbm_bursts* b, uint32_t first, uint32_t count, const pbm_launch* in,
                 hipStream_t s) {
    pbm_handle* h = b->h;
    if ((uint64_t)first + count > b->n_bursts) return pbmi::fail(PBM_ERR_RANGE, "burst range beyond the trace");
    pbm_launch L{};
    if (in) L = *in;
    if (L.block == 0) L.block = 256;
    if (L.block % 64 != 0 || L.block > 1024)
        return pbmi::fail(PBM_ERR_INVALID, "launch.block must be a multiple of 64, at most 1024");
    if (L.waves_per_cu == 0) L.waves_per_cu = 8;
    if (L.cluster_words == 0) L.cluster_words = 64;
    if (L.cluster_words != 64 && L.cluster_words != 128 && L.cluster_words != 256)
        return pbmi::fail(PBM_ERR_INVALID, "launch.cluster_words must be 64, 128 or 256");
    if (L.mode < PBM_MODE_PLAIN || L.mode > PBM_MODE_SERIAL)
        return pbmi::fail(PBM_ERR_INVALID, "unknown launch mode");
    // events armed by pbm_time_next_kernel time the persistent kernel, and
    // are dropped by a phase-kernel run
    struct Disarm {
        pbm_handle* h;
        ~Disarm() { h->t_start = h->t_stop = nullptr; }
    } disarm{h};
    const uint32_t spans = L.cluster_words / 64u;
    pbm_ctx c = h->ctx;
    if (L.walk_budget) c.walk_budget = L.walk_budget;
    if (L.flags & PBM_LAUNCH_NO_COUNTERS) c.counters = nullptr;
    const BurstDev& d = b->d;
    if (L.flags & PBM_LAUNCH_PERSISTENT) {
        if (L.mode != PBM_MODE_BATCHED)
            return pbmi::fail(PBM_ERR_INVALID, "a persistent launch runs in batched mode");
        if (count == 0) return PBM_OK;
        const uint32_t last = first + count;
        const uint32_t a0 = b->alloc_off[first], a1 = b->alloc_off[last];
        const uint32_t i0 = b->item_off[first], i1 = b->item_off[last];
        uint32_t pb = (a1 - a0 + 255u) / 256u;
        pb = pb < 1u ? 1u : (pb > 2048u ? 2048u : pb);
        hipLaunchKernelGGL(k_bursts_prepare, dim3(pb), dim3(256), 0, s, d, a0, a1, first, last);
        if (i1 > i0) {
            uint64_t waves = (uint64_t)h->n_cu * L.waves_per_cu;
            if (waves > i1 - i0) waves = i1 - i0;
            // run-ahead floor: a quarter of the pool unreserved
            const long long ahead_min = (L.flags & PBM_LAUNCH_ORDERED)
                                            ? LLONG_MAX
                                            : (long long)((h->ctx.n_pages + 3u) / 4u);
            pbmi::launch_timed(h, k_bursts_persistent, dim3((uint32_t)waves), dim3(64), s, c, d,
                               i0, i1, first, spans, ahead_min);
        }
        PBM_HIP(hipGetLastError());
        return PBM_OK;
    }
    for (uint32_t k = first; k < first + count; ++k) {
        const uint32_t f0 = b->free_off[k], f1 = b->free_off[k + 1];
        const uint32_t a0 = b->alloc_off[k], a1 = b->alloc_off[k + 1];
        if (f1 > f0) {
            // serial: one wave.  Frees of distinct runs commute, so the map
            // and the counter after the phase are those of any order.
            dim3 grid(1), blk(64);
            if (L.mode == PBM_MODE_PLAIN) {
                blk = dim3(L.block);
                grid = dim3((f1 - f0 + L.block - 1) / L.block);
            } else if (L.mode == PBM_MODE_BATCHED) {
                blk = dim3(L.block);
                uint64_t thr = (uint64_t)h->n_cu * L.waves_per_cu * 64u;
                if (thr > f1 - f0) thr = f1 - f0;
                grid = dim3((uint32_t)((thr + L.block - 1) / L.block));
            }
            hipLaunchKernelGGL(k_burst_free, grid, blk, 0, s, c, d, f0, f1);
        }
        if (a1 > a0) {
            const uint32_t chunks = (a1 - a0 + 63u) / 64u;
            dim3 grid(1), blk(64);
            if (L.mode == PBM_MODE_PLAIN) {
                blk = dim3(L.block);
                grid = dim3((chunks * 64u + L.block - 1) / L.block);
            } else if (L.mode == PBM_MODE_BATCHED) {
                blk = dim3(L.block);
                uint64_t waves = (uint64_t)h->n_cu * L.waves_per_cu;
                if (waves > chunks) waves = chunks;
                grid = dim3((uint32_t)((waves * 64u + L.block - 1) / L.block));
            }
            hipLaunchKernelGGL(k_burst_alloc, grid, blk, 0, s, c, d, a0, a1, spans);
        }
    }
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_bursts_run(pbm_bursts* b, uint32_t first, uint32_t count,
                              const pbm_launch* launch, void* stream) {
    if (!b) return pbmi::fail(PBM_ERR_INVALID, "null bursts");
    PBM_CHECK_HANDLE(b->h);
    return burst_launch(b, first, count, launch, pbmi::as_stream(stream));
}

extern "C" int pbm_bursts_graph(pbm_bursts* b, uint32_t first, uint32_t count,
                                const pbm_launch* launch, pbm_graph** out) {
    if (!b || !out) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output graph");
    PBM_CHECK_HANDLE(b->h);
    *out = nullptr;
    if (count == 0) return pbmi::fail(PBM_ERR_INVALID, "a graph needs count > 0");
    hipStream_t st = nullptr;
    PBM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    pbm_graph* g = new pbm_graph();
    g->device = b->h->device;
    int rc = PBM_OK;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = pbmi::hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    } else {
        rc = burst_launch(b, first, count, launch, st);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(st, &graph);
        g->graph = graph;
        if (!rc && e != hipSuccess) rc = pbmi::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        if (!rc) {
            e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = pbmi::hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
        }
    }
    (void)hipStreamDestroy(st);
    if (rc) {
        pbm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return PBM_OK;
}

extern "C" int pbm_bursts_read_ptrs(pbm_bursts* b, uint32_t first, uint32_t n, uint64_t* host,
                                    void* stream) {
    if (!b || (n && !host)) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output");
    PBM_CHECK_HANDLE(b->h);
    if ((uint64_t)first + n > b->n_allocs) return pbmi::fail(PBM_ERR_RANGE, "beyond the trace's requests");
    if (n == 0) return PBM_OK;
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(host, b->d.table + first, (size_t)n * 8u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_bursts_reset(pbm_bursts* b, void* stream) {
    if (!b) return pbmi::fail(PBM_ERR_INVALID, "null bursts");
    PBM_CHECK_HANDLE(b->h);
    hipStream_t s = pbmi::as_stream(stream);
    if (b->n_allocs) {
        uint32_t blocks = (b->n_allocs + 255u) / 256u;
        if (blocks > 4096u) blocks = 4096u;
        hipLaunchKernelGGL(k_zero_u64, dim3(blocks), dim3(256), 0, s, b->d.table, b->n_allocs);
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, (uint64_t*)b->d.ctl, 8u);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

extern "C" int pbm_bursts_wait_timeouts(pbm_bursts* b, unsigned long long* out, void* stream) {
    if (!b || !out) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output");
    PBM_CHECK_HANDLE(b->h);
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(out, b->d.ctl + 1, 8u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_bursts_autotune(pbm_bursts* b, uint32_t count, uint32_t flags,
                                   pbm_tune_result* out, void* stream) {
    if (!b || !out) return pbmi::fail(PBM_ERR_INVALID, "null bursts or output");
    pbm_handle* h = b->h;
    PBM_CHECK_HANDLE(h);
    if (count == 0 || count > b->n_bursts) return pbmi::fail(PBM_ERR_RANGE, "count must be in [1, n_bursts]");
    hipStream_t s = pbmi::as_stream(stream);
    const uint32_t L = h->ctx.n_words;
    std::vector<uint64_t> map(L);
    int rc = pbm_snapshot_bitmap(h, map.data(), L, stream);
    if (rc) return rc;
    std::vector<pbm_launch> space;
    const uint32_t wpcs[] = {1, 2, 4, 8, 16};
    const uint32_t budgets[] = {4, 16, 64};
    const uint32_t widths[] = {64, 128, 256};
    for (uint32_t wpc : wpcs)
        for (uint32_t bud : budgets)
            for (uint32_t cw : widths)
                for (int pers = 0; pers < ((flags & PBM_TUNE_NO_PERSISTENT) ? 1 : 2); ++pers) {
                    pbm_launch l{};
                    l.mode = PBM_MODE_BATCHED;
                    l.waves_per_cu = wpc;
                    l.walk_budget = bud;
                    l.cluster_words = cw;
                    l.flags = PBM_LAUNCH_NO_COUNTERS | (pers ? PBM_LAUNCH_PERSISTENT : 0u);
                    space.push_back(l);
                }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PBM_HIP(hipEventCreate(&e0));
    PBM_HIP(hipEventCreate(&e1));
    const int reps = (flags & PBM_TUNE_QUICK) ? 1 : 3;
    pbm_tune_result res{};
    // median time of launch l over the first n bursts, from the saved state
    auto time_of = [&](const pbm_launch& l, uint32_t n, int r_count, float* out_ms) -> int {
        std::vector<float> t;
        for (int r = 0; r < r_count; ++r) {
            int e = pbm_restore_bitmap(h, map.data(), L, stream);
            if (!e) e = pbm_bursts_reset(b, stream);
            if (e) return e;
            (void)hipEventRecord(e0, s);
            e = pbm_bursts_run(b, 0, n, &l, stream);
            (void)hipEventRecord(e1, s);
            if (e) return e;
            if (hipEventSynchronize(e1) != hipSuccess) return pbmi::fail(PBM_ERR_HIP, "bursts autotune run");
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        *out_ms = t[t.size() / 2];
        return PBM_OK;
    };
    std::vector<std::pair<float, size_t>> ranked;
    for (size_t i = 0; i < space.size() && !rc; ++i) {
        float ms = 0;
        rc = time_of(space[i], count, reps, &ms);
        if (!rc) ranked.push_back({ms, i});
    }
    if (!rc) {
        std::sort(ranked.begin(), ranked.end());
        size_t pick = ranked[0].second;
        res.ms_alloc = ranked[0].first;
        // the three fastest again over up to four times as many bursts:
        // later bursts run on a fuller pool, where more waves in flight
        // wait on the page-budget window, so a short sample can rank them
        // the other way round
        const uint32_t count2 = count > b->n_bursts / 4u ? b->n_bursts : 4u * count;
        if (count2 > count) {
            float best = 1e30f;
            for (size_t k = 0; k < ranked.size() && k < 3u && !rc; ++k) {
                float ms = 0;
                rc = time_of(space[ranked[k].second], count2, reps > 1 ? 2 : 1, &ms);
                if (!rc && ms < best) {
                    best = ms;
                    pick = ranked[k].second;
                    res.ms_alloc = ranked[k].first;
                    res.ms_free = ms;
                }
            }
        }
        res.get_launch = space[pick];
        res.free_launch = space[pick];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) rc = pbm_restore_bitmap(h, map.data(), L, stream);
    if (!rc) rc = pbm_bursts_reset(b, stream);
    if (rc) return rc;
    res.candidates = (uint32_t)space.size();
    *out = res;
    return PBM_OK;
}
