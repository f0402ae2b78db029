// Continuous-batching serving loop: a device-resident slot table, arrival
// queue and per-slot page lists; one step = admit, grow, release.
#include <hip/hip_runtime.h>
#include <cstring>
#include <vector>
#include "pbm_internal.hpp"

using pbm::u64;

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kAdmitWindow = 1024;  // waiting sequences examined per step
constexpr uint32_t kStepRing = 1u << 16; // start times kept for the last 65,536 steps
constexpr uint32_t kWaitUnknown = 0xFFFFFFFFu;

// Words of the control block.
enum : uint32_t {
    kStep = 0,    // steps run
    kHead,        // next sequence to admit
    kTail,        // sequences submitted
    kActive,      // slots admitted at the last grow (finished ones included)
    kPromised,    // pages promised at admission, summed over every admitted sequence
    kTop,         // free-slot stack height
    kDone,        // sequences that finished at a grow of an even step (done list 0)
    kAdmitted,
    kFinished,
    kAllocs,
    kFrees,
    kAllocFail,
    kPayloadErr,
    kLatCount,
    kDoneOdd,     // the same for odd steps (done list 1)
    kCtlWords = 16
};

struct ServeDev {
    pbm_ctx c;
    uint32_t* arr_step;        // [max_seqs] arrival step
    uint32_t* target;          // [max_seqs] target length
    uint32_t* key;             // [max_seqs] payload key
    uint32_t* res_admit;       // [max_seqs]
    uint32_t* res_finish;      // [max_seqs]
    uint32_t* res_len;         // [max_seqs]
    u64* res_sum;              // [max_seqs]
    uint32_t* slot_seq;        // [max_slots] sequence in the slot or kEmpty
    uint32_t* slot_len;        // [max_slots]
    uint32_t* slot_target;     // [max_slots]
    int32_t* pages;            // [max_slots * max_pages]
    uint32_t* stack;           // [max_slots] free slots
    uint32_t* done;            // [2][2 * max_slots] (slot, sequence) finished at the grow of a
                               // step, one list per step parity
    u64* ctl;                  // [kCtlWords]
    uint32_t* lat;             // [lat_cap] wall-clock ticks per allocation
    u64* seq_visible;          // [max_seqs] wall clock when the sequence was submitted
    u64* step_tick;            // [kStepRing] wall clock at the start of step s (s mod ring)
    uint32_t* res_wait;        // [max_seqs] ticks from arrival to admission
    uint32_t lat_cap;
    uint32_t max_slots;
    uint32_t max_pages;
    uint32_t max_seqs;
    uint32_t flags;
    u64 seed;
};

__device__ __forceinline__ u64 ld_ctl(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The done list of step s and its length word: two lists by parity, so the
// grow of step s can list its finished sequences while the admission of
// step s reads those of step s - 1.
__device__ __forceinline__ uint32_t done_word(u64 s) { return (s & 1u) ? kDoneOdd : kDone; }
// A done-list entry: slot, sequence, pages held (the release frees them
// without reading the slot's length).
constexpr uint32_t kDoneWords = 3;
__device__ __forceinline__ uint32_t* done_list(const ServeDev& d, u64 s) {
    return d.done + (size_t)(s & 1u) * kDoneWords * d.max_slots;
}

// Admission (one workgroup).  First the bookkeeping of the last step: the
// sequences on its done list finished, and their slots go back on the
// free-slot stack.  Then
// waiting sequences are admitted in arrival order while their arrival step
// has come, a slot is free and the pool's unreserved pages cover the
// sequence's target.  Admission reserves every admitted sequence's whole
// target in the pool's reservation counter at once, so the counter holds
// free pages less the pages still owed to admitted sequences (the oracle's
// free - owed), and the grow takes pages already reserved.
//
// Two round trips: the control block and the pool counter, then everything
// the admissions read (the finished slots, the stack below them, the
// waiting sequences' arrival steps, targets and submission stamps).  The
// slots a step pops are the last finished ones first, then the stack from
// its old top down, so both come in with the same loads.
__device__ void serve_admit(const ServeDev& d) {
    __shared__ u64 s_ctl[kCtlWords];
    __shared__ long long s_free;
    __shared__ uint32_t s_arr[kAdmitWindow];
    __shared__ uint32_t s_tgt[kAdmitWindow];
    __shared__ uint32_t s_slot[kAdmitWindow];  // the slot the i-th admission takes
    __shared__ u64 s_vis[kAdmitWindow];
    __shared__ uint32_t s_m;
    __shared__ long long s_add;
    __shared__ u64 s_now;
    const uint32_t t = threadIdx.x;
    if (t < kCtlWords) s_ctl[t] = ld_ctl(&d.ctl[t]);
    if (t == 0) {
        s_free = __hip_atomic_load(d.c.free_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_now = wall_clock64();
    }
    __syncthreads();
    const u64 step = s_ctl[kStep] + 1u;
    const u64 now = s_now;
    const u64 head = s_ctl[kHead];
    const u64 tail = s_ctl[kTail];
    const u64 done_n = s_ctl[done_word(step - 1u)];
    const uint32_t* done = done_list(d, step - 1u);
    const u64 top_old = s_ctl[kTop];
    const u64 top = top_old + done_n;
    const u64 active = s_ctl[kActive] - done_n;
    const uint32_t win = (uint32_t)(tail - head < kAdmitWindow ? tail - head : kAdmitWindow);
    for (u64 i = t; i < done_n; i += blockDim.x) {
        const uint32_t sl = done[kDoneWords * i];
        d.stack[top_old + i] = sl;
        const u64 j = done_n - 1u - i;  // popped j-th
        if (j < win) s_slot[j] = sl;
    }
    for (uint32_t i = t; i < win; i += blockDim.x) {
        if (i >= done_n && i < top) s_slot[i] = d.stack[top_old - 1u - (i - done_n)];
        s_arr[i] = d.arr_step[head + i];
        s_tgt[i] = d.target[head + i];
        s_vis[i] = d.seq_visible[head + i];
    }
    __syncthreads();
    if (t < 64u) {
        // the first 64 waiting sequences are decided at once: sequence i is
        // admitted while its arrival step has come, a slot is free and the
        // targets up to i fit the unreserved pages.  Arrival steps do not
        // decrease and targets are positive, so the admitted ones are a
        // prefix; past 64, one at a time
        const long long budget = s_free;
        const u64 slots = (u64)d.max_slots - active;
        long long pre = t < win ? (long long)s_tgt[t] : 0;
#pragma unroll
        for (unsigned o = 1; o < 64u; o <<= 1) {
            const long long v = (long long)pbm::shfl64((u64)pre, t >= o ? t - o : t);
            if (t >= o) pre += v;
        }
        const bool ok = t < win && (u64)s_arr[t] <= step && t < slots && pre <= budget;
        const u64 okm = pbm::ballot(ok);
        uint32_t m = okm == pbm::kFull ? 64u : (uint32_t)__ffsll((long long)~okm) - 1u;
        long long add = m ? (long long)pbm::readlane64((u64)pre, m - 1u) : 0;
        if (t == 0) {
            if (m == 64u) {
                long long left = budget - add;
                while (m < win && (u64)s_arr[m] <= step && m < slots && left >= (long long)s_tgt[m]) {
                    left -= s_tgt[m];
                    add += s_tgt[m];
                    ++m;
                }
            }
            if (add) pbm::atom_add_ll(d.c.free_count, -add);
            s_m = m;
            s_add = add;
        }
    }
    __syncthreads();
    const uint32_t m = s_m;
    for (uint32_t i = t; i < m; i += blockDim.x) {
        const uint32_t slot = s_slot[i];
        const uint32_t seq = (uint32_t)(head + i);
        d.slot_seq[slot] = seq;
        d.slot_len[slot] = 0;
        d.slot_target[slot] = s_tgt[i];
        d.res_admit[seq] = (uint32_t)step;
        d.res_sum[seq] = 0;
        // arrival: the later of its submission and the start of its arrival
        // step (known for the last kStepRing steps; step 0 is before the
        // first step)
        const uint32_t as = s_arr[i];
        bool known = true;
        u64 a = 0;
        if (as == step) a = now;
        else if (as > 0u && step - as < kStepRing) a = d.step_tick[as % kStepRing];
        else if (as > 0u) known = false;
        const u64 vis = s_vis[i];
        const u64 arrive = a > vis ? a : vis;
        d.res_wait[seq] = !known ? kWaitUnknown
                          : (now - arrive > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)(now - arrive));
    }
    if (t == 0) {
        d.step_tick[step % kStepRing] = now;
        d.ctl[kStep] = step;
        d.ctl[kHead] = head + m;
        d.ctl[kActive] = active + m;
        d.ctl[kPromised] = s_ctl[kPromised] + (u64)s_add;
        d.ctl[kTop] = top - m;
        d.ctl[done_word(step - 1u)] = 0;  // the list of step + 1
        d.ctl[kAdmitted] = s_ctl[kAdmitted] + m;
    }
}

__global__ __launch_bounds__(1024) void k_serve_admit(ServeDev d) { serve_admit(d); }

// Grow: every admitted slot gets one page.  A sequence that reaches its
// target is recorded as finished and its slot listed for release.  The
// step's tallies are summed over the workgroup in LDS and added to the
// control block with one atomic per word per workgroup.
//
// The pages were reserved when the sequences were admitted, so the
// getPage calls reserve from a count in LDS that grants every request (it
// starts far above any workgroup's requests) instead of the pool's
// reservation counter; what a failed call gives back there stays reserved
// for the sequence's next step.
constexpr long long kGrowUnits = 1ll << 40;

struct GrowLds {
    uint32_t sum[12];  // [0] finished, [1] allocations, [2] failed, [3..10] pool counters
                       // (Stats order), [11] payload errors
    u64 base[2];       // first done-list and latency entries of the workgroup
    long long units;   // the getPage calls' reservation count
};

// One slot's page for step `step`: written to its page list (and its
// payload), the length stored; a slot that reaches its target is emptied
// (seq becomes kEmpty) and its sequence's finish recorded.
template <int S>
__device__ __forceinline__ void grow_slot(const ServeDev& d, const pbm_ctx& cl, uint32_t slot,
                                          u64 step, uint32_t& seq, uint32_t& len, uint32_t tgt,
                                          pbm::Stats& st, bool& fin, bool& got, uint32_t& ticks) {
    using namespace pbm;
    const int32_t seed = request_seed(d.c.seed, (u64)seq * d.max_pages + len);
    const u64 t0 = wall_clock64();
    const int32_t id = get_page_of<S>(cl, slot, seed, nullptr, &st);
    const u64 t1 = wall_clock64();
    ticks = t1 - t0 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)(t1 - t0);
    if (id < 0) return;
    got = true;
    d.pages[(u64)slot * d.max_pages + len] = id;
    if ((d.flags & PBM_SERVE_PAYLOAD) && d.c.pool)
        *(u64*)page_address(d.c, id) = serve_payload(d.seed, d.key[seq], len);
    ++len;
    d.slot_len[slot] = len;
    if (len == tgt) {
        fin = true;
        d.res_finish[seq] = (uint32_t)step;
        d.res_len[seq] = len;
        d.slot_seq[slot] = kEmpty;
        seq = kEmpty;
    }
}

// Control word of tally lane `lane` (nullptr: none): [0] finished at the
// grow of step `step` (the done list's base), [1] finished in all, [2]
// allocations, [3] latency samples (the latency list's base), [4] failed
// allocations, [5] frees, [6] payload errors, [8..15] the pool counters
// (Stats order).
__device__ __forceinline__ u64* tally_word(const ServeDev& d, unsigned lane, u64 step) {
    switch (lane) {
        case 0: return &d.ctl[done_word(step)];
        case 1: return &d.ctl[kFinished];
        case 2: return &d.ctl[kAllocs];
        case 3: return (d.flags & PBM_SERVE_LATENCY) ? &d.ctl[kLatCount] : nullptr;
        case 4: return &d.ctl[kAllocFail];
        case 5: return &d.ctl[kFrees];
        case 6: return &d.ctl[kPayloadErr];
        default:
            return (lane >= 8u && lane < 16u && d.c.counters) ? (u64*)d.c.counters + (lane - 8u)
                                                              : nullptr;
    }
}

// The workgroup's tallies of one grow (every thread calls it): finished
// sequences go on the done list, latency samples on the latency list, and
// the counts into the control block, one atomic per word.  g.sum must be
// zero and visible to the workgroup.
__device__ void grow_tally(const ServeDev& d, GrowLds& g, const pbm::Stats& st, u64 bad,
                           bool act, bool got, bool fin, uint32_t ticks, uint32_t slot,
                           uint32_t seq, uint32_t len, u64 step) {
    using namespace pbm;
    const unsigned lane = lane_id();
    const u64 F = ballot(fin);
    const u64 G = ballot(got);
    const u64 X = ballot(act && !got);
    // this wave's place among the workgroup's finished and granted lanes
    uint32_t wf = 0, wg = 0;
    if (lane == 0) {
        if (F) wf = atomicAdd(&g.sum[0], (uint32_t)__popcll(F));
        if (G) wg = atomicAdd(&g.sum[1], (uint32_t)__popcll(G));
        if (X) atomicAdd(&g.sum[2], (uint32_t)__popcll(X));
    }
    wf = readlane(wf, 0);
    wg = readlane(wg, 0);
    uint32_t v[8] = {st.allocs, st.frees, st.failures, st.steps,
                     st.probes, st.atomics, st.fallbacks, st.errors};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint32_t x = v[i];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
        if (lane == 0 && x) atomicAdd(&g.sum[3 + i], x);
    }
    {
        uint32_t x = (uint32_t)bad;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
        if (lane == 0 && x) atomicAdd(&g.sum[11], x);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // one lane per control word (tally_word)
        const uint32_t nf = g.sum[0], ng = g.sum[1], nx = g.sum[2];
        u64* dst = tally_word(d, lane, step);
        u64 add = 0;
        switch (lane) {
            case 0: case 1: add = nf; break;
            case 2: case 3: add = ng; break;
            case 4: add = nx; break;
            case 5: add = g.sum[4]; break;
            case 6: add = g.sum[11]; break;
            default:
                if (lane >= 8 && lane < 16) add = g.sum[3 + (lane - 8u)];
        }
        u64 old = 0;
        if (dst && add) old = atom_add_u64(dst, add);
        if (lane == 0) g.base[0] = old;
        if (lane == 3) g.base[1] = old;
    }
    __syncthreads();
    if (fin) {
        const u64 k = g.base[0] + wf + rank_in(F);
        uint32_t* done = done_list(d, step);
        done[kDoneWords * k] = slot;
        done[kDoneWords * k + 1u] = seq;
        done[kDoneWords * k + 2u] = len;
    }
    if ((d.flags & PBM_SERVE_LATENCY) && got) {
        const u64 k = g.base[1] + wg + rank_in(G);
        if (k < d.lat_cap) d.lat[k] = ticks;
    }
}

template <int S>
__device__ void serve_grow(const ServeDev& d, uint32_t slot) {
    using namespace pbm;
    __shared__ GrowLds g;
    if (threadIdx.x < 12) g.sum[threadIdx.x] = 0;
    if (threadIdx.x == 0) g.units = kGrowUnits;
    __syncthreads();
    pbm_ctx cl = d.c;
    cl.free_count = &g.units;
    const u64 step = ld_ctl(&d.ctl[kStep]);
    Stats st;
    // the slot's sequence, length and target are loaded together (one
    // round trip; an empty slot's length and target go unused)
    uint32_t seq = kEmpty, len = 0, tgt = 0;
    if (slot < d.max_slots) {
        seq = d.slot_seq[slot];
        len = d.slot_len[slot];
        tgt = d.slot_target[slot];
    }
    const bool act = seq != kEmpty;
    const uint32_t fseq = seq;
    bool fin = false, got = false;
    uint32_t ticks = 0;
    if (act) grow_slot<S>(d, cl, slot, step, seq, len, tgt, st, fin, got, ticks);
    grow_tally(d, g, st, 0ull, act, got, fin, ticks, slot, fseq, len, step);
}

template <int S>
__global__ __launch_bounds__(256) void k_serve_grow(ServeDev d) {
    serve_grow<S>(d, blockIdx.x * blockDim.x + threadIdx.x);
}

// Frees the page lists of n finished sequences, cut into 64-page chunks
// dealt round-robin to waves wave, wave + n_waves, ...; entry(e, slot, seq,
// len) gives the e-th sequence (a wave's lanes ask for 64 at once).  A
// chunk's payloads are read back, checked and summed into the sequence's
// checksum, then its pages go back to the pool through cl, whose
// reservation counter takes the credit.
//
// With a bitmap strategy a wave takes kFreeChunks of its chunks at a time:
// their page ids load together and their atomicAnds go out together, so a
// wave with many chunks (a workgroup of the loop freeing its own long
// sequences) waits two round trips per kFreeChunks chunks, not per chunk.
constexpr uint32_t kFreeChunks = 4;

template <int S, typename Entry>
__device__ void free_finished(const ServeDev& d, const pbm_ctx& cl, uint32_t n, Entry entry,
                              uint32_t wave, uint32_t n_waves, pbm::Stats& st, u64& bad) {
    using namespace pbm;
    constexpr bool kBits = S != PBM_LINKED_LIST && S != PBM_OUROBOROS;
    constexpr uint32_t U = kBits ? kFreeChunks : 1u;
    const unsigned lane = lane_id();
    const bool payload = (d.flags & PBM_SERVE_PAYLOAD) && d.c.pool;
    for (uint32_t g0 = 0; g0 < n; g0 += 64u) {
        const uint32_t e = g0 + lane;
        uint32_t slot = 0, seq = 0, len = 0;
        if (e < n) entry(e, slot, seq, len);
        const uint32_t chunks = (len + 63u) / 64u;
        uint32_t incl = chunks;
#pragma unroll
        for (unsigned s = 1; s < 64u; s <<= 1) {
            const uint32_t v = shfl32(incl, lane >= s ? lane - s : lane);
            if (lane >= s) incl += v;
        }
        const uint32_t total = readlane(incl, 63u);
        for (uint32_t c0 = wave; c0 < total; c0 += n_waves * U) {
            int32_t id[U];
            uint32_t cj[U], cseq[U];
            bool valid[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t ci = c0 + u * n_waves;  // the same for the whole wave
                id[u] = -1;
                cj[u] = 0;
                cseq[u] = 0;
                valid[u] = false;
                if (ci < total) {
                    const unsigned src = (unsigned)__ffsll((long long)ballot(incl > ci)) - 1u;
                    const uint32_t c_slot = readlane(slot, src);
                    const uint32_t c_len = readlane(len, src);
                    const uint32_t c_first = readlane(incl - chunks, src);
                    cseq[u] = readlane(seq, src);
                    cj[u] = (ci - c_first) * 64u + lane;
                    valid[u] = cj[u] < c_len;
                    if (valid[u]) id[u] = d.pages[(u64)c_slot * d.max_pages + cj[u]];
                }
            }
            if (payload) {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    if (c0 + u * n_waves >= total) continue;
                    u64 sum = 0;
                    if (valid[u]) {
                        const u64 v = *(const u64*)page_address(d.c, id[u]);
                        sum = v;
                        bad += v != serve_payload(d.seed, d.key[cseq[u]], cj[u]);
                    }
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) sum += shfl64(sum, lane ^ (unsigned)o);
                    if (lane == 0) atom_add_u64(&d.res_sum[cseq[u]], sum);
                }
                // the read-back completes before the pages go back: a page
                // freed with its load still in flight could be taken and
                // written by a concurrent grow (the loop's workgroups free
                // while others grow) before the load is served
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            }
            if constexpr (kBits) {
                // free_page, the chunks' atomics in flight together
                u64 m[U], o[U];
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    m[u] = 0;
                    o[u] = 0;
                    if (valid[u] && (uint32_t)id[u] < d.c.n_pages) {
                        m[u] = 1ull << ((uint32_t)id[u] & 63u);
                        o[u] = atom_and(&d.c.bitmap[(uint32_t)id[u] >> 6], ~m[u]);
                        st.atomics++;
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const bool ok = (o[u] & m[u]) != 0ull;
                    if (valid[u]) {
                        if (ok) st.frees++;
                        else st.errors++;
                    }
                    release_one(cl, ok);
                }
            } else {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    if (valid[u]) free_page_of<S>(cl, id[u], &st);
            }
        }
    }
}

// Release: the page lists of the sequences that finished at this step's
// grow (the done list), dealt over the waves of the grid.
//
// The freed pages are credited to the pool's reservation counter once per
// workgroup, at the end: the frees count them into LDS instead of one
// atomic per wave at the pool counter's one address.
template <int S>
__device__ void serve_release(const ServeDev& d, uint32_t wave, uint32_t n_waves) {
    using namespace pbm;
    __shared__ uint32_t s_sum[9];
    __shared__ long long s_credit;
    if (threadIdx.x < 9) s_sum[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_credit = 0;
    __syncthreads();
    pbm_ctx cl = d.c;
    cl.free_count = &s_credit;
    const unsigned lane = lane_id();
    // the step, both parities' counts and both done lists' first 64 entries
    // load together (one round trip); the step's parity picks the list
    const u64 step = ld_ctl(&d.ctl[kStep]);
    const u64 dn[2] = {ld_ctl(&d.ctl[kDone]), ld_ctl(&d.ctl[kDoneOdd])};
    uint32_t pre[2][kDoneWords] = {};
    if (lane < d.max_slots) {
#pragma unroll
        for (uint32_t p = 0; p < 2; ++p)
#pragma unroll
            for (uint32_t w = 0; w < kDoneWords; ++w)
                pre[p][w] = done_list(d, p)[kDoneWords * lane + w];
    }
    const uint32_t done_n = (uint32_t)dn[step & 1u];
    const uint32_t* done = done_list(d, step);
    const uint32_t par = (uint32_t)(step & 1u);
    const uint32_t p_slot = par ? pre[1][0] : pre[0][0];
    const uint32_t p_seq = par ? pre[1][1] : pre[0][1];
    const uint32_t p_len = par ? pre[1][2] : pre[0][2];
    Stats st;
    u64 bad = 0;
    free_finished<S>(
        d, cl, done_n,
        [&](uint32_t e, uint32_t& slot, uint32_t& seq, uint32_t& len) {
            if (e < 64u) {  // e is the lane's own: its entry is loaded
                slot = p_slot;
                seq = p_seq;
                len = p_len;
            } else {
                slot = done[kDoneWords * e];
                seq = done[kDoneWords * e + 1u];
                len = done[kDoneWords * e + 2u];
            }
        },
        wave, n_waves, st, bad);
    // tallies summed over the workgroup, then one atomic per word: [0]
    // payload errors, [1..8] the pool counters (Stats order; frees is also
    // the control block's count)
    uint32_t v[9] = {(uint32_t)bad, st.allocs, st.frees, st.failures, st.steps,
                     st.probes, st.atomics, st.fallbacks, st.errors};
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        uint32_t x = v[i];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
        if (lane == 0 && x) atomicAdd(&s_sum[i], x);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        u64* dst = nullptr;
        u64 add = 0;
        if (lane == 0) {
            dst = &d.ctl[kPayloadErr];
            add = s_sum[0];
        } else if (lane == 1) {
            dst = &d.ctl[kFrees];
            add = s_sum[2];
        } else if (lane >= 8 && lane < 16 && d.c.counters) {
            dst = (u64*)d.c.counters + (lane - 8u);
            add = s_sum[1 + (lane - 8u)];
        }
        if (dst && add) atom_add_u64(dst, add);
        if (lane == 2 && s_credit) atom_add_ll(d.c.free_count, s_credit);
    }
}

template <int S>
__global__ __launch_bounds__(256) void k_serve_release(ServeDev d) {
    serve_release<S>(d, (blockIdx.x * blockDim.x + threadIdx.x) >> 6,
                     (gridDim.x * blockDim.x) >> 6);
}

// ---- the device-resident loop: n_steps steps in one kernel
//
// One 1,024-thread workgroup per 1,024 slots, and one more that admits.
// Each step the admitting workgroup admits, and every slot workgroup grows
// its slots and frees the pages of its own slots' finished sequences
// (their 64-page chunks dealt over its sixteen waves).
//
// A slot workgroup keeps its slots' state in registers from step to step.
// It grows the slots admitted before this step, lists the sequences that
// this finishes on the step's done list (admission reads the last step's:
// one list per step parity) and their latency samples, and frees their
// pages, all while admission runs: admission writes only empty slots, and
// the pages were reserved at admission.  The credit of those frees waits
// in LDS until admission is visible, so admission counts exactly the pages
// the last step's releases gave back.  Then the workgroup grows the slots
// admitted now, lists and frees what that finishes (a sequence of one page
// only), and credits the pool.  A new slot's first latency sample is
// listed with the next step's; the counts are summed per thread and
// written once, at the end of the launch.  No slot workgroup waits for
// another's grow; admission waits for every workgroup's step.
//
// The phases follow one another through counters in `sync`: [0] steps
// admitted, [2] workgroup steps done; [3] counts waits that gave up.  A
// workgroup makes its writes visible before it counts itself done (every
// wave waits for its stores, then the barrier, then one device-scope
// release and the count), and acquires after it sees what it waited for
// (one acquire, then the barrier).  Admission waits for every slot
// workgroup, so all of them must be resident at once: the launch checks
// that the device holds them, and every wait is bounded (kLoopWaitTicks)
// and counted when it gives up, so a device shared with other work cannot
// hang the loop, though its results are then not valid.
constexpr u64 kLoopWaitTicks = 200000000ull;  // 2 s of the 100 MHz wall clock

// Thread 0 waits for the word to reach `target` and acquires for the
// workgroup: its acquire invalidates this CU's L1 and this XCD's L2, and the
// barrier orders every wave's later reads after it.
__device__ inline void loop_wait_ge(u64* sync, uint32_t word, u64 target) {
    if (threadIdx.x == 0) {
        const u64* p = sync + word;
        if (ld_ctl(p) < target) {
            const u64 t0 = wall_clock64();
            for (uint32_t spins = 1;; ++spins) {
                if (spins < 16u) __builtin_amdgcn_s_sleep(2);
                else __builtin_amdgcn_s_sleep(16);
                if (ld_ctl(p) >= target) break;
                if ((spins & 15u) == 0u && wall_clock64() - t0 > kLoopWaitTicks) {
                    pbm::atom_add_u64(sync + 3, 1ull);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Every wave waits until its own memory operations have completed (so the
// workgroup's writes are all in this XCD's L2), then thread 0 releases for
// the workgroup (writing the L2's dirty lines back) and counts.
__device__ inline void loop_signal(u64* sync, uint32_t word, bool store, u64 value) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (store) __hip_atomic_store(sync + word, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else pbm::atom_add_u64(sync + word, value);
    }
}

constexpr uint32_t kLoopThreads = 1024;  // slots per workgroup of the loop

// A slot workgroup's list of its sequences that finished this step, for
// its own release.
struct LoopLds {
    uint32_t n;
    uint32_t slot[kLoopThreads], seq[kLoopThreads], len[kLoopThreads];
    long long credit;  // pages freed this step, not yet credited to the pool
};

// Lists the lanes' finished sequences (every thread calls it).
__device__ __forceinline__ void loop_list(LoopLds& L, bool fin, uint32_t slot, uint32_t seq,
                                          uint32_t len) {
    const u64 F = pbm::ballot(fin);
    uint32_t b = 0;
    if (pbm::lane_id() == 0 && F) b = atomicAdd(&L.n, (uint32_t)__popcll(F));
    b = pbm::readlane(b, 0);
    if (fin) {
        const uint32_t k = b + pbm::rank_in(F);
        L.slot[k] = slot;
        L.seq[k] = seq;
        L.len[k] = len;
    }
}

// Positions in a device list for the workgroup's items: one atomic on the
// list's length word `len` for the whole workgroup, none when it has no
// item (every thread calls it, with k items of its own).  Returns the
// position of this thread's first item.  s.n is zero on entry.
struct TakeLds {
    uint32_t n;
    u64 base;
};

__device__ u64 take_positions(u64* len, uint32_t k, TakeLds& s) {
    using namespace pbm;
    const unsigned lane = lane_id();
    uint32_t inc = k;
#pragma unroll
    for (unsigned o = 1; o < 64u; o <<= 1) {
        const uint32_t v = shfl32(inc, lane >= o ? lane - o : lane);
        if (lane >= o) inc += v;
    }
    const uint32_t wtot = readlane(inc, 63u);
    uint32_t wbase = 0;
    if (lane == 0 && wtot) wbase = atomicAdd(&s.n, wtot);
    wbase = readlane(wbase, 0);
    __syncthreads();
    if (threadIdx.x == 0) s.base = s.n ? atom_add_u64(len, (u64)s.n) : 0ull;
    __syncthreads();
    return s.base + wbase + (inc - k);
}

template <int S>
__global__ __launch_bounds__(kLoopThreads) void k_serve_loop(ServeDev d, uint32_t n_steps,
                                                             u64* sync) {
    using namespace pbm;
    constexpr int R = (S == PBM_LINKED_LIST || S == PBM_OUROBOROS) ? S : PBM_RW_BM;
    constexpr uint32_t kWaves = kLoopThreads / 64u;
    const u64 P = gridDim.x - 1u;  // slot workgroups; the last workgroup admits
    if (blockIdx.x == P) {
        for (uint32_t k = 0; k < n_steps; ++k) {
            // the last step's pages are all back before admission counts them
            if (k > 0) loop_wait_ge(sync, 2, (u64)k * P);
            serve_admit(d);
            loop_signal(sync, 0, true, (u64)k + 1u);
        }
        return;
    }
    __shared__ GrowLds g;
    __shared__ LoopLds L;
    // [0] finished held sequences, [1] latency samples, [2] sequences that
    // finished at their first page
    __shared__ TakeLds tk[3];
    if (threadIdx.x == 0) {
        g.units = kGrowUnits;
        L.n = 0;
        L.credit = 0;
    }
    pbm_ctx cl = d.c;  // the grow's: pages reserved at admission
    cl.free_count = &g.units;
    pbm_ctx cr = d.c;  // the release's: credit held in LDS
    cr.free_count = &L.credit;
    const uint32_t slot = blockIdx.x * kLoopThreads + threadIdx.x;
    const uint32_t wave = threadIdx.x >> 6;
    const bool lat = (d.flags & PBM_SERVE_LATENCY) != 0;
    // the launch's counts, per thread, written once at its end
    Stats st;
    u64 bad = 0;
    uint32_t c_alloc = 0, c_fin = 0, c_fail = 0;
    // a new slot's first latency sample, listed with the next step's
    bool pend = false;
    uint32_t pend_ticks = 0;
    // the slot as this thread left it at its last grow (kEmpty: load it
    // once admission is visible)
    uint32_t seq = kEmpty, len = 0, tgt = 0;
    u64 step = 0;
    for (uint32_t k = 0; k < n_steps; ++k) {
        if (threadIdx.x < 3u) tk[threadIdx.x].n = 0;
        bool fin = false, got = false;
        uint32_t ticks = 0, fseq = kEmpty;
        // slots admitted before this step, while admission runs
        const bool held = seq != kEmpty;  // never at k = 0
        const u64 now_step = step + 1u;   // this step, for k > 0
        if (held) {
            fseq = seq;
            grow_slot<S>(d, cl, slot, now_step, seq, len, tgt, st, fin, got, ticks);
            c_alloc += got ? 1u : 0u;
            c_fail += got ? 0u : 1u;
            c_fin += fin ? 1u : 0u;
        }
        loop_list(L, fin, slot, fseq, len);
        __syncthreads();
        const uint32_t n_held = L.n;
        if (k > 0) {
            // also while admission runs: they go on this step's done list
            // (admission reads the last step's) and the latency list
            const u64 dp = take_positions(&d.ctl[done_word(now_step)], fin ? 1u : 0u, tk[0]);
            if (fin) {
                uint32_t* dl = done_list(d, now_step);
                dl[kDoneWords * dp] = slot;
                dl[kDoneWords * dp + 1u] = fseq;
                dl[kDoneWords * dp + 2u] = len;
            }
            if (lat) {
                u64 lp = take_positions(&d.ctl[kLatCount], (pend ? 1u : 0u) + (got ? 1u : 0u),
                                        tk[1]);
                if (pend) {
                    if (lp < d.lat_cap) d.lat[lp] = pend_ticks;
                    ++lp;
                }
                if (got && lp < d.lat_cap) d.lat[lp] = ticks;
                pend = false;
            }
        }
        free_finished<R>(
            d, cr, n_held,
            [](uint32_t e, uint32_t& sl, uint32_t& sq, uint32_t& ln) {
                sl = L.slot[e];
                sq = L.seq[e];
                ln = L.len[e];
            },
            wave, kWaves, st, bad);
        // slots admitted now
        loop_wait_ge(sync, 0, (u64)k + 1u);
        uint32_t nseq = kEmpty, nlen = 0, ntgt = 0;
        u64 cur = now_step;
        if (k == 0) cur = ld_ctl(&d.ctl[kStep]);
        if (!held && slot < d.max_slots) {
            nseq = d.slot_seq[slot];
            nlen = d.slot_len[slot];
            ntgt = d.slot_target[slot];
        }
        step = cur;
        bool fin_new = false;
        if (nseq != kEmpty) {
            bool got_new = false;
            uint32_t ticks_new = 0;
            seq = nseq;
            len = nlen;
            tgt = ntgt;
            fseq = seq;
            grow_slot<S>(d, cl, slot, step, seq, len, tgt, st, fin_new, got_new, ticks_new);
            c_alloc += got_new ? 1u : 0u;
            c_fail += got_new ? 0u : 1u;
            c_fin += fin_new ? 1u : 0u;
            if (got_new) {
                pend = true;
                pend_ticks = ticks_new;
            }
        }
        loop_list(L, fin_new, slot, fseq, len);
        __syncthreads();
        free_finished<R>(
            d, cr, L.n - n_held,
            [n_held](uint32_t e, uint32_t& sl, uint32_t& sq, uint32_t& ln) {
                sl = L.slot[n_held + e];
                sq = L.seq[n_held + e];
                ln = L.len[n_held + e];
            },
            wave, kWaves, st, bad);
        // sequences that finished now (at their first page; at k = 0 every
        // sequence grows here): no atomic when there is none
        const u64 dp2 = take_positions(&d.ctl[done_word(step)], fin_new ? 1u : 0u, tk[2]);
        if (fin_new) {
            uint32_t* dl = done_list(d, step);
            dl[kDoneWords * dp2] = slot;
            dl[kDoneWords * dp2 + 1u] = fseq;
            dl[kDoneWords * dp2 + 2u] = len;
        }
        // the step's freed pages back to the pool, after admission read it
        if (threadIdx.x == 0) {
            if (L.credit) atom_add_ll(d.c.free_count, L.credit);
            L.credit = 0;
            L.n = 0;
        }
        loop_signal(sync, 2, false, 1ull);
    }
    // the launch's last latency samples, then its counts: summed over the
    // workgroup, one atomic per control word
    __syncthreads();
    if (threadIdx.x < 12u) g.sum[threadIdx.x] = 0;
    if (threadIdx.x == 0) tk[1].n = 0;
    __syncthreads();
    if (lat) {
        const u64 lp = take_positions(&d.ctl[kLatCount], pend ? 1u : 0u, tk[1]);
        if (pend && lp < d.lat_cap) d.lat[lp] = pend_ticks;
    }
    const unsigned lane = lane_id();
    uint32_t v[12] = {c_fin, c_alloc, c_fail, st.allocs, st.frees, st.failures, st.steps,
                      st.probes, st.atomics, st.fallbacks, st.errors, (uint32_t)bad};
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        uint32_t x = v[i];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
        if (lane == 0 && x) atomicAdd(&g.sum[i], x);
    }
    __syncthreads();
    if (threadIdx.x < 16u) {
        u64* dst = nullptr;
        u64 add = 0;
        switch (threadIdx.x) {
            case 0: dst = &d.ctl[kFinished]; add = g.sum[0]; break;
            case 1: dst = &d.ctl[kAllocs]; add = g.sum[1]; break;
            case 2: dst = &d.ctl[kAllocFail]; add = g.sum[2]; break;
            case 3: dst = &d.ctl[kFrees]; add = g.sum[4]; break;
            case 4: dst = &d.ctl[kPayloadErr]; add = g.sum[11]; break;
            default:
                if (threadIdx.x >= 8u && d.c.counters) {
                    dst = (u64*)d.c.counters + (threadIdx.x - 8u);
                    add = g.sum[3 + (threadIdx.x - 8u)];
                }
        }
        if (dst && add) atom_add_u64(dst, add);
    }
}

// Before a loop launch: its phase counters start at zero.
__global__ void k_serve_loop_prep(u64* sync) {
    if (threadIdx.x < 3) sync[threadIdx.x] = 0;
}

// The j-th slot handed out: slots go round the 1,024-slot groups (0, 1024,
// 2048, ..., 1, 1025, ...), so the running sequences spread evenly over
// the device-resident loop's workgroups instead of filling the first ones.
__device__ __forceinline__ uint32_t slot_order(uint32_t j, uint32_t max_slots) {
    const uint32_t G = max_slots / 1024u, rem = max_slots % 1024u;
    uint32_t r, q;
    if (j < rem * (G + 1u)) {
        r = j / (G + 1u);
        q = j % (G + 1u);
    } else {
        const uint32_t jj = j - rem * (G + 1u);
        r = rem + jj / G;
        q = jj % G;
    }
    return q * 1024u + r;
}

__global__ void k_serve_init(ServeDev d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.max_slots) {
        d.slot_seq[i] = kEmpty;
        d.slot_len[i] = 0;
        // popped from the top: stack[max_slots - 1] goes first
        d.stack[i] = slot_order(d.max_slots - 1u - i, d.max_slots);
    }
    if (i < kCtlWords) d.ctl[i] = i == kTop ? (u64)d.max_slots : 0ull;
}

// Takes sequences first .. tail-1 from the server's pinned host mirror
// (read by the device directly, so a submission is this one kernel on the
// stream, no copies), stamps when they became visible, and makes them
// visible to admission: the steps after this kernel on the stream see them.
__global__ void k_serve_take(ServeDev d, const uint32_t* m_arr, const uint32_t* m_tgt,
                             const uint32_t* m_key, u64 first, u64 tail) {
    const u64 now = wall_clock64();
    for (u64 i = first + blockIdx.x * blockDim.x + threadIdx.x; i < tail;
         i += (u64)gridDim.x * blockDim.x) {
        d.arr_step[i] = m_arr[i];
        d.target[i] = m_tgt[i];
        d.key[i] = m_key[i];
        d.seq_visible[i] = now;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) d.ctl[kTail] = tail;
}

// Gives the pool back what a server holds when it is destroyed: the pages
// of every running sequence (its page list, through the strategy's
// freePage) and the pages promised at admission and not yet granted.  One
// wave per 64 slots; the sequences of a wave's slots are freed one after
// another, 64 pages at a time.
template <int S>
__global__ __launch_bounds__(256) void k_serve_drop(ServeDev d) {
    using namespace pbm;
    Stats st;
    const unsigned lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    for (u64 g0 = (u64)wave * 64u; g0 < d.max_slots; g0 += (u64)n_waves * 64u) {
        const uint32_t slot = (uint32_t)g0 + lane;
        uint32_t len = 0;
        if (slot < d.max_slots && d.slot_seq[slot] != kEmpty) len = d.slot_len[slot];
        for (u64 act = ballot(len > 0u); act; act &= act - 1ull) {
            const unsigned q = (unsigned)__ffsll((long long)act) - 1u;
            const uint32_t qs = readlane(slot, q), ql = readlane(len, q);
            for (uint32_t j = lane; j < ql; j += 64u)
                free_page_of<S>(d.c, d.pages[(u64)qs * d.max_pages + j], &st);
        }
    }
    flush_stats(d.c, st);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const u64 owed = d.ctl[kPromised] - d.ctl[kAllocs];
        if (owed) atom_add_ll(d.c.free_count, (long long)owed);
    }
}

}  // namespace

struct pbm_server {
    pbm_handle* h = nullptr;
    pbm_serve_config cfg{};
    ServeDev d{};
    uint32_t submitted = 0;
    uint32_t last_arrival = 0;
    uint32_t release_blocks = 1;
    std::vector<void*> mem;
    // pinned host mirror of arr_step, target and key: a submit writes into
    // it and launches one kernel that reads it from the device, without
    // waiting for the stream; no sequence id is ever written twice, so an
    // entry a launch still has to read is never overwritten
    uint32_t* h_mirror = nullptr;  // [3 * max_seqs]
    uint32_t* d_mirror = nullptr;  // the same memory as the device addresses it
    u64* d_sync = nullptr;         // device-resident loop: phase counters, give-ups
    bool ready = false;            // created and initialised
};

extern "C" int pbm_serve_destroy(pbm_server* s) {
    if (!s) return PBM_OK;
    if (s->h) (void)pbmi::use_device(s->h);
    (void)hipDeviceSynchronize();
    if (s->ready) {
        const dim3 g((s->d.max_slots + 255u) / 256u < 1024u ? (s->d.max_slots + 255u) / 256u : 1024u);
        if (s->h->cfg.strategy == PBM_LINKED_LIST)
            hipLaunchKernelGGL(k_serve_drop<PBM_LINKED_LIST>, g, dim3(256), 0, nullptr, s->d);
        else if (s->h->cfg.strategy == PBM_OUROBOROS)
            hipLaunchKernelGGL(k_serve_drop<PBM_OUROBOROS>, g, dim3(256), 0, nullptr, s->d);
        else
            hipLaunchKernelGGL(k_serve_drop<PBM_RW_BM>, g, dim3(256), 0, nullptr, s->d);
        (void)hipDeviceSynchronize();
    }
    for (void* p : s->mem)
        if (p) (void)hipFree(p);
    if (s->h_mirror) (void)hipHostFree(s->h_mirror);
    delete s;
    return PBM_OK;
}

extern "C" int pbm_serve_create(pbm_handle* h, const pbm_serve_config* cfg, pbm_server** out) {
    PBM_CHECK_HANDLE(h);
    if (!cfg || !out) return pbmi::fail(PBM_ERR_INVALID, "null serve config or output");
    *out = nullptr;
    if (cfg->max_slots == 0 || cfg->max_slots > (1u << 22))
        return pbmi::fail(PBM_ERR_RANGE, "max_slots must be in [1, 2^22]");
    if (cfg->max_pages == 0 || cfg->max_seqs == 0)
        return pbmi::fail(PBM_ERR_INVALID, "max_pages and max_seqs must be > 0");
    if (cfg->flags & ~(PBM_SERVE_PAYLOAD | PBM_SERVE_LATENCY))
        return pbmi::fail(PBM_ERR_INVALID, "unknown serve flags");
    if ((cfg->flags & PBM_SERVE_PAYLOAD) && (!h->d_pool || h->cfg.page_size < 8))
        return pbmi::fail(PBM_ERR_INVALID, "PBM_SERVE_PAYLOAD needs a backed pool");
    pbm_server* s = new pbm_server();
    s->h = h;
    s->cfg = *cfg;
    ServeDev& d = s->d;
    d.c = h->ctx;
    d.max_slots = cfg->max_slots;
    d.max_pages = cfg->max_pages;
    d.max_seqs = cfg->max_seqs;
    d.flags = cfg->flags;
    d.seed = cfg->seed;
    d.lat_cap = (cfg->flags & PBM_SERVE_LATENCY) ? cfg->lat_capacity : 0u;
    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8u);
        if (e != hipSuccess) return pbmi::hip_fail(e, "hipMalloc(serve)", __FILE__, __LINE__);
        s->mem.push_back(*p);
        return PBM_OK;
    };
    const size_t ns = cfg->max_seqs, nl = cfg->max_slots;
    int rc = PBM_OK;
    void* p = nullptr;
#define PBM_SALLOC(field, type, count)                          \
    if (!rc) {                                                  \
        rc = alloc(&p, (size_t)(count) * sizeof(type));         \
        if (!rc) d.field = (type*)p;                            \
    }
    PBM_SALLOC(arr_step, uint32_t, ns)
    PBM_SALLOC(target, uint32_t, ns)
    PBM_SALLOC(key, uint32_t, ns)
    PBM_SALLOC(res_admit, uint32_t, ns)
    PBM_SALLOC(res_finish, uint32_t, ns)
    PBM_SALLOC(res_len, uint32_t, ns)
    PBM_SALLOC(res_sum, u64, ns)
    PBM_SALLOC(slot_seq, uint32_t, nl)
    PBM_SALLOC(slot_len, uint32_t, nl)
    PBM_SALLOC(slot_target, uint32_t, nl)
    PBM_SALLOC(pages, int32_t, nl * cfg->max_pages)
    PBM_SALLOC(stack, uint32_t, nl)
    PBM_SALLOC(done, uint32_t, 2u * kDoneWords * nl)
    PBM_SALLOC(ctl, u64, kCtlWords)
    PBM_SALLOC(lat, uint32_t, d.lat_cap)
    PBM_SALLOC(seq_visible, u64, ns)
    PBM_SALLOC(step_tick, u64, kStepRing)
    PBM_SALLOC(res_wait, uint32_t, ns)
#undef PBM_SALLOC
    if (!rc) {
        rc = alloc(&p, 4u * sizeof(u64));
        if (!rc) s->d_sync = (u64*)p;
    }
    if (!rc) {
        hipError_t he = hipHostMalloc((void**)&s->h_mirror, 3u * ns * 4u, hipHostMallocDefault);
        if (he != hipSuccess) rc = pbmi::hip_fail(he, "hipHostMalloc(serve mirror)", __FILE__, __LINE__);
        if (!rc) {
            he = hipHostGetDevicePointer((void**)&s->d_mirror, s->h_mirror, 0);
            if (he != hipSuccess)
                rc = pbmi::hip_fail(he, "hipHostGetDevicePointer(serve mirror)", __FILE__, __LINE__);
        }
    }
    if (rc) {
        pbm_serve_destroy(s);
        return rc;
    }
    hipError_t e;
    if ((e = hipMemset((void*)d.res_admit, 0, ns * 4u)) != hipSuccess ||
        (e = hipMemset((void*)d.res_finish, 0, ns * 4u)) != hipSuccess ||
        (e = hipMemset((void*)d.res_len, 0, ns * 4u)) != hipSuccess ||
        (e = hipMemset((void*)d.res_sum, 0, ns * 8u)) != hipSuccess ||
        (e = hipMemset((void*)d.res_wait, 0xFF, ns * 4u)) != hipSuccess ||
        (e = hipMemset((void*)d.step_tick, 0, (size_t)kStepRing * 8u)) != hipSuccess ||
        (e = hipMemset((void*)s->d_sync, 0, 4u * sizeof(u64))) != hipSuccess) {
        pbm_serve_destroy(s);
        return pbmi::hip_fail(e, "hipMemset(serve)", __FILE__, __LINE__);
    }
    const uint32_t n_init = cfg->max_slots > kCtlWords ? cfg->max_slots : kCtlWords;
    hipLaunchKernelGGL(k_serve_init, dim3((n_init + 255u) / 256u), dim3(256), 0, nullptr, d);
    if ((e = hipDeviceSynchronize()) != hipSuccess) {
        pbm_serve_destroy(s);
        return pbmi::hip_fail(e, "serve init", __FILE__, __LINE__);
    }
    s->release_blocks = (uint32_t)h->n_cu;  // four waves per CU
    s->ready = true;
    *out = s;
    return PBM_OK;
}

extern "C" int pbm_serve_submit(pbm_server* s, uint32_t n, const uint32_t* arrival_step,
                                const uint32_t* target, const uint32_t* key, void* stream) {
    if (!s) return pbmi::fail(PBM_ERR_INVALID, "null server");
    PBM_CHECK_HANDLE(s->h);
    if (n == 0) return PBM_OK;
    if (!arrival_step || !target) return pbmi::fail(PBM_ERR_INVALID, "null submit arrays");
    if ((uint64_t)s->submitted + n > s->cfg.max_seqs)
        return pbmi::fail(PBM_ERR_RANGE, "more sequences than max_seqs");
    uint32_t last = s->last_arrival;
    for (uint32_t i = 0; i < n; ++i) {
        if (arrival_step[i] < last) return pbmi::fail(PBM_ERR_INVALID, "arrival steps must not decrease");
        last = arrival_step[i];
        if (target[i] == 0 || target[i] > s->cfg.max_pages)
            return pbmi::fail(PBM_ERR_RANGE, "target length must be in [1, max_pages]");
    }
    hipStream_t st = pbmi::as_stream(stream);
    const size_t ns = s->cfg.max_seqs, at = s->submitted;
    uint32_t* m_arr = s->h_mirror + at;
    uint32_t* m_tgt = s->h_mirror + ns + at;
    uint32_t* m_key = s->h_mirror + 2u * ns + at;
    std::memcpy(m_arr, arrival_step, (size_t)n * 4u);
    std::memcpy(m_tgt, target, (size_t)n * 4u);
    for (uint32_t i = 0; i < n; ++i) m_key[i] = key ? key[i] : s->submitted + i;
    // the device reads the mirror directly; the host's writes above come
    // before the launch that reads them
    const uint32_t blocks = n > 65536u ? 256u : (n + 255u) / 256u;
    hipLaunchKernelGGL(k_serve_take, dim3(blocks), dim3(256), 0, st, s->d, s->d_mirror,
                       s->d_mirror + ns, s->d_mirror + 2u * ns, (u64)at, (u64)at + n);
    PBM_HIP(hipGetLastError());
    s->submitted += n;
    s->last_arrival = last;
    return PBM_OK;
}

extern "C" int pbm_serve_read_waits(pbm_server* s, uint32_t first, uint32_t n, uint32_t* ticks,
                                    void* stream) {
    if (!s || (n && !ticks)) return pbmi::fail(PBM_ERR_INVALID, "null server or output");
    PBM_CHECK_HANDLE(s->h);
    if ((uint64_t)first + n > s->cfg.max_seqs)
        return pbmi::fail(PBM_ERR_RANGE, "wait range beyond max_seqs");
    if (!n) return PBM_OK;
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(ticks, s->d.res_wait + first, (size_t)n * 4u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

namespace {

// A launch that records e0 / e1 at the kernel's own start and end when
// they are given (a plain launch otherwise, which graph capture takes).
template <typename K, typename... A>
void launch_ev(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t e0, hipEvent_t e1, A... a) {
    if (e0 || e1) hipExtLaunchKernelGGL(kernel, g, b, 0, st, e0, e1, 0, a...);
    else hipLaunchKernelGGL(kernel, g, b, 0, st, a...);
}

// One step: admit, grow, release.  With ev (six events), each kernel
// records ev[2k] and ev[2k+1] at its own start and end.
int launch_step(pbm_server* s, hipStream_t st, hipEvent_t* ev = nullptr) {
    const ServeDev& d = s->d;
    hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr;
    launch_ev(k_serve_admit, dim3(1), dim3(1024), st, e0, e1, d);
    const dim3 g((d.max_slots + 255u) / 256u), b(256);
    e0 = ev ? ev[2] : nullptr;
    e1 = ev ? ev[3] : nullptr;
    switch (s->h->cfg.strategy) {
#define PBM_GROW_CASE(S) \
    case S: launch_ev(k_serve_grow<S>, g, b, st, e0, e1, d); break;
        PBM_GROW_CASE(PBM_RW_BM)
        PBM_GROW_CASE(PBM_COLLAB_RW_BM)
        PBM_GROW_CASE(PBM_CRW_BM)
        PBM_GROW_CASE(PBM_RW_MALLOC)
        PBM_GROW_CASE(PBM_LINKED_LIST)
        PBM_GROW_CASE(PBM_SINGLE_CLOCK)
        PBM_GROW_CASE(PBM_DIST_CLOCK)
        PBM_GROW_CASE(PBM_RW_BASIC)
        PBM_GROW_CASE(PBM_CRW_BASIC)
        PBM_GROW_CASE(PBM_COLLAB_RW_BASIC)
        PBM_GROW_CASE(PBM_OUROBOROS)
#undef PBM_GROW_CASE
        default: return pbmi::fail(PBM_ERR_UNSUPPORTED, "no serving launch for this strategy");
    }
    const dim3 rg(s->release_blocks), rb(256);
    e0 = ev ? ev[4] : nullptr;
    e1 = ev ? ev[5] : nullptr;
    if (s->h->cfg.strategy == PBM_LINKED_LIST)
        launch_ev(k_serve_release<PBM_LINKED_LIST>, rg, rb, st, e0, e1, d);
    else if (s->h->cfg.strategy == PBM_OUROBOROS)
        launch_ev(k_serve_release<PBM_OUROBOROS>, rg, rb, st, e0, e1, d);
    else
        launch_ev(k_serve_release<PBM_RW_BM>, rg, rb, st, e0, e1, d);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_serve_run(pbm_server* s, uint32_t n_steps, void* stream) {
    if (!s) return pbmi::fail(PBM_ERR_INVALID, "null server");
    PBM_CHECK_HANDLE(s->h);
    hipStream_t st = pbmi::as_stream(stream);
    for (uint32_t k = 0; k < n_steps; ++k) {
        int rc = launch_step(s, st);
        if (rc) return rc;
    }
    return PBM_OK;
}

namespace {

template <int S>
int launch_loop_s(pbm_server* s, uint32_t n_steps, hipStream_t st) {
    // the slot workgroups and the admitting one
    const uint32_t groups = (s->d.max_slots + kLoopThreads - 1u) / kLoopThreads + 1u;
    int per_cu = 0;
    PBM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_serve_loop<S>, kLoopThreads,
                                                         0));
    if ((uint64_t)per_cu * (uint64_t)s->h->n_cu < groups)
        return pbmi::fail(PBM_ERR_UNSUPPORTED,
                          "the device-resident loop needs every workgroup of the slot table "
                          "resident at once; this device holds fewer");
    hipLaunchKernelGGL(k_serve_loop_prep, dim3(1), dim3(64), 0, st, s->d_sync);
    hipLaunchKernelGGL(k_serve_loop<S>, dim3(groups), dim3(kLoopThreads), 0, st, s->d, n_steps,
                       s->d_sync);
    PBM_HIP(hipGetLastError());
    return PBM_OK;
}

}  // namespace

extern "C" int pbm_serve_run_loop(pbm_server* s, uint32_t n_steps, void* stream) {
    if (!s) return pbmi::fail(PBM_ERR_INVALID, "null server");
    PBM_CHECK_HANDLE(s->h);
    if (n_steps == 0) return PBM_OK;
    hipStream_t st = pbmi::as_stream(stream);
    switch (s->h->cfg.strategy) {
#define PBM_LOOP_CASE(S) \
    case S: return launch_loop_s<S>(s, n_steps, st);
        PBM_LOOP_CASE(PBM_RW_BM)
        PBM_LOOP_CASE(PBM_COLLAB_RW_BM)
        PBM_LOOP_CASE(PBM_CRW_BM)
        PBM_LOOP_CASE(PBM_RW_MALLOC)
        PBM_LOOP_CASE(PBM_LINKED_LIST)
        PBM_LOOP_CASE(PBM_SINGLE_CLOCK)
        PBM_LOOP_CASE(PBM_DIST_CLOCK)
        PBM_LOOP_CASE(PBM_RW_BASIC)
        PBM_LOOP_CASE(PBM_CRW_BASIC)
        PBM_LOOP_CASE(PBM_COLLAB_RW_BASIC)
        PBM_LOOP_CASE(PBM_OUROBOROS)
#undef PBM_LOOP_CASE
        default: return pbmi::fail(PBM_ERR_UNSUPPORTED, "no serving loop for this strategy");
    }
}

extern "C" int pbm_serve_loop_timeouts(pbm_server* s, unsigned long long* out, void* stream) {
    if (!s || !out) return pbmi::fail(PBM_ERR_INVALID, "null server or output");
    PBM_CHECK_HANDLE(s->h);
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(out, s->d_sync + 3, 8u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_serve_run_timed(pbm_server* s, uint32_t n_steps, float* kernel_ms,
                                   void* stream) {
    if (!s || !kernel_ms) return pbmi::fail(PBM_ERR_INVALID, "null server or output");
    PBM_CHECK_HANDLE(s->h);
    if (n_steps == 0 || n_steps > 4096u)
        return pbmi::fail(PBM_ERR_RANGE, "pbm_serve_run_timed runs 1 .. 4096 steps");
    hipStream_t st = pbmi::as_stream(stream);
    std::vector<hipEvent_t> ev((size_t)n_steps * 6u, nullptr);
    int rc = PBM_OK;
    for (hipEvent_t& e : ev) {
        const hipError_t he = hipEventCreate(&e);
        if (he != hipSuccess) {
            rc = pbmi::hip_fail(he, "hipEventCreate", __FILE__, __LINE__);
            break;
        }
    }
    for (uint32_t k = 0; k < n_steps && !rc; ++k) rc = launch_step(s, st, &ev[(size_t)k * 6u]);
    double sum[3] = {0, 0, 0};
    if (!rc) {
        const hipError_t he = hipStreamSynchronize(st);
        if (he != hipSuccess) rc = pbmi::hip_fail(he, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    for (uint32_t k = 0; k < n_steps && !rc; ++k)
        for (int j = 0; j < 3 && !rc; ++j) {
            float ms = 0;
            const hipError_t he =
                hipEventElapsedTime(&ms, ev[(size_t)k * 6u + 2u * j], ev[(size_t)k * 6u + 2u * j + 1u]);
            if (he != hipSuccess) rc = pbmi::hip_fail(he, "hipEventElapsedTime", __FILE__, __LINE__);
            sum[j] += ms;
        }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int j = 0; j < 3; ++j) kernel_ms[j] = (float)(sum[j] / n_steps);
    return PBM_OK;
}

namespace {

// Captures n_steps steps, as step kernels or as one device-resident loop
// launch, into a HIP graph.
int capture_steps(pbm_server* s, uint32_t n_steps, bool loop, pbm_graph** out) {
    if (!s || !out) return pbmi::fail(PBM_ERR_INVALID, "null server or output graph");
    PBM_CHECK_HANDLE(s->h);
    *out = nullptr;
    if (n_steps == 0) return pbmi::fail(PBM_ERR_INVALID, "a graph needs n_steps > 0");
    hipStream_t st = nullptr;
    PBM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    pbm_graph* g = new pbm_graph();
    g->device = s->h->device;
    int rc = PBM_OK;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = pbmi::hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    } else {
        if (loop) rc = pbm_serve_run_loop(s, n_steps, st);
        for (uint32_t k = 0; !loop && k < n_steps && !rc; ++k) rc = launch_step(s, st);
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

}  // namespace

extern "C" int pbm_serve_graph(pbm_server* s, uint32_t n_steps, pbm_graph** out) {
    return capture_steps(s, n_steps, false, out);
}

extern "C" int pbm_serve_graph_loop(pbm_server* s, uint32_t n_steps, pbm_graph** out) {
    return capture_steps(s, n_steps, true, out);
}

extern "C" int pbm_serve_read_stats(pbm_server* s, pbm_serve_stats* out, void* stream) {
    if (!s || !out) return pbmi::fail(PBM_ERR_INVALID, "null server or output");
    PBM_CHECK_HANDLE(s->h);
    u64 ctl[kCtlWords];
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(ctl, s->d.ctl, sizeof(ctl), hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    out->step = ctl[kStep];
    out->submitted = ctl[kTail];
    out->admitted = ctl[kAdmitted];
    out->finished = ctl[kFinished];
    out->active = ctl[kActive] - ctl[(ctl[kStep] & 1u) ? kDoneOdd : kDone];
    out->reserved_pages = ctl[kPromised] - ctl[kAllocs];
    out->allocs = ctl[kAllocs];
    out->frees = ctl[kFrees];
    out->alloc_failures = ctl[kAllocFail];
    out->payload_errors = ctl[kPayloadErr];
    out->lat_samples = ctl[kLatCount];
    return PBM_OK;
}

extern "C" int pbm_serve_read_results(pbm_server* s, uint32_t first, uint32_t n,
                                      uint32_t* admit_step, uint32_t* finish_step,
                                      uint32_t* length, unsigned long long* checksum,
                                      void* stream) {
    if (!s) return pbmi::fail(PBM_ERR_INVALID, "null server");
    PBM_CHECK_HANDLE(s->h);
    if ((uint64_t)first + n > s->cfg.max_seqs)
        return pbmi::fail(PBM_ERR_RANGE, "result range beyond max_seqs");
    hipStream_t st = pbmi::as_stream(stream);
    if (admit_step)
        PBM_HIP(hipMemcpyAsync(admit_step, s->d.res_admit + first, (size_t)n * 4u,
                               hipMemcpyDeviceToHost, st));
    if (finish_step)
        PBM_HIP(hipMemcpyAsync(finish_step, s->d.res_finish + first, (size_t)n * 4u,
                               hipMemcpyDeviceToHost, st));
    if (length)
        PBM_HIP(hipMemcpyAsync(length, s->d.res_len + first, (size_t)n * 4u,
                               hipMemcpyDeviceToHost, st));
    if (checksum)
        PBM_HIP(hipMemcpyAsync(checksum, s->d.res_sum + first, (size_t)n * 8u,
                               hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    return PBM_OK;
}

extern "C" int pbm_serve_read_latency(pbm_server* s, uint32_t* ticks, uint32_t n, uint32_t* got,
                                      void* stream) {
    if (!s || !got) return pbmi::fail(PBM_ERR_INVALID, "null server or output");
    PBM_CHECK_HANDLE(s->h);
    *got = 0;
    u64 cnt = 0;
    hipStream_t st = pbmi::as_stream(stream);
    PBM_HIP(hipMemcpyAsync(&cnt, s->d.ctl + kLatCount, 8u, hipMemcpyDeviceToHost, st));
    PBM_HIP(hipStreamSynchronize(st));
    u64 k = cnt < s->d.lat_cap ? cnt : s->d.lat_cap;
    if (k > n) k = n;
    if (k && !ticks) return pbmi::fail(PBM_ERR_INVALID, "null latency buffer");
    if (k) {
        PBM_HIP(hipMemcpyAsync(ticks, s->d.lat, (size_t)k * 4u, hipMemcpyDeviceToHost, st));
        PBM_HIP(hipStreamSynchronize(st));
    }
    *got = (uint32_t)k;
    return PBM_OK;
}

extern "C" int pbm_serve_clock_hz(pbm_server* s, unsigned long long* hz) {
    if (!s || !hz) return pbmi::fail(PBM_ERR_INVALID, "null server or output");
    PBM_CHECK_HANDLE(s->h);
    int khz = 0;
    PBM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->h->device));
    *hz = (unsigned long long)khz * 1000ull;
    return PBM_OK;
}
