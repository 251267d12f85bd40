// ns_overlap.hip -- the merge of an overlapped run.
//
// The step time of the binned integrator is set by a few particles whose
// adaptive chains run to the work cap (capped Levi-Civita plunges, about
// 2^17 dependent micro-steps on one wave), while the rest of the chip idles.
// An overlapped run integrates those particles on a side stream (the late
// kernel) and meanwhile sorts the next step's radii on the main stream, with
// the late particles' keys replaced by a sentinel above every radius key, so
// that the sort leaves them last in their halo's segment (the whole array
// when there are no halos), in rank order (the sort is stable).  Once the
// late kernel is done, the merge inserts them where a stable sort of the true
// keys puts them, halo by halo:
//   * a late particle of halo h at old rank t with key k goes after every
//     early element of h whose (key, old rank) is below (k, t) -- lo, a
//     search over h's sorted early elements -- and after the late particles
//     of h below it; the merged order is then the stable sort of all the
//     keys, bit for bit;
//   * an early element of h at local sorted position e moves up by the number
//     of late particles of h inserted at or before e.
// Which particles are deferred: those whose previous step was slow
// (slow_region: capped or deep, else among the innermost ranks of their halo,
// else costly in force evaluations; the capped chains persist from step to
// step), found through the slow list
// the integrate kernels append to.  The choice affects only the time, never
// the results.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_kernels.h"
#include "ns_radix.h"

namespace ns {

namespace {

constexpr int kPlanThreads = 1024;  // == kSlowCap: one thread per entry
constexpr int kSearchThreads = 256;
constexpr int kShiftThreads = 256;
constexpr int kShiftItems = 16;
constexpr int kShiftTile = kShiftThreads * kShiftItems;
constexpr int kShiftPairs = kSlowCap / kShiftThreads;  // insertion pairs per thread, at most
static_assert(kSlowCap % kShiftThreads == 0, "whole insertion pairs per thread");
constexpr int kShiftHalos = 64;  // halo bounds of a shift tile held in LDS
static_assert(kPlanThreads == kSlowCap, "one plan thread per slow-list entry");

// The halos as the merge sees them: halo h holds positions
// [start(h), start(h + 1)); without halos, one halo of all n.
struct Halos {
    const int64_t *off;  // nseg + 1 offsets, or null
    int32_t nseg;
    uint32_t n;
    __device__ inline uint32_t start(uint32_t h) const { return off ? (uint32_t)off[h] : (h ? n : 0u); }
    __device__ inline uint32_t size(uint32_t h) const { return start(h + 1) - start(h); }
    __device__ inline uint32_t count() const { return off ? (uint32_t)nseg : 1u; }
    // halo of position p < n
    __device__ inline uint32_t of(uint32_t p) const
    {
        if (!off)
            return 0;
        uint32_t lo = 0, hi = (uint32_t)nseg;  // off[lo] <= p < off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint32_t)off[mid] <= p)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    }
};

__device__ inline uint64_t hl_pack(uint32_t h, uint32_t lo) { return ((uint64_t)h << 32) | lo; }

// Entries of the ascending array a[0, m) that are <= x (upper bound).
__device__ inline uint32_t count_le(const uint64_t *a, uint32_t m, uint64_t x)
{
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Entries of the ascending array a[0, m) that are < x (lower bound).
__device__ inline uint32_t count_lt(const uint64_t *a, uint32_t m, uint64_t x)
{
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_overlap_reset(uint32_t *__restrict__ ctl)
{
    if (threadIdx.x < 8)
        ctl[threadIdx.x] = 0;
}

// The slow list of the step that left the level records lvl (one per
// position of the current order, so every entry is a distinct position).
__global__ void k_overlap_seed(OverlapBufs o, Halos g, const uint8_t *__restrict__ lvl)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = t < g.n;  // every lane reaches the pushes
    const uint8_t rec = in ? lvl[t] : (uint8_t)0;
    const int reg = in ? slow_region(rec, t - g.start(g.of(t))) : -1;  // no force evaluations here: never class 3
#pragma unroll
    for (int r = 0; r < kSlowRegions; r++)
        slow_push(o.ctl, o.slow, o.slow_rec, r, reg == r, t, (uint32_t)rec);
}

template <typename S>
__device__ inline typename KeyOf<S>::type key_at(const S *__restrict__ r, int64_t p)
{
    return radius_key(r[p]);
}

// Whether the early element at sorted position e is below (k, t) in (key,
// old rank) order.
template <typename S>
__device__ inline bool below(const uint32_t *__restrict__ sorted, const S *__restrict__ r, uint32_t e,
                             typename KeyOf<S>::type k, uint32_t t)
{
    const uint32_t p = sorted[e];
    const typename KeyOf<S>::type km = key_at(r, p);
    return km < k || (km == k && p < t);
}

// Early elements at sorted positions [base, base + E) below (k, t), found by
// one wave: each round its 64 lanes probe 64 evenly spaced positions of the
// range that holds the answer (the probes below form a prefix, being
// sorted), so that 1.2e7 elements take four rounds of two dependent loads.
// Wave-uniform.
template <typename S>
__device__ inline uint32_t wave_count_below(const uint32_t *__restrict__ sorted, const S *__restrict__ r,
                                            uint32_t base, uint32_t E, typename KeyOf<S>::type k, uint32_t t)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = E;  // the count lies in [lo, hi]
    while (hi > lo) {
        const uint32_t stride = (hi - lo + 63) / 64;
        const uint32_t p = lo + lane * stride;
        const bool b = p < hi && below(sorted, r, base + p, k, t);
        const uint32_t nb = (uint32_t)__popcll(__ballot(b));
        if (nb == 0) {
            hi = lo;
        } else {  // between the last probe below and the next probe (or hi)
            const uint32_t next = lo + nb * stride;
            lo = lo + (nb - 1) * stride + 1;
            hi = next < hi ? next : hi;
        }
    }
    return lo;
}

__device__ inline uint32_t late_count(const OverlapBufs &o, uint32_t n)
{
    uint32_t K = o.ctl[1];
    if (K > (uint32_t)kSlowCap)
        K = kSlowCap;
    return K > n ? n : K;
}

// Slow-list entries selected for deferral (k_overlap_order).
__device__ inline uint32_t selected_count(const OverlapBufs &o)
{
    const uint32_t m = o.ctl[4];
    return m > (uint32_t)kSlowCap ? (uint32_t)kSlowCap : m;
}

// Cost order of the deferred ranks (the record word of slow_word): class 2
// first (capped chains first, then by level), then class 1 (by level), then
// class 3 (by spent bin), so that class 3 fills only the places the others
// leave.
constexpr int kCostBins = 130;
__device__ inline uint32_t slow_cost(uint32_t word, int cls)
{
    const uint32_t level = word & NS_LVL_MASK, bin = (word >> 8) & 63u;
    if (cls == 2)
        return 65u + ((word & NS_LVL_NOCONV) ? 64u : level);
    if (cls == 1)
        return 32u + (level < 32u ? level : 32u);
    return bin < 31u ? bin : 31u;
}

// The selection's entries per thread of k_overlap_order: entry u * kPlanThreads
// + tid of the slow list for u < kSelPer, in region region_of_block(u).
constexpr int kSelPer = kSlowLists / kPlanThreads;
static_assert(kSlowLists % kPlanThreads == 0 && kPlanThreads == kSlowCap, "whole blocks of entries per thread");
__host__ __device__ constexpr int region_of_block(int u)
{
    int g = 0;
    while (g + 1 < kSlowRegions && slow_region_first(g + 1) <= (uint32_t)u * kPlanThreads)
        g++;
    return g;
}

// Late particles of halo h in the (halo, key, rank)-sorted late list
// late_h[0, K): [first, first + count).
__device__ inline void late_run(const uint32_t *late_h, uint32_t K, uint32_t h, uint32_t &first, uint32_t &count)
{
    uint32_t lo = 0, hi = K;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (late_h[mid] < h)
            lo = mid + 1;
        else
            hi = mid;
    }
    first = lo;
    hi = K;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (late_h[mid] <= h)
            lo = mid + 1;
        else
            hi = mid;
    }
    count = lo - first;
}

// One block.  The ctl[1] late particles (old ranks def_pos[j]) in (halo,
// key, old rank) order by a bitonic sort of one entry per thread held in
// registers: stages whose partner (tid ^ j) lies in the same wave exchange
// through shuffles, the others (j >= 64, ten stages for 1024 entries)
// through two alternating LDS buffers, one barrier each.  Into late_h /
// late_t / late_slot[r] for the r-th.  And the slow-list entries to defer:
// the costliest kSlowCap of all regions by a counting sort on the cost (the
// order within a cost is free: it decides only which wave runs which
// particle), into sel / ctl[4].
template <typename S>
__global__ __launch_bounds__(kPlanThreads) void k_overlap_order(OverlapBufs o, Halos g, const S *__restrict__ r)
{
    using KeyT = typename KeyOf<S>::type;
    __shared__ uint32_t s_h[2][kSlowCap];
    __shared__ KeyT s_key[2][kSlowCap];
    __shared__ uint32_t s_old[2][kSlowCap];
    __shared__ uint32_t s_slot[2][kSlowCap];
    __shared__ uint32_t s_bin[kCostBins];  // cost histogram, then the first slot of each cost
    __shared__ uint32_t s_wtot[3];
    const uint32_t tid = threadIdx.x;
    const uint32_t K = late_count(o, g.n);
    if (tid < (uint32_t)kCostBins)
        s_bin[tid] = 0;
    uint32_t P = 1;  // padded above every real entry to a power of two
    while (P < K)
        P <<= 1;
    // entry tid; padding entries sort above every real one, and are distinct
    // from each other (so that no comparison meets a tie)
    uint32_t h = ~0u, t = ~0u - tid, slot = tid;
    KeyT key = ~KeyT(0);
    if (tid < K)
        t = o.def_pos[tid];
    // the selection's inputs (entry u * kPlanThreads + tid of the slow list,
    // and the region counts): loaded now, used after the sort
    uint32_t rec[kSelPer], ent[kSelPer], cnt[kSlowRegions];
#pragma unroll
    for (int u = 0; u < kSelPer; u++) {
        rec[u] = o.slow_rec[(uint32_t)u * kPlanThreads + tid];
        ent[u] = o.slow[(uint32_t)u * kPlanThreads + tid];
    }
#pragma unroll
    for (int r = 0; r < kSlowRegions; r++)
        cnt[r] = o.ctl[slow_region_ctl(r)];
    if (tid < K) {
        h = g.of(t);
        key = key_at(r, t);
    }
    int buf = 0;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            uint32_t hb, tb, sb;
            KeyT kb;
            if (j >= 64) {
                s_h[buf][tid] = h;
                s_key[buf][tid] = key;
                s_old[buf][tid] = t;
                s_slot[buf][tid] = slot;
                lds_barrier();
                const uint32_t l = tid ^ j;
                hb = s_h[buf][l];
                kb = s_key[buf][l];
                tb = s_old[buf][l];
                sb = s_slot[buf][l];
                buf ^= 1;  // the next LDS stage writes the other buffer
            } else {
                hb = __shfl_xor(h, (int)j, 64);
                kb = __shfl_xor(key, (int)j, 64);
                tb = __shfl_xor(t, (int)j, 64);
                sb = __shfl_xor(slot, (int)j, 64);
            }
            const bool b_above = hb > h || (hb == h && (kb > key || (kb == key && tb > t)));
            // ascending where bit k of tid is clear: the lower index of the
            // pair keeps the smaller entry there, the upper one the larger
            const bool keep_min = ((tid & j) == 0) == ((tid & k) == 0);
            if (keep_min != b_above) {
                h = hb;
                key = kb;
                t = tb;
                slot = sb;
            }
        }
    }
    if (tid < K) {
        o.late_h[tid] = h;
        o.late_t[tid] = t;
        o.late_slot[tid] = slot;
    }
    lds_barrier();  // the cleared cost bins (no LDS stage ran when P < 128)
    // selection: the entries each region holds (its count, at most its capacity)
    uint32_t cost[kSelPer], within[kSelPer], m_all = 0;
#pragma unroll
    for (int r = 0; r < kSlowRegions; r++)
        m_all += cnt[r] < slow_region_cap(r) ? cnt[r] : slow_region_cap(r);
#pragma unroll
    for (int u = 0; u < kSelPer; u++) {
        const int r = region_of_block(u);
        const uint32_t k = (uint32_t)u * kPlanThreads + tid - slow_region_first(r);  // index within its region
        const bool have = k < cnt[r] && k < slow_region_cap(r);
        cost[u] = have ? slow_cost(rec[u], slow_region_class(r)) : 0u;
        within[u] = have ? atomicAdd(&s_bin[cost[u]], 1u) : ~0u;
    }
    __syncthreads();
    {   // first slot of each cost, costliest first: an exclusive scan over
        // the bins from the top, thread q taking bin kCostBins - 1 - q
        static_assert(kCostBins <= 192, "three waves scan the cost bins");
        const uint32_t lane = tid & 63, wave = tid >> 6;
        const int bq = kCostBins - 1 - (int)tid;
        const uint32_t x = bq >= 0 ? s_bin[bq] : 0u;
        uint32_t incl = x;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off)
                incl += y;
        }
        if (lane == 63 && wave < 3)
            s_wtot[wave] = incl;
        __syncthreads();
        const uint32_t before = (wave > 0 ? s_wtot[0] : 0u) + (wave > 1 ? s_wtot[1] : 0u);
        if (bq >= 0)
            s_bin[bq] = before + incl - x;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSelPer; u++) {
        if (within[u] == ~0u)
            continue;
        const uint32_t slot = s_bin[cost[u]] + within[u];
        if (slot < (uint32_t)kSlowCap)
            o.sel[slot] = ent[u];
    }
    if (tid == 0)
        o.ctl[4] = m_all < (uint32_t)kSlowCap ? m_all : (uint32_t)kSlowCap;
}

// One wave per query: wave w < kSlowCap is the w-th late particle (its
// insertion point among its halo's early elements), wave kSlowCap + q is
// selected entry q when it is an early element (its position among its
// halo's early elements).  Results in found[w].
template <typename S>
__global__ __launch_bounds__(kSearchThreads) void k_overlap_search(OverlapBufs o, Halos g,
                                                                   const uint32_t *__restrict__ sorted,
                                                                   const S *__restrict__ r)
{
    const uint32_t w = blockIdx.x * (kSearchThreads / 64) + (threadIdx.x >> 6);
    const uint32_t K = late_count(o, g.n), m = selected_count(o);
    uint32_t t, h;
    if (w < (uint32_t)kSlowCap) {
        if (w >= K)
            return;
        t = o.late_t[w];
        h = o.late_h[w];
    } else {
        const uint32_t q = w - (uint32_t)kSlowCap;
        if (q >= m)
            return;
        t = o.sel[q];
        if (t & kDeferBit)
            return;
        h = g.of(t);
    }
    uint32_t first, kh;
    late_run(o.late_h, K, h, first, kh);
    const uint32_t c = wave_count_below(sorted, r, g.start(h), g.size(h) - kh, key_at(r, t), t);
    if ((threadIdx.x & 63) == 0)
        o.found[w] = c;
}

// One block, after k_overlap_order and k_overlap_search.  Writes the late
// particles' merged positions into perm and their (halo, insertion point)
// pairs, ascending, into ins; the selected entries (the previous step's
// slow list, costliest first) at their new ranks, with their old positions,
// into def_pos / def_src; their count into ctl[1], the late count into
// ctl[2]; clears the slow list's region counts.
__global__ __launch_bounds__(kPlanThreads) void k_overlap_plan(OverlapBufs o, Halos g, unsigned long long *chain)
{
    // the previous late launch has ended (same stream): fold its longest chain
    if (chain && threadIdx.x == 0) {
        chain[kChainSum] += chain[kChainMax];
        chain[kChainMax] = 0;
    }
    __shared__ uint64_t s_ins[kSlowCap];  // (halo, insertion point), ascending
    __shared__ uint32_t s_pos[kSlowCap];  // merged position of late slot j
    __shared__ uint32_t s_t[kSlowCap];    // old rank of late slot j
    const uint32_t tid = threadIdx.x;
    const uint32_t K = late_count(o, g.n), m = selected_count(o);
    if (tid < K) {  // the tid-th late particle in (halo, key, old rank) order
        const uint32_t h = o.late_h[tid], t = o.late_t[tid], j = o.late_slot[tid];
        uint32_t first, kh;
        late_run(o.late_h, K, h, first, kh);
        const uint32_t lo = o.found[tid];
        const uint32_t pos = g.start(h) + lo + (tid - first);
        s_pos[j] = pos;
        s_t[j] = t;
        s_ins[tid] = hl_pack(h, lo);
        o.ins[tid] = hl_pack(h, lo);
        o.perm[pos] = t;
    }
    __syncthreads();
    if (tid < m) {
        const uint32_t v = o.sel[tid];
        uint32_t pos, src;
        if (v & kDeferBit) {  // a late particle of this merge
            pos = s_pos[v & kIndexMask];
            src = s_t[v & kIndexMask];
        } else {  // an early element, moved past its halo's late particles inserted at or before it
            const uint32_t h = g.of(v), e = o.found[kSlowCap + tid];
            pos = g.start(h) + e + count_le(s_ins, K, hl_pack(h, e)) - count_lt(s_ins, K, hl_pack(h, 0));
            src = v;
        }
        o.def_pos[tid] = pos;
        o.def_src[tid] = src;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int r = 0; r < kSlowRegions; r++)
            o.ctl[slow_region_ctl(r)] = 0;
        o.ctl[1] = m;
        o.ctl[2] = K;
    }
}

// Early elements to their merged positions.  Sorted position p of halo h
// with local index e = p - start(h) < size(h) - (late particles of h) moves
// to p + (late particles of h inserted at or before e); the positions past
// that are h's late particles, placed by the plan.  Each block takes
// kShiftTile positions; the bounds of their halos and those halos'
// insertion pairs are staged in LDS (a tile spanning more than kShiftHalos
// halos searches the halo bounds in global memory).
__global__ __launch_bounds__(kShiftThreads) void k_overlap_shift(OverlapBufs o, Halos g,
                                                                const uint32_t *__restrict__ sorted)
{
    __shared__ uint64_t s_all[kSlowCap];
    __shared__ uint32_t s_start[kShiftHalos + 1];
    __shared__ uint32_t s_before[kShiftHalos], s_kh[kShiftHalos];  // each staged halo's run of pairs
    __shared__ uint32_t s_wlo[kShiftHalos], s_whi[kShiftHalos];    // ... and its pairs inside the tile
    __shared__ uint32_t s_rng[4];  // first halo, halo count, first insertion pair, pair count
    uint32_t K = o.ctl[2];
    K = K < (uint32_t)kSlowCap ? K : (uint32_t)kSlowCap;
    const uint32_t b0 = blockIdx.x * (uint32_t)kShiftTile;
    if (b0 >= g.n)
        return;
    const uint32_t b1 = b0 + (uint32_t)kShiftTile < g.n ? b0 + (uint32_t)kShiftTile : g.n;
    // every insertion pair, then the tile's sorted positions, all loads in
    // flight at once (one round trip); the pairs go to LDS, so that the
    // tile's range of pairs is found by searches in LDS rather than chains
    // of dependent global loads
    uint64_t pr[kShiftPairs];
#pragma unroll
    for (int u = 0; u < kShiftPairs; u++) {
        const uint32_t q = threadIdx.x + (uint32_t)u * kShiftThreads;
        pr[u] = q < K ? o.ins[q] : ~0ull;
    }
    uint32_t val[kShiftItems];
#pragma unroll
    for (int it = 0; it < kShiftItems; it++) {
        const uint32_t p = b0 + (uint32_t)it * kShiftThreads + threadIdx.x;
        val[it] = p < b1 ? sorted[p] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kShiftPairs; u++) {
        const uint32_t q = threadIdx.x + (uint32_t)u * kShiftThreads;
        if (q < K)
            s_all[q] = pr[u];
    }
    if (threadIdx.x < 2)
        s_rng[threadIdx.x] = 0;
    lds_barrier();  // the positions' loads stay in flight
    if (!g.off) {
        // one halo: pairs (0, e), e ascending.  The tile's pairs are those
        // with b0 <= e <= b1 - 1, [wlo, whi), counted by all threads at once
        uint32_t clo = 0, chi = 0;
#pragma unroll
        for (int u = 0; u < kShiftPairs; u++) {
            const uint32_t q = threadIdx.x + (uint32_t)u * kShiftThreads;
            const uint32_t e = (uint32_t)pr[u];
            clo += q < K && e < b0 ? 1u : 0u;
            chi += q < K && e <= b1 - 1 ? 1u : 0u;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            clo += __shfl_xor(clo, off, 64);
            chi += __shfl_xor(chi, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&s_rng[0], clo);
            atomicAdd(&s_rng[1], chi);
        }
        lds_barrier();
        const uint32_t wlo = s_rng[0], whi = s_rng[1], E = g.n - K;  // E: the early elements
#pragma unroll
        for (int it = 0; it < kShiftItems; it++) {
            const uint32_t p = b0 + (uint32_t)it * kShiftThreads + threadIdx.x;
            if (p >= b1 || p >= E)
                break;
            uint32_t lo = wlo, hi = whi;  // the pairs with e <= p number lo
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((uint32_t)s_all[mid] <= p)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            o.perm[p + lo] = val[it];
        }
        return;
    }
    if (threadIdx.x == 0) {
        const uint32_t h0 = g.of(b0), h1 = g.of(b1 - 1);
        const uint32_t i0 = count_lt(s_all, K, hl_pack(h0, 0));
        const uint32_t i1 = count_le(s_all, K, hl_pack(h1, ~0u));
        s_rng[0] = h0;
        s_rng[1] = h1 - h0 + 1;
        s_rng[2] = i0;
        s_rng[3] = i1 - i0;
    }
    __syncthreads();
    const uint32_t h0 = s_rng[0], nh = s_rng[1], i0 = s_rng[2], ni = s_rng[3];
    const bool staged = nh <= (uint32_t)kShiftHalos;
    const uint64_t *s_ins = s_all + i0;  // the tile's pairs
    if (staged)
        for (uint32_t q = threadIdx.x; q <= nh; q += kShiftThreads)
            s_start[q] = g.start(h0 + q);
    __syncthreads();
    if (staged) {
        for (uint32_t q = threadIdx.x; q < nh; q += kShiftThreads) {
            const uint32_t before = count_lt(s_ins, ni, hl_pack(h0 + q, 0));
            const uint32_t kh = count_le(s_ins, ni, hl_pack(h0 + q, ~0u)) - before;
            s_before[q] = before;
            s_kh[q] = kh;
            // the halo's pairs before the tile's first local index ef number
            // wlo; [wlo, whi) are those up to its last one el: an element
            // searches only those (usually none)
            const uint32_t hs = s_start[q], he = s_start[q + 1];
            uint32_t wlo = 0, whi = 0;
            if (he > hs) {
                const uint32_t ef = b0 > hs ? b0 - hs : 0;
                const uint32_t el = (b1 < he ? b1 : he) - 1 - hs;
                wlo = count_lt(s_ins + before, kh, hl_pack(h0 + q, ef));
                whi = count_le(s_ins + before, kh, hl_pack(h0 + q, el));
            }
            s_wlo[q] = wlo;
            s_whi[q] = whi;
        }
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < kShiftItems; it++) {
        const uint32_t p = b0 + (uint32_t)it * kShiftThreads + threadIdx.x;
        if (p >= b1)
            break;
        uint32_t hs, he, before, kh, wlo, whi;
        if (staged) {
            uint32_t lo = 0, hi = nh;  // s_start[lo] <= p < s_start[hi]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_start[mid] <= p)
                    lo = mid;
                else
                    hi = mid;
            }
            hs = s_start[lo];
            he = s_start[lo + 1];
            before = s_before[lo];
            kh = s_kh[lo];
            wlo = s_wlo[lo];
            whi = s_whi[lo];
        } else {
            const uint32_t h = g.of(p);
            hs = g.start(h);
            he = g.start(h + 1);
            before = count_lt(s_ins, ni, hl_pack(h, 0));
            kh = count_le(s_ins, ni, hl_pack(h, ~0u)) - before;
            wlo = 0;
            whi = kh;
        }
        const uint32_t e = p - hs;
        if (e < he - hs - kh) {
            // insertion points of this halo at or before e: its run is
            // s_ins[before, before + kh), of which those before wlo lie
            // below e and those from whi on above it
            uint32_t lo = wlo, hi = whi;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((uint32_t)s_ins[before + mid] <= e)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            o.perm[p + lo] = val[it];
        }
    }
}

template <typename S>
__global__ __launch_bounds__(kPlanThreads) void k_overlap_fix_keys(OverlapBufs o, const S *__restrict__ r,
                                                                   void *__restrict__ keys, int64_t n,
                                                                   const QEntry *__restrict__ qtab, uint32_t stride)
{
    const uint32_t j = threadIdx.x;
    if (j < o.ctl[1]) {
        const uint32_t t = o.def_pos[j];
        if constexpr (sizeof(S) == 8) {
            if (qtab) {
                static_cast<uint32_t *>(keys)[t] = quantile_key(qtab, stride, t, radius_key(r[t]));
                return;
            }
        }
        store_key(keys, n, t, radius_key(r[t]));
    }
}

Halos halos_of(const int64_t *seg_off, int nseg, int64_t n)
{
    Halos g;
    g.off = nseg > 0 ? seg_off : nullptr;
    g.nseg = nseg;
    g.n = (uint32_t)n;
    return g;
}

template <typename S>
hipError_t plan_typed(const OverlapBufs &o, const Halos &g, const uint32_t *sorted, const S *r,
                      unsigned long long *chain, hipStream_t stream)
{
    hipLaunchKernelGGL(k_overlap_search<S>, dim3(2 * kSlowCap / (kSearchThreads / 64)), dim3(kSearchThreads), 0,
                       stream, o, g, sorted, r);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_overlap_plan, dim3(1), dim3(kPlanThreads), 0, stream, o, g, chain);
    return hipGetLastError();
}

} // namespace

hipError_t overlap_reset(const OverlapBufs &o, const uint8_t *lvl, const int64_t *seg_off, int nseg, int64_t n,
                         hipStream_t stream)
{
    hipLaunchKernelGGL(k_overlap_reset, dim3(1), dim3(64), 0, stream, o.ctl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !lvl || n <= 0)
        return e;
    hipLaunchKernelGGL(k_overlap_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, o,
                       halos_of(seg_off, nseg, n), lvl);
    return hipGetLastError();
}

hipError_t overlap_order(const OverlapBufs &o, const void *r, bool f32, const int64_t *seg_off, int nseg, int64_t n,
                         hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    if (n > (int64_t)kIndexMask)
        return hipErrorInvalidValue;
    const Halos g = halos_of(seg_off, nseg, n);
    if (f32)
        hipLaunchKernelGGL(k_overlap_order<float>, dim3(1), dim3(kPlanThreads), 0, stream, o, g,
                           static_cast<const float *>(r));
    else
        hipLaunchKernelGGL(k_overlap_order<double>, dim3(1), dim3(kPlanThreads), 0, stream, o, g,
                           static_cast<const double *>(r));
    return hipGetLastError();
}

hipError_t overlap_plan(const OverlapBufs &o, const uint32_t *sorted, const void *r, bool f32, const int64_t *seg_off,
                        int nseg, int64_t n, unsigned long long *chain, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    if (n > (int64_t)kIndexMask)
        return hipErrorInvalidValue;
    const Halos g = halos_of(seg_off, nseg, n);
    return f32 ? plan_typed<float>(o, g, sorted, static_cast<const float *>(r), chain, stream)
               : plan_typed<double>(o, g, sorted, static_cast<const double *>(r), chain, stream);
}

hipError_t overlap_place(const OverlapBufs &o, const uint32_t *sorted, const int64_t *seg_off, int nseg, int64_t n,
                         hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_overlap_shift, dim3((unsigned)((n + kShiftTile - 1) / kShiftTile)), dim3(kShiftThreads), 0,
                       stream, o, halos_of(seg_off, nseg, n), sorted);
    return hipGetLastError();
}

hipError_t overlap_fix_keys(const OverlapBufs &o, const void *r, bool f32, void *keys, int64_t n,
                            const QEntry *qtab, uint32_t q_stride, hipStream_t stream)
{
    if (f32)
        hipLaunchKernelGGL(k_overlap_fix_keys<float>, dim3(1), dim3(kPlanThreads), 0, stream, o,
                           static_cast<const float *>(r), keys, n, nullptr, 0u);
    else
        hipLaunchKernelGGL(k_overlap_fix_keys<double>, dim3(1), dim3(kPlanThreads), 0, stream, o,
                           static_cast<const double *>(r), keys, n, qtab, q_stride);
    return hipGetLastError();
}

} // namespace ns
