// ns_overlap.hip -- the merge of an overlapped run.
//
// The step time of the binned integrator is set by a few particles whose
// adaptive chains run to the work cap (capped Levi-Civita plunges, about
// 2^17 dependent micro-steps on one wave), while the rest of the chip idles.
// An overlapped run integrates those particles on a side stream (the late
// kernel) and meanwhile sorts the next step's radii on the main stream, with
// the late particles' keys replaced by a sentinel above every radius key, so
// that the sort leaves them last, in rank order (the sort is stable).  Once
// the late kernel is done, the merge inserts them where a stable sort of the
// true keys puts them:
//   * a late particle at old rank t with key k goes after every early element
//     whose (key, old rank) is below (k, t) -- lo, a binary search over the
//     sorted early elements -- and after the late particles below it; the
//     merged order is then the stable sort of all the keys, bit for bit;
//   * an early element at sorted position e moves up by the number of late
//     particles inserted at or before e.
// Which particles are deferred: those whose previous step was slow
// (slow_record: capped, deep, or among the innermost ranks; the capped chains
// persist from step to step), found through the slow list the integrate
// kernels append to.  The choice affects only the time, never the results.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

namespace {

constexpr int kPlanThreads = 1024;  // == kSlowCap: one thread per entry
constexpr int kSearchThreads = 256;
constexpr int kShiftThreads = 256;
constexpr int kShiftItems = 4;
static_assert(kPlanThreads == kSlowCap, "one plan thread per slow-list entry");

__global__ void k_overlap_reset(uint32_t *__restrict__ ctl)
{
    if (threadIdx.x < 4)
        ctl[threadIdx.x] = 0;
}

// The slow list of the step that left the level records lvl (one per
// position of the current order, so every entry is a distinct position).
__global__ void k_overlap_seed(OverlapBufs o, const uint8_t *__restrict__ lvl, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && slow_record(lvl[t], t)) {
        const uint32_t k = atomicAdd(o.ctl, 1u);
        if (k < (uint32_t)kSlowCap) {
            o.slow[k] = t;
            o.slow_rec[k] = lvl[t];
        }
    }
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

// Early elements (sorted positions [0, E)) below (k, t), found by one wave:
// each round its 64 lanes probe 64 evenly spaced positions of the range that
// holds the answer (the probes below form a prefix, being sorted), so that
// 1.2e7 elements take four rounds of two dependent loads.  Wave-uniform.
template <typename S>
__device__ inline uint32_t wave_count_below(const uint32_t *__restrict__ sorted, const S *__restrict__ r, uint32_t E,
                                            typename KeyOf<S>::type k, uint32_t t)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = E;  // the count lies in [lo, hi]
    while (hi > lo) {
        const uint32_t stride = (hi - lo + 63) / 64;
        const uint32_t p = lo + lane * stride;
        const bool b = p < hi && below(sorted, r, p, k, t);
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

// One wave per query: wave w < kSlowCap is late slot w (its insertion
// point), wave kSlowCap + q is slow-list entry q when it is an early element
// (its sorted position).  Results in found[w].
template <typename S>
__global__ __launch_bounds__(kSearchThreads) void k_overlap_search(OverlapBufs o, const uint32_t *__restrict__ sorted,
                                                                   const S *__restrict__ r, uint32_t n)
{
    const uint32_t w = blockIdx.x * (kSearchThreads / 64) + (threadIdx.x >> 6);
    uint32_t K = o.ctl[1];
    if (K > (uint32_t)kSlowCap)
        K = kSlowCap;
    if (K > n)
        K = n;
    uint32_t m = o.ctl[0];
    if (m > (uint32_t)kSlowCap)
        m = kSlowCap;
    uint32_t t;
    if (w < (uint32_t)kSlowCap) {
        if (w >= K)
            return;
        t = o.def_pos[w];
    } else {
        const uint32_t q = w - (uint32_t)kSlowCap;
        if (q >= m)
            return;
        t = o.slow[q];
        if (t & kDeferBit)
            return;
    }
    const uint32_t c = wave_count_below(sorted, r, n - K, key_at(r, t), t);
    if ((threadIdx.x & 63) == 0)
        o.found[w] = c;
}

// Cost order of the deferred ranks: capped chains first, then by level.
constexpr int kCostBins = 65;
__device__ inline uint32_t slow_cost(uint32_t rec)
{
    return (rec & NS_LVL_NOCONV) ? 64u : (rec & NS_LVL_MASK);
}

// One block, after k_overlap_search.  ctl[1] late particles (old ranks
// def_pos[j]) were sorted last; the previous step's slow list (ctl[0]
// entries) names the ranks to defer in the coming step.  Ranks the late
// particles among themselves by a bitonic sort on (key, old rank) and writes
// their merged positions into perm and their insertion points (ascending)
// into ins_lo; orders the coming step's deferred ranks costliest first by a
// counting sort on the cost (the order within a cost is free: it decides
// only which wave runs which particle) and writes them with their old
// positions into def_pos / def_src; their count into ctl[1], the late count
// into ctl[2]; clears ctl[0].
template <typename S>
__global__ __launch_bounds__(kPlanThreads) void k_overlap_plan(OverlapBufs o, const S *__restrict__ r, uint32_t n)
{
    using KeyT = typename KeyOf<S>::type;
    __shared__ KeyT s_key[kSlowCap];        // late keys, sorted in place
    __shared__ uint32_t s_old[kSlowCap];    // their old ranks, sorted alongside
    __shared__ uint32_t s_slot[kSlowCap];   // their late slots, sorted alongside
    __shared__ uint32_t s_t[kSlowCap];      // old rank of late slot j
    __shared__ uint32_t s_pos[kSlowCap];    // merged position of late slot j
    __shared__ uint32_t s_lo[kSlowCap];     // insertion points, ascending
    __shared__ uint32_t s_bin[kCostBins];   // cost histogram, then the first slot of each cost
    const uint32_t tid = threadIdx.x;
    uint32_t K = o.ctl[1];
    if (K > (uint32_t)kSlowCap)
        K = kSlowCap;
    if (K > n)
        K = n;
    uint32_t m = o.ctl[0];
    if (m > (uint32_t)kSlowCap)
        m = kSlowCap;
    // late particles, padded above every real (key, rank) to a power of two
    uint32_t P = 1;
    while (P < K)
        P <<= 1;
    if (tid < K) {
        const uint32_t t = o.def_pos[tid];
        s_key[tid] = key_at(r, t);
        s_old[tid] = t;
        s_t[tid] = t;
    } else {
        s_key[tid] = ~KeyT(0);
        s_old[tid] = ~0u;
    }
    s_slot[tid] = tid;
    if (tid < (uint32_t)kCostBins)
        s_bin[tid] = 0;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t l = tid ^ j;
            if (tid < P && l > tid) {
                const KeyT ka = s_key[tid], kb = s_key[l];
                const uint32_t ta = s_old[tid], tb = s_old[l];
                const bool a_above = ka > kb || (ka == kb && ta > tb);
                if (a_above == ((tid & k) == 0)) {  // ascending where bit k of tid is clear
                    s_key[tid] = kb;
                    s_key[l] = ka;
                    s_old[tid] = tb;
                    s_old[l] = ta;
                    const uint32_t x = s_slot[tid];
                    s_slot[tid] = s_slot[l];
                    s_slot[l] = x;
                }
            }
            __syncthreads();
        }
    }
    if (tid < K) {  // the tid-th late particle in (key, old rank) order
        const uint32_t j = s_slot[tid];
        const uint32_t lo = o.found[j];
        s_pos[j] = lo + tid;
        s_lo[tid] = lo;
        o.ins_lo[tid] = lo;
        o.perm[lo + tid] = s_old[tid];
    }
    uint32_t v = 0, c = 0, within = 0;
    if (tid < m) {
        v = o.slow[tid];
        c = slow_cost(o.slow_rec[tid]);
        within = atomicAdd(&s_bin[c], 1u);
    }
    __syncthreads();
    if (tid == 0) {  // first slot of each cost, costliest first
        uint32_t acc = 0;
        for (int b = kCostBins - 1; b >= 0; b--) {
            const uint32_t x = s_bin[b];
            s_bin[b] = acc;
            acc += x;
        }
    }
    __syncthreads();
    if (tid < m) {
        uint32_t pos, src;
        if (v & kDeferBit) {  // a late particle of this merge
            pos = s_pos[v & kIndexMask];
            src = s_t[v & kIndexMask];
        } else {  // an early element, moved past the late particles inserted at or before it
            const uint32_t e = o.found[kSlowCap + tid];
            uint32_t lo = 0, hi = K;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_lo[mid] <= e)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            pos = e + lo;
            src = v;
        }
        const uint32_t slot = s_bin[c] + within;
        o.def_pos[slot] = pos;
        o.def_src[slot] = src;
    }
    __syncthreads();
    if (tid == 0) {
        o.ctl[0] = 0;
        o.ctl[1] = m;
        o.ctl[2] = K;
    }
}

// Early elements to their merged positions: sorted[e] -> perm[e + shift(e)],
// shift(e) = late particles with insertion point <= e.
__global__ __launch_bounds__(kShiftThreads) void k_overlap_shift(OverlapBufs o, const uint32_t *__restrict__ sorted,
                                                                uint32_t n)
{
    __shared__ uint32_t s_lo[kSlowCap];
    uint32_t K = o.ctl[2];
    const uint32_t E = n - K;
    for (uint32_t q = threadIdx.x; q < K; q += kShiftThreads)
        s_lo[q] = o.ins_lo[q];
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)(kShiftThreads * kShiftItems) + threadIdx.x;
#pragma unroll
    for (int it = 0; it < kShiftItems; it++) {
        const uint32_t e = base + (uint32_t)it * kShiftThreads;
        if (e >= E)
            break;
        uint32_t lo = 0, hi = K;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_lo[mid] <= e)
                lo = mid + 1;
            else
                hi = mid;
        }
        o.perm[e + lo] = sorted[e];
    }
}

// The coming step's deferred ranks, flagged in the merged order.
__global__ __launch_bounds__(kPlanThreads) void k_overlap_flag(OverlapBufs o)
{
    const uint32_t j = threadIdx.x;
    if (j < o.ctl[1])
        o.perm[o.def_pos[j]] |= kDeferBit;
}

template <typename S>
__global__ __launch_bounds__(kPlanThreads) void k_overlap_fix_keys(OverlapBufs o, const S *__restrict__ r,
                                                                   typename KeyOf<S>::type *__restrict__ keys)
{
    const uint32_t j = threadIdx.x;
    if (j < o.ctl[1]) {
        const uint32_t t = o.def_pos[j];
        keys[t] = radius_key(r[t]);
    }
}

template <typename S>
hipError_t plan_typed(const OverlapBufs &o, const uint32_t *sorted, const S *r, int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_overlap_search<S>, dim3(2 * kSlowCap / (kSearchThreads / 64)), dim3(kSearchThreads), 0,
                       stream, o, sorted, r, (uint32_t)n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_overlap_plan<S>, dim3(1), dim3(kPlanThreads), 0, stream, o, r, (uint32_t)n);
    return hipGetLastError();
}

} // namespace

hipError_t overlap_reset(const OverlapBufs &o, const uint8_t *lvl, int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_overlap_reset, dim3(1), dim3(64), 0, stream, o.ctl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !lvl || n <= 0)
        return e;
    hipLaunchKernelGGL(k_overlap_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, o, lvl,
                       (uint32_t)n);
    return hipGetLastError();
}

hipError_t overlap_plan(const OverlapBufs &o, const uint32_t *sorted, const void *r, bool f32, int64_t n,
                        hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    if (n > (int64_t)kIndexMask)
        return hipErrorInvalidValue;
    return f32 ? plan_typed<float>(o, sorted, static_cast<const float *>(r), n, stream)
               : plan_typed<double>(o, sorted, static_cast<const double *>(r), n, stream);
}

hipError_t overlap_place(const OverlapBufs &o, const uint32_t *sorted, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    const int64_t per_block = (int64_t)kShiftThreads * kShiftItems;
    hipLaunchKernelGGL(k_overlap_shift, dim3((unsigned)((n + per_block - 1) / per_block)), dim3(kShiftThreads), 0,
                       stream, o, sorted, (uint32_t)n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_overlap_flag, dim3(1), dim3(kPlanThreads), 0, stream, o);
    return hipGetLastError();
}

hipError_t overlap_fix_keys(const OverlapBufs &o, const void *r, bool f32, void *keys, hipStream_t stream)
{
    if (f32)
        hipLaunchKernelGGL(k_overlap_fix_keys<float>, dim3(1), dim3(kPlanThreads), 0, stream, o,
                           static_cast<const float *>(r), static_cast<uint32_t *>(keys));
    else
        hipLaunchKernelGGL(k_overlap_fix_keys<double>, dim3(1), dim3(kPlanThreads), 0, stream, o,
                           static_cast<const double *>(r), static_cast<uint64_t *>(keys));
    return hipGetLastError();
}

} // namespace ns
