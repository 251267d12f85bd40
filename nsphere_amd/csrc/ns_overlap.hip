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
// Which particles are deferred: those whose previous step was slow -- ended
// flagged NS_LVL_NOCONV (the chains that hit the cap, which persist from step
// to step) or reached level kSlowLevel (a particle plunging towards the
// centre climbs through the deep leapfrog levels before its chain is
// capped) -- found through the slow list the integrate kernels append to.
// The choice affects only the time, never the results.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

namespace {

constexpr int kPlanThreads = 1024;  // == kSlowCap: one thread per entry
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
    if (t < n && slow_record(lvl[t])) {
        const uint32_t k = atomicAdd(o.ctl, 1u);
        if (k < (uint32_t)kSlowCap)
            o.slow[k] = t;
    }
}

template <typename S>
__device__ inline typename KeyOf<S>::type key_at(const S *__restrict__ r, int64_t p)
{
    return radius_key(r[p]);
}

// Early elements of the sort (positions [0, E)) strictly below (k, t) in
// (key, old rank) order: the sort output is ascending in that order.
template <typename S>
__device__ inline uint32_t count_below(const uint32_t *__restrict__ sorted, const S *__restrict__ r, uint32_t E,
                                       typename KeyOf<S>::type k, uint32_t t)
{
    uint32_t lo = 0, hi = E;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        const uint32_t p = sorted[mid];
        const typename KeyOf<S>::type km = key_at(r, p);
        if (km < k || (km == k && p < t))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// One block.  ctl[1] late particles (old ranks def_pos[j]) were sorted last;
// the previous step's slow list (ctl[0] entries) names the ranks to defer in
// the coming step.  Writes the late particles' merged positions into perm,
// their insertion points (ascending) into ins_lo, the coming step's deferred
// ranks into def_pos / ctl[1] and the late count into ctl[2]; clears ctl[0].
template <typename S>
__global__ __launch_bounds__(kPlanThreads) void k_overlap_plan(OverlapBufs o, const uint32_t *__restrict__ sorted,
                                                               const S *__restrict__ r, uint32_t n)
{
    using KeyT = typename KeyOf<S>::type;
    __shared__ KeyT s_key[kSlowCap];
    __shared__ uint32_t s_old[kSlowCap];
    __shared__ uint32_t s_pos[kSlowCap];   // merged position of late slot j
    __shared__ uint32_t s_lo[kSlowCap];    // insertion points, ascending
    const uint32_t tid = threadIdx.x;
    uint32_t K = o.ctl[1];
    if (K > (uint32_t)kSlowCap)
        K = kSlowCap;
    if (K > n)
        K = n;
    const uint32_t E = n - K;
    uint32_t m = o.ctl[0];
    if (m > (uint32_t)kSlowCap)
        m = kSlowCap;
    if (tid < K) {
        const uint32_t t = o.def_pos[tid];
        s_old[tid] = t;
        s_key[tid] = key_at(r, t);
    }
    __syncthreads();
    if (tid < K) {
        const KeyT k = s_key[tid];
        const uint32_t t = s_old[tid];
        const uint32_t lo = count_below(sorted, r, E, k, t);
        uint32_t rank = 0;  // late particles below this one
        for (uint32_t q = 0; q < K; q++)
            rank += (s_key[q] < k || (s_key[q] == k && s_old[q] < t)) ? 1u : 0u;
        s_pos[tid] = lo + rank;
        s_lo[rank] = lo;
        o.ins_lo[rank] = lo;
        o.perm[lo + rank] = t;
    }
    __syncthreads();
    if (tid < m) {
        const uint32_t v = o.slow[tid];
        uint32_t pos;
        if (v & kDeferBit) {
            pos = s_pos[v & kIndexMask];  // a late particle of this merge
        } else {
            // an early element: its sorted position, moved past the late
            // particles inserted at or before it
            const uint32_t e = count_below(sorted, r, E, key_at(r, v), v);
            uint32_t lo = 0, hi = K;  // late particles with insertion point <= e
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_lo[mid] <= e)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            pos = e + lo;
        }
        o.def_pos[tid] = pos;
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
hipError_t merge_typed(const OverlapBufs &o, const uint32_t *sorted, const S *r, int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_overlap_plan<S>, dim3(1), dim3(kPlanThreads), 0, stream, o, sorted, r, (uint32_t)n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const int64_t per_block = (int64_t)kShiftThreads * kShiftItems;
    hipLaunchKernelGGL(k_overlap_shift, dim3((unsigned)((n + per_block - 1) / per_block)), dim3(kShiftThreads), 0,
                       stream, o, sorted, (uint32_t)n);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_overlap_flag, dim3(1), dim3(kPlanThreads), 0, stream, o);
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

hipError_t overlap_merge(const OverlapBufs &o, const uint32_t *sorted, const void *r, bool f32, int64_t n,
                         hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    if (n > (int64_t)kIndexMask)
        return hipErrorInvalidValue;
    return f32 ? merge_typed<float>(o, sorted, static_cast<const float *>(r), n, stream)
               : merge_typed<double>(o, sorted, static_cast<const double *>(r), n, stream);
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
