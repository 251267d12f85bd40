// ns_radix.h -- wave-level helpers of the radix and bucket sorts (gfx950,
// wave64), shared by ns_sort.hip and ns_qsort.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ns_device.h"

namespace ns {

// The deferral sentinel: the 32-bit sort key of a particle whose step is
// still running on the second stream (all ones, above every radius and
// quantile key).  Its tie keys are never read: the late kernel may still be
// writing those particles' radii.
constexpr uint32_t kSentinelWord = 0xFFFFFFFFu;

__device__ inline uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes (within `valid`) whose 8-bit digit equals this lane's: one ballot per
// digit bit, each folded in as peers &= (ballot ^ ~bit-mask) on both 32-bit
// halves -- a single v_bitop3 per half on gfx950.
__device__ inline uint64_t match_digit(uint32_t d, uint64_t valid)
{
    uint32_t lo = (uint32_t)valid, hi = (uint32_t)(valid >> 32);
    const uint32_t nd = ~d;
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
        const uint32_t nm = (uint32_t)__builtin_amdgcn_sbfe((int)nd, b, 1);  // ~0 where bit b of d is 0
        const uint64_t bb = __ballot(nm == 0);                                 // lanes whose bit b is 1
        lo &= (uint32_t)bb ^ nm;
        hi &= (uint32_t)(bb >> 32) ^ nm;
    }
    return ((uint64_t)hi << 32) | lo;
}

// Barrier for LDS traffic only: the workgroup's LDS writes before it are seen
// by all its waves after it, while their global loads stay in flight (a
// __syncthreads waits for those too).
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive scan of x over the first 256 threads (one per digit; the first
// four waves).  Every thread of the workgroup must call it (it synchronises);
// threads past 256 get an unspecified value.
__device__ inline uint32_t digit_exclusive_scan(uint32_t x, uint32_t *s_wave)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += y;
    }
    if (lane == 63 && wave < 4)
        s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < 4; w++)
        before += (w < wave) ? s_wave[w] : 0u;
    __syncthreads();
    return before + incl - x;
}

} // namespace ns
