// ns_qsort.hip -- the sort between steps as a bucket sort on quantile keys.
//
// Replaces the reference's radial sort between steps (sort_particles_with_alg,
// nsphere.c:8605-8652, quadsort's stable order, nsphere.c:8356-8359) for
// float64 radii that the integrate has keyed with 24-bit quantile keys: the
// bucket of the new radius among kQBuckets quantiles of the step's start
// radii, then kQSubBits of position inside the bucket (ns_device.h).  The
// key is a non-decreasing function of the radius key, and between steps a
// particle moves a few buckets at most, so a workgroup's 3072 consecutive
// positions hold keys of a handful of buckets.  Two kernels:
//   1. k_qs_scatter: a workgroup ranks its keys per bucket in LDS (a window
//      of buckets around the tile's expected first bucket), reserves one
//      range per (tile, bucket) with one device atomic on the bucket's
//      counter, and writes (key, position) into the bucket's region of cap
//      entries (past cap: an overflow list).  The order inside a region is
//      the reservations' order; nothing after depends on it.
//   2. k_qs_bucket: one workgroup per bucket sums the counters of the
//      buckets below its own (its offset) and orders its elements by
//      (sub-bin, radius key, position): an LDS counting sort on the sub-bin,
//      then each run of equal sub-bins ranked by the radius keys of r and
//      the positions of its elements, and a coalesced write.  As the sub-bin
//      is a non-decreasing function of the radius key within a bucket, the
//      result is the stable sort of the radii (ties in input order), however
//      the scatter ordered a region.  A bucket larger than the workgroup
//      holds, or with a run of equal sub-bins longer than kQsMaxRun, is
//      sorted by its workgroup with 8-bit LSD passes on the position and then
//      the radius key, in global memory.  The last workgroup to finish clears
//      the counters for the next sort.
// Deferred particles carry the all-ones key and form bucket kQBuckets, last,
// ordered by position alone: their radii may still be being written.
//
// Bytes per key: scatter 4 read + 8 written, bucket 8 read + 4 written plus
// the radii of the elements that share a sub-bin (about half): about 28 B,
// against 72.9 B for three LSD passes and the fix-up of equal keys.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ns_device.h"
#include "ns_radix.h"

namespace ns {

namespace {

constexpr int kQsThreads = 256;
constexpr int kQsRounds = 3;                               // uint4 loads per thread
constexpr int kQsTile = kQsThreads * 4 * kQsRounds;        // 3072 positions per workgroup
constexpr int kQsWin = 512;                                // buckets counted in LDS per tile
constexpr int kQsWinBelow = 128;                           // ... below the tile's expected first bucket
constexpr uint32_t kQsSent = kQsWin;                       // window index of the sentinel bucket
constexpr uint32_t kQsMaxRun = 128;                        // longest run of equal sub-bins a workgroup ranks
constexpr int kQsSubs = 1 << kQSubBits;                    // sub-bins per bucket
constexpr int kQsBucketThreads = 1024;
constexpr int kQsCapSmall = kQsBucketThreads * 4;          // region capacities: the two bucket kernels
constexpr int kQsCapLarge = kQsBucketThreads * 16;

// ctl words
constexpr int kQcOvf = 0;     // overflow entries
constexpr int kQcTicket = 1;  // bucket workgroups done
constexpr int kQcWords = 4;
constexpr int kQCntWords = kQBuckets + 4;  // counters, padded to whole uint4

// Bucket of a key: kQBuckets for the sentinel; clamped below it for any other
// word (quantile keys stay at or below kQKeyMax), so that no key can address
// past the counters.
__device__ inline uint32_t qs_bucket(uint32_t q)
{
    if (q == kSentinelWord)
        return (uint32_t)kQBuckets;
    const uint32_t b = q >> kQSubBits;
    return b < (uint32_t)kQBuckets ? b : (uint32_t)kQBuckets - 1;
}

// Window index of key q for a tile whose window starts at bucket wb0:
// [0, kQsWin) a window bucket, kQsSent the sentinel, larger: outside.
__device__ inline uint32_t qs_win(uint32_t q, int wb0)
{
    const uint32_t b = qs_bucket(q);
    if (b == (uint32_t)kQBuckets)
        return kQsSent;
    const uint32_t d = b - (uint32_t)wb0;
    return d < (uint32_t)kQsWin ? d : kQsSent + 1;
}

// 1. Scatter into bucket regions.  Round rd of thread t holds positions
// p0 + rd * 1024 + 4t .. + 3, all four rounds loaded before any is ranked.
// A wave ranks its 256 keys of a round bucket by bucket: the first unranked
// key's window bucket, the wave's keys of that bucket (one ballot per key of
// a lane), one LDS atomic for all of them, and slots in (key of the lane,
// lane) order, so that each of the four stores of a round writes consecutive
// entries of a region.  A round takes as many of these as it holds window
// buckets (two or three between steps).
__global__ __launch_bounds__(kQsThreads) void k_qs_scatter(const uint32_t *__restrict__ keys, int64_t n,
                                                           uint32_t stride, uint32_t cap, uint2 *__restrict__ elem,
                                                           uint2 *__restrict__ ovf, uint32_t *__restrict__ cnt,
                                                           uint32_t *__restrict__ ctl)
{
    __shared__ uint32_t s_cnt[kQsWin + 1];  // count per window bucket (+ sentinel), then its reserved offset
    const int tid = threadIdx.x, lane = tid & 63;
    const uint64_t lt = lanemask_lt();
    const int64_t p0 = (int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * kQsTile;
    const int64_t eb = p0 / stride;
    int wb0 = eb < kQsWinBelow ? 0 : (int)(eb - kQsWinBelow);
    wb0 = wb0 < kQBuckets - kQsWin ? wb0 : kQBuckets - kQsWin;
    for (int i = tid; i <= kQsWin; i += kQsThreads)
        s_cnt[i] = 0;
    uint32_t key[kQsRounds][4], slot[kQsRounds][4];
#pragma unroll
    for (int rd = 0; rd < kQsRounds; rd++) {
        const int64_t pb = p0 + (int64_t)rd * (kQsThreads * 4) + 4 * tid;
        if (pb + 3 < n) {
            const uint4 v = *reinterpret_cast<const uint4 *>(keys + pb);
            key[rd][0] = v.x;
            key[rd][1] = v.y;
            key[rd][2] = v.z;
            key[rd][3] = v.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                key[rd][j] = pb + j < n ? keys[pb + j] : kSentinelWord;
        }
    }
    __syncthreads();
#pragma unroll
    for (int rd = 0; rd < kQsRounds; rd++) {
        const int64_t pb = p0 + (int64_t)rd * (kQsThreads * 4) + 4 * tid;
        uint64_t pend[4];
        uint32_t win[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            win[j] = pb + j < n ? qs_win(key[rd][j], wb0) : kQsSent + 1;
            slot[rd][j] = 0;
            pend[j] = __ballot(win[j] <= kQsSent);
        }
        while (pend[0] | pend[1] | pend[2] | pend[3]) {  // uniform
            const int j0 = pend[0] ? 0 : (pend[1] ? 1 : (pend[2] ? 2 : 3));
            const int l0 = __ffsll((unsigned long long)pend[j0]) - 1;
            const uint32_t wj = j0 == 0 ? win[0] : (j0 == 1 ? win[1] : (j0 == 2 ? win[2] : win[3]));
            const uint32_t u = (uint32_t)__builtin_amdgcn_readlane((int)wj, l0);
            uint64_t m[4];
            uint32_t before[4], total = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                m[j] = __ballot(win[j] == u) & pend[j];
                before[j] = total;
                total += (uint32_t)__popcll(m[j]);
                pend[j] &= ~m[j];
            }
            uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(&s_cnt[u], total);
            base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((m[j] >> lane) & 1)
                    slot[rd][j] = base + before[j] + (uint32_t)__popcll(m[j] & lt);
        }
    }
    __syncthreads();
    for (int i = tid; i <= kQsWin; i += kQsThreads) {
        const uint32_t c = s_cnt[i];
        if (c)
            s_cnt[i] = atomicAdd(&cnt[i == kQsWin ? kQBuckets : wb0 + i], c);
    }
    __syncthreads();
#pragma unroll
    for (int rd = 0; rd < kQsRounds; rd++) {
        const int64_t pb = p0 + (int64_t)rd * (kQsThreads * 4) + 4 * tid;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (pb + j >= n)
                continue;
            const uint32_t w = qs_win(key[rd][j], wb0);
            const uint32_t b = qs_bucket(key[rd][j]);
            // outside the window (a particle that moved far): a device atomic of its own
            const uint32_t off = w <= kQsSent ? s_cnt[w] + slot[rd][j] : atomicAdd(&cnt[b], 1u);
            const uint2 e = make_uint2(key[rd][j], (uint32_t)(pb + j));
            if (off < cap)
                elem[(size_t)b * cap + off] = e;
            else
                ovf[atomicAdd(&ctl[kQcOvf], 1u)] = e;
        }
    }
}

// Sum of x over the workgroup (all threads call it; every thread gets it).
template <int WAVES>
__device__ inline uint32_t block_sum(uint32_t x, uint32_t *s_red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        x += __shfl_xor(x, off, 64);
    if (lane == 0)
        s_red[wave] = x;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++)
        t += s_red[w];
    __syncthreads();
    return t;
}

// The bucket's positions, gathered from its region and the overflow list
// into A[0, M) in any order, sorted stably by 8-bit LSD passes between A and
// B (global memory, this workgroup's slices): first on the position, then,
// unless the radius keys are then already non-decreasing, on the radius key
// (deferred particles: the position alone).  A pass whose digit is the same
// for every element is skipped.  The result ends in A.  LDS: lds, at least
// (2 + THREADS / 64) * kRadix + 8 words.
template <int THREADS>
__device__ void qs_fallback(uint32_t b, uint32_t M, const uint2 *__restrict__ elem, uint32_t cap,
                            const uint2 *__restrict__ ovf, uint32_t K, const double *__restrict__ r, int64_t n,
                            uint32_t *__restrict__ A, uint32_t *__restrict__ B, uint32_t *lds)
{
    constexpr int WAVES = THREADS / 64;
    uint32_t *s_tot = lds, *s_off = lds + kRadix, *s_wcnt = lds + 2 * kRadix;  // s_wcnt[w * kRadix + d]
    uint32_t *s_wave = lds + (2 + WAVES) * kRadix, *s_fill = s_wave + 4, *s_flag = s_wave + 5;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lt = lanemask_lt();
    const bool sentinel = b == (uint32_t)kQBuckets;
    int pos_passes = 1;
    while (pos_passes < 4 && ((uint64_t)(n - 1) >> (8 * pos_passes)) != 0)
        pos_passes++;
    const uint32_t mr = M < cap ? M : cap;
    for (uint32_t e = tid; e < mr; e += THREADS)
        A[e] = elem[(size_t)b * cap + e].y;
    if (tid == 0)
        *s_fill = mr;
    __syncthreads();
    if (M > mr)
        for (uint32_t k = tid; k < K; k += THREADS) {
            const uint2 v = ovf[k];
            if (qs_bucket(v.x) == b)
                A[atomicAdd(s_fill, 1u)] = v.y;
        }
    uint32_t *src = A, *dst = B;
    const int passes = pos_passes + (sentinel ? 0 : 8);
    for (int p = 0; p < passes; p++) {
        __syncthreads();  // every read of the previous pass's flag, offsets and output is done
        const bool by_pos = p < pos_passes;
        const int shift = 8 * (by_pos ? p : p - pos_passes);
        if (p == pos_passes) {  // radius keys already non-decreasing: done
            if (tid == 0)
                *s_flag = 0;
            __syncthreads();
            bool desc = false;
            for (uint32_t e = tid + 1; e < M; e += THREADS)
                desc |= radius_key(r[src[e]]) < radius_key(r[src[e - 1]]);
            if (desc)
                *s_flag = 1;
            __syncthreads();
            if (!*s_flag)
                break;
            __syncthreads();
        }
        for (int d = tid; d < kRadix; d += THREADS)
            s_tot[d] = 0;
        if (tid == 0)
            *s_flag = 0;
        __syncthreads();
        for (uint32_t e = tid; e < M; e += THREADS) {
            const uint32_t x = src[e];
            const uint32_t dg = (uint32_t)((by_pos ? (uint64_t)x : radius_key(r[x])) >> shift) & (kRadix - 1);
            atomicAdd(&s_tot[dg], 1u);
        }
        __syncthreads();
        if (tid < kRadix && s_tot[tid] == M)
            *s_flag = 1;  // one digit for all: the pass would not move anything
        const uint32_t exd = digit_exclusive_scan(tid < kRadix ? s_tot[tid] : 0u, s_wave);
        if (tid < kRadix)
            s_off[tid] = exd;
        if (*s_flag)
            continue;  // uniform: written before the scan's barriers
        for (int k = tid; k < WAVES * kRadix; k += THREADS)
            s_wcnt[k] = 0;
        __syncthreads();
        // chunks of THREADS in input order: stable ranks from wave ballots
        // and per-wave digit counts
        for (uint32_t c0 = 0; c0 < M; c0 += THREADS) {
            const uint32_t e = c0 + tid;
            const bool valid = e < M;
            const uint32_t x = valid ? src[e] : 0u;
            const uint32_t dg =
                valid ? (uint32_t)((by_pos ? (uint64_t)x : radius_key(r[x])) >> shift) & (kRadix - 1) : 0u;
            const uint64_t peers = match_digit(dg, __ballot(valid));
            const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : lane;
            if (valid && lane == leader)
                s_wcnt[wave * kRadix + dg] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (tid < kRadix) {
                uint32_t run = 0;
                for (int w = 0; w < WAVES; w++) {
                    const uint32_t cw = s_wcnt[w * kRadix + tid];
                    s_wcnt[w * kRadix + tid] = run;
                    run += cw;
                }
                s_tot[tid] = run;
            }
            __syncthreads();
            if (valid)
                dst[s_off[dg] + s_wcnt[wave * kRadix + dg] + (uint32_t)__popcll(peers & lt)] = x;
            __syncthreads();
            if (tid < kRadix) {
                s_off[tid] += s_tot[tid];
                for (int w = 0; w < WAVES; w++)
                    s_wcnt[w * kRadix + tid] = 0;
            }
            __syncthreads();
        }
        uint32_t *t = src;
        src = dst;
        dst = t;
    }
    __syncthreads();
    if (src != A)
        for (uint32_t e = tid; e < M; e += THREADS)
            A[e] = src[e];
}

// 2. One workgroup per bucket (the sentinel bucket kQBuckets among them).
// Everything it reads first comes in one round trip: the counters (its
// offset is the sum of those below it, its count its own) and its whole
// region (entries past the count are ignored).  It then counts itself done
// (all its counter reads have arrived), so the last workgroup to do so may
// clear the counters for the next sort while the others work on.  Its
// elements k * THREADS + tid: LDS counts per sub-bin (a slot each from the
// atomic) and, at the same time, their radius keys gathered from r; the
// counts' exclusive scan; the positions in sub-bin order; then each element's
// rank in its run of equal sub-bins by (radius key, position).  STAGE: the
// elements of runs of two or more put their radius keys in LDS beside their
// positions and rank against those; otherwise (the larger regions, whose keys
// would not fit the LDS) each comparison gathers the other element's radius.
// Buckets past the capacity or the run limit take qs_fallback.
template <int ITEMS, bool STAGE>
__global__ __launch_bounds__(kQsBucketThreads) void k_qs_bucket(const uint2 *__restrict__ elem, uint32_t cap,
                                                               const uint2 *__restrict__ ovf,
                                                               uint32_t *__restrict__ cnt,
                                                               const double *__restrict__ r, int64_t n,
                                                               uint32_t *__restrict__ out,
                                                               uint32_t *__restrict__ scratch,
                                                               uint32_t *__restrict__ ctl)
{
    constexpr int THREADS = kQsBucketThreads;
    constexpr int CAP = THREADS * ITEMS;
    constexpr int PER = kQsSubs / THREADS;  // sub-bins per thread in the scan
    constexpr int WAVES = THREADS / 64;
    static_assert(kQsSubs % THREADS == 0 && kQBuckets == 4 * THREADS, "counter and sub-bin layout");
    constexpr int kSubWords = kQsSubs, kPosWords = CAP, kRkWords = STAGE ? 2 * CAP : 2;
    constexpr int kFbWords = (2 + WAVES) * kRadix + 8;
    constexpr int kWords = kSubWords + kPosWords + kRkWords > kFbWords ? kSubWords + kPosWords + kRkWords : kFbWords;
    __shared__ __align__(16) uint32_t s_lds[kWords];
    __shared__ uint32_t s_w[WAVES], s_m[WAVES];
    __shared__ uint32_t s_mk[2];  // the bucket's count, the overflow count
    __shared__ uint32_t s_last;
    uint32_t *s_sub = s_lds;               // counts, then first slot of each sub-bin
    uint32_t *s_pos = s_lds + kSubWords;   // positions in sub-bin order, then in final order
    uint64_t *s_rk = reinterpret_cast<uint64_t *>(s_lds + kSubWords + kPosWords);  // radius keys in sub-bin order
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t b = xcd_contiguous_block(blockIdx.x, gridDim.x);
    const bool sentinel = b == (uint32_t)kQBuckets;
    const uint4 c4 = reinterpret_cast<const uint4 *>(cnt)[tid];  // buckets 4 tid .. 4 tid + 3
    if (tid == 0) {
        s_mk[0] = cnt[b];
        s_mk[1] = ctl[kQcOvf];
    }
    const uint2 *src = elem + (size_t)b * cap;
    uint2 v[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
        v[k] = src[k * THREADS + tid];  // inside the region (cap == CAP)
    for (int i = tid; i < kQsSubs; i += THREADS)
        s_sub[i] = 0;
    const uint32_t t4 = 4u * (uint32_t)tid;
    const uint32_t below = (t4 < b ? c4.x : 0u) + (t4 + 1 < b ? c4.y : 0u) + (t4 + 2 < b ? c4.z : 0u) +
                           (t4 + 3 < b ? c4.w : 0u);
    const uint32_t lo = block_sum<WAVES>(below, s_w);  // its barriers also publish s_mk and the cleared s_sub
    const uint32_t M = s_mk[0], K = s_mk[1];
    if (tid == 0)
        s_last = atomicAdd(&ctl[kQcTicket], 1u) == gridDim.x - 1 ? 1u : 0u;
    bool fallback = M > (uint32_t)CAP;
    if (M > 0 && !fallback) {
        uint32_t pos[ITEMS], sub[ITEMS], slot[ITEMS];
        double rr[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t e = (uint32_t)(k * THREADS + tid);
            pos[k] = sub[k] = slot[k] = 0;
            rr[k] = 0.0;
            if (e < M) {
                pos[k] = v[k].y;
                // deferred particles: a sub-bin that grows with the position
                sub[k] = sentinel ? (uint32_t)(((uint64_t)pos[k] << kQSubBits) / (uint64_t)n)
                                  : (v[k].x & (kQsSubs - 1));
                if (!sentinel)
                    rr[k] = r[pos[k]];  // the radii arrive while the counts are taken and scanned
            }
        }
#pragma unroll
        for (int k = 0; k < ITEMS; k++)
            if ((uint32_t)(k * THREADS + tid) < M)
                slot[k] = atomicAdd(&s_sub[sub[k]], 1u);
        lds_barrier();
        uint32_t c[PER], sum = 0, mx = 0;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            c[j] = s_sub[tid * PER + j];
            sum += c[j];
            mx = c[j] > mx ? c[j] : mx;
        }
        uint32_t incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            const uint32_t z = __shfl_xor(mx, off, 64);
            if (lane >= off)
                incl += y;
            mx = z > mx ? z : mx;
        }
        if (lane == 63)
            s_w[wave] = incl;
        if (lane == 0)
            s_m[wave] = mx;
        lds_barrier();
        uint32_t ex = incl - sum, maxrun = 0;
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            ex += w < wave ? s_w[w] : 0u;
            maxrun = s_m[w] > maxrun ? s_m[w] : maxrun;
        }
#pragma unroll
        for (int j = 0; j < PER; j++) {
            s_sub[tid * PER + j] = ex;
            ex += c[j];
        }
        fallback = !sentinel && maxrun > kQsMaxRun;  // uniform: every thread read the same s_m
        lds_barrier();
        if (!fallback) {
            // run of each element: [st, en) in sub-bin order
            uint32_t st[ITEMS], en[ITEMS];
            uint64_t rk[ITEMS];
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                st[k] = en[k] = 0;
                rk[k] = sentinel ? 0ull : radius_key(rr[k]);
                if ((uint32_t)(k * THREADS + tid) < M) {
                    st[k] = s_sub[sub[k]];
                    en[k] = sub[k] + 1 < (uint32_t)kQsSubs ? s_sub[sub[k] + 1] : M;
                    s_pos[st[k] + slot[k]] = pos[k];
                    if (STAGE && en[k] - st[k] > 1)
                        s_rk[st[k] + slot[k]] = rk[k];
                }
            }
            lds_barrier();
            uint32_t fin[ITEMS];
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                fin[k] = st[k];
                if (en[k] - st[k] < 2)
                    continue;
                uint32_t rank = 0;
                for (uint32_t m = st[k]; m < en[k]; m++) {
                    const uint32_t pm = s_pos[m];
                    uint64_t km = 0;
                    if constexpr (STAGE)
                        km = s_rk[m];
                    else
                        km = sentinel ? 0ull : radius_key(r[pm]);
                    rank += (km < rk[k] || (km == rk[k] && pm < pos[k])) ? 1u : 0u;
                }
                fin[k] = st[k] + rank;
            }
            lds_barrier();
#pragma unroll
            for (int k = 0; k < ITEMS; k++)
                if ((uint32_t)(k * THREADS + tid) < M)
                    s_pos[fin[k]] = pos[k];
            lds_barrier();
            for (uint32_t e = tid; e < M; e += THREADS)
                out[lo + e] = s_pos[e];
        }
    }
    if (fallback)  // uniform
        qs_fallback<THREADS>(b, M, elem, cap, ovf, K, r, n, out + lo, scratch + lo, s_lds);
    __syncthreads();  // s_last
    if (s_last) {  // every other workgroup has read the counters
        for (int i = tid; i < kQCntWords; i += THREADS)
            cnt[i] = 0;
        if (tid == 0) {
            ctl[kQcOvf] = 0;
            ctl[kQcTicket] = 0;
        }
    }
}

// Region capacity for n keys: the bucket kernel's, when a bucket (s =
// quantile_stride(n) keys at the start of the step, the last one n -
// (kQBuckets - 1) s) fits it with room to grow; 0 when none does.
int qsort_cap(int64_t n)
{
    if (n < kQMinN || n > (int64_t)0xFFFFFFFFu)
        return 0;
    const int64_t s = quantile_stride(n);
    const int64_t need = std::max(s + s / 4, n - (int64_t)(kQBuckets - 1) * s + 256);
    if (need <= kQsCapSmall)
        return kQsCapSmall;
    if (need <= kQsCapLarge)
        return kQsCapLarge;
    return 0;
}

template <typename T>
hipError_t qs_alloc(T **p, size_t count)
{
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

} // namespace

void qsort_release(QSortBufs &b)
{
    void *bufs[] = {b.elem, b.ovf, b.cnt, b.ctl};
    for (void *p : bufs)
        if (p)
            (void)hipFree(p);
    b = QSortBufs();
}

hipError_t qsort_prepare(QSortBufs &b, int64_t n)
{
    if (b.n == n)
        return hipSuccess;  // prepared (cap 0: no bucket sort for this n)
    qsort_release(b);
    const int cap = qsort_cap(n);
    if (cap == 0) {
        b.n = n;
        return hipSuccess;
    }
    hipError_t e = qs_alloc(&b.elem, (size_t)(kQBuckets + 1) * cap);
    if (e == hipSuccess) e = qs_alloc(&b.ovf, (size_t)n);
    if (e == hipSuccess) e = qs_alloc(&b.cnt, (size_t)kQCntWords);
    if (e == hipSuccess) e = qs_alloc(&b.ctl, (size_t)kQcWords);
    if (e == hipSuccess) e = hipMemset(b.cnt, 0, (size_t)kQCntWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b.ctl, 0, (size_t)kQcWords * sizeof(uint32_t));
    if (e != hipSuccess) {
        qsort_release(b);
        return e;
    }
    b.cap = cap;
    b.n = n;
    return hipSuccess;
}

hipError_t qsort_perm(QSortBufs &b, const uint32_t *keys, int64_t n, const double *r, uint32_t *perm_out,
                      uint32_t *scratch, hipStream_t stream, Profiler *prof)
{
    if (b.cap == 0 || b.n != n || n < kQMinN)
        return hipErrorInvalidValue;
    if (prof)
        prof->begin(K_QS_SCATTER);
    hipLaunchKernelGGL(k_qs_scatter, dim3((unsigned)((n + kQsTile - 1) / kQsTile)), dim3(kQsThreads), 0, stream,
                       keys, n, quantile_stride(n), (uint32_t)b.cap, b.elem, b.ovf, b.cnt, b.ctl);
    if (prof)
        prof->end(K_QS_SCATTER);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    if (prof)
        prof->begin(K_QS_BUCKET);
    if (b.cap == kQsCapSmall)
        hipLaunchKernelGGL((k_qs_bucket<kQsCapSmall / kQsBucketThreads, true>), dim3(kQBuckets + 1),
                           dim3(kQsBucketThreads), 0, stream, b.elem, (uint32_t)b.cap, b.ovf, b.cnt, r, n, perm_out,
                           scratch, b.ctl);
    else
        hipLaunchKernelGGL((k_qs_bucket<kQsCapLarge / kQsBucketThreads, false>), dim3(kQBuckets + 1),
                           dim3(kQsBucketThreads), 0, stream, b.elem, (uint32_t)b.cap, b.ovf, b.cnt, r, n, perm_out,
                           scratch, b.ctl);
    e = hipGetLastError();
    if (prof)
        prof->end(K_QS_BUCKET);
    return e;
}

} // namespace ns
