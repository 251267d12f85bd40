// ns_sort.hip -- stable LSD radix sort of radius keys for MI355X (gfx950).
//
// Replaces the reference's radial sort (sort_particles_with_alg,
// nsphere.c:8605-8652, with the stable quadsort of nsphere.c:8356-8359 and the
// comparator of nsphere.c:8297-8306).  An LSD radix sort is stable, so keys
// that tie keep their input (previous-step) order -- the order quadsort gives.
//
// Design (one HIP stream, wave64), for 64-bit keys (float64 radii, 8 passes)
// or 32-bit keys (float32 radii, 4 passes).  Every pass scatters tiles of
// THREADS * ITEMS keys: a workgroup ranks its keys stably per digit with wave
// ballots (a 64-lane match of the 8 digit bits), finds the tile's global
// offset of every digit, then stages keys and payload -- one after the
// other, through one LDS buffer -- in digit order and writes them out so that
// each digit run leaves as one coalesced segment.  Pass 0 synthesises the
// payload (input positions) instead of reading it; the last pass writes only
// the payload (the permutation).  The tile offsets come from one of:
//   * reduce-then-scan (default): per pass, a count kernel writes every
//     tile's digit counts, one workgroup per digit scans them over the
//     tiles, and the scatter kernel reads its offsets -- one extra read of
//     the keys per pass, no workgroup ever waits for another;
//   * onesweep: one upfront kernel builds all the digit histograms, and
//     per pass a single scatter kernel takes tiles by atomic ticket,
//     publishes its digit counts and resolves its offsets by decoupled
//     lookback over earlier tiles.  At N ~ 1e7 about 1300 tiles are in
//     flight at once and the lookback walks dominate (measured ~95 us per
//     32-bit pass against ~60 us for the scatter of reduce-then-scan).
// Algorithm, block size and keys per thread are a SortConfig.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_radix.h"

namespace ns {

namespace {

// Lookback word: high 32 bits = epoch << 2 | status, low 32 bits = count.
constexpr uint32_t kStatusAggregate = 1;
constexpr uint32_t kStatusPrefix = 2;

__device__ inline uint64_t lb_pack(uint32_t epoch, uint32_t status, uint32_t count)
{
    return ((uint64_t)((epoch << 2) | status) << 32) | count;
}

// The lookback word carries its whole message (status and count), so it needs
// no ordering against other memory: relaxed agent-scope atomics keep it
// coherent across the eight XCD L2s (sc1 accesses) without the L2 writeback
// and invalidate that release/acquire would issue on every access.
__device__ inline void lb_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t lb_load(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename S>
__global__ void k_radius_keys(const S *__restrict__ r, void *__restrict__ keys, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        store_key(keys, n, i, radius_key(r[i]));
}

// Id tiebreak: the ids as 32-bit keys, then the radius keys gathered through
// the id order (with that order as the first pass's payload).
__global__ void k_id_keys(const uint32_t *__restrict__ id, uint32_t *__restrict__ keys, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        keys[i] = id[i];
}

template <typename S>
__global__ void k_radius_keys_gather(const S *__restrict__ r, const uint32_t *__restrict__ order,
                                     typename KeyOf<S>::type *__restrict__ keys, uint32_t *__restrict__ vals,
                                     int64_t n)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        const uint32_t p = order[j];
        keys[j] = radius_key(r[p]);
        vals[j] = p;
    }
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, int64_t n, uint64_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void k_sort_zero(uint32_t *__restrict__ hist, int nhist, uint32_t *__restrict__ counters,
                            int ncounters, uint32_t *__restrict__ zero_word)
{
    for (int i = threadIdx.x; i < nhist; i += blockDim.x)
        hist[i] = 0;
    for (int i = threadIdx.x; i < ncounters; i += blockDim.x)
        counters[i] = 0;
    if (zero_word && threadIdx.x == 0)
        *zero_word = 0;
}

// All digit histograms in one read of the keys.  Each thread takes
// kHistItems consecutive keys and adds runs of equal digits with one LDS
// atomic, which keeps the nearly-constant high digits of nearly-sorted radii
// from serialising on one bank.
template <typename KeyT>
__global__ __launch_bounds__(kHistThreads) void k_sort_histogram(const KeyT *__restrict__ keys, int64_t n,
                                                                  uint32_t *__restrict__ hist)
{
    constexpr int kPasses = (int)sizeof(KeyT);
    __shared__ uint32_t s_hist[kPasses * kRadix];
    for (int i = threadIdx.x; i < kPasses * kRadix; i += kHistThreads)
        s_hist[i] = 0;
    __syncthreads();

    const int64_t per_block = (int64_t)kHistThreads * kHistItems;
    for (int64_t base = (int64_t)blockIdx.x * per_block; base < n; base += (int64_t)gridDim.x * per_block) {
        const int64_t first = base + (int64_t)threadIdx.x * kHistItems;
        KeyT k[kHistItems];
        int cnt = 0;
        if (first + kHistItems <= n) {
            constexpr int kPerVec = 16 / (int)sizeof(KeyT);
            const uint4 *src = reinterpret_cast<const uint4 *>(keys + first);
#pragma unroll
            for (int j = 0; j < kHistItems / kPerVec; j++) {
                const uint4 t = src[j];
                const KeyT *tk = reinterpret_cast<const KeyT *>(&t);
#pragma unroll
                for (int e = 0; e < kPerVec; e++)
                    k[j * kPerVec + e] = tk[e];
            }
            cnt = kHistItems;
        } else {
#pragma unroll
            for (int j = 0; j < kHistItems; j++)
                k[j] = (first + j < n) ? keys[first + j] : (KeyT)0;
            cnt = first < n ? (int)(n - first) : 0;
        }
        if (cnt > 0) {
#pragma unroll
            for (int p = 0; p < kPasses; p++) {
                uint32_t cur = (uint32_t)(k[0] >> (kRadixBits * p)) & (kRadix - 1);
                uint32_t run = 1;
#pragma unroll
                for (int j = 1; j < kHistItems; j++) {
                    if (j < cnt) {
                        const uint32_t d = (uint32_t)(k[j] >> (kRadixBits * p)) & (kRadix - 1);
                        if (d == cur) {
                            run++;
                        } else {
                            atomicAdd(&s_hist[p * kRadix + cur], run);
                            cur = d;
                            run = 1;
                        }
                    }
                }
                atomicAdd(&s_hist[p * kRadix + cur], run);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kPasses * kRadix; i += kHistThreads) {
        const uint32_t c = s_hist[i];
        if (c)
            atomicAdd(&hist[i], c);
    }
}

// Reduce-then-scan, step 1: the pass digit's count in every tile, written
// tile-major (counts[tile * kRadix + d], one 1 KB row per tile).  A wave
// whose 64 keys share one digit (common for the near-constant high bytes of
// nearly sorted radii) adds them with one LDS atomic instead of 64 colliding
// ones; other waves add per lane.
template <typename KeyT, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_rts_upsweep(const KeyT *__restrict__ keys, int64_t n, int shift,
                                                         uint32_t *__restrict__ counts, int64_t tiles,
                                                         const int64_t *__restrict__ tile_start,
                                                         uint32_t *__restrict__ zero_word)
{
    constexpr int WAVES = THREADS / 64;
    __shared__ uint32_t s_cnt[WAVES][kRadix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (zero_word && blockIdx.x == 0 && tid == 0)
        *zero_word = 0;  // a counter of a later kernel of the same sort
    for (int i = tid; i < WAVES * kRadix; i += THREADS)
        (&s_cnt[0][0])[i] = 0;
    __syncthreads();

    const int64_t tile = blockIdx.x;
    int64_t start = tile * (THREADS * ITEMS), end = start + THREADS * ITEMS;
    if (tile_start) {  // segmented: tiles never straddle a segment boundary
        start = tile_start[tile];
        end = tile_start[tile + 1];
    }
    if (end > n)
        end = n;
    // Counting ignores order, so a full unsegmented tile is read 16 bytes per
    // lane (the tile start is a multiple of the tile, hence aligned); a
    // segment's or the last, partial tile key by key.
    constexpr int VEC = 16 / (int)sizeof(KeyT);
    constexpr bool kVector = ITEMS % VEC == 0;
    KeyT key[ITEMS];
    bool full = false;
    if constexpr (kVector)
        full = !tile_start && end - start == (int64_t)THREADS * ITEMS;
    if (full) {
        typedef KeyT vec_t __attribute__((ext_vector_type(VEC)));
        const vec_t *src = reinterpret_cast<const vec_t *>(keys + start);
#pragma unroll
        for (int j = 0; j < ITEMS / VEC; j++) {
            const vec_t x = src[j * THREADS + tid];  // kept in cache: the scatter re-reads the keys
#pragma unroll
            for (int k = 0; k < VEC; k++)
                key[j * VEC + k] = x[k];
        }
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            if (__ballot(d != d0) == 0) {
                if (lane == 0)
                    atomicAdd(&s_cnt[wave][d0], 64u);
            } else {
                atomicAdd(&s_cnt[wave][d], 1u);
            }
        }
    } else {
        const int64_t wave_base = start + (int64_t)wave * (ITEMS * 64);
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const int64_t idx = wave_base + (int64_t)j * 64 + lane;
            key[j] = idx < end ? keys[idx] : (KeyT)0;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            // valid lanes are a prefix of the wave (lane 0 first)
            const bool valid = wave_base + (int64_t)j * 64 + lane < end;
            const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));  // whole wave: no divergence yet
            if (__ballot(valid && d != d0) == 0) {
                if (lane == 0 && valid)
                    atomicAdd(&s_cnt[wave][d0], nvalid);
            } else if (valid) {
                atomicAdd(&s_cnt[wave][d], 1u);
            }
        }
    }
    __syncthreads();
    if (tid < kRadix) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < WAVES; w++)
            c += s_cnt[w][tid];
        counts[tile * kRadix + tid] = c;
    }
}

// Reduce-then-scan, step 2: one workgroup per digit turns its column of tile
// counts into exclusive tile offsets (in place) and writes the digit total.
// Workgroups are dealt to the eight XCDs round-robin; workgroup b takes digit
// (b % 8) * 32 + b / 8, so that the 16 digits of a 64-byte line of a tile row
// are all read and written through one XCD's L2.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;

__global__ __launch_bounds__(kScanThreads) void k_rts_scan(uint32_t *__restrict__ counts, int64_t tiles,
                                                           uint32_t *__restrict__ totals)
{
    constexpr int WAVES = kScanThreads / 64;
    __shared__ uint32_t s_wave[WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int d = (blockIdx.x & 7) * (kRadix / 8) + (blockIdx.x >> 3);
    uint32_t *col = counts + d;  // entry t at col[t * kRadix]
    uint32_t carry = 0;
    for (int64_t base = 0; base < tiles; base += kScanThreads * kScanItems) {
        const int64_t first = base + (int64_t)tid * kScanItems;
        uint32_t x[kScanItems];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            x[k] = first + k < tiles ? col[(first + k) * kRadix] : 0u;
            sum += x[k];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            if (lane >= off)
                incl += y;
        }
        if (lane == 63)
            s_wave[wave] = incl;
        __syncthreads();
        uint32_t before = carry, chunk = 0;
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            const uint32_t t = s_wave[w];
            before += w < wave ? t : 0u;
            chunk += t;
        }
        __syncthreads();
        uint32_t run = before + incl - sum;
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            if (first + k < tiles)
                col[(first + k) * kRadix] = run;
            run += x[k];
        }
        carry += chunk;
    }
    if (tid == 0)
        totals[d] = carry;
}

// Segmented sort, after k_rts_scan: for segment s (one workgroup, a thread
// per digit) the digit counts of its tiles are the scanned column values from
// its first tile to the next segment's first; base[s][d] = segment start +
// the segment's exclusive digit prefix - the scanned value at its first tile,
// so that base[s][d] + scanned[t][d] is the global offset of tile t's digit d.
__global__ __launch_bounds__(kRadix) void k_seg_bases(const uint32_t *__restrict__ scanned,
                                                       const uint32_t *__restrict__ totals,
                                                       const uint32_t *__restrict__ seg_first,
                                                       const uint32_t *__restrict__ seg_start, int nseg,
                                                       uint32_t *__restrict__ seg_base)
{
    __shared__ uint32_t s_wave[4];
    const int s = blockIdx.x, d = threadIdx.x;
    const uint32_t t0 = seg_first[s], t1 = seg_first[s + 1];
    const uint32_t first = scanned[(int64_t)t0 * kRadix + d];
    const uint32_t next = (s + 1 < nseg) ? scanned[(int64_t)t1 * kRadix + d] : totals[d];
    const uint32_t ex = digit_exclusive_scan(next - first, s_wave);
    seg_base[(int64_t)s * kRadix + d] = seg_start[s] + ex - first;
}

// One scatter pass over digit `shift`.  FIRST: payload = input position.
// LAST: keys are not written.  LOOKBACK (onesweep): tiles are taken by
// ticket and each tile's global digit offsets come from decoupled lookback
// over earlier tiles, with hist_pass the pass's digit histogram.  Otherwise
// (reduce-then-scan) tile = blockIdx.x, the offsets are read from
// tile_offsets[tile * kRadix + d] and hist_pass holds the digit totals.
template <typename KeyT, int THREADS, int ITEMS, bool FIRST, bool LAST, bool LOOKBACK>
__global__ __launch_bounds__(THREADS) void k_radix_scatter(
    const KeyT *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, KeyT *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, int64_t n, int shift, const uint32_t *__restrict__ hist_pass,
    uint64_t *__restrict__ lookback, uint32_t *__restrict__ tile_counter, uint32_t epoch,
    const uint32_t *__restrict__ tile_offsets, int64_t tiles, SegView seg)
{
    constexpr int WAVES = THREADS / 64;
    constexpr int TILE = THREADS * ITEMS;
    // 32-bit keys, and 64-bit keys in tiles of up to 4096: keys and payload
    // staged together and written in one sweep (a slot's digit comes from its
    // staged key); larger 64-bit tiles: keys and their digits first, then the
    // payload in the key buffer (LDS for both at once would leave one block
    // per CU)
    constexpr bool kOneSweep = sizeof(KeyT) == 4 || TILE <= 4096;
    static_assert(THREADS >= kRadix, "one thread per digit");
    __shared__ KeyT s_stage[TILE];                     // keys in digit order
    __shared__ uint32_t s_pay1[kOneSweep ? TILE : 1];  // payload in digit order (one sweep)
    __shared__ uint8_t s_dig[kOneSweep ? 1 : TILE];    // digit of each staged slot (two sweeps)
    __shared__ uint32_t s_wcnt[WAVES][kRadix];
    __shared__ uint32_t s_off[kRadix];      // global base minus tile-local start, per digit
    __shared__ uint32_t s_lstart[kRadix];   // tile-local digit start
    __shared__ uint32_t s_dbase[kRadix];    // global exclusive digit base
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;

    // reduce-then-scan: the tile is the block's own, so its scanned digit
    // offsets are requested now and arrive while the keys are ranked
    uint32_t pre_excl = 0;
    if (!LOOKBACK && tid < kRadix) {
        const uint32_t t = xcd_contiguous_block(blockIdx.x, gridDim.x);
        pre_excl = tile_offsets[(int64_t)t * kRadix + tid];
        if (seg.tile_start)
            pre_excl += seg.base[(int64_t)seg.tile_seg[t] * kRadix + tid];
    }
    if (tid == 0)
        s_tile = LOOKBACK ? atomicAdd(tile_counter, 1u) : xcd_contiguous_block(blockIdx.x, gridDim.x);
    for (int i = tid; i < WAVES * kRadix; i += THREADS)
        (&s_wcnt[0][0])[i] = 0;
    {
        // segmented: the digit bases are in seg.base, per segment
        const uint32_t b = digit_exclusive_scan((tid < kRadix && !seg.tile_start) ? hist_pass[tid] : 0u, s_wave);
        if (tid < kRadix)
            s_dbase[tid] = b;
    }
    __syncthreads();

    const uint32_t tile = s_tile;
    int64_t tile_base = (int64_t)tile * TILE, tile_end = tile_base + TILE;
    if (seg.tile_start) {
        tile_base = seg.tile_start[tile];
        tile_end = seg.tile_start[tile + 1];
    }
    if (tile_end > n)
        tile_end = n;
    const int64_t wave_base = tile_base + (int64_t)wave * (ITEMS * 64);

    KeyT key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t pos[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int64_t idx = wave_base + (int64_t)j * 64 + lane;
        const bool valid = idx < tile_end;
        key[j] = valid ? keys_in[idx] : (KeyT)~(KeyT)0;
        if (FIRST)
            val[j] = (uint32_t)idx;
        else
            val[j] = valid ? vals_in[idx] : 0u;
    }

    // Stable per-wave ranking: items of one wave are visited in index order
    // (round j covers indices wave_base + 64 j + lane).
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int64_t idx = wave_base + (int64_t)j * 64 + lane;
        const bool valid = idx < tile_end;
        const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
        const uint64_t peers = match_digit(d, __ballot(valid));
        const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : lane;
        uint32_t old = 0;
        if (valid && lane == leader)
            old = atomicAdd(&s_wcnt[wave][d], (uint32_t)__popcll(peers));
        old = __shfl(old, leader, 64);
        pos[j] = old + (uint32_t)__popcll(peers & lt);  // rank within the wave, for now
    }
    __syncthreads();

    // Per digit (thread tid < 256 owns digit tid): wave-exclusive offsets,
    // tile total, decoupled lookback for the tile's global offset.
    uint32_t total = 0, excl = 0;
    if (tid < kRadix) {
        const int d = tid;
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            const uint32_t c = s_wcnt[w][d];
            s_wcnt[w][d] = total;
            total += c;
        }
        if (!LOOKBACK) {
            excl = pre_excl;
        } else if (tile == 0) {
            lb_store(lookback + d, lb_pack(epoch, kStatusPrefix, total));
        } else {
            uint64_t *mine = lookback + (size_t)tile * kRadix + d;
            lb_store(mine, lb_pack(epoch, kStatusAggregate, total));
            int64_t t = (int64_t)tile - 1;
            while (true) {
                const uint64_t wv = lb_load(lookback + (size_t)t * kRadix + d);
                const uint32_t hi = (uint32_t)(wv >> 32);
                const uint32_t status = ((hi >> 2) == epoch) ? (hi & 3u) : 0u;
                if (status == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (uint32_t)wv;
                if (status == kStatusPrefix)
                    break;
                t--;
            }
            lb_store(mine, lb_pack(epoch, kStatusPrefix, excl + total));
        }
    }
    {
        const uint32_t lstart = digit_exclusive_scan(total, s_wave);
        if (tid < kRadix) {
            s_lstart[tid] = lstart;
            s_off[tid] = s_dbase[tid] + excl - lstart;
        }
    }
    __syncthreads();

    // slot in the tile, digit order
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
        pos[j] += s_lstart[d] + s_wcnt[wave][d];
    }
    const int count = (int)(tile_end - tile_base);

    if constexpr (kOneSweep) {
        // keys and payload through LDS, out in coalesced digit runs
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const int64_t idx = wave_base + (int64_t)j * 64 + lane;
            if (idx < tile_end) {
                s_stage[pos[j]] = key[j];
                s_pay1[pos[j]] = val[j];
            }
        }
        __syncthreads();
        for (int j = tid; j < count; j += THREADS) {
            const KeyT k = s_stage[j];
            const uint32_t dst = s_off[(uint32_t)(k >> shift) & (kRadix - 1)] + (uint32_t)j;
            if (!LAST)
                keys_out[dst] = k;
            vals_out[dst] = s_pay1[j];
        }
    } else {
        if (!LAST) {
#pragma unroll
            for (int j = 0; j < ITEMS; j++) {
                const int64_t idx = wave_base + (int64_t)j * 64 + lane;
                if (idx < tile_end) {
                    s_stage[pos[j]] = key[j];
                    s_dig[pos[j]] = (uint8_t)((key[j] >> shift) & (kRadix - 1));
                }
            }
            __syncthreads();
            for (int j = tid; j < count; j += THREADS)
                keys_out[s_off[s_dig[j]] + (uint32_t)j] = s_stage[j];
            __syncthreads();
        }
        uint32_t *s_pay = reinterpret_cast<uint32_t *>(s_stage);
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const int64_t idx = wave_base + (int64_t)j * 64 + lane;
            if (idx < tile_end) {
                s_pay[pos[j]] = val[j];
                if (LAST)
                    s_dig[pos[j]] = (uint8_t)((key[j] >> shift) & (kRadix - 1));
            }
        }
        __syncthreads();
        for (int j = tid; j < count; j += THREADS)
            vals_out[s_off[s_dig[j]] + (uint32_t)j] = s_pay[j];
    }
}

template <typename KeyT, int THREADS, int ITEMS, bool FIRST, bool LAST>
hipError_t launch_pass(SortWork &w, int src, uint32_t *vals_out, int64_t n, int pass, hipStream_t stream,
                       Profiler *prof, uint32_t *zero_word)
{
    constexpr int TILE = THREADS * ITEMS;
    const int64_t tiles = (n + TILE - 1) / TILE;
    const KeyT *kin = reinterpret_cast<const KeyT *>(w.keys[src]);
    KeyT *kout = LAST ? nullptr : reinterpret_cast<KeyT *>(w.keys[src ^ 1]);
    const uint32_t *vin = FIRST ? nullptr : w.vals[src];
    const int shift = pass * kRadixBits;
    hipError_t e;
    if (w.config.alg == kSortOnesweep) {
        const uint32_t epoch = w.epoch++;
        if (prof)
            prof->begin(K_SORT_PASS);
        hipLaunchKernelGGL((k_radix_scatter<KeyT, THREADS, ITEMS, FIRST, LAST, true>), dim3((unsigned)tiles),
                           dim3(THREADS), 0, stream, kin, vin, kout, vals_out, n, shift, w.hist + pass * kRadix,
                           w.lookback, w.counters + pass, epoch, nullptr, tiles, SegView{});
        if (prof)
            prof->end(K_SORT_PASS);
        return hipGetLastError();
    }
    // reduce-then-scan: the tile counts live in the lookback buffer (which
    // holds at least sort_tiles(n) * kRadix 64-bit words) or, segmented, in
    // the segment table; the digit totals in the first kRadix words of hist.
    const bool segmented = !w.seg.sizes.empty();
    const int64_t ntiles = segmented ? w.seg.tiles : tiles;
    uint32_t *counts = segmented ? w.seg.counts : reinterpret_cast<uint32_t *>(w.lookback);
    SegView sv;
    if (segmented) {
        sv.tile_start = w.seg.tile_start;
        sv.tile_seg = w.seg.tile_seg;
        sv.base = w.seg.base;
    }
    if (prof)
        prof->begin(K_SORT_UPSWEEP);
    hipLaunchKernelGGL((k_rts_upsweep<KeyT, THREADS, ITEMS>), dim3((unsigned)ntiles), dim3(THREADS), 0, stream,
                       kin, n, shift, counts, ntiles, sv.tile_start, zero_word);
    if (prof)
        prof->end(K_SORT_UPSWEEP);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if (prof)
        prof->begin(K_SORT_SCAN);
    hipLaunchKernelGGL(k_rts_scan, dim3(kRadix), dim3(kScanThreads), 0, stream, counts, ntiles, w.hist);
    if (segmented)
        hipLaunchKernelGGL(k_seg_bases, dim3((unsigned)w.seg.sizes.size()), dim3(kRadix), 0, stream, counts, w.hist,
                           w.seg.seg_first, w.seg.seg_start, (int)w.seg.sizes.size(), w.seg.base);
    if (prof)
        prof->end(K_SORT_SCAN);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if (prof)
        prof->begin(K_SORT_PASS);
    hipLaunchKernelGGL((k_radix_scatter<KeyT, THREADS, ITEMS, FIRST, LAST, false>), dim3((unsigned)ntiles),
                       dim3(THREADS), 0, stream, kin, vin, kout, vals_out, n, shift, w.hist, nullptr, nullptr, 0u,
                       counts, ntiles, sv);
    if (prof)
        prof->end(K_SORT_PASS);
    return hipGetLastError();
}

// have_vals: the first pass reads its payload from vals[src] instead of
// generating the input positions; keep_keys: the last pass writes its keys
// too (the sorted keys stay in keys[src ^ 1]); zero_word: cleared by the
// first pass for a later kernel of the same sort.
struct PassFlags {
    bool have_vals;
    bool keep_keys;
    uint32_t *zero_word;
};

template <typename KeyT, int THREADS, int ITEMS>
hipError_t launch_pass_cfg(SortWork &w, int src, uint32_t *vals_out, int64_t n, int pass, int passes,
                           hipStream_t stream, Profiler *prof, PassFlags f)
{
    const bool first = pass == 0 && !f.have_vals;
    const bool last = pass == passes - 1 && !f.keep_keys;
    uint32_t *zw = pass == 0 ? f.zero_word : nullptr;
    if (first && last)
        return launch_pass<KeyT, THREADS, ITEMS, true, true>(w, src, vals_out, n, pass, stream, prof, zw);
    if (first)
        return launch_pass<KeyT, THREADS, ITEMS, true, false>(w, src, vals_out, n, pass, stream, prof, zw);
    if (last)
        return launch_pass<KeyT, THREADS, ITEMS, false, true>(w, src, vals_out, n, pass, stream, prof, zw);
    return launch_pass<KeyT, THREADS, ITEMS, false, false>(w, src, vals_out, n, pass, stream, prof, zw);
}

template <typename KeyT>
hipError_t launch_pass_any(SortWork &w, int src, uint32_t *vals_out, int64_t n, int pass, int passes,
                           hipStream_t stream, Profiler *prof, PassFlags f)
{
    const SortConfig c = w.config;
    if (c.threads == 512 && c.items == 16)
        return launch_pass_cfg<KeyT, 512, 16>(w, src, vals_out, n, pass, passes, stream, prof, f);
    if (c.threads == 256 && c.items == 24)
        return launch_pass_cfg<KeyT, 256, 24>(w, src, vals_out, n, pass, passes, stream, prof, f);
    if (c.threads == 512 && c.items == 12)
        return launch_pass_cfg<KeyT, 512, 12>(w, src, vals_out, n, pass, passes, stream, prof, f);
    if (c.threads == 512 && c.items == 8)
        return launch_pass_cfg<KeyT, 512, 8>(w, src, vals_out, n, pass, passes, stream, prof, f);
    if (c.threads == 1024 && c.items == 6)
        return launch_pass_cfg<KeyT, 1024, 6>(w, src, vals_out, n, pass, passes, stream, prof, f);
    return launch_pass_cfg<KeyT, 256, 16>(w, src, vals_out, n, pass, passes, stream, prof, f);
}

// (Re)builds the device tile table of the segments for tile size `tile`:
// every segment is cut into ceil(size / tile) tiles of its own.
hipError_t seg_table_build(SegTable &t, int tile, int64_t n)
{
    const int nseg = (int)t.sizes.size();
    std::vector<int64_t> start;
    std::vector<uint32_t> tseg, first(nseg + 1), sstart(nseg + 1);
    int64_t pos = 0;
    for (int s = 0; s < nseg; s++) {
        first[s] = (uint32_t)tseg.size();
        sstart[s] = (uint32_t)pos;
        for (int64_t k = 0; k < t.sizes[s]; k += tile) {
            start.push_back(pos + k);
            tseg.push_back((uint32_t)s);
        }
        pos += t.sizes[s];
    }
    first[nseg] = (uint32_t)tseg.size();
    sstart[nseg] = (uint32_t)pos;
    start.push_back(n);
    const int64_t tiles = (int64_t)tseg.size();
    if (pos != n || tiles > t.capacity)
        return hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipMemcpy(t.tile_start, start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return e;
    if ((e = hipMemcpy(t.tile_seg, tseg.data(), tseg.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return e;
    if ((e = hipMemcpy(t.seg_first, first.data(), first.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return e;
    if ((e = hipMemcpy(t.seg_start, sstart.data(), sstart.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return e;
    t.tiles = tiles;
    t.tile = tile;
    return hipSuccess;
}

template <typename KeyT>
hipError_t sort_typed(SortWork &w, int64_t n, uint32_t *perm_out, hipStream_t stream, Profiler *prof,
                      int kPasses = (int)sizeof(KeyT), bool use_segments = true, bool have_vals = false,
                      bool keep_keys = false, uint32_t *zero_word = nullptr)
{
    const PassFlags flags{have_vals, keep_keys, zero_word};
    hipError_t e;
    // the segment table is set aside for an unsegmented sort of this call
    std::vector<int64_t> parked;
    if (!use_segments)
        parked.swap(w.seg.sizes);
    struct Restore {
        SortWork &w;
        std::vector<int64_t> &v;
        ~Restore()
        {
            if (!v.empty())
                w.seg.sizes.swap(v);
        }
    } restore{w, parked};
    if (!w.seg.sizes.empty()) {
        if (w.config.alg != kSortReduceScan)
            return hipErrorNotSupported;
        const int tile = w.config.threads * w.config.items;
        if (w.seg.tile != tile && (e = seg_table_build(w.seg, tile, n)) != hipSuccess)
            return e;
    }
    if (w.config.alg == kSortOnesweep) {
        // The epoch lives in 30 bits of the lookback word; wrap by clearing
        // (a fill kernel, not a memset node, so that the sort stays
        // capturable).  Epoch 0 also marks words left by reduce-then-scan.
        if (w.epoch + kPasses >= (1u << 30) || w.epoch == 0) {
            const int64_t words = sort_tiles(n) * kRadix;
            hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, stream,
                               w.lookback, words, (uint64_t)0);
            if ((e = hipGetLastError()) != hipSuccess)
                return e;
            w.epoch = 1;
        }
        if (prof)
            prof->begin(K_SORT_ZERO);
        hipLaunchKernelGGL(k_sort_zero, dim3(1), dim3(256), 0, stream, w.hist, kPasses * kRadix, w.counters,
                           kPasses, zero_word);
        if (prof)
            prof->end(K_SORT_ZERO);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        const int64_t per_block = (int64_t)kHistThreads * kHistItems;
        int64_t blocks = (n + per_block - 1) / per_block;
        if (blocks > 2048)
            blocks = 2048;
        if (prof)
            prof->begin(K_SORT_HIST);
        hipLaunchKernelGGL(k_sort_histogram<KeyT>, dim3((unsigned)blocks), dim3(kHistThreads), 0, stream,
                           reinterpret_cast<const KeyT *>(w.keys[0]), n, w.hist);
        if (prof)
            prof->end(K_SORT_HIST);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
    } else {
        w.epoch = 0;  // the tile counts overwrite lookback words
    }
    int src = 0;
    for (int p = 0; p < kPasses; p++) {
        uint32_t *vout = (p == kPasses - 1) ? perm_out : w.vals[src ^ 1];
        if ((e = launch_pass_any<KeyT>(w, src, vout, n, p, kPasses, stream, prof, flags)) != hipSuccess)
            return e;
        src ^= 1;
    }
    return hipSuccess;
}

// ------------------------------------------------- 64-bit keys in two planes
//
// A 64-bit radius key is stored as a high and a low 32-bit plane (store_key).
// The stable sort on the whole key is the stable sort on the high plane (four
// 8-bit passes) followed by a stable sort on the low plane within each run of
// equal high halves: runs are contiguous after the first sort and keep their
// input order, so ordering a run by (low half, position) orders it by (key,
// input position).  Between steps the radii of N = 1.2e7 particles share
// their high half (sign, exponent and 20 mantissa bits) in runs of about
// three, so the second stage is one light kernel instead of four more
// passes.  Runs that a block cannot hold (longer than its look-ahead, or any
// length on adversarial input) go to a kernel that sorts each on one
// workgroup.

constexpr int kLongThreads = 1024;
constexpr int kLongBlocks = 64;
constexpr int kLongRunCounter = 48;  // SortWork::counters word: runs left to k_split_long_runs

// First index s of seg_start[0..nseg) with seg_start[s] >= x (nseg if none).
__device__ inline int first_start_at_or_after(const uint32_t *seg_start, int nseg, int64_t x)
{
    int lo = 0, hi = nseg;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)seg_start[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Orders the runs of equal high halves that start in a wave's tile.  hs: the
// high plane in high-half order; perm_hi: the input position of each (the
// first sort's permutation); lo: the low plane by input position; out: the
// final permutation.  A run is also cut where a segment starts.
//
// Each wave works alone (no workgroup barrier): it owns kFxTile positions and
// reads one row of 64 more ahead, so that a run starting in its tile and
// crossing the tile's end is still its own.  Run starts are one 64-bit ballot
// per row; from them every lane finds the start and the end of its run with
// bit scans and two uniform tables of the last and next start of each row.
// Low halves are gathered only for positions that share their run; a run in
// low-half order already (runs of one, exact ties, the key sentinel of
// deferred particles) is copied, the others rank each element against their
// run in LDS.  A run that reaches past the read-ahead row is recorded (its
// first position) for k_split_long_runs.
constexpr int kFxRows = 8;
constexpr int kFxTile = 64 * kFxRows;   // positions whose runs one wave orders
constexpr int kFxAll = kFxRows + 1;     // rows read, the last one ahead of the tile
constexpr int kFxWaves = 4;             // independent waves per workgroup
constexpr int kFxNone = 1 << 30;        // "no run start" in the uniform tables

// Sources of the key that orders a run of equal sort keys, by input
// position: the low plane of two-plane 64-bit keys, or the radius key of the
// float64 radius itself (after a sort on quantile keys).
struct TieLowPlane {
    using T = uint32_t;
    const uint32_t *lo;
    __device__ T operator()(uint32_t p) const { return lo[p]; }
};
struct TieRadius {
    using T = uint64_t;
    const double *r;
    __device__ T operator()(uint32_t p) const { return radius_key(r[p]); }
};

template <class Lo>
__global__ __launch_bounds__(64 * kFxWaves) void k_split_fixup(const uint32_t *__restrict__ hs, Lo lo,
                                                                 const uint32_t *__restrict__ perm_hi,
                                                                 uint32_t *__restrict__ out, int64_t n,
                                                                 const uint32_t *__restrict__ seg_start, int nseg,
                                                                 uint32_t *__restrict__ long_runs,
                                                                 uint32_t *__restrict__ long_count)
{
    using T = typename Lo::T;
    __shared__ T s_l[kFxWaves][kFxAll * 64];
    __shared__ uint32_t s_uns[kFxWaves][kFxAll * 16];       // bytes by run start: the run has a descent
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t w0 = ((int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * kFxWaves + wv) * kFxTile;
    if (w0 >= n)
        return;  // the whole wave
    uint8_t *uns = reinterpret_cast<uint8_t *>(s_uns[wv]);
    uint32_t h[kFxAll], p[kFxAll];
#pragma unroll
    for (int r = 0; r < kFxAll; r++) {
        const int64_t j = w0 + r * 64 + lane;
        h[r] = j < n ? hs[j] : 0u;
        p[r] = j < n ? perm_hi[j] : 0u;
    }
    for (int k = lane; k < kFxAll * 16; k += 64)
        s_uns[wv][k] = 0;
    const int64_t after = w0 + kFxAll * 64;  // the position past the rows
    const uint32_t h_before = w0 > 0 ? hs[w0 - 1] : 0u;
    const uint32_t h_after = after < n ? hs[after] : 0u;
    // run starts: bit `lane` of row r; positions past n count as starts
    uint64_t M[kFxAll + 1];
    uint32_t carry = h_before;
#pragma unroll
    for (int r = 0; r < kFxAll; r++) {
        const int64_t j = w0 + r * 64 + lane;
        uint32_t prev = __shfl_up(h[r], 1, 64);
        if (lane == 0)
            prev = carry;
        carry = __shfl(h[r], 63, 64);
        M[r] = __ballot(j >= n || j == 0 || h[r] != prev);
    }
    M[kFxAll] = (after >= n || h_after != carry) ? 1ull : 0ull;
    if (seg_start) {  // segment starts inside the rows (and right after them)
        for (int sg = first_start_at_or_after(seg_start, nseg, w0); sg < nseg && (int64_t)seg_start[sg] <= after;
             sg++) {
            const int64_t q = (int64_t)seg_start[sg] - w0;
#pragma unroll
            for (int r = 0; r <= kFxAll; r++)
                if ((q >> 6) == r)
                    M[r] |= 1ull << (q & 63);
        }
    }
    // uniform tables: the last start at or before the end of row r, the first
    // start at or after the beginning of row r (window positions)
    int lastS[kFxAll], nextS[kFxAll + 1];
    int run_last = -1;
#pragma unroll
    for (int r = 0; r < kFxAll; r++) {
        if (M[r])
            run_last = r * 64 + 63 - __clzll((long long)M[r]);
        lastS[r] = run_last;
    }
    nextS[kFxAll] = (M[kFxAll] & 1) ? kFxAll * 64 : kFxNone;
#pragma unroll
    for (int r = kFxAll - 1; r >= 0; r--)
        nextS[r] = M[r] ? r * 64 + __ffsll((long long)M[r]) - 1 : nextS[r + 1];
    const uint64_t le = (2ull << lane) - 1;  // bits 0..lane
    // low halves (shared runs only), run bounds, descents
    int ra[kFxAll], re[kFxAll];
    T lprev_row = 0;
#pragma unroll
    for (int r = 0; r < kFxAll; r++) {
        const int64_t j = w0 + r * 64 + lane;
        const bool hd = (M[r] >> lane) & 1;
        const bool next_hd = lane < 63 ? ((M[r] >> (lane + 1)) & 1) : (M[r + 1] & 1);
        const T l = (j < n && !(hd && next_hd) && h[r] != kSentinelWord) ? lo(p[r]) : (T)0;
        s_l[wv][r * 64 + lane] = l;
        T lprev = __shfl_up(l, 1, 64);
        if (lane == 0)
            lprev = lprev_row;
        lprev_row = __shfl(l, 63, 64);
        const uint64_t mle = M[r] & le, mgt = M[r] & ~le;
        const int a = mle ? r * 64 + 63 - __clzll((long long)mle) : (r > 0 ? lastS[r - 1] : -1);
        const int e = mgt ? r * 64 + __ffsll((long long)mgt) - 1 : nextS[r + 1];
        ra[r] = a;
        re[r] = e;
        if (j < n && !hd && l < lprev && a >= 0)
            uns[a] = 1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int r = 0; r < kFxAll; r++) {
        const int jl = r * 64 + lane;
        const int a = ra[r], e = re[r];
        if (w0 + jl >= n || a < 0 || a >= kFxTile)
            continue;  // past the end, or a run of an earlier or the next wave
        if (e == kFxNone) {  // reaches past the rows read
            if (jl == a)
                long_runs[atomicAdd(long_count, 1u)] = (uint32_t)(w0 + a);
            continue;
        }
        int rank = jl - a;
        if (uns[a]) {  // rank by (tie key, position) within the run
            const T lk = s_l[wv][jl];
            rank = 0;
            for (int q = a; q < e; q++) {
                const T lq = s_l[wv][q];
                rank += (lq < lk || (lq == lk && q < jl)) ? 1 : 0;
            }
        }
        out[w0 + a + rank] = p[r];
    }
}

// The runs k_split_fixup left: one workgroup per run (a stride of runs per
// workgroup).  The run's end is found by a scan; a run already in tie-key
// order (exact ties, the deferral sentinel, the usual cases) keeps its order;
// otherwise it is sorted stably by its tie keys, 8-bit passes (four for the
// low plane, eight for a radius key) between the run's own slices of out and
// p1, ending in out.  Each pass gathers the tie keys through the permutation
// it reads, so the run needs no key staging.
template <class Lo>
__global__ __launch_bounds__(kLongThreads) void k_split_long_runs(
    const uint32_t *__restrict__ hs, Lo lo, const uint32_t *__restrict__ perm_hi, uint32_t *__restrict__ out,
    int64_t n, const uint32_t *__restrict__ seg_start, int nseg, const uint32_t *__restrict__ long_runs,
    const uint32_t *__restrict__ long_count, uint32_t *__restrict__ p1)
{
    using T = typename Lo::T;
    constexpr int WAVES = kLongThreads / 64;
    constexpr int kPasses = (int)sizeof(T);
    static_assert(kPasses % 2 == 0, "an even pass count ends in out");
    __shared__ uint32_t s_wcnt[WAVES][kRadix];
    __shared__ uint32_t s_off[kRadix], s_tot[kRadix];
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_stop;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nruns = *long_count;
    const uint64_t lt = lanemask_lt();
    for (uint32_t ri = blockIdx.x; ri < nruns; ri += gridDim.x) {
        const int64_t a = long_runs[ri];
        int64_t limit = n;
        if (seg_start) {
            const int sg = first_start_at_or_after(seg_start, nseg, a + 1);
            if (sg < nseg)
                limit = seg_start[sg];
        }
        const uint32_t h = hs[a];
        int64_t e = limit;
        for (int64_t base = a + 1; base < limit; base += kLongThreads) {
            if (tid == 0)
                s_stop = kLongThreads;
            __syncthreads();
            const int64_t k = base + tid;
            if (k < limit && hs[k] != h)
                atomicMin(&s_stop, (uint32_t)tid);
            __syncthreads();
            const uint32_t st = s_stop;
            __syncthreads();
            if (st < (uint32_t)kLongThreads) {
                e = base + st;
                break;
            }
        }
        const int64_t m = e - a;
        for (int64_t k = tid; k < m; k += kLongThreads)
            out[a + k] = perm_hi[a + k];
        if (h == kSentinelWord)
            continue;  // exact ties: input order
        bool unsorted = false;
        for (int64_t k = tid + 1; k < m; k += kLongThreads)
            unsorted |= lo(perm_hi[a + k]) < lo(perm_hi[a + k - 1]);
        if (!__syncthreads_or(unsorted))
            continue;  // in order already: out holds it
        for (int pass = 0; pass < kPasses; pass++) {
            const int shift = 8 * pass;
            const uint32_t *sp = (pass & 1) ? p1 : out;
            uint32_t *dp = (pass & 1) ? out : p1;
            for (int i = tid; i < WAVES * kRadix; i += kLongThreads)
                (&s_wcnt[0][0])[i] = 0;
            if (tid < kRadix)
                s_tot[tid] = 0;
            __syncthreads();
            for (int64_t k = tid; k < m; k += kLongThreads)
                atomicAdd(&s_tot[(uint32_t)(lo(sp[a + k]) >> shift) & (kRadix - 1)], 1u);
            __syncthreads();
            const uint32_t ex = digit_exclusive_scan(tid < kRadix ? s_tot[tid] : 0u, s_wave);
            if (tid < kRadix)
                s_off[tid] = ex;
            __syncthreads();
            // chunks of kLongThreads in input order: stable ranks from wave
            // ballots and per-wave digit counts
            for (int64_t base = 0; base < m; base += kLongThreads) {
                const int64_t k = base + tid;
                const bool valid = k < m;
                const uint32_t pv = valid ? sp[a + k] : 0u;
                const uint32_t d = valid ? (uint32_t)(lo(pv) >> shift) & (kRadix - 1) : 0u;
                const uint64_t peers = match_digit(d, __ballot(valid));
                const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : lane;
                if (valid && lane == leader)
                    s_wcnt[wave][d] = (uint32_t)__popcll(peers);
                __syncthreads();
                if (tid < kRadix) {
                    uint32_t run = 0;
                    for (int w = 0; w < WAVES; w++) {
                        const uint32_t c = s_wcnt[w][tid];
                        s_wcnt[w][tid] = run;
                        run += c;
                    }
                    s_tot[tid] = run;
                }
                __syncthreads();
                if (valid) {
                    const uint32_t pos = s_off[d] + s_wcnt[wave][d] + (uint32_t)__popcll(peers & lt);
                    dp[a + pos] = pv;
                }
                __syncthreads();
                if (tid < kRadix) {
                    s_off[tid] += s_tot[tid];
                    for (int w = 0; w < WAVES; w++)
                        s_wcnt[w][tid] = 0;
                }
                __syncthreads();
            }
        }
    }
}

// The run stage after the passes: hs = the sorted sort keys, perm_hi = the
// passes' permutation (no alias of out or p1).
template <class Lo>
hipError_t order_runs(SortWork &w, int64_t n, const uint32_t *hs, Lo lo, const uint32_t *perm_hi,
                      uint32_t *perm_out, uint32_t *p1, uint32_t *count, hipStream_t stream, Profiler *prof)
{
    const uint32_t *seg_start = w.seg.sizes.empty() ? nullptr : w.seg.seg_start;
    const int nseg = (int)w.seg.sizes.size();
    // the tile counts are spent: their buffer lists the long runs (at most
    // one per wave tile of kFxTile positions).  Onesweep's lookback words share
    // that buffer, so the next onesweep sort must clear it first (epoch 0).
    uint32_t *long_runs = reinterpret_cast<uint32_t *>(w.lookback);
    w.epoch = 0;
    if (prof)
        prof->begin(K_SORT_FIXUP);
    const int64_t waves = (n + kFxTile - 1) / kFxTile;
    hipLaunchKernelGGL(k_split_fixup<Lo>, dim3((unsigned)((waves + kFxWaves - 1) / kFxWaves)), dim3(64 * kFxWaves),
                       0, stream, hs, lo, perm_hi, perm_out, n, seg_start, nseg, long_runs, count);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_split_long_runs<Lo>, dim3(kLongBlocks), dim3(kLongThreads), 0, stream, hs, lo, perm_hi,
                           perm_out, n, seg_start, nseg, long_runs, count, p1);
        e = hipGetLastError();
    }
    if (prof)
        prof->end(K_SORT_FIXUP);
    return e;
}

// The stable sort of the two-plane 64-bit keys in keys[0] (see above).
hipError_t sort_split64(SortWork &w, int64_t n, uint32_t *perm_out, hipStream_t stream, Profiler *prof)
{
    uint32_t *count = w.counters + kLongRunCounter;
    uint32_t *perm_hi = w.vals[0];
    // four passes on the high plane; the last writes its keys, so the sorted
    // high halves end in keys[0] (passes alternate keys[0] -> keys[1] -> ...),
    // and the low plane at keys[0] + n is never written
    hipError_t e = sort_typed<uint32_t>(w, n, perm_hi, stream, prof, 4, true, false, true, count);
    if (e != hipSuccess)
        return e;
    const uint32_t *hs = reinterpret_cast<const uint32_t *>(w.keys[0]);
    return order_runs(w, n, hs, TieLowPlane{hs + n}, perm_hi, perm_out, w.vals[1], count, stream, prof);
}

} // namespace

hipError_t sort_prepare(SortWork &w, int64_t n)
{
    if (w.seg.sizes.empty() || w.config.alg != kSortReduceScan)
        return hipSuccess;
    const int tile = w.config.threads * w.config.items;
    return w.seg.tile == tile ? hipSuccess : seg_table_build(w.seg, tile, n);
}

void seg_table_release(SegTable &t)
{
    void *bufs[] = {t.tile_start, t.tile_seg, t.seg_first, t.seg_start, t.base, t.counts};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    t.tile_start = nullptr;
    t.tile_seg = t.seg_first = t.seg_start = t.base = t.counts = nullptr;
    t.capacity = t.tiles = 0;
    t.tile = 0;
}

bool sort_config_supported(SortConfig c)
{
    if (c.alg != kSortReduceScan && c.alg != kSortOnesweep)
        return false;
    return (c.threads == 256 && (c.items == 16 || c.items == 24)) ||
           (c.threads == 512 && (c.items == 8 || c.items == 12 || c.items == 16)) ||
           (c.threads == 1024 && c.items == 6);
}

hipError_t launch_radius_keys(const double *r, uint64_t *keys, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_radius_keys<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, r, keys, n);
    return hipGetLastError();
}

hipError_t launch_radius_keys(const float *r, uint32_t *keys, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_radius_keys<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, r, keys, n);
    return hipGetLastError();
}

hipError_t radix_sort_perm(SortWork &w, int64_t n, int key_bits, uint32_t *perm_out, hipStream_t stream,
                           Profiler *prof)
{
    if (n <= 0)
        return hipSuccess;
    if (key_bits == 32)
        return sort_typed<uint32_t>(w, n, perm_out, stream, prof);
    return sort_split64(w, n, perm_out, stream, prof);
}

hipError_t radix_sort_perm_q24(SortWork &w, int64_t n, const double *r, uint32_t *perm_out, hipStream_t stream,
                               Profiler *prof, bool buckets)
{
    if (n <= 0)
        return hipSuccess;
    if (buckets && w.qs.cap > 0 && w.qs.n == n)
        return qsort_perm(w.qs, reinterpret_cast<const uint32_t *>(w.keys[0]), n, r, perm_out, w.vals[0], stream,
                          prof);
    uint32_t *count = w.counters + kLongRunCounter;
    // three passes on the low 24 bits; the last keeps its keys, which end in
    // keys[1] (keys[0] -> keys[1] -> keys[0] -> keys[1]), and its payload
    // lands in vals[1] (vals[1] -> vals[0] -> vals[1])
    uint32_t *perm_hi = w.vals[1];
    hipError_t e = sort_typed<uint32_t>(w, n, perm_hi, stream, prof, 3, true, false, true, count);
    if (e != hipSuccess)
        return e;
    const uint32_t *hs = reinterpret_cast<const uint32_t *>(w.keys[1]);
    return order_runs(w, n, hs, TieRadius{r}, perm_hi, perm_out, w.vals[0], count, stream, prof);
}

hipError_t radix_sort_perm_ids(SortWork &w, int64_t n, const void *r, bool f32, const uint32_t *id,
                               uint32_t *order, uint32_t *perm_out, hipStream_t stream, Profiler *prof)
{
    if (n <= 0)
        return hipSuccess;
    hipError_t e;
    const dim3 grid((unsigned)((n + 255) / 256));
    // 1. the ids, stably (a permutation within each segment)
    hipLaunchKernelGGL(k_id_keys, grid, dim3(256), 0, stream, id, reinterpret_cast<uint32_t *>(w.keys[0]), n);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if ((e = sort_typed<uint32_t>(w, n, order, stream, prof)) != hipSuccess)
        return e;
    // 2. the radius keys in id order, that order as the payload; a stable sort
    //    on them leaves equal radii in id order
    if (f32)
        hipLaunchKernelGGL(k_radius_keys_gather<float>, grid, dim3(256), 0, stream, static_cast<const float *>(r),
                           order, reinterpret_cast<uint32_t *>(w.keys[0]), w.vals[0], n);
    else
        hipLaunchKernelGGL(k_radius_keys_gather<double>, grid, dim3(256), 0, stream, static_cast<const double *>(r),
                           order, w.keys[0], w.vals[0], n);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    if (f32)
        return sort_typed<uint32_t>(w, n, perm_out, stream, prof, 4, true, true);
    return sort_typed<uint64_t>(w, n, perm_out, stream, prof, 8, true, true);
}

hipError_t radix_sort_perm_u8(SortWork &w, int64_t n, uint32_t *perm_out, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    if (w.config.alg != kSortReduceScan)
        return hipErrorNotSupported;
    return sort_typed<uint32_t>(w, n, perm_out, stream, nullptr, 1, false);
}

} // namespace ns
