// ns_sort.hip -- stable LSD radix sort of radius keys for MI355X (gfx950).
//
// Replaces the reference's radial sort (sort_particles_with_alg,
// nsphere.c:8605-8652, with the stable quadsort of nsphere.c:8356-8359 and the
// comparator of nsphere.c:8297-8306).  An LSD radix sort is stable, so keys
// that tie keep their input (previous-step) order -- the order quadsort gives.
//
// Design (one HIP stream, wave64), for 64-bit keys (float64 radii, 8 passes)
// or 32-bit keys (float32 radii, 4 passes):
//   * one upfront pass reads every key once and builds all the 8-bit digit
//     histograms (LDS-privatised, run-length-combined counters);
//   * one "onesweep" kernel per pass: a workgroup takes a tile of
//     THREADS * ITEMS keys by atomic ticket, ranks its keys stably per digit
//     with wave ballots (a 64-lane match of the 8 digit bits), publishes its
//     digit counts and resolves its global digit offsets by decoupled
//     lookback over earlier tiles, then stages keys and payload -- one after
//     the other, through one LDS buffer -- in digit order and writes them out
//     so that each digit run leaves as one coalesced segment;
//   * pass 0 synthesises the payload (input positions) instead of reading
//     it, and the last pass writes only the payload (the permutation).
// Block size and keys per thread are a SortConfig (the autotuner's choice).
#include <hip/hip_runtime.h>

#include "ns_device.h"

namespace ns {

namespace {

// Lookback word: high 32 bits = epoch << 2 | status, low 32 bits = count.
constexpr uint32_t kStatusAggregate = 1;
constexpr uint32_t kStatusPrefix = 2;
// Predecessor tiles read per lookback round (each read is a ~0.4 us coherent
// load; walks of 30-80 tiles are typical with ~1300 tiles in flight).
constexpr int kLookbackWindow = 16;

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

__device__ inline uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
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

template <typename S>
__global__ void k_radius_keys(const S *__restrict__ r, typename KeyOf<S>::type *__restrict__ keys, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        keys[i] = radius_key(r[i]);
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, int64_t n, uint64_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void k_sort_zero(uint32_t *__restrict__ hist, int nhist, uint32_t *__restrict__ counters,
                            int ncounters)
{
    for (int i = threadIdx.x; i < nhist; i += blockDim.x)
        hist[i] = 0;
    for (int i = threadIdx.x; i < ncounters; i += blockDim.x)
        counters[i] = 0;
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

// One onesweep pass over digit `shift`.  FIRST: payload = input position.
// LAST: keys are not written.
template <typename KeyT, int THREADS, int ITEMS, bool FIRST, bool LAST>
__global__ __launch_bounds__(THREADS) void k_onesweep(
    const KeyT *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, KeyT *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, int64_t n, int shift, const uint32_t *__restrict__ hist_pass,
    uint64_t *__restrict__ lookback, uint32_t *__restrict__ tile_counter, uint32_t epoch)
{
    constexpr int WAVES = THREADS / 64;
    constexpr int TILE = THREADS * ITEMS;
    static_assert(THREADS >= kRadix, "one thread per digit");
    __shared__ KeyT s_stage[TILE];          // keys in digit order, then the payload
    __shared__ uint8_t s_dig[TILE];         // digit of each staged slot
    __shared__ uint32_t s_wcnt[WAVES][kRadix];
    __shared__ uint32_t s_off[kRadix];      // global base minus tile-local start, per digit
    __shared__ uint32_t s_lstart[kRadix];   // tile-local digit start
    __shared__ uint32_t s_dbase[kRadix];    // global exclusive digit base
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;

    if (tid == 0)
        s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < WAVES * kRadix; i += THREADS)
        (&s_wcnt[0][0])[i] = 0;
    {
        const uint32_t b = digit_exclusive_scan(tid < kRadix ? hist_pass[tid] : 0u, s_wave);
        if (tid < kRadix)
            s_dbase[tid] = b;
    }
    __syncthreads();

    const uint32_t tile = s_tile;
    const int64_t tile_base = (int64_t)tile * TILE;
    const int64_t wave_base = tile_base + (int64_t)wave * (ITEMS * 64);

    KeyT key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t pos[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int64_t idx = wave_base + (int64_t)j * 64 + lane;
        const bool valid = idx < n;
        key[j] = valid ? keys_in[idx] : (KeyT)~(KeyT)0;
        if (FIRST)
            val[j] = (uint32_t)idx;
        else
            val[j] = valid ? vals_in[idx] : 0u;
    }

    // Stable per-wave ranking: items of one wave are visited in index order
    // (round j covers indices wave_base + 64 j + lane).  Three batched phases
    // -- all ballots, then all LDS counter updates (one per digit group and
    // round, issued back to back; LDS applies one wave's operations in
    // order), then all broadcasts -- so the rounds' LDS round trips overlap.
    const uint64_t lt = lanemask_lt();
    uint32_t below[ITEMS];  // same-digit lanes below this one in the round
    uint32_t group[ITEMS];  // same-digit lanes in the round
    uint32_t leader[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int64_t idx = wave_base + (int64_t)j * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        below[j] = (uint32_t)__popcll(peers & lt);
        group[j] = valid ? (uint32_t)__popcll(peers) : 0u;
        leader[j] = valid ? (uint32_t)(__ffsll((unsigned long long)peers) - 1) : (uint32_t)lane;
    }
    uint32_t old[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        old[j] = 0;
        if (group[j] != 0 && (uint32_t)lane == leader[j])
            old[j] = atomicAdd(&s_wcnt[wave][(uint32_t)(key[j] >> shift) & (kRadix - 1)], group[j]);
    }
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        pos[j] = __shfl(old[j], (int)leader[j], 64) + below[j];  // rank within the wave, for now
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
        uint64_t *mine = lookback + (size_t)tile * kRadix + d;
        if (tile == 0) {
            lb_store(mine, lb_pack(epoch, kStatusPrefix, total));
        } else {
            lb_store(mine, lb_pack(epoch, kStatusAggregate, total));
            // Windowed lookback: the words of the kLookbackWindow previous
            // tiles are loaded together and consumed in order -- aggregates
            // summed until an inclusive prefix is met; a tile that has not
            // published yet ends the window and is read again.
            int64_t t = (int64_t)tile - 1;
            while (true) {
                uint64_t wv[kLookbackWindow];
#pragma unroll
                for (int k = 0; k < kLookbackWindow; k++)
                    wv[k] = (t - k >= 0) ? lb_load(lookback + (size_t)(t - k) * kRadix + d) : 0ull;
                int used = 0;
                bool found = false;
#pragma unroll
                for (int k = 0; k < kLookbackWindow; k++) {
                    if (found || used < k)
                        continue;
                    const uint32_t hi = (uint32_t)(wv[k] >> 32);
                    const uint32_t status = ((hi >> 2) == epoch) ? (hi & 3u) : 0u;
                    if (status == 0)
                        continue;
                    excl += (uint32_t)wv[k];
                    used = k + 1;
                    found = status == kStatusPrefix;
                }
                if (found)
                    break;
                if (used == 0)
                    __builtin_amdgcn_s_sleep(1);
                t -= used;
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

#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint32_t d = (uint32_t)(key[j] >> shift) & (kRadix - 1);
        pos[j] += s_lstart[d] + s_wcnt[wave][d];  // slot in the tile, digit order
    }
    const int64_t remain = n - tile_base;
    const int count = remain < TILE ? (int)remain : TILE;

    if (!LAST) {
        // keys through LDS in digit order, then out in coalesced digit runs
#pragma unroll
        for (int j = 0; j < ITEMS; j++) {
            const int64_t idx = wave_base + (int64_t)j * 64 + lane;
            if (idx < n) {
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
        if (idx < n) {
            s_pay[pos[j]] = val[j];
            if (LAST)
                s_dig[pos[j]] = (uint8_t)((key[j] >> shift) & (kRadix - 1));
        }
    }
    __syncthreads();
    for (int j = tid; j < count; j += THREADS)
        vals_out[s_off[s_dig[j]] + (uint32_t)j] = s_pay[j];
}

template <typename KeyT, int THREADS, int ITEMS, bool FIRST, bool LAST>
hipError_t launch_pass(SortWork &w, int src, uint32_t *vals_out, int64_t n, int pass, hipStream_t stream)
{
    constexpr int TILE = THREADS * ITEMS;
    const int64_t tiles = (n + TILE - 1) / TILE;
    const uint32_t epoch = w.epoch++;
    hipLaunchKernelGGL((k_onesweep<KeyT, THREADS, ITEMS, FIRST, LAST>), dim3((unsigned)tiles), dim3(THREADS), 0,
                       stream, reinterpret_cast<const KeyT *>(w.keys[src]), FIRST ? nullptr : w.vals[src],
                       LAST ? nullptr : reinterpret_cast<KeyT *>(w.keys[src ^ 1]), vals_out, n,
                       pass * kRadixBits, w.hist + pass * kRadix, w.lookback, w.counters + pass, epoch);
    return hipGetLastError();
}

template <typename KeyT, int THREADS, int ITEMS>
hipError_t launch_pass_cfg(SortWork &w, int src, uint32_t *vals_out, int64_t n, int pass, int passes,
                           hipStream_t stream)
{
    if (pass == 0)
        return launch_pass<KeyT, THREADS, ITEMS, true, false>(w, src, vals_out, n, pass, stream);
    if (pass == passes - 1)
        return launch_pass<KeyT, THREADS, ITEMS, false, true>(w, src, vals_out, n, pass, stream);
    return launch_pass<KeyT, THREADS, ITEMS, false, false>(w, src, vals_out, n, pass, stream);
}

template <typename KeyT>
hipError_t launch_pass_any(SortWork &w, int src, uint32_t *vals_out, int64_t n, int pass, int passes,
                           hipStream_t stream)
{
    const SortConfig c = w.config;
    if (c.threads == 512 && c.items == 16)
        return launch_pass_cfg<KeyT, 512, 16>(w, src, vals_out, n, pass, passes, stream);
    if (c.threads == 256 && c.items == 24)
        return launch_pass_cfg<KeyT, 256, 24>(w, src, vals_out, n, pass, passes, stream);
    if (c.threads == 512 && c.items == 12)
        return launch_pass_cfg<KeyT, 512, 12>(w, src, vals_out, n, pass, passes, stream);
    return launch_pass_cfg<KeyT, 256, 16>(w, src, vals_out, n, pass, passes, stream);
}

template <typename KeyT>
hipError_t sort_typed(SortWork &w, int64_t n, uint32_t *perm_out, hipStream_t stream, Profiler *prof)
{
    constexpr int kPasses = (int)sizeof(KeyT);
    hipError_t e;
    // The epoch lives in 30 bits of the lookback word; wrap by clearing (a
    // fill kernel, not a memset node, so that the sort stays capturable).
    if (w.epoch + kPasses >= (1u << 30) || w.epoch == 0) {
        const int64_t words = sort_tiles(n) * kRadix;
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, stream, w.lookback,
                           words, (uint64_t)0);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        w.epoch = 1;
    }
    if (prof)
        prof->begin(K_SORT_ZERO);
    hipLaunchKernelGGL(k_sort_zero, dim3(1), dim3(256), 0, stream, w.hist, kPasses * kRadix, w.counters, kPasses);
    if (prof)
        prof->end(K_SORT_ZERO);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    {
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
    }
    int src = 0;
    for (int p = 0; p < kPasses; p++) {
        uint32_t *vout = (p == kPasses - 1) ? perm_out : w.vals[src ^ 1];
        if (prof)
            prof->begin(K_SORT_PASS);
        e = launch_pass_any<KeyT>(w, src, vout, n, p, kPasses, stream);
        if (prof)
            prof->end(K_SORT_PASS);
        if (e != hipSuccess)
            return e;
        src ^= 1;
    }
    return hipSuccess;
}

} // namespace

bool sort_config_supported(SortConfig c)
{
    return (c.threads == 256 && (c.items == 16 || c.items == 24)) ||
           (c.threads == 512 && (c.items == 12 || c.items == 16));
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
    return sort_typed<uint64_t>(w, n, perm_out, stream, prof);
}

} // namespace ns
