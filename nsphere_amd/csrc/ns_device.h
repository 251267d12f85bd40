// ns_device.h -- device-side helpers and kernel launchers shared by the
// HIP translation units of nsphere_amd (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "nsphere_gpu.h"

namespace ns {

// ---------------------------------------------------------------- sort sizes
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kKeyPasses = 64 / kRadixBits;
constexpr int kSortThreads = 256;                      // 4 wave64 per workgroup
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;                         // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;   // 4096 keys per tile
constexpr int kHistThreads = 256;
constexpr int kHistItems = 16;

inline int64_t sort_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }

// MI355X deals a launch's workgroups to its eight XCDs round-robin (block b
// to XCD b % 8), and each XCD has its own L2.  This bijection of [0, nblocks)
// gives XCD x one contiguous range of block indices instead, so that
// neighbouring blocks -- which gather from overlapping windows, or write the
// two halves of one cache line -- meet in the same L2.
__device__ inline uint32_t xcd_contiguous_block(uint32_t b, uint32_t nblocks)
{
    const uint32_t q = nblocks / 8, r = nblocks % 8;
    const uint32_t x = b % 8, j = b / 8;
    return x * q + (x < r ? x : r) + j;
}

// Order-preserving 64-bit radius key: unsigned order of the key equals the
// order of compare_particles (nsphere.c:8297-8306) on non-NaN radii; -0 and
// +0 map to one key (they compare equal there), every NaN maps to one key
// after +inf.
__host__ __device__ inline uint64_t radius_key(double r)
{
    uint64_t u;
    if (r != r) {
        u = 0x7FF8000000000000ull;
    } else if (r == 0.0) {
        u = 0;
    } else {
        u = (uint64_t)__builtin_bit_cast(uint64_t, r);
    }
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// The same total order on float32 radii (the state of the mixed-precision
// form): a 32-bit key, so the sort needs four 8-bit passes instead of eight.
__host__ __device__ inline uint32_t radius_key(float r)
{
    uint32_t u;
    if (r != r) {
        u = 0x7FC00000u;
    } else if (r == 0.0f) {
        u = 0;
    } else {
        u = (uint32_t)__builtin_bit_cast(uint32_t, r);
    }
    return (u >> 31) ? ~u : (u | 0x80000000u);
}

// Radius keys of n particles in a key buffer.  A 32-bit key (float32 radii)
// is stored at [i].  A 64-bit key (float64 radii) is stored as two 32-bit
// planes, the high half at [i] and the low half at [n + i] (32-bit words):
// the radix passes sort the high plane as 32-bit keys and the runs of equal
// high halves are then ordered by the low plane, which the passes leave
// untouched.
__device__ inline void store_key(void *keys, int64_t n, int64_t i, uint64_t k)
{
    uint32_t *w = static_cast<uint32_t *>(keys);
    w[i] = (uint32_t)(k >> 32);
    w[n + i] = (uint32_t)k;
}
__device__ inline void store_key(void *keys, int64_t, int64_t i, uint32_t k)
{
    static_cast<uint32_t *>(keys)[i] = k;
}

// ------------------------------------------------------------ quantile keys
//
// Between steps the float64 radii are sorted on a 24-bit key instead of
// their 64 bits: the bucket of the radius in a table of kQBuckets quantiles
// of the step's start radii (the radius keys of ranks 0, s, 2s, ... for the
// stride s = quantile_stride(n)), then kQSubBits of linear position inside the
// bucket.  The key is a non-decreasing function of the radius key, so a
// stable sort on it followed by a stable sort of each run of equal keys on
// the full radius key is the stable sort on the radius.  Quantile buckets
// hold about s particles each and the sub-bins under one particle each on
// average, so three 8-bit passes sort it and about half of the positions
// share their key with another.
constexpr int kQBuckets = 4096;
constexpr int kQSubBits = 12;
constexpr uint32_t kQKeyMax = 0xFFFFFEu;  // keys stay below the deferral sentinel's low 24 bits
constexpr int64_t kQMinN = 65536;         // smaller sorts keep the 64-bit keys
constexpr int kQDBits = 20;               // the sub-bin offset is scaled to under 2^20 first

// Start ranks per quantile bucket: ceil(n / kQBuckets) when the last
// bucket's first rank still lies below n (the last bucket then holds no more
// than the others), else n / kQBuckets (the last bucket takes the rest).
__host__ __device__ inline uint32_t quantile_stride(int64_t n)
{
    const int64_t c = (n + kQBuckets - 1) / kQBuckets;
    return (uint32_t)((int64_t)(kQBuckets - 1) * c <= n - 1 ? c : n / kQBuckets);
}

// One bucket: k = radius key of its first quantile; ms = (shift << 56) | mul,
// sub-bin = (((key - k) >> shift) * mul) >> 32 (integers only, so the map is
// exactly monotone).
struct QEntry {
    uint64_t k;
    uint64_t ms;
};

// Bucket parameters from the radius keys of its quantile (k0) and of the next
// one (k1; for the last bucket the largest start radius, inclusive).
__host__ __device__ inline uint64_t quantile_ms(uint64_t k0, uint64_t k1, bool last)
{
    const uint64_t span = k1 - k0 + (last ? 1u : 0u);  // keys k0 .. k0 + span - 1
    if (span == 0)
        return 0;
    int t = 0;
    while (((span - 1) >> t) >= ((uint64_t)1 << kQDBits))
        t++;
    const uint64_t d = ((span - 1) >> t) + 1;  // offsets after the shift: 0 .. d - 1
    const uint64_t mul = ((uint64_t)1 << (32 + kQSubBits)) / d;
    return ((uint64_t)t << 56) | mul;
}

// Largest j in [0, n) with j == 0 or key_at(j) <= k, for non-decreasing
// key_at: a galloping search from the guess b (then bisection), so that a
// key near its guess costs two or three reads.
template <class KeyAt>
__device__ inline int gallop_bucket(KeyAt key_at, int n, int b, uint64_t k)
{
    b = b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
    int lo, hi;
    if (b > 0 && key_at(b) > k) {
        hi = b;
        lo = b - 1;
        int step = 1;
        while (lo > 0 && key_at(lo) > k) {
            hi = lo;
            step <<= 1;
            lo = b - step > 0 ? b - step : 0;
        }
    } else {
        lo = b;
        hi = b + 1;
        int step = 1;
        while (hi < n && key_at(hi) <= k) {
            lo = hi;
            step <<= 1;
            hi = b + step < n ? b + step : n;
        }
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (key_at(mid) <= k)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Buckets staged in LDS by a kernel: k[j] = key of bucket w0 + j for j < wn,
// k[wn] = key of bucket w0 + wn (all ones past the last bucket); ms[j] the
// bucket's sub-bin parameters (null: read from the table).
struct QWin {
    const uint64_t *k = nullptr;
    const uint64_t *ms = nullptr;
    int w0 = 0, wn = 0;
};

// Stages buckets [w0, w0 + wn) (and the key after them) of q into k (wn + 1
// words) and ms (wn words, or null), with all threads of the block; the
// caller synchronises the block before the first search.
__device__ inline QWin qwin_stage(const QEntry *__restrict__ q, int w0, int wn, uint64_t *k, uint64_t *ms)
{
    // rounds of kPer independent loads per thread, then their stores, so a
    // thread waits for one round trip per round (the whole table: two)
    constexpr int kPer = 9;
    const int nt = (int)blockDim.x;
    for (int base = 0; base <= wn; base += kPer * nt) {
        uint64_t kk[kPer], mm[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int j = base + u * nt + (int)threadIdx.x;
            kk[u] = (j <= wn && w0 + j < kQBuckets) ? q[w0 + j].k : ~0ull;
            mm[u] = (ms && j < wn) ? q[w0 + j].ms : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int j = base + u * nt + (int)threadIdx.x;
            if (j <= wn)
                k[j] = kk[u];
            if (ms && j < wn)
                ms[j] = mm[u];
        }
    }
    QWin w;
    w.k = k;
    w.ms = ms;
    w.w0 = w0;
    w.wn = wn;
    return w;
}

// 24-bit quantile key of radius key k for the particle of start rank `rank`
// (the search starts at the bucket of its start radius): in the staged
// buckets when they hold k's bucket, else in the table.
__device__ inline uint32_t quantile_key(const QEntry *__restrict__ q, uint32_t stride, int64_t rank, uint64_t k,
                                        const QWin &w = QWin())
{
    const int guess = (int)(rank / (int64_t)stride);
    int b;
    uint64_t base, ms;
    if (w.k && (w.w0 == 0 || w.k[0] <= k) && k < w.k[w.wn]) {
        const int j = gallop_bucket([&](int x) { return w.k[x]; }, w.wn, guess - w.w0, k);
        b = w.w0 + j;
        base = w.k[j];
        ms = w.ms ? w.ms[j] : q[b].ms;
    } else {
        b = gallop_bucket([&](int x) { return q[x].k; }, kQBuckets, guess, k);
        base = q[b].k;
        ms = q[b].ms;
    }
    uint32_t sub = 0;
    if (k > base) {
        uint64_t d = (k - base) >> (ms >> 56);
        const uint64_t dmax = ((uint64_t)1 << kQDBits) - 1;
        d = d < dmax ? d : dmax;  // only the last bucket reaches past its span
        const uint64_t s = (d * (ms & (((uint64_t)1 << 56) - 1))) >> 32;
        sub = s < ((1u << kQSubBits) - 1) ? (uint32_t)s : ((1u << kQSubBits) - 1);
    }
    const uint32_t key = ((uint32_t)b << kQSubBits) | sub;
    return key < kQKeyMax ? key : kQKeyMax;
}

template <typename S>
struct KeyOf;
template <>
struct KeyOf<double> {
    using type = uint64_t;
};
template <>
struct KeyOf<float> {
    using type = uint32_t;
};

// Radix-pass algorithm: reduce-then-scan (per pass a count kernel, a scan of
// the tile counts and a scatter kernel; no inter-workgroup waiting) or
// onesweep (one scatter kernel per pass with decoupled lookback).
constexpr int kSortReduceScan = 0;
constexpr int kSortOnesweep = 1;

// Sort configuration (the autotuner's choice): the pass algorithm and the
// tile shape, workgroup size and keys per thread (tile = product, at least
// kSortTile so that the lookback buffer's tile count bounds it).
struct SortConfig {
    int alg = kSortReduceScan;
    int threads = 512;  // 512 x 12: fastest measured shape for both key widths
    int items = 12;     // at N = 1.2e7 on MI355X, with either pass algorithm
};
bool sort_config_supported(SortConfig c);

// Segmented sort: the tile table handed to the reduce-then-scan kernels.
// Tiles never straddle a segment boundary; tile t covers
// [tile_start[t], tile_start[t + 1]) of segment tile_seg[t], and
// base[s * kRadix + d] turns a tile's scanned digit count into its global
// offset (k_seg_bases).  All null: one unsegmented array.
struct SegView {
    const int64_t *tile_start = nullptr;
    const uint32_t *tile_seg = nullptr;
    const uint32_t *base = nullptr;
};

// Segments of the sort (halos of an ensemble): sizes on the host, the tile
// table of the last tile size on the device.
struct SegTable {
    std::vector<int64_t> sizes;     // particles per segment (empty: unsegmented)
    int tile = 0;                   // tile size the device table was built for
    int64_t tiles = 0;
    int64_t *tile_start = nullptr;  // tiles + 1
    uint32_t *tile_seg = nullptr;   // tiles
    uint32_t *seg_first = nullptr;  // nseg + 1: first tile of each segment
    uint32_t *seg_start = nullptr;  // nseg + 1: first element of each segment
    uint32_t *base = nullptr;       // nseg * kRadix
    uint32_t *counts = nullptr;     // tiles * kRadix (32-bit tile counts)
    int64_t capacity = 0;           // tiles the device table can hold
};

// Device workspace of the bucket sort of quantile keys (ns_qsort.hip),
// allocated by qsort_prepare for one n; cap == 0: not allocated (or n too
// large for a bucket to fit a workgroup), the LSD passes sort quantile keys.
struct QSortBufs {
    uint2 *elem = nullptr;     // (kQBuckets + 1) regions of cap (key, position) pairs, one per bucket
    uint2 *ovf = nullptr;      // n pairs: the elements past their region's cap
    uint32_t *cnt = nullptr;   // kQBuckets + 1 bucket counters (zero between sorts)
    uint32_t *ctl = nullptr;   // overflow count, bucket workgroups done (zero between sorts)
    int cap = 0;               // elements per region (the bucket kernel's capacity)
    int64_t n = 0;
};

// Device workspace of one radix sort.  Key buffers hold 64-bit keys, or
// 32-bit keys in their first half.
struct SortWork {
    SortConfig config;
    uint64_t *keys[2];   // ping-pong key buffers (keys[0] holds the input)
    uint32_t *vals[2];   // ping-pong payload buffers
    uint32_t *hist;      // kKeyPasses * kRadix digit counts (onesweep) or the
                         // pass's digit totals (reduce-then-scan)
    uint64_t *lookback;  // sort_tiles(n) * kRadix decoupled-lookback words, or
                         // the reduce-then-scan tile counts (32-bit)
    uint32_t *counters;  // kKeyPasses tile tickets (+ spare)
    uint32_t epoch;      // lookback epoch of the next pass (host-owned)
    SegTable seg;        // segments (reduce-then-scan only)
    QSortBufs qs;        // bucket sort of quantile keys
};
// Frees the device part of a segment table.
void seg_table_release(SegTable &t);
// The host-side set-up that the first segmented sort of n keys would do (the
// segment table for the current tile), done ahead: a graph capture must not
// meet its synchronous copies.
hipError_t sort_prepare(SortWork &w, int64_t n);

// Launchers (ns_sort.hip).  All asynchronous on `stream`.
hipError_t launch_radius_keys(const double *r, uint64_t *keys, int64_t n, hipStream_t stream);
hipError_t launch_radius_keys(const float *r, uint32_t *keys, int64_t n, hipStream_t stream);
// Sorts the key_bits-wide keys in keys[0], stably: 32-bit keys, or 64-bit
// keys in two planes (store_key).  On return perm_out[k] = input position of
// rank k (perm_out aliases no vals buffer).  The 64-bit sort uses both key
// buffers, both payload buffers, the tile-count buffer and counters[48].
struct Profiler;
// One stable pass on the low byte of 32-bit keys in keys[0], ignoring any
// segments (reduce-then-scan only): the order of a small integer label.
hipError_t radix_sort_perm_u8(SortWork &w, int64_t n, uint32_t *perm_out, hipStream_t stream);
// Stable sort on (radius key, id): the ids first, then the radius keys of r
// (float32 or float64) gathered through the id order; order: n scratch
// entries (no alias of the sort's buffers or of perm_out).
hipError_t radix_sort_perm_ids(SortWork &w, int64_t n, const void *r, bool f32, const uint32_t *id,
                               uint32_t *order, uint32_t *perm_out, hipStream_t stream, Profiler *prof);
hipError_t radix_sort_perm(SortWork &w, int64_t n, int key_bits, uint32_t *perm_out, hipStream_t stream,
                           Profiler *prof);
// Sorts the 24-bit quantile keys in keys[0] (quantile_key of the radii r,
// n >= kQMinN) stably, then orders each run of equal keys by the radius key
// of r and input position: the stable sort on r.  Uses both key and payload
// buffers, the tile-count buffer and counters[48].
hipError_t radix_sort_perm_q24(SortWork &w, int64_t n, const double *r, uint32_t *perm_out, hipStream_t stream,
                               Profiler *prof, bool buckets = true);
// The bucket sort's workspace for n keys (ns_qsort.hip): allocated, and its
// counters cleared, once per n; a no-op when n is too large for a bucket to
// fit a workgroup.  Synchronous: call it outside any capture.
hipError_t qsort_prepare(QSortBufs &b, int64_t n);
void qsort_release(QSortBufs &b);
// The stable sort of the quantile keys in keys (n >= kQMinN, b prepared for
// n, deferred particles at the all-ones key): perm_out[k] = input position of
// rank k by (radius key of r, position), the deferred last by position.
// scratch: n entries no other argument aliases.
hipError_t qsort_perm(QSortBufs &b, const uint32_t *keys, int64_t n, const double *r, uint32_t *perm_out,
                      uint32_t *scratch, hipStream_t stream, Profiler *prof);

// Kernel-family ids for profiling.
enum KernelId : int {
    K_KEYS = 0,
    K_SORT_ZERO,
    K_SORT_HIST,
    K_SORT_PASS,
    K_GATHER,
    K_INTEGRATE,
    K_INVMAP,
    K_ENERGY,
    K_TIEFIX,
    K_SORT_UPSWEEP,
    K_SORT_SCAN,
    K_INT_BULK,
    K_INT_DEEP,
    K_MASS_SCAN,
    K_MLP,
    K_APD,
    K_POTENTIAL,
    K_DENSITY,
    K_RANKFILES,
    K_INT_LATE,
    K_MERGE,
    K_PLACE,
    K_SORT_FIXUP,
    K_INT_LC,
    K_QS_SCATTER,
    K_QS_BUCKET,
    K_ORDER,
    K_COUNT
};
const char *kernel_name(int id);

// Per-kernel-family timing with HIP events on the launching stream.  begin()
// and end() bracket one launch (on `stream`, or on s when given); fold()
// waits for recorded pairs and adds their durations.  Disabled (no events)
// unless `on`, and only for the families whose bit (1 << id) is in mask.
struct Profiler {
    bool on = false;
    uint32_t mask = ~0u;
    hipStream_t stream = nullptr;
    int64_t launches[K_COUNT] = {};
    double ms[K_COUNT] = {};
    void begin(int kid, hipStream_t s = nullptr);
    void end(int kid, hipStream_t s = nullptr);
    hipError_t fold();
    void reset();
    void release();

  private:
    struct Pending {
        int kid;
        hipEvent_t a, b;
    };
    Pending *pend_ = nullptr;
    int npend_ = 0, cap_ = 0;
    hipEvent_t open_ = nullptr;
    int open_kid_ = -1;
    hipEvent_t *pool_ = nullptr;
    int npool_ = 0, pool_cap_ = 0;
    hipEvent_t take();
};

} // namespace ns
