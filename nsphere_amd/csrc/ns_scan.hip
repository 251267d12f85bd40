// ns_scan.hip -- enclosed mass M(r) for per-particle masses, MI355X (gfx950).
//
// The reference has equal masses and takes M(r) from the rank,
// ((double)i / npts) * HALO_MASS (nsphere.c:513, 5815), and so does this
// library unless the caller sets per-particle masses (no reference function
// has them).  Then the enclosed mass of the particle at rank i is the
// exclusive prefix sum of the masses in rank order, computed in one pass:
//   * a workgroup takes a tile of kScanTile masses by ticket, gathered
//     through the sort's permutation when one is given; each wave owns 1024
//     consecutive masses as 16 rows of 64 (every load and store coalesced),
//     scans each row with wave shuffles and carries the row totals, then a
//     four-wave combine finishes the tile -- in float64 for float64 masses,
//     or for float32 masses in float32 (hi, lo) pairs with two-sum
//     compensation (Knuth's TwoSum, renormalised after every addition);
//   * the tile's aggregate, as a 64-bit fixed-point integer, is published
//     and the tile's exclusive prefix is found by decoupled lookback, one
//     wave reading 64 predecessor tiles per round.  Integer partial sums do
//     not depend on how far a walk had to go, so M(r) is the same bit for bit
//     on every run.  The lookback epoch and the tile ticket live in device
//     memory and a one-block kernel advances them before each scan, so the
//     scan replays correctly from a captured HIP graph.
// menc[i] = (double)(tile prefix) / scale + (within-tile exclusive prefix);
// menc[n] = the total, menc[n + 1] = the mass of the particle at rank 0.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kScanWaves = kScanThreads / 64;
constexpr uint32_t kFlagAggregate = 1;
constexpr uint32_t kFlagPrefix = 2;

// float32 pair hi + lo with |lo| <= ulp(hi) / 2 (exact TwoSum needs
// -ffp-contract=off, which the library is built with)
struct F2 {
    float hi, lo;
};

__device__ inline F2 two_sum(float a, float b)
{
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ inline F2 f2_add(F2 a, F2 b)
{
    const F2 s = two_sum(a.hi, b.hi);
    const float lo = s.lo + (a.lo + b.lo);
    const float hi = s.hi + lo;  // fast renormalisation: |lo| small against s.hi
    return {hi, lo - (hi - s.hi)};
}

// Accumulator per mass type.
template <typename M>
struct Acc;

template <>
struct Acc<double> {
    using T = double;
    __device__ static T zero() { return 0.0; }
    __device__ static T of(double m) { return m; }
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static T shfl_up(T x, int d) { return __shfl_up(x, d, 64); }
    __device__ static T shfl(T x, int l) { return __shfl(x, l, 64); }
    __device__ static double value(T x) { return x; }
    __device__ static int64_t fixed(T x, double scale) { return (int64_t)(x * scale); }
};

template <>
struct Acc<float> {
    using T = F2;
    __device__ static T zero() { return {0.0f, 0.0f}; }
    __device__ static T of(float m) { return {m, 0.0f}; }
    __device__ static T add(T a, T b) { return f2_add(a, b); }
    __device__ static T shfl_up(T x, int d) { return {__shfl_up(x.hi, d, 64), __shfl_up(x.lo, d, 64)}; }
    __device__ static T shfl(T x, int l) { return {__shfl(x.hi, l, 64), __shfl(x.lo, l, 64)}; }
    __device__ static double value(T x) { return (double)x.hi + (double)x.lo; }
    __device__ static int64_t fixed(T x, double scale)
    {
        return (int64_t)((double)x.hi * scale) + (int64_t)((double)x.lo * scale);
    }
};

__device__ inline void word_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t word_load(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A tile's 64-bit value goes out as two self-describing words, each
// (tag << 32 | one half); a reader accepts a pair only when both tags carry
// this launch's epoch and the same status, i.e. come from one publication.
__device__ inline void publish(uint64_t *words, int64_t tile, uint32_t epoch, uint32_t status, uint64_t x)
{
    const uint64_t tag = (uint64_t)((epoch << 2) | status) << 32;
    word_store(words + 2 * tile, tag | (x >> 32));
    word_store(words + 2 * tile + 1, tag | (x & 0xFFFFFFFFull));
}

// ctl[0] = tile ticket, ctl[1] = lookback epoch (30 bits; on wrap-around the
// words are cleared and the epochs restart).
template <typename M>
__global__ __launch_bounds__(kScanThreads) void k_mass_scan(const M *__restrict__ mass_in,
                                                             const uint32_t *__restrict__ perm,
                                                             M *__restrict__ mass_out, double *__restrict__ menc,
                                                             int64_t n, double scale, uint64_t *__restrict__ words,
                                                             uint32_t *__restrict__ ctl)
{
    uint32_t *ticket = ctl;
    const uint32_t epoch = ctl[1];
    using A = Acc<M>;
    using T = typename A::T;
    __shared__ T s_wave[kScanWaves];
    __shared__ uint32_t s_tile;
    __shared__ int64_t s_prefix;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0)
        s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t wave_base = tile * kScanTile + (int64_t)wave * (kScanItems * 64);

    // row j of this wave: items wave_base + 64 j + lane
    M m[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        const int64_t idx = wave_base + (int64_t)j * 64 + lane;
        M x = (M)0;
        if (idx < n)
            x = perm ? mass_in[perm[idx]] : mass_in[idx];
        m[j] = x;
        if (mass_out && idx < n)
            mass_out[idx] = x;
    }
    // exclusive prefix of every item within the wave: per row a shuffle
    // scan, plus the carried totals of the rows before
    T loc[kScanItems];
    T carry = A::zero();
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        const T x = A::of(m[j]);
        T incl = x;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const T y = A::shfl_up(incl, off);
            if (lane >= off)
                incl = A::add(y, incl);
        }
        const T ex = A::shfl_up(incl, 1);
        loc[j] = lane == 0 ? carry : A::add(carry, ex);
        const T row = A::shfl(incl, 63);
        carry = A::add(carry, row);
    }
    if (lane == 0)
        s_wave[wave] = carry;
    __syncthreads();
    T before = A::zero(), total = A::zero();
#pragma unroll
    for (int w = 0; w < kScanWaves; w++) {
        if (w < wave)
            before = A::add(before, s_wave[w]);
        total = A::add(total, s_wave[w]);
    }

    // decoupled lookback on fixed-point tile aggregates (wave 0)
    if (wave == 0) {
        const uint64_t agg = (uint64_t)A::fixed(total, scale);
        uint64_t acc = 0;
        if (tile == 0) {
            if (lane == 0)
                publish(words, 0, epoch, kFlagPrefix, agg);
        } else {
            if (lane == 0)
                publish(words, tile, epoch, kFlagAggregate, agg);
            int64_t base = tile - 1;
            while (true) {
                const int64_t t = base - lane;
                uint32_t st = kFlagPrefix;  // before tile 0: an empty prefix
                uint64_t val = 0;
                if (t >= 0) {
                    const uint64_t w0 = word_load(words + 2 * t), w1 = word_load(words + 2 * t + 1);
                    const uint32_t g0 = (uint32_t)(w0 >> 32), g1 = (uint32_t)(w1 >> 32);
                    st = (g0 == g1 && (g0 >> 2) == epoch) ? (g0 & 3u) : 0u;
                    val = ((w0 & 0xFFFFFFFFull) << 32) | (w1 & 0xFFFFFFFFull);
                }
                const uint64_t not_ready = __ballot(st == 0);
                const uint64_t prefix = __ballot(st == kFlagPrefix);
                const int ready = not_ready ? __ffsll((unsigned long long)not_ready) - 1 : 64;
                const uint64_t run = ready >= 64 ? ~0ull : ((1ull << ready) - 1);
                const uint64_t pre = prefix & run;
                // lanes taken this round: up to the nearest prefix, else the
                // whole published run
                const int take = pre ? __ffsll((unsigned long long)pre) : ready;
                uint64_t part = lane < take ? val : 0;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1)
                    part += __shfl_xor(part, off, 64);
                acc += part;
                if (pre)
                    break;
                if (take == 0)
                    __builtin_amdgcn_s_sleep(1);
                base -= take;
            }
            if (lane == 0)
                publish(words, tile, epoch, kFlagPrefix, acc + agg);
        }
        if (lane == 0)
            s_prefix = (int64_t)acc;
    }
    __syncthreads();
    const double tile_base = (double)s_prefix / scale;

#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        const int64_t idx = wave_base + (int64_t)j * 64 + lane;
        if (idx < n)
            menc[idx] = tile_base + A::value(A::add(before, loc[j]));
        if (idx == n - 1)  // the total, after the last item
            menc[n] = tile_base + A::value(A::add(A::add(before, loc[j]), A::of(m[j])));
    }
    if (tile == 0 && tid == 0)
        menc[n + 1] = (double)m[0];
}

// Before each scan: ticket to 0, epoch + 1 (clearing the words and
// restarting at 1 when the 30-bit epoch would wrap).
__global__ __launch_bounds__(1024) void k_scan_begin(uint32_t *ctl, uint64_t *words, int64_t nwords)
{
    const uint32_t e = ctl[1];  // read by every thread before anyone writes
    __syncthreads();
    const bool wrap = e + 1 >= (1u << 30) - 1;
    if (wrap)
        for (int64_t i = threadIdx.x; i < nwords; i += blockDim.x)
            words[i] = 0;
    if (threadIdx.x == 0) {
        ctl[0] = 0;
        ctl[1] = wrap ? 1u : e + 1;
    }
}

// ---- dynamic potential (PsiA, nsphere.c:7698-7738) -------------------------
//
// psi_i = G ((i + 1) dM / r_i + dM sum_{j > i} 1 / r_j) vel_conv_sq over the
// radii in rank order, psi0 = G dM sum_j (1 / r_j) vel_conv_sq: a suffix sum
// of 1/r, done as reduce-then-scan over tiles of kScanTile in a fixed order,
// so that the result is the same on every run.

// Exclusive suffix over a block's threads (the sum of x over the threads
// above this one) and the block total, by wave shuffles and the wave totals,
// in a fixed order.  s_w: WAVES doubles.
template <int WAVES>
__device__ inline double block_suffix_exclusive(double x, double *s_w, double &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double inc = x;  // sum over lanes >= lane
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_down(inc, off, 64);
        if (lane + off < 64)
            inc += y;
    }
    const double up = __shfl_down(inc, 1, 64);
    const double excl_wave = lane < 63 ? up : 0.0;
    if (lane == 0)
        s_w[wave] = inc;
    __syncthreads();
    double after = 0.0;
    total = 0.0;
    for (int w = WAVES - 1; w >= 0; w--) {
        const double t = s_w[w];
        if (w > wave)
            after += t;
        total += t;
    }
    __syncthreads();  // s_w may be reused
    return after + excl_wave;
}

// The tile's sum of 1/r: coalesced reads (thread t takes items t, t + 256,
// ...), then the block's fixed-order total.
template <typename S>
__global__ __launch_bounds__(kScanThreads) void k_psi_tiles(const S *__restrict__ r, int64_t n,
                                                            double *__restrict__ tile_sum)
{
    __shared__ double s_w[kScanWaves];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t i = base + (int64_t)k * kScanThreads + tid;
        if (i < n)
            acc += 1.0 / (double)r[i];
    }
    double total;
    (void)block_suffix_exclusive<kScanWaves>(acc, s_w, total);
    if (tid == 0)
        tile_sum[blockIdx.x] = total;
}

// One workgroup: exclusive suffix over the tile sums (in place) and the total.
__global__ __launch_bounds__(1024) void k_psi_top(double *__restrict__ tile_sum, int64_t tiles,
                                                  double *__restrict__ total)
{
    __shared__ double s_w[16];
    const int tid = threadIdx.x;
    const int64_t per = (tiles + 1023) / 1024;
    const int64_t lo = (int64_t)tid * per, hi = lo + per < tiles ? lo + per : tiles;
    double acc = 0.0;
    for (int64_t t = hi - 1; t >= lo; t--)
        acc += tile_sum[t];
    double all;
    double run = block_suffix_exclusive<16>(acc, s_w, all);
    if (tid == 0)
        *total = all;
    for (int64_t t = hi - 1; t >= lo; t--) {
        const double c = tile_sum[t];
        tile_sum[t] = run;
        run += c;
    }
}

// Each tile again: its threads' partials suffix-scanned across the block,
// then each thread walks its items from the last, so item i sees
// E_i = sum_{j > i} 1 / r_j; psi_i in the reference's operation order.
template <typename S>
__global__ __launch_bounds__(kScanThreads) void k_psi_apply(const S *__restrict__ r, int64_t n,
                                                            const double *__restrict__ tile_after, double dm,
                                                            double g, double vcs, double *__restrict__ psi)
{
    __shared__ double s_w[kScanWaves];
    // radii, then psi, of the tile; one pad slot per 16 so that the blocked
    // access (stride kScanItems) does not fall on one bank
    __shared__ double s_r[kScanTile + kScanTile / 16];
    auto at = [](int j) { return j + (j >> 4); };
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    // coalesced in, then each thread takes kScanItems consecutive items
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t i = base + (int64_t)k * kScanThreads + tid;
        s_r[at(k * kScanThreads + tid)] = i < n ? (double)r[i] : 1.0;
    }
    __syncthreads();
    const int own = tid * kScanItems;
    double inv[kScanItems];
    double acc = 0.0;
    for (int k = kScanItems - 1; k >= 0; k--) {
        inv[k] = base + own + k < n ? 1.0 / s_r[at(own + k)] : 0.0;
        acc += inv[k];
    }
    double total;
    double after = tile_after[blockIdx.x] + block_suffix_exclusive<kScanWaves>(acc, s_w, total);
    for (int k = kScanItems - 1; k >= 0; k--) {
        const int64_t i = base + own + k;
        const double M_r = (double)(i + 1) * dm;
        const double outer = dm * after;  // 0 for the last rank
        double v = g * (M_r / s_r[at(own + k)] + outer);
        v *= vcs;
        s_r[at(own + k)] = v;  // only this thread reads or writes its own slots here
        after += inv[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int64_t i = base + (int64_t)k * kScanThreads + tid;
        if (i < n)
            psi[i] = s_r[at(k * kScanThreads + tid)];
    }
}

} // namespace

int64_t mass_scan_words(int64_t n) { return 2 * ((n + kScanTile - 1) / kScanTile) + 2; }

template <typename M>
hipError_t launch_mass_scan(MassScan &w, const M *mass_in, const uint32_t *perm, M *mass_out, double *menc,
                            int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipError_t e;
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_begin, dim3(1), dim3(1024), 0, stream, w.ctl, w.words, mass_scan_words(n));
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL((k_mass_scan<M>), dim3((unsigned)tiles), dim3(kScanThreads), 0, stream, mass_in, perm,
                       mass_out, menc, n, w.scale, w.words, w.ctl);
    return hipGetLastError();
}

template hipError_t launch_mass_scan<double>(MassScan &, const double *, const uint32_t *, double *, double *,
                                             int64_t, hipStream_t);
hipError_t launch_potential(const void *r, bool f32, int64_t n, double dm, double g, double vcs, double *psi,
                            double *work, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    double *tile_sum = work, *total = work + tiles;
    if (f32)
        hipLaunchKernelGGL(k_psi_tiles<float>, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream,
                           static_cast<const float *>(r), n, tile_sum);
    else
        hipLaunchKernelGGL(k_psi_tiles<double>, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream,
                           static_cast<const double *>(r), n, tile_sum);
    hipLaunchKernelGGL(k_psi_top, dim3(1), dim3(1024), 0, stream, tile_sum, tiles, total);
    if (!psi)  // the total alone
        return hipGetLastError();
    if (f32)
        hipLaunchKernelGGL(k_psi_apply<float>, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream,
                           static_cast<const float *>(r), n, tile_sum, dm, g, vcs, psi);
    else
        hipLaunchKernelGGL(k_psi_apply<double>, dim3((unsigned)tiles), dim3(kScanThreads), 0, stream,
                           static_cast<const double *>(r), n, tile_sum, dm, g, vcs, psi);
    return hipGetLastError();
}

int64_t potential_work_doubles(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

template hipError_t launch_mass_scan<float>(MassScan &, const float *, const uint32_t *, float *, double *, int64_t,
                                            hipStream_t);

} // namespace ns
