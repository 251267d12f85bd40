// ns_density.hip -- density estimate of the sorted radii (nsphere.c:7236-7686,
// with fft_gaussian_convolution, nsphere.c:2901-3060).
//
// The reference builds the density of a snapshot as follows:
//   1. radii sorted, enclosed mass M_i = (i + 1) dM; if some radius does not
//      exceed its predecessor, keep only the first of each run of equal
//      radii (the strictly increasing subsequence);
//   2. decimate to min(ceil(n / 100), 30000) points taken at round(i step);
//      nudge non-increasing decimated radii up by 1e-12;
//   3. a natural cubic spline of M(r) through them, evaluated on 131072
//      points uniform in log10 r from the first decimated radius to 1.04
//      times the last (the mass held constant outside the nodes);
//   4. dM/dr by central differences, floored at 1e-10, ends copied;
//   5. smoothed by a one-sided Gaussian of width 0.08 (n / 1e4)^-0.4 in
//      log10 r, normalised to unit sum, applied as a linear (zero-padded)
//      convolution result[i] = sum_{k <= i} kernel[k] density[i - k], floored
//      at 1e-10;
//   6. a natural cubic spline of the smoothed grid, evaluated at every radius
//      (end values outside the grid).
// Here the per-point work runs on the device: the strictly-increasing filter
// and its compaction, the decimation gather, the two spline evaluations (a
// binary search per point), the differences, the kernel and its sum, and the
// convolution -- directly, over the taps where the Gaussian is not below
// 1e-300 of its peak, which is the zero-padded linear convolution the
// reference computes with an FFT, without the FFT's round-off floor.  The two
// tridiagonal solves of the natural splines (30000 and 131072 unknowns, a
// sequential recurrence) and the grid of log-spaced radii run on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "nsphere_gpu.h"
#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

namespace {

constexpr int kDenThreads = 256;
constexpr int kGrid = 131072;
constexpr int kMaxDecimated = 30000;

template <typename S>
__device__ inline double rad(const S *r, int64_t i)
{
    return (double)r[i];
}

// flag[i] = 1 if radius i starts a new value (strictly above its predecessor)
template <typename S>
__global__ void k_den_flags(const S *__restrict__ r, int64_t n, uint32_t *__restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flag[i] = (i == 0 || rad(r, i) > rad(r, i - 1)) ? 1u : 0u;
}

// Exclusive scan of flags in tiles of 256 x 16: tile counts, one-block scan of
// the counts, then each tile's positions; kept items are written compacted.
__global__ __launch_bounds__(kDenThreads) void k_den_tile_counts(const uint32_t *__restrict__ flag, int64_t n,
                                                                 uint32_t *__restrict__ counts)
{
    __shared__ uint32_t s_sum[kDenThreads];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kDenThreads * 16;
    uint32_t c = 0;
    for (int k = 0; k < 16; k++) {
        const int64_t i = base + (int64_t)k * kDenThreads + tid;
        if (i < n)
            c += flag[i];
    }
    s_sum[tid] = c;
    __syncthreads();
    for (int w = kDenThreads / 2; w > 0; w >>= 1) {
        if (tid < w)
            s_sum[tid] += s_sum[tid + w];
        __syncthreads();
    }
    if (tid == 0)
        counts[blockIdx.x] = s_sum[0];
}

__global__ __launch_bounds__(1024) void k_den_scan_counts(uint32_t *__restrict__ counts, int64_t tiles,
                                                          uint32_t *__restrict__ total)
{
    __shared__ uint32_t s_part[1024];
    const int tid = threadIdx.x;
    const int64_t per = (tiles + 1023) / 1024;
    const int64_t lo = (int64_t)tid * per, hi = lo + per < tiles ? lo + per : tiles;
    uint32_t acc = 0;
    for (int64_t t = lo; t < hi; t++)
        acc += counts[t];
    s_part[tid] = acc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan (integers: exact)
        const uint32_t y = tid >= off ? s_part[tid - off] : 0u;
        __syncthreads();
        s_part[tid] += y;
        __syncthreads();
    }
    uint32_t run = s_part[tid] - acc;
    for (int64_t t = lo; t < hi; t++) {
        const uint32_t c = counts[t];
        counts[t] = run;
        run += c;
    }
    if (tid == 1023)
        *total = s_part[1023];
}

// Compaction: kept radius i (with its mass (i + 1) dM) goes to slot
// tile_offset + rank among the tile's kept items, in index order.
template <typename S>
__global__ __launch_bounds__(kDenThreads) void k_den_compact(const S *__restrict__ r, const uint32_t *__restrict__ flag,
                                                             int64_t n, const uint32_t *__restrict__ offsets,
                                                             double dm, double *__restrict__ r_out,
                                                             double *__restrict__ m_out)
{
    __shared__ uint32_t s_flag[kDenThreads * 16];
    __shared__ uint32_t s_sum[kDenThreads];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kDenThreads * 16;
    for (int k = 0; k < 16; k++) {
        const int64_t i = base + (int64_t)k * kDenThreads + tid;
        s_flag[k * kDenThreads + tid] = i < n ? flag[i] : 0u;
    }
    __syncthreads();
    // each thread takes 16 consecutive items
    uint32_t c = 0;
    for (int k = 0; k < 16; k++)
        c += s_flag[tid * 16 + k];
    s_sum[tid] = c;
    __syncthreads();
    for (int off = 1; off < kDenThreads; off <<= 1) {
        const uint32_t y = tid >= off ? s_sum[tid - off] : 0u;
        __syncthreads();
        s_sum[tid] += y;
        __syncthreads();
    }
    uint32_t slot = offsets[blockIdx.x] + s_sum[tid] - c;
    for (int k = 0; k < 16; k++) {
        const int64_t i = base + tid * 16 + k;
        if (i < n && s_flag[tid * 16 + k]) {
            r_out[slot] = rad(r, i);
            m_out[slot] = (double)(i + 1) * dm;
            slot++;
        }
    }
}

// Decimation: point k takes filtered index (int) round(k step), clamped.
__global__ void k_den_decimate(const double *__restrict__ rf, const double *__restrict__ mf, int64_t count, int dec,
                               double step, double *__restrict__ r_dec, double *__restrict__ m_dec)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dec)
        return;
    int64_t idx = (int64_t)round((double)k * step);
    if (idx >= count)
        idx = count - 1;
    r_dec[k] = rf[idx];
    m_dec[k] = mf[idx];
}

// Natural cubic spline through (x[0..m-1], y) with second-derivative
// coefficients c (c[0] = c[m-1] = 0), evaluated at t inside [x[0], x[m-1]]:
// the interval by binary search (x[lo] <= t < x[lo + 1]; the last node falls
// in the last interval), then y_lo + d (b + d (c_lo + d e)).
__device__ inline double spline_eval(const double *__restrict__ x, const double *__restrict__ y,
                                     const double *__restrict__ c, int m, double t)
{
    int lo = 0, hi = m - 1;
    while (hi > lo + 1) {
        const int mid = (hi + lo) >> 1;
        if (x[mid] > t)
            hi = mid;
        else
            lo = mid;
    }
    const double h = x[lo + 1] - x[lo], dy = y[lo + 1] - y[lo], delx = t - x[lo];
    const double ci = c[lo], cn = c[lo + 1];
    const double b = dy / h - h * (cn + 2.0 * ci) / 3.0;
    const double d = (cn - ci) / (3.0 * h);
    return y[lo] + delx * (b + delx * (ci + delx * d));
}

// Mass on the grid: constant outside the decimated nodes.
__global__ void k_den_mass_grid(const double *__restrict__ r_grid, const double *__restrict__ r_dec,
                                const double *__restrict__ m_dec, const double *__restrict__ c_dec, int dec,
                                double *__restrict__ mass_grid)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGrid)
        return;
    const double t = r_grid[i];
    double m;
    if (t < r_dec[0])
        m = m_dec[0];
    else if (t > r_dec[dec - 1])
        m = m_dec[dec - 1];
    else
        m = dec > 1 ? spline_eval(r_dec, m_dec, c_dec, dec, t) : m_dec[0];
    mass_grid[i] = m;
}

// dM/dr by central differences, floored; the ends copy their neighbours
// (a second launch of one thread handles the two ends after the interior).
__global__ void k_den_diff(const double *__restrict__ r_grid, const double *__restrict__ mass_grid,
                           double *__restrict__ den)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 1 || i >= kGrid - 1)
        return;
    const double dr = r_grid[i + 1] - r_grid[i - 1];
    double v;
    if (dr > 1e-15)
        v = (mass_grid[i + 1] - mass_grid[i - 1]) / dr;
    else
        v = 0.0;  // coincident grid points do not occur on a strictly increasing grid
    den[i] = v < 1e-10 ? 1e-10 : v;
}

__global__ void k_den_ends(double *__restrict__ den)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        den[0] = den[1];
        den[kGrid - 1] = den[kGrid - 2];
    }
}

// Gaussian kernel at x = k dlog, and its sum in a fixed order (one block).
__global__ void k_den_kernel(double dlog, double sigma, double *__restrict__ ker)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kGrid)
        return;
    const double x = (double)k * dlog;
    ker[k] = (1.0 / (sigma * sqrt(2.0 * 3.14159265358979323846))) * exp(-0.5 * (x / sigma) * (x / sigma));
}

__global__ __launch_bounds__(1024) void k_den_kernel_sum(const double *__restrict__ ker, double *__restrict__ sum)
{
    __shared__ double s_part[1024];
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int k = tid; k < kGrid; k += 1024)
        acc += ker[k];
    s_part[tid] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (tid < w)
            s_part[tid] += s_part[tid + w];
        __syncthreads();
    }
    if (tid == 0)
        *sum = s_part[0];
}

// result[i] = sum_{k = 0}^{min(i, taps - 1)} (ker[k] / norm) den[i - k],
// floored at 1e-10.  One wave per output point; lanes stride the taps and
// the wave sum is taken in a fixed order.
__global__ __launch_bounds__(kDenThreads) void k_den_convolve(const double *__restrict__ ker,
                                                              const double *__restrict__ sum,
                                                              const double *__restrict__ den, int taps,
                                                              double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= kGrid)
        return;
    const double norm = *sum;
    const int kmax = i < taps - 1 ? i : taps - 1;
    double acc = 0.0;
    for (int k = lane; k <= kmax; k += 64)
        acc += (ker[k] / norm) * den[i - k];
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off, 64);
    if (lane == 0)
        out[i] = acc < 1e-10 ? 1e-10 : acc;
}

// The smoothed density at every radius (end values outside the grid).
template <typename S>
__global__ void k_den_eval(const S *__restrict__ r, int64_t n, const double *__restrict__ r_grid,
                           const double *__restrict__ sm, const double *__restrict__ c_grid,
                           double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double t = rad(r, i);
    double v;
    if (t <= r_grid[0])
        v = sm[0];
    else if (t >= r_grid[kGrid - 1])
        v = sm[kGrid - 1];
    else
        v = spline_eval(r_grid, sm, c_grid, kGrid, t);
    out[i] = v;
}

// Second-derivative coefficients of the natural cubic spline through
// (x, y), m points: the symmetric tridiagonal system of the interior
// coefficients (diagonal 2 (h_i + h_{i+1}), off-diagonal h_{i+1}, right-hand
// side 3 (dy_{i+1} / h_{i+1} - dy_i / h_i)) solved by LDL^T elimination.
void natural_spline_coeffs(const std::vector<double> &x, const std::vector<double> &y, std::vector<double> &c)
{
    const size_t m = x.size();
    c.assign(m, 0.0);
    if (m < 3)
        return;
    const size_t k = m - 2;
    std::vector<double> diag(k), off(k), g(k);
    for (size_t i = 0; i < k; i++) {
        const double h_i = x[i + 1] - x[i], h_ip1 = x[i + 2] - x[i + 1];
        const double ydiff_i = y[i + 1] - y[i], ydiff_ip1 = y[i + 2] - y[i + 1];
        const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
        const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
        off[i] = h_ip1;
        diag[i] = 2.0 * (h_ip1 + h_i);
        g[i] = 3.0 * (ydiff_ip1 * g_ip1 - ydiff_i * g_i);
    }
    if (k == 1) {
        c[1] = g[0] / diag[0];
        return;
    }
    std::vector<double> gamma(k), alpha(k), z(k);
    alpha[0] = diag[0];
    gamma[0] = off[0] / alpha[0];
    for (size_t i = 1; i + 1 < k; i++) {
        alpha[i] = diag[i] - off[i - 1] * gamma[i - 1];
        gamma[i] = off[i] / alpha[i];
    }
    alpha[k - 1] = diag[k - 1] - off[k - 2] * gamma[k - 2];
    z[0] = g[0];
    for (size_t i = 1; i < k; i++)
        z[i] = g[i] - gamma[i - 1] * z[i - 1];
    std::vector<double> xs(k);
    xs[k - 1] = z[k - 1] / alpha[k - 1];
    for (size_t i = k - 1; i-- > 0;)
        xs[i] = z[i] / alpha[i] - gamma[i] * xs[i + 1];
    for (size_t i = 0; i < k; i++)
        c[i + 1] = xs[i];
}

} // namespace

int64_t density_grid_size() { return kGrid; }

hipError_t density_estimate(const void *r_dev, bool f32, int64_t n, double halo_mass, double *out,
                            ns_density_info *info, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipError_t e = hipSuccess;
    const double dm = halo_mass / (double)n;  // deltaM, nsphere.c:5179
    const int64_t tiles = (n + kDenThreads * 16 - 1) / (kDenThreads * 16);
    std::vector<void *> bufs;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (e == hipSuccess)
            e = hipMalloc(&p, bytes);
        if (p)
            bufs.push_back(p);
        return p;
    };
    uint32_t *flag = (uint32_t *)alloc((size_t)n * sizeof(uint32_t));
    uint32_t *counts = (uint32_t *)alloc((size_t)(tiles + 1) * sizeof(uint32_t));
    double *rf = (double *)alloc((size_t)n * sizeof(double));
    double *mf = (double *)alloc((size_t)n * sizeof(double));
    double *r_dec = (double *)alloc(kMaxDecimated * sizeof(double));
    double *m_dec = (double *)alloc(kMaxDecimated * sizeof(double));
    double *c_dec = (double *)alloc(kMaxDecimated * sizeof(double));
    double *r_grid = (double *)alloc(kGrid * sizeof(double));
    double *mass_grid = (double *)alloc(kGrid * sizeof(double));
    double *den = (double *)alloc(kGrid * sizeof(double));
    double *ker = (double *)alloc(kGrid * sizeof(double));
    double *sm = (double *)alloc(kGrid * sizeof(double));
    double *c_grid = (double *)alloc(kGrid * sizeof(double));
    double *ksum = (double *)alloc(sizeof(double));
    auto done = [&]() {
        for (void *p : bufs)
            (void)hipFree(p);
        return e;
    };
    if (e != hipSuccess)
        return done();
    const unsigned gn = (unsigned)((n + kDenThreads - 1) / kDenThreads);
    // 1. strictly increasing subsequence, compacted with its masses
    if (f32)
        hipLaunchKernelGGL(k_den_flags<float>, dim3(gn), dim3(kDenThreads), 0, stream, (const float *)r_dev, n, flag);
    else
        hipLaunchKernelGGL(k_den_flags<double>, dim3(gn), dim3(kDenThreads), 0, stream, (const double *)r_dev, n,
                           flag);
    hipLaunchKernelGGL(k_den_tile_counts, dim3((unsigned)tiles), dim3(kDenThreads), 0, stream, flag, n, counts);
    hipLaunchKernelGGL(k_den_scan_counts, dim3(1), dim3(1024), 0, stream, counts, tiles, counts + tiles);
    if (f32)
        hipLaunchKernelGGL(k_den_compact<float>, dim3((unsigned)tiles), dim3(kDenThreads), 0, stream,
                           (const float *)r_dev, flag, n, counts, dm, rf, mf);
    else
        hipLaunchKernelGGL(k_den_compact<double>, dim3((unsigned)tiles), dim3(kDenThreads), 0, stream,
                           (const double *)r_dev, flag, n, counts, dm, rf, mf);
    uint32_t kept = 0;
    if ((e = hipGetLastError()) == hipSuccess)
        e = hipMemcpyAsync(&kept, counts + tiles, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return done();
    // 2. decimation
    const int64_t count = kept;
    int dec = (int)std::ceil((double)n / 100.0);
    if (dec > kMaxDecimated)
        dec = kMaxDecimated;
    if (dec <= 0)
        dec = 1;
    if (count <= dec) {
        dec = (int)count;
        e = hipMemcpyAsync(r_dec, rf, (size_t)dec * sizeof(double), hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(m_dec, mf, (size_t)dec * sizeof(double), hipMemcpyDeviceToDevice, stream);
    } else {
        const double step = (double)(count - 1) / (dec - 1);
        hipLaunchKernelGGL(k_den_decimate, dim3((unsigned)((dec + 255) / 256)), dim3(256), 0, stream, rf, mf, count,
                           dec, step, r_dec, m_dec);
        e = hipGetLastError();
    }
    std::vector<double> hr(dec), hm(dec);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hr.data(), r_dec, (size_t)dec * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hm.data(), m_dec, (size_t)dec * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return done();
    for (int i = 1; i < dec; i++)
        if (hr[i] <= hr[i - 1])
            hr[i] = hr[i - 1] + 1e-12;
    // (a natural cubic spline needs three nodes: with fewer the reference's
    // spline allocation fails and the snapshot is skipped)
    if (dec < 3 || !(hr[0] > 0.0) || !(hr[dec - 1] > hr[0]) || !std::isfinite(hr[0]) ||
        !std::isfinite(hr[dec - 1])) {
        e = hipErrorInvalidValue;  // the reference skips the snapshot here
        return done();
    }
    // 3. the log-spaced grid (host libm, as the reference) and the mass spline
    const double min_r = hr[0], max_r = hr[dec - 1] * 1.04;
    const double log_min = log10(min_r), log_max = log10(max_r);
    const double dlog = (log_max - log_min) / (kGrid - 1);
    std::vector<double> hg(kGrid);
    for (int i = 0; i < kGrid; i++)
        hg[i] = pow(10.0, log_min + i * dlog);
    for (int i = 1; i < kGrid; i++)
        if (hg[i] <= hg[i - 1])
            hg[i] = hg[i - 1] + 1e-12;
    std::vector<double> hc;
    natural_spline_coeffs(hr, hm, hc);
    e = hipMemcpyAsync(r_dec, hr.data(), (size_t)dec * sizeof(double), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c_dec, hc.data(), (size_t)dec * sizeof(double), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r_grid, hg.data(), kGrid * sizeof(double), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess)
        return done();
    const unsigned gg = (unsigned)((kGrid + 255) / 256);
    hipLaunchKernelGGL(k_den_mass_grid, dim3(gg), dim3(256), 0, stream, r_grid, r_dec, m_dec, c_dec, dec, mass_grid);
    // 4. differences
    hipLaunchKernelGGL(k_den_diff, dim3(gg), dim3(256), 0, stream, r_grid, mass_grid, den);
    hipLaunchKernelGGL(k_den_ends, dim3(1), dim3(64), 0, stream, den);
    // 5. smoothing
    const double sigma = 0.08 * pow((double)n / 10000.0, -0.40);
    // taps past which the Gaussian is below 1e-300 of its peak: exp(-x^2/2) < 1e-300 for x > 37.2
    int taps = (int)std::ceil(37.2 * sigma / dlog) + 1;
    if (taps > kGrid)
        taps = kGrid;
    hipLaunchKernelGGL(k_den_kernel, dim3(gg), dim3(256), 0, stream, dlog, sigma, ker);
    hipLaunchKernelGGL(k_den_kernel_sum, dim3(1), dim3(1024), 0, stream, ker, ksum);
    hipLaunchKernelGGL(k_den_convolve, dim3((unsigned)((int64_t)kGrid * 64 / kDenThreads)), dim3(kDenThreads), 0,
                       stream, ker, ksum, den, taps, sm);
    // 6. the density spline and its values at the radii
    std::vector<double> hs(kGrid);
    if ((e = hipGetLastError()) == hipSuccess)
        e = hipMemcpyAsync(hs.data(), sm, kGrid * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return done();
    std::vector<double> hcg;
    natural_spline_coeffs(hg, hs, hcg);
    e = hipMemcpyAsync(c_grid, hcg.data(), kGrid * sizeof(double), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess)
        return done();
    if (f32)
        hipLaunchKernelGGL(k_den_eval<float>, dim3(gn), dim3(256), 0, stream, (const float *)r_dev, n, r_grid, sm,
                           c_grid, out);
    else
        hipLaunchKernelGGL(k_den_eval<double>, dim3(gn), dim3(256), 0, stream, (const double *)r_dev, n, r_grid, sm,
                           c_grid, out);
    if ((e = hipGetLastError()) == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (info) {
        info->kept = count;
        info->decimated = dec;
        info->grid = kGrid;
        info->taps = taps;
        info->sigma = sigma;
        info->r_min = min_r;
        info->r_max = max_r;
    }
    return done();
}

} // namespace ns
