// ns_rank.hip -- the Rank_Mass_Rad_VRad files of a snapshot
// (nsphere.c:7043-7231, 7698-7850, 7911-7989).
//
// The reference post-processes a snapshot's records (id order: rank, float32
// r, v_r, ell) into two binary files written by fprintf_bin (int32 for %d,
// float32 for every %f):
//   unsorted, one row per id:       rank, (rank + 1) dM, r, v_r, PsiA(r), E, ell
//   sorted by r (ties in id order): i, (i + 1) dM, r, v_r, PsiA(r), E, ell, density
// where PsiA is the method-A potential of the sorted radii (suffix sum of
// 1/r) put on a linear spline through (0, PsiA0), (r_i, psi_i) whose nodes
// are forced strictly increasing by x_i = x_{i-1} + 1e-5 wherever r_i does
// not exceed the previous node (nsphere.c:7760-7769), E = PsiA - ell^2 /
// (2 r^2) - v_r^2 / 2, and density the snapshot density estimate
// (ns_density.hip).  Here the records, the stable radius sort, the suffix
// scan, the spline lookups, the energies, the density and the packing of the
// rows run on the device; the node recurrence (a sequential chain: a run of
// forced nodes can carry far past the radius that caused it) runs on the host.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "nsphere_gpu.h"
#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

namespace {

constexpr int kRankThreads = 256;

// Sorted columns: record perm[i] (records in id order).
__global__ void k_rank_gather(const int4 *__restrict__ rec, const uint32_t *__restrict__ perm, int64_t n,
                              float *__restrict__ rs, float *__restrict__ vs, float *__restrict__ ls)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int4 x = rec[perm[i]];
    rs[i] = __int_as_float(x.y);
    vs[i] = __int_as_float(x.z);
    ls[i] = __int_as_float(x.w);
}

__global__ void k_rank_keys(const int4 *__restrict__ rec, int64_t n, uint32_t *__restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        keys[i] = radius_key(__int_as_float(rec[i].y));
}

// The linear spline through the nodes (x[0..m-1], y) at t: the interval
// x[lo] <= t < x[lo + 1] by binary search (the last node in the last
// interval), then y_lo + (t - x_lo) / dx (y_hi - y_lo).
__device__ inline double linear_eval(const double *__restrict__ x, const double *__restrict__ y, int64_t m, double t)
{
    int64_t lo = 0, hi = m - 1;
    while (hi > lo + 1) {
        const int64_t mid = (hi + lo) >> 1;
        if (x[mid] > t)
            hi = mid;
        else
            lo = mid;
    }
    const double dx = x[lo + 1] - x[lo];
    return dx > 0.0 ? y[lo] + (t - x[lo]) / dx * (y[lo + 1] - y[lo]) : 0.0;
}

__device__ inline void put_f(unsigned char *row, int k, double v)
{
    const float f = (float)v;
    memcpy(row + 4 * k, &f, 4);
}

// Unsorted rows (id order, 28 bytes) and sorted rows (32 bytes).
__global__ void k_rank_rows(const int4 *__restrict__ rec, const float *__restrict__ rs, const float *__restrict__ vs,
                            const float *__restrict__ ls, const double *__restrict__ dens, const double *__restrict__ x,
                            const double *__restrict__ y, int64_t n, double dm, unsigned char *__restrict__ un,
                            unsigned char *__restrict__ so)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (un) {
        const int4 q = rec[i];
        const double rr = (double)__int_as_float(q.y), vr = (double)__int_as_float(q.z),
                     l = (double)__int_as_float(q.w);
        const double psi = linear_eval(x, y, n + 1, rr);
        const double e = psi - (l * l / (2.0 * rr * rr)) - 0.5 * vr * vr;
        unsigned char *row = un + 28 * i;
        memcpy(row, &q.x, 4);
        put_f(row, 1, (double)(q.x + 1) * dm);
        put_f(row, 2, rr);
        put_f(row, 3, vr);
        put_f(row, 4, psi);
        put_f(row, 5, e);
        put_f(row, 6, l);
    }
    if (so) {
        const double rr = (double)rs[i], vr = (double)vs[i], l = (double)ls[i];
        const double psi = linear_eval(x, y, n + 1, rr);
        const double e = psi - (l * l / (2.0 * rr * rr)) - 0.5 * vr * vr;
        unsigned char *row = so + 32 * i;
        const int rank = (int)i;
        memcpy(row, &rank, 4);
        put_f(row, 1, (double)(rank + 1) * dm);
        put_f(row, 2, rr);
        put_f(row, 3, vr);
        put_f(row, 4, psi);
        put_f(row, 5, e);
        put_f(row, 6, l);
        put_f(row, 7, dens[i]);
    }
}

} // namespace

hipError_t rank_files(SortWork &sw, const void *r, const void *v, const void *l, bool f32, const uint32_t *id,
                      int64_t n, double G, double vcs, double halo_mass, const char *unsorted_path,
                      const char *sorted_path, hipStream_t stream, int *io_failed)
{
    *io_failed = 0;
    if (n <= 0)
        return hipSuccess;
    hipError_t e = hipSuccess;
    std::vector<void *> bufs;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (e == hipSuccess)
            e = hipMalloc(&p, bytes);
        if (p)
            bufs.push_back(p);
        return p;
    };
    auto done = [&]() {
        for (void *p : bufs)
            (void)hipFree(p);
        return e;
    };
    const double dm = halo_mass / (double)n;  // deltaM, nsphere.c:5179
    int4 *rec = (int4 *)alloc((size_t)n * 16);
    uint32_t *perm = (uint32_t *)alloc((size_t)n * 4);
    float *rs = (float *)alloc((size_t)n * 4), *vs = (float *)alloc((size_t)n * 4), *ls = (float *)alloc((size_t)n * 4);
    double *psi = (double *)alloc((size_t)n * 8);
    double *work = (double *)alloc((size_t)potential_work_doubles(n) * 8);
    double *x = (double *)alloc((size_t)(n + 1) * 8), *y = (double *)alloc((size_t)(n + 1) * 8);
    double *dens = (double *)alloc((size_t)n * 8);
    unsigned char *un = unsorted_path ? (unsigned char *)alloc((size_t)n * 28) : nullptr;
    unsigned char *so = sorted_path ? (unsigned char *)alloc((size_t)n * 32) : nullptr;
    if (e != hipSuccess)
        return done();
    const dim3 grid((unsigned)((n + kRankThreads - 1) / kRankThreads));
    // records in id order, then the stable sort of their float32 radii (ties
    // stay in id order, as the reference's merge-sorting qsort leaves them)
    if ((e = launch_apd_records(r, v, l, f32, id, n, rec, stream)) != hipSuccess)
        return done();
    hipLaunchKernelGGL(k_rank_keys, grid, dim3(kRankThreads), 0, stream, rec, n,
                       reinterpret_cast<uint32_t *>(sw.keys[0]));
    if ((e = hipGetLastError()) != hipSuccess)
        return done();
    if ((e = radix_sort_perm(sw, n, 32, perm, stream, nullptr)) != hipSuccess)
        return done();
    hipLaunchKernelGGL(k_rank_gather, grid, dim3(kRankThreads), 0, stream, rec, perm, n, rs, vs, ls);
    // PsiA of the sorted radii; the density estimate of the same radii
    if ((e = hipGetLastError()) != hipSuccess || (e = launch_potential(rs, true, n, dm, G, vcs, psi, work, stream)) !=
                                                     hipSuccess)
        return done();
    if (sorted_path) {
        if ((e = density_estimate(rs, true, n, halo_mass, dens, nullptr, stream)) != hipSuccess)
            return done();
    }
    // spline nodes: (0, PsiA0), then the sorted radii forced strictly increasing
    std::vector<float> hr((size_t)n);
    std::vector<double> hx((size_t)n + 1), hy((size_t)n + 1);
    double total = 0.0;
    e = hipMemcpyAsync(hr.data(), rs, (size_t)n * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hy.data() + 1, psi, (size_t)n * 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&total, work + potential_work_doubles(n) - 1, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess)
        return done();
    hy[0] = G * dm * total * vcs;  // PsiA0, nsphere.c:7733-7738
    hx[0] = 0.0;
    for (int64_t i = 0; i < n; i++)
        hx[(size_t)i + 1] = (double)hr[(size_t)i];
    for (int64_t i = 1; i <= n; i++)
        if (hx[(size_t)i] <= hx[(size_t)i - 1])
            hx[(size_t)i] = hx[(size_t)i - 1] + 1e-5;
    e = hipMemcpyAsync(x, hx.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(y, hy.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess)
        return done();
    hipLaunchKernelGGL(k_rank_rows, grid, dim3(kRankThreads), 0, stream, rec, rs, vs, ls, dens, x, y, n, dm, un, so);
    if ((e = hipGetLastError()) != hipSuccess)
        return done();
    // rows to the files
    for (int which = 0; which < 2; which++) {
        const char *path = which ? sorted_path : unsorted_path;
        unsigned char *rows = which ? so : un;
        if (!path)
            continue;
        const size_t bytes = (size_t)n * (which ? 32 : 28);
        std::vector<unsigned char> h(bytes);
        if ((e = hipMemcpyAsync(h.data(), rows, bytes, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
            return done();
        FILE *f = fopen(path, "wb");
        if (!f || fwrite(h.data(), 1, bytes, f) != bytes) {
            if (f)
                fclose(f);
            *io_failed = 1;
            return done();
        }
        if (fclose(f) != 0) {
            *io_failed = 1;
            return done();
        }
    }
    return done();
}

} // namespace ns
