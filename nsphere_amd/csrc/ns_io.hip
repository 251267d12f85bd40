// ns_io.hip -- device side of the snapshot path.
//
// The reference's per-dtwrite block (nsphere.c:6452-6471) fills, for every
// original particle id k, the record (rank, (float) r, (float) v_r,
// (float) ell) of the particle that id k has become, with rank =
// inverse_map[k], the particle's position in the rank-ordered arrays; blocks
// of such snapshots go to all_particle_data<suffix>.dat as 16-byte records
// (int32 + 3 float32), snapshot-major (append_all_particle_data_chunk_to_file,
// nsphere.c:1295-1324).  Here one kernel writes the records of a snapshot
// straight into id order: position i of the resident arrays holds the
// particle of rank i, so record id[i] = (i, r[i], v[i], ell[i]) -- a coalesced
// read and one 16-byte store per particle, no inverse map needed.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_ic_sampler.h"
#include "ns_kernels.h"

namespace ns {

namespace {

constexpr int kIoThreads = 256;

template <typename S>
__global__ __launch_bounds__(kIoThreads) void k_apd_records(const S *__restrict__ r, const S *__restrict__ v,
                                                            const S *__restrict__ l, const uint32_t *__restrict__ id,
                                                            int64_t n, int4 *__restrict__ rec)
{
    const int64_t i = (int64_t)xcd_contiguous_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t k = id[i];
    if (k >= (uint64_t)n)  // ids are a permutation of 0..n-1; others have no record
        return;
    int4 out;
    out.x = (int)i;
    out.y = __float_as_int((float)r[i]);
    out.z = __float_as_int((float)v[i]);
    out.w = __float_as_int((float)l[i]);
    rec[k] = out;
}

// r, v_r, ell and position of the particles with the given ids (tracked
// particles), through the inverse map inv[id] = position.
template <typename S>
__global__ void k_gather_ids(const S *__restrict__ r, const S *__restrict__ v, const S *__restrict__ l,
                             const int32_t *__restrict__ inv, int64_t n, const uint32_t *__restrict__ ids, int k,
                             double *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k)
        return;
    const uint32_t q = ids[j];
    const int32_t p = q < (uint64_t)n ? inv[q] : -1;
    const bool ok = p >= 0 && p < n;
    out[4 * j + 0] = ok ? (double)r[p] : 0.0;
    out[4 * j + 1] = ok ? (double)v[p] : 0.0;
    out[4 * j + 2] = ok ? (double)l[p] : 0.0;
    out[4 * j + 3] = ok ? (double)p : -1.0;
}

// Seeded initial conditions drawn on the device (ns_ic_sampler.h, the same
// draws as the host generator).
__global__ __launch_bounds__(kIoThreads) void k_generate_ic(int kind, uint64_t seed, double rmax, double lowl_frac,
                                                            double lowl_scale, int64_t n, double *__restrict__ r,
                                                            double *__restrict__ v, double *__restrict__ l,
                                                            uint32_t *__restrict__ id)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    ic::sample_particle(kind, seed, (uint64_t)i, rmax, lowl_frac, lowl_scale, r[i], v[i], l[i]);
    id[i] = (uint32_t)i;
}

// An initial-conditions file's records (r, |v|, ell, id, mu per particle,
// read_initial_conditions, nsphere.c:2445-2487) into the state, with the
// reference's conversion after reading (nsphere.c:4908-4920):
// v_r = (|v| mu) kmsec_to_kpcmyr, ell = ell kmsec_to_kpcmyr, in that order.
__global__ __launch_bounds__(kIoThreads) void k_ic_records(const double *__restrict__ rec, int64_t n, double conv,
                                                           double *__restrict__ r, double *__restrict__ v,
                                                           double *__restrict__ l, uint32_t *__restrict__ id)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *p = rec + 5 * i;
    r[i] = p[0];
    double vr = p[1];
    vr *= p[4];
    vr *= conv;
    v[i] = vr;
    l[i] = p[2] * conv;
    id[i] = (uint32_t)p[3];
}

} // namespace

hipError_t launch_generate_ic(int kind, uint64_t seed, double rmax, double lowl_frac, double lowl_scale, int64_t n,
                              double *r, double *v, double *l, uint32_t *id, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_generate_ic, dim3((unsigned)((n + kIoThreads - 1) / kIoThreads)), dim3(kIoThreads), 0,
                       stream, kind, seed, rmax, lowl_frac, lowl_scale, n, r, v, l, id);
    return hipGetLastError();
}

hipError_t launch_ic_records(const double *rec, int64_t n, double conv, double *r, double *v, double *l,
                             uint32_t *id, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ic_records, dim3((unsigned)((n + kIoThreads - 1) / kIoThreads)), dim3(kIoThreads), 0,
                       stream, rec, n, conv, r, v, l, id);
    return hipGetLastError();
}

hipError_t launch_apd_records(const void *r, const void *v, const void *l, bool f32, const uint32_t *id, int64_t n,
                              void *rec, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    const dim3 grid((unsigned)((n + kIoThreads - 1) / kIoThreads));
    int4 *out = static_cast<int4 *>(rec);
    if (f32)
        hipLaunchKernelGGL(k_apd_records<float>, grid, dim3(kIoThreads), 0, stream, static_cast<const float *>(r),
                           static_cast<const float *>(v), static_cast<const float *>(l), id, n, out);
    else
        hipLaunchKernelGGL(k_apd_records<double>, grid, dim3(kIoThreads), 0, stream, static_cast<const double *>(r),
                           static_cast<const double *>(v), static_cast<const double *>(l), id, n, out);
    return hipGetLastError();
}

hipError_t launch_gather_ids(const void *r, const void *v, const void *l, bool f32, const int32_t *inv, int64_t n,
                             const uint32_t *ids, int k, double *out, hipStream_t stream)
{
    if (k <= 0)
        return hipSuccess;
    const dim3 grid((unsigned)((k + 63) / 64));
    if (f32)
        hipLaunchKernelGGL(k_gather_ids<float>, grid, dim3(64), 0, stream, static_cast<const float *>(r),
                           static_cast<const float *>(v), static_cast<const float *>(l), inv, n, ids, k, out);
    else
        hipLaunchKernelGGL(k_gather_ids<double>, grid, dim3(64), 0, stream, static_cast<const double *>(r),
                           static_cast<const double *>(v), static_cast<const double *>(l), inv, n, ids, k, out);
    return hipGetLastError();
}

} // namespace ns
