// ns_integrate.hip -- fused gather + force + adaptive leapfrog / Levi-Civita
// kernel for the method-5 step (nsphere.c:5796-5838), plus the permutation
// gather of sort_particles, the inverse map (nsphere.c:5789-5794) and the
// float64 energy reduction.
//
// One thread per rank i.  The thread gathers its particle through the sort
// permutation (the state of the previous step is nearly sorted, so the gather
// is nearly sequential), writes ell and id to the new rank-ordered buffers,
// integrates (r, v_r) at frozen rank i, writes the result, the radius key of
// the new r (input of the next step's sort) and a one-byte level record.
//
// Arithmetic follows the reference's operation order exactly (the library is
// compiled with -ffp-contract=off; division and sqrt are IEEE on gfx950), so
// this kernel reproduces the strict-IEEE CPU restatement bit for bit.  Two
// exact rewrites remove redundant work without changing a single result:
//   * the fine pass at level k (4N+1 micro-steps) is the coarse pass at level
//     k+1 (2(2N)+1 micro-steps): same step sizes, same operations, so each
//     level costs one new pass (nsphere.c:1670-1716);
//   * in a Levi-Civita micro-step the force at rho_next is the next
//     micro-step's force at rho_curr (nsphere.c:2008-2031), so it is
//     evaluated once.
#include <hip/hip_runtime.h>

#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

namespace {

struct Phys {
    double gm;  // -(vel_conv_sq G) (i/npts) M : gravity numerator at frozen rank
    double l2;  // ell * ell
    bool idg;   // identity gravity
};

// gravitational_force + effective_angular_force, nsphere.c:502-535.
__device__ inline double accel(const Phys &ph, double r)
{
    const double r2 = r * r;
    const double f = ph.idg ? 0.0 : ph.gm / r2;
    return f + ph.l2 / (r2 * r);
}

// forceLCfun, nsphere.c:1777-1787 with nsphere.c:563-599.
__device__ inline double force_lc(const Phys &ph, double rho)
{
    const double rr = rho * rho;
    const double g = ph.idg ? 0.0 : ph.gm / rr;
    return g + ph.l2 / (rr * rho * rho);
}

struct Budget {
    int64_t left;  // force evaluations still allowed for this particle
    bool hit;
    __device__ inline bool take(int64_t n)
    {
        if (n > left) {
            hit = true;
            return false;
        }
        left -= n;
        return true;
    }
};

// doMicroLeapfrog, nsphere.c:1550-1607, for N subdivisions; fine = the
// (4N+1)-step form.
__device__ inline void micro_leap(const Phys &ph, double r_in, double v_in, double h, int64_t N,
                                  bool fine, double &r_out, double &v_out)
{
    const double Nd = (double)N;
    const double hk = fine ? h / (4.0 * Nd) : h / (2.0 * Nd);
    const double ms = fine ? h / (2.0 * Nd) : h / (1.0 * Nd);
    const int64_t pairs = fine ? 2 * N : N;
    double r = r_in, v = v_in;
    v += hk * accel(ph, r);
    for (int64_t pp = 1; pp <= pairs - 1; pp++) {
        r += ms * v;
        v += ms * accel(ph, r);
    }
    r += ms * v;
    v += hk * accel(ph, r);
    r_out = r;
    v_out = v;
}

__device__ inline void leap_output(int out_type, double rc, double vc, double rf, double vf,
                                   double &r_out, double &v_out)
{
    if (out_type == 0) {
        r_out = rc;
        v_out = vc;
    } else if (out_type == 1) {
        r_out = rf;
        v_out = vf;
    } else {
        r_out = 4.0 * rf - 3.0 * rc;
        v_out = 4.0 * vf - 3.0 * vc;
    }
}

__device__ inline int level_of(int64_t N)
{
    int k = 63 - __clzll((unsigned long long)N);  // N is a power of two
    return k > NS_LVL_MASK ? NS_LVL_MASK : k;
}

// doAdaptiveFullLeap, nsphere.c:1648-1735.
__device__ int adaptive_leap(const Phys &ph, const StepArgs &a, double r_in, double v_in,
                             double &r_out, double &v_out, Budget &bud)
{
    const double h = a.dt;
    double rc = r_in, vc = v_in, rf = r_in, vf = v_in;
    int64_t N = 1;
    int last = 0;
    if (N <= a.max_subdiv) {
        if (!bud.take(2)) {
            r_out = r_in;
            v_out = v_in;
            return NS_LVL_NOCONV | NS_LVL_MASK;
        }
        micro_leap(ph, r_in, v_in, h, N, false, rc, vc);
    }
    while (N <= a.max_subdiv) {
        if (!bud.take(2 * N + 1)) {
            leap_output(a.out_type, rc, vc, rf, vf, r_out, v_out);
            return last | NS_LVL_NOCONV;
        }
        micro_leap(ph, r_in, v_in, h, N, true, rf, vf);
        last = level_of(N);
        const double rd = fabs(rf - rc) / (fabs(rf) + 1.0e-30);
        const double vd = fabs(vf - vc) / (fabs(vf) + 1.0e-30);
        if ((rd < a.radius_tol) && (vd < a.velocity_tol)) {
            leap_output(a.out_type, rc, vc, rf, vf, r_out, v_out);
            return last;
        }
        N *= 2;
        if (N <= a.max_subdiv) {  // fine(N) is coarse(2N)
            rc = rf;
            vc = vf;
        }
    }
    leap_output(a.out_type, rc, vc, rf, vf, r_out, v_out);
    return last | NS_LVL_NOCONV;
}

// doMicroLeviCivita, nsphere.c:1990-2036 (force at rho_next reused).
__device__ inline void micro_lc(const Phys &ph, double rho_in, double v_in, double t_in,
                                int64_t sub, double h_tau, double &rho_out, double &v_out,
                                double &t_out)
{
    const double dtau = h_tau / (double)sub;
    double rho = rho_in, v = v_in, t = t_in;
    double f = force_lc(ph, rho);
    for (int64_t ss = 0; ss < sub; ss++) {
        const double v_half = v + 0.5 * dtau * f;
        const double rho_next = rho + dtau * (0.5 * rho * v_half);
        const double f2 = force_lc(ph, rho_next);
        const double v_next = v_half + 0.5 * dtau * f2;
        const double rho_mid = 0.5 * (rho + rho_next);
        const double dt_phys = dtau * (rho_mid * rho_mid);
        t += dt_phys;
        rho = rho_next;
        v = v_next;
        f = f2;
    }
    rho_out = rho;
    v_out = v;
    t_out = t;
}

// doSingleTauStepAdaptiveLeviCivita, nsphere.c:2077-2186.  The fine pass
// (4N+1) of level k is the coarse pass (2(2N)+1) of level k+1.
__device__ int single_tau_lc(const Phys &ph, const StepArgs &a, double rho_in, double v_in,
                             double t_in, double h, double &rho_out, double &v_out, double &t_out,
                             Budget &bud)
{
    int64_t N = 1;
    int last = 0;
    double rhoC = 0, vC = 0, tC = 0, rhoF = 0, vF = 0, tF = 0;
    if (N <= a.max_subdiv) {
        if (!bud.take(2 * N + 2)) {
            rho_out = rho_in;
            v_out = v_in;
            t_out = t_in;
            return NS_LVL_NOCONV | NS_LVL_MASK;
        }
        micro_lc(ph, rho_in, v_in, t_in, 2 * N + 1, h, rhoC, vC, tC);
    }
    while (N <= a.max_subdiv) {
        if (!bud.take(4 * N + 2)) {
            rho_out = rhoC;
            v_out = vC;
            t_out = tC;
            return last | NS_LVL_NOCONV;
        }
        micro_lc(ph, rho_in, v_in, t_in, 4 * N + 1, h, rhoF, vF, tF);
        last = level_of(N);
        const double rF = rhoF * rhoF;
        const double rhodif = fabs(rhoF - rhoC) / (fabs(rF) + 1.0e-30);
        const double vdif = fabs(vF - vC) / (fabs(vF) + 1.0e-30);
        if ((rhodif < a.radius_tol) && (vdif < a.velocity_tol)) {
            if (a.out_type == 0) {
                rho_out = rhoC;
                v_out = vC;
                t_out = tC;
            } else if (a.out_type == 1) {
                rho_out = rhoF;
                v_out = vF;
                t_out = tF;
            } else {
                double rho_rich = 4.0 * rhoF - 3.0 * rhoC;
                const double v_rich = 4.0 * vF - 3.0 * vC;
                const double t_rich = 4.0 * tF - 3.0 * tC;
                if (rho_rich < 0.0)
                    rho_rich = 0.0;
                rho_out = rho_rich;
                v_out = v_rich;
                t_out = t_rich;
            }
            return last;
        }
        N *= 2;
        if (N <= a.max_subdiv) {
            rhoC = rhoF;
            vC = vF;
            tC = tF;
        }
    }
    // Not converged: the fine pass at the overflowed N.
    if (!bud.take(4 * N + 2)) {
        rho_out = rhoF;
        v_out = vF;
        t_out = tF;
        return last | NS_LVL_NOCONV;
    }
    micro_lc(ph, rho_in, v_in, t_in, 4 * N + 1, h, rho_out, v_out, t_out);
    return level_of(N) | NS_LVL_NOCONV;
}

// doAdaptiveFullLeviCivita, nsphere.c:2231-2334.
__device__ int adaptive_lc(const Phys &ph, const StepArgs &a, double r_in, double v_in,
                           double &r_out, double &v_out, Budget &bud)
{
    int rec = NS_LVL_LC;
    if (r_in < 1.0e-30) {
        r_out = r_in;
        v_out = v_in;
        return rec;
    }
    const double dt = a.dt;
    double rho_current = sqrt(r_in);
    double v_current = v_in;
    double t_cur = 0.0;
    double delta_tau;
    if (r_in > 1.0e-30 && a.N_taumin > 0)
        delta_tau = dt / (2.0 * r_in * (double)a.N_taumin);
    else
        delta_tau = dt / 100.0;
    int64_t step_count = 0;
    const int64_t step_max = 100000000;
    while (true) {
        if (t_cur >= dt)
            break;
        double rho_next, v_next, t_next;
        const int lv = single_tau_lc(ph, a, rho_current, v_current, t_cur, delta_tau, rho_next,
                                     v_next, t_next, bud);
        if ((lv & NS_LVL_MASK) > (rec & NS_LVL_MASK))
            rec = (rec & ~NS_LVL_MASK) | (lv & NS_LVL_MASK);
        rec |= (lv & NS_LVL_NOCONV);
        if (t_next > dt) {
            double alpha = 0.0;
            if (fabs(t_next - t_cur) > 1.0e-30)
                alpha = (dt - t_cur) / (t_next - t_cur);
            const double rho_final = rho_current + alpha * (rho_next - rho_current);
            const double v_final = v_current + alpha * (v_next - v_current);
            r_out = rho_final * rho_final;
            v_out = v_final;
            return rec;
        }
        rho_current = rho_next;
        v_current = v_next;
        t_cur = t_next;
        step_count++;
        if (step_count > step_max || bud.hit) {
            r_out = rho_current * rho_current;
            v_out = v_current;
            return rec | NS_LVL_NOCONV;
        }
    }
    r_out = rho_current * rho_current;
    v_out = v_current;
    return rec;
}

template <bool GATHER>
__global__ __launch_bounds__(256) void k_integrate_f64(StepArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    const int64_t src = GATHER ? (int64_t)a.perm[i] : i;
    const double r = a.r0[src];
    const double v = a.v0[src];
    const double ell = a.l0[src];
    if (GATHER) {
        a.l1[i] = ell;
        a.id1[i] = a.id0[src];
    }

    // r_crit switch, nsphere.c:5803-5828.
    double r_crit = 0.0;
    if (fabs(ell) > 1.0e-30) {
        double M_enc;
        if (i == 0)
            M_enc = 0.1 * (1.0 / (double)a.n) * a.halo_mass;
        else
            M_enc = ((double)i / (double)a.n) * a.halo_mass;
        const double grav_part = (a.vel_conv_sq * a.G) * M_enc;
        r_crit = (ell * ell) * a.alpha / grav_part;
    }
    Phys ph;
    ph.idg = a.identity_gravity != 0;
    ph.gm = -(a.vel_conv_sq * a.G) * ((double)i / (double)a.n) * a.halo_mass;
    ph.l2 = ell * ell;
    Budget bud{a.budget, false};
    double r_new, v_new;
    int rec;
    if (r > 1.0e-30 && r < r_crit)
        rec = adaptive_lc(ph, a, r, v, r_new, v_new, bud);
    else
        rec = adaptive_leap(ph, a, r, v, r_new, v_new, bud);
    if (bud.hit)
        atomicAdd(a.budget_hits, 1ull);
    a.r1[i] = r_new;
    a.v1[i] = v_new;
    if (a.keys)
        a.keys[i] = radius_key(r_new);
    if (a.lvl)
        a.lvl[i] = (uint8_t)rec;
}

template <typename T>
__global__ void k_gather(const T *__restrict__ src, const uint32_t *__restrict__ perm, T *__restrict__ dst,
                         int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[perm[i]];
}

__global__ void k_gather_state(const double *__restrict__ r0, const double *__restrict__ v0,
                               const double *__restrict__ l0, const uint32_t *__restrict__ id0,
                               const uint32_t *__restrict__ perm, double *__restrict__ r1,
                               double *__restrict__ v1, double *__restrict__ l1,
                               uint32_t *__restrict__ id1, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t p = perm[i];
        r1[i] = r0[p];
        v1[i] = v0[p];
        l1[i] = l0[p];
        id1[i] = id0[p];
    }
}

__global__ void k_inverse_map(const uint32_t *__restrict__ id, int32_t *__restrict__ inv, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t k = id[i];
        if (k < (uint64_t)n)
            inv[k] = (int32_t)i;
    }
}

constexpr int kEnergyThreads = 256;

// Per-block float64 partial sums of K and W in a fixed tree order.  Rank i
// is the position in radius order: particle perm[i] (perm may be null when
// the arrays are already sorted).
__global__ __launch_bounds__(kEnergyThreads) void k_energy_partial(
    const double *__restrict__ r, const double *__restrict__ v, const double *__restrict__ ell,
    const uint32_t *__restrict__ perm, int64_t n, double m, double gpre, double halo_mass, int idg,
    double *__restrict__ partial)
{
    __shared__ double sk[kEnergyThreads], sw[kEnergyThreads];
    double k = 0.0, w = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kEnergyThreads;
    for (int64_t i = (int64_t)blockIdx.x * kEnergyThreads + threadIdx.x; i < n; i += stride) {
        const int64_t p = perm ? (int64_t)perm[i] : i;
        const double rr = r[p], vv = v[p], ll = ell[p];
        k += m * 0.5 * (vv * vv + (ll * ll) / (rr * rr));
        if (!idg) {
            const double menc = ((double)i / (double)n) * halo_mass;
            w -= gpre * menc * m / rr;
        }
    }
    sk[threadIdx.x] = k;
    sw[threadIdx.x] = w;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sk[threadIdx.x] += sk[threadIdx.x + s];
            sw[threadIdx.x] += sw[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sk[0];
        partial[2 * blockIdx.x + 1] = sw[0];
    }
}

__global__ void k_energy_final(const double *__restrict__ partial, int nblocks, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double k = 0.0, w = 0.0;
        for (int b = 0; b < nblocks; b++) {
            k += partial[2 * b];
            w += partial[2 * b + 1];
        }
        out[0] = k;
        out[1] = w;
    }
}

} // namespace

hipError_t launch_integrate(const StepArgs &a, bool gather, hipStream_t stream)
{
    if (a.n <= 0)
        return hipSuccess;
    const int threads = 256;
    const dim3 grid((unsigned)((a.n + threads - 1) / threads));
    if (gather)
        hipLaunchKernelGGL(k_integrate_f64<true>, grid, dim3(threads), 0, stream, a);
    else
        hipLaunchKernelGGL(k_integrate_f64<false>, grid, dim3(threads), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_gather_state(const double *r0, const double *v0, const double *l0, const uint32_t *id0,
                               const uint32_t *perm, double *r1, double *v1, double *l1, uint32_t *id1,
                               int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    const int threads = 256;
    hipLaunchKernelGGL(k_gather_state, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                       stream, r0, v0, l0, id0, perm, r1, v1, l1, id1, n);
    return hipGetLastError();
}

hipError_t launch_inverse_map(const uint32_t *id, int32_t *inv, int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    const int threads = 256;
    hipLaunchKernelGGL(k_inverse_map, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                       stream, id, inv, n);
    return hipGetLastError();
}

int energy_blocks(int64_t n)
{
    int64_t b = (n + kEnergyThreads - 1) / kEnergyThreads;
    if (b > 1024)
        b = 1024;
    return b < 1 ? 1 : (int)b;
}

hipError_t launch_energy(const double *r, const double *v, const double *ell, const uint32_t *perm,
                         int64_t n, double m, double gpre, double halo_mass, int idg, double *partial,
                         double *out, hipStream_t stream)
{
    const int blocks = energy_blocks(n);
    hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(kEnergyThreads), 0, stream, r, v, ell, perm, n,
                       m, gpre, halo_mass, idg, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, stream, partial, blocks, out);
    return hipGetLastError();
}

} // namespace ns
