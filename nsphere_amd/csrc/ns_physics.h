// ns_physics.h -- per-particle physics of the method-5 step, device side.
//
// Restates, for one particle at frozen rank, the reference's force helpers
// (nsphere.c:502-599, 1757-1787), doMicroLeapfrog / doAdaptiveFullLeap
// (nsphere.c:1550-1735) and the adaptive Levi-Civita family
// (nsphere.c:1990-2334).  Two arithmetic forms:
//   STRICT (FAST = false): the reference's operation order with IEEE division
//     and no contraction (the library is compiled with -ffp-contract=off), so
//     results equal the strict CPU restatement bit for bit;
//   FAST: one v_rcp_f64 per force evaluation refined by one Newton step, the
//     kick coefficient folded into the force constants, fused multiply-adds
//     and (Levi-Civita) rescaled state, so that a micro-step is seven (leap-
//     frog) or nine (Levi-Civita) FP64 operations -- a lone wave's chain runs
//     at about 3 ns per operation on MI355X, so the count sets the tail of
//     the step; same algorithm and control flow, results within a few tens
//     of ulp per force of the strict form.
// Exact rewrites used by both forms (no result changes):
//   * fine pass at level k (4N+1 micro-steps) == coarse pass at level k+1
//     (2(2N)+1 micro-steps): same step sizes and operations;
//   * a Levi-Civita micro-step's force at rho_next is the next micro-step's
//     force at rho_curr.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ns_device.h"
#include "ns_kernels.h"

namespace ns {

// Returned by a capped leapfrog loop that must continue in the deep queue.
constexpr int kEscalate = -1;

struct Phys {
    double gm;  // -(vel_conv_sq G) (i/npts) M : gravity numerator at frozen rank
    double l2;  // ell * ell
    bool idg;   // identity gravity (nsphere.c:476)
};

// The rank, particle count and total mass the force sees at global rank i:
// the whole array, or in an ensemble the halo holding rank i (a binary search
// over the halo offsets).
struct RankInfo {
    int64_t i, n;
    double M;
};

__device__ inline RankInfo rank_info(const StepArgs &a, int64_t i)
{
    if (a.nseg <= 0)
        return {i, a.n, a.halo_mass};
    int lo = 0, hi = a.nseg;  // seg_off[0] = 0 <= i < seg_off[nseg] = n
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.seg_off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return {i - a.seg_off[lo], a.seg_off[lo + 1] - a.seg_off[lo], a.seg_mass[lo]};
}

// Enclosed mass at rank i: the reference's (i / npts) M (nsphere.c:513), or
// with per-particle masses the scanned sum of the masses below rank i.
__device__ inline double enclosed_mass(const StepArgs &a, const RankInfo &k)
{
    return a.menc ? a.menc[k.i] : ((double)k.i / (double)k.n) * k.M;
}

// What a particle's force needs of its rank: its halo's rank and count, and
// the rank fraction i / npts of the reference's enclosed mass (one division,
// shared by the force constants and the r_crit switch).
struct RankMass {
    RankInfo k;
    double frac;  // (double) i / (double) npts
};

__device__ inline RankMass rank_mass(const StepArgs &a, int64_t i)
{
    const RankInfo k = rank_info(a, i);
    return {k, a.menc ? 0.0 : (double)k.i / (double)k.n};
}

__device__ inline Phys make_phys(const StepArgs &a, int64_t i, const RankMass &rm, double ell)
{
    Phys ph;
    ph.idg = a.identity_gravity != 0;
    // the reference's association, ((-(vel_conv_sq G)) (i / npts)) M
    ph.gm = a.menc ? -(a.vel_conv_sq * a.G) * a.menc[i] : -(a.vel_conv_sq * a.G) * rm.frac * rm.k.M;
    ph.l2 = ell * ell;
    return ph;
}

__device__ inline Phys make_phys(const StepArgs &a, int64_t i, double ell)
{
    return make_phys(a, i, rank_mass(a, i), ell);
}

// r_crit switch of nsphere.c:5803-5828.
__device__ inline bool takes_levi_civita(const StepArgs &a, const RankMass &rm, double r, double ell)
{
    double r_crit = 0.0;
    if (fabs(ell) > 1.0e-30) {
        // rank 0: a tenth of one particle's mass (with masses: of its own)
        double M_enc;
        if (rm.k.i == 0)
            M_enc = a.menc ? 0.1 * a.menc[a.n + 1] : 0.1 * (1.0 / (double)rm.k.n) * rm.k.M;
        else
            M_enc = a.menc ? a.menc[rm.k.i] : rm.frac * rm.k.M;  // enclosed_mass
        const double grav_part = (a.vel_conv_sq * a.G) * M_enc;
        r_crit = (ell * ell) * a.alpha / grav_part;
    }
    return r > 1.0e-30 && r < r_crit;
}

__device__ inline bool takes_levi_civita(const StepArgs &a, int64_t i, double r, double ell)
{
    return takes_levi_civita(a, rank_mass(a, i), r, ell);
}

// Whether a falling leapfrog particle's straight path (r + s v dt, at its
// enclosed mass now) reaches the Levi-Civita regime within kSoonSteps
// steps.  Leapfrog particles at level kSoonLevel or more that do turn into
// Levi-Civita chains of thousands of micro-steps a tau step on their way to
// pericentre (in a cusp); the deferral of the next step ranks them beside
// the deepest ones.
constexpr double kSoonSteps = 2.0;
constexpr int kSoonLevel = 6;
__device__ inline bool lc_soon(const StepArgs &a, const RankMass &rm, double r, double v, double ell)
{
    if (!(v < 0.0))
        return false;
    const double ahead = fma(kSoonSteps * a.dt, v, r);
    return ahead <= 1.0e-30 || takes_levi_civita(a, rm, ahead, ell);
}

// h / 2^e for the power-of-two divisors of the leapfrog passes: exact, as
// the division it replaces (no result here comes near the subnormals).
__device__ inline double div_pow2(double h, int e) { return ldexp(h, -e); }

// log2 of a power of two
__device__ inline int ilog2(int64_t N) { return 63 - __clzll((unsigned long long)N); }

// The reference's convergence test (nsphere.c:1700-1712): relative change
// of the fine over the coarse pass below both tolerances.  The fast form
// multiplies instead of dividing, which differs from it only where a
// quotient rounds across the tolerance.
template <bool FAST>
__device__ inline bool converged(const StepArgs &a, double rc, double vc, double rf, double vf)
{
    if (FAST)
        return fabs(rf - rc) < a.radius_tol * (fabs(rf) + 1.0e-30) &&
               fabs(vf - vc) < a.velocity_tol * (fabs(vf) + 1.0e-30);
    const double rd = fabs(rf - rc) / (fabs(rf) + 1.0e-30);
    const double vd = fabs(vf - vc) / (fabs(vf) + 1.0e-30);
    return (rd < a.radius_tol) && (vd < a.velocity_tol);
}

// Fast-form reciprocal: v_rcp_f64 (y = (1 + d)/x, |d| < 2^-24 on gfx950)
// and one Newton step, y + y (1 - x y): error d^2, within 11 ulp of the
// correctly rounded 1/x over 4M values spanning 2^-40..2^40.
__device__ inline double fast_rcp(double x)
{
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}

// gravitational_force + effective_angular_force (nsphere.c:502-535).
__device__ inline double accel_strict(const Phys &ph, double r)
{
    const double r2 = r * r;
    const double f = ph.idg ? 0.0 : ph.gm / r2;
    return f + ph.l2 / (r2 * r);
}

// forceLCfun (nsphere.c:1777-1787 with 563-599): -GM/rho^2 + ell^2/rho^4.
__device__ inline double force_lc_strict(const Phys &ph, double rho)
{
    const double rr = rho * rho;
    const double g = ph.idg ? 0.0 : ph.gm / rr;
    return g + ph.l2 / (rr * rho * rho);
}

// A kick of coefficient kk.  The fast form folds kk into the force
// constants: kk accel(r) = (lk / r + gk) / r^2, lk = kk ell^2, gk = kk gm
// (and kk forceLC(rho) = (lk / rho^2 + gk) / rho^2 with kk = dtau).
struct Kick {
    double kk, lk, gk;
};

__device__ inline Kick make_kick(const Phys &ph, double kk)
{
    return {kk, ph.l2 * kk, (ph.idg ? 0.0 : ph.gm) * kk};
}

template <bool FAST>
__device__ inline void kick(const Phys &ph, const Kick &c, double r, double &v)
{
    if constexpr (!FAST) {
        v += c.kk * accel_strict(ph, r);
    } else {
        const double inv = fast_rcp(r);
        v = fma(fma(c.lk, inv, c.gk), inv * inv, v);
    }
}

template <bool FAST>
__device__ inline void drift_kick(const Phys &ph, double ms, const Kick &c, double &r, double &v)
{
    if constexpr (!FAST) {
        r += ms * v;
        v += c.kk * accel_strict(ph, r);
    } else {
        r = fma(ms, v, r);
        kick<true>(ph, c, r, v);
    }
}

// n (the same on every active lane) calls of f, four per loop trip, with a
// scalar trip counter: a lone wave's chain pays for every loop instruction.
__device__ inline int64_t uniform_i64(int64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <typename F>
__device__ inline void repeat_uniform(int64_t n, F &&f)
{
    // The trip count is wave-uniform.  It runs down in 32-bit chunks so that
    // the loop test stays on the scalar unit: there is no 64-bit scalar
    // less-than compare, and a vector compare would cost the lone wave of a
    // long chain one more vector issue per pass.
    int64_t m = uniform_i64(n);
    while (m > 0) {
        uint32_t c = m > (int64_t)0x40000000 ? 0x40000000u : (uint32_t)m;
        m -= (int64_t)c;
        for (; c >= 4; c -= 4) {
            f();
            f();
            f();
            f();
        }
        for (; c > 0; c--)
            f();
    }
}

// Force-evaluation budget per particle and step: a guard against inputs on
// which the reference would loop for hours (NaN state, r -> 0 in the LC
// branch); the default is far above any converging particle's cost.
struct Budget {
    int64_t left;
    bool hit;
    int64_t chain = 0;  // micro-steps the wave Levi-Civita form ran one after another
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

// doMicroLeapfrog, nsphere.c:1550-1607: N subdivisions, fine = (4N+1) form.
template <bool FAST>
__device__ inline void micro_leap(const Phys &ph, double r_in, double v_in, double h, int64_t N, bool fine,
                                  double &r_out, double &v_out)
{
    // h / (4N), h / (2N) (fine) or h / (2N), h / N (coarse); N = 2^e
    const int e = ilog2(N) + (fine ? 1 : 0);
    const double hk = div_pow2(h, e + 1);
    const double ms = div_pow2(h, e);
    const int64_t pairs = fine ? 2 * N : N;
    const Kick kh = make_kick(ph, hk), km = make_kick(ph, ms);
    double r = r_in, v = v_in;
    kick<FAST>(ph, kh, r, v);
    for (int64_t pp = 1; pp <= pairs - 1; pp++)
        drift_kick<FAST>(ph, ms, km, r, v);
    drift_kick<FAST>(ph, ms, kh, r, v);
    r_out = r;
    v_out = v;
}

__device__ inline void leap_output(int out_type, double rc, double vc, double rf, double vf, double &r_out,
                                   double &v_out)
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
    const int k = 63 - __clzll((unsigned long long)N);  // N is a power of two
    return k > NS_LVL_MASK ? NS_LVL_MASK : k;
}

// doAdaptiveFullLeap, nsphere.c:1648-1735, one lane.  With cap_N > 0 the
// loop stops before a level whose N exceeds cap_N and returns kEscalate.
template <bool FAST>
__device__ int adaptive_leap(const Phys &ph, const StepArgs &a, double r_in, double v_in, double &r_out,
                             double &v_out, Budget &bud, int64_t cap_N)
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
        micro_leap<FAST>(ph, r_in, v_in, h, N, false, rc, vc);
    }
    while (N <= a.max_subdiv) {
        if (cap_N > 0 && N > cap_N)
            return kEscalate;
        if (!bud.take(2 * N + 1)) {
            leap_output(a.out_type, rc, vc, rf, vf, r_out, v_out);
            return last | NS_LVL_NOCONV;
        }
        micro_leap<FAST>(ph, r_in, v_in, h, N, true, rf, vf);
        last = level_of(N);
        if (converged<FAST>(a, rc, vc, rf, vf)) {
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

// doAdaptiveFullLeap, wave-cooperative form (all 64 lanes call it with the
// same particle).  Lane k runs the fine pass at N = 2^k (k = 0..kmax, kmax =
// floor(log2 max_subdiv)); lane 63 runs the coarse pass at N = 1.  All lanes
// step in lockstep, so after 2^(k+1) iterations lanes k-1 and k hold coarse
// and fine of level k and the reference's test for level k is evaluated; the
// first level that passes is returned, exactly as the sequential loop would
// return it, after 2^(K+1) iterations instead of about 2^(K+2).
template <bool FAST>
__device__ int adaptive_leap_wave(const Phys &ph, const StepArgs &a, double r_in, double v_in, double &r_out,
                                  double &v_out, Budget &bud)
{
    const int lane = threadIdx.x & 63;
    const double h = a.dt;
    if (a.max_subdiv < 1) {
        leap_output(a.out_type, r_in, v_in, r_in, v_in, r_out, v_out);
        return NS_LVL_NOCONV;
    }
    int kmax = 63 - __clzll((unsigned long long)a.max_subdiv);
    if (kmax > 62)
        kmax = 62;
    const bool coarse_lane = lane == 63;
    const bool active = coarse_lane || lane <= kmax;
    const int64_t N = coarse_lane ? 1 : ((int64_t)1 << (lane <= kmax ? lane : 0));
    // h / (2N), h / N (coarse) or h / (4N), h / (2N) (fine); N = 2^e
    const int e = coarse_lane ? 0 : (lane <= kmax ? lane : 0) + 1;
    const double hk = div_pow2(h, e + 1);
    const double ms = div_pow2(h, e);
    const int64_t pairs = coarse_lane ? N : 2 * N;
    const Kick kh = make_kick(ph, hk), km = make_kick(ph, ms);
    double r = r_in, v = v_in;
    if (!bud.take(2)) {
        r_out = r_in;
        v_out = v_in;
        return NS_LVL_NOCONV | NS_LVL_MASK;
    }
    if (active)
        kick<FAST>(ph, kh, r, v);
    int64_t t = 0;
    double prev_r = 0.0, prev_v = 0.0;  // fine of the previous level (shifted coarse)
    for (int k = 0; k <= kmax; k++) {
        const int64_t Nk = (int64_t)1 << k;
        if (!bud.take(2 * Nk + 1)) {
            if (k == 0) {
                const double rc = __shfl(r, 63, 64), vc = __shfl(v, 63, 64);
                leap_output(a.out_type, rc, vc, r_in, v_in, r_out, v_out);
                return NS_LVL_NOCONV;
            }
            leap_output(a.out_type, prev_r, prev_v, prev_r, prev_v, r_out, v_out);
            return (k - 1) | NS_LVL_NOCONV;
        }
        // Iterations [2^k, 2^(k+1)) belong to lanes k..kmax, whose passes all
        // run through them; the lane mask is set once per level, the last
        // iteration closes lane k's pass with its half kick.  Level 0 also
        // carries the one-iteration coarse pass of lane 63.
        const int64_t target = (int64_t)2 << k;
        if (k == 0) {
            for (; t < target; t++)
                if (active && t < pairs)
                    drift_kick<FAST>(ph, ms, (t < pairs - 1) ? km : kh, r, v);
        } else {
            if (lane >= k && lane <= kmax) {
                repeat_uniform(target - 1 - t, [&]() { drift_kick<FAST>(ph, ms, km, r, v); });
                drift_kick<FAST>(ph, ms, lane == k ? kh : km, r, v);
            }
            t = target;
        }
        const int src_c = (k == 0) ? 63 : k - 1;
        const double rc = __shfl(r, src_c, 64), vc = __shfl(v, src_c, 64);
        const double rf = __shfl(r, k, 64), vf = __shfl(v, k, 64);
        if (converged<FAST>(a, rc, vc, rf, vf)) {
            leap_output(a.out_type, rc, vc, rf, vf, r_out, v_out);
            return level_of(Nk);
        }
        if (k == kmax) {
            leap_output(a.out_type, rc, vc, rf, vf, r_out, v_out);
            return level_of(Nk) | NS_LVL_NOCONV;
        }
        prev_r = rf;
        prev_v = vf;
    }
    // unreachable: the loop returns at k == kmax
    r_out = r_in;
    v_out = v_in;
    return NS_LVL_NOCONV;
}

// Fast-form state of one Levi-Civita pass (doMicroLeviCivita's loop).  The
// half-kicked velocity vh is carried as u = hd vh (hd = dtau/2), so that the
// drift is rho' = rho + rho u; the kick constants are pre-scaled so that
// g inv2 = hd dtau forceLC(rho); the physical time is accumulated as
// T = sum (rho + rho')^2, t = t_in + dtau T / 4 at the end.  The closing
// half-kick of one micro-step and the opening one of the next are one kick.
// inv2 = (1/rho)^2 from the refined reciprocal of rho: nine operations per
// micro-step, the same count as the reciprocal of rho^2, but measured 5 %
// faster on one wave (23.3 against 24.6 ns per micro-step).
struct LcFast {
    double hd, lh, gh;  // dtau/2, hd dtau ell^2, hd dtau gm
    double ihd;         // 1 / hd
    double rho, u, u_prev, T;
    double g, inv2;     // force terms at the current rho
    double inv;         // 1 / rho (set by init and step, read by step_seeded)

    // init with the dtau-dependent constants (LcPassConst) and 1/rho_in^2
    // given: the same operations as init
    __device__ inline void init_from(double hd_, double lh_, double gh_, double ihd_, double rho_in, double inv2_in,
                                     double v_in)
    {
        hd = hd_;
        lh = lh_;
        gh = gh_;
        ihd = ihd_;
        rho = rho_in;
        T = 0.0;
        inv2 = inv2_in;  // inv stays unset: step, not step_seeded, comes first
        g = fma(lh, inv2, gh);
        u = fma(0.5 * g, inv2, hd * v_in);
        u_prev = u;
    }
    __device__ inline void init(const Phys &ph, double dtau, double rho_in, double v_in)
    {
        hd = 0.5 * dtau;
        ihd = 1.0 / hd;
        const double sc = hd * dtau;
        lh = ph.l2 * sc;
        gh = (ph.idg ? 0.0 : ph.gm) * sc;
        rho = rho_in;
        T = 0.0;
        inv = fast_rcp(rho);
        inv2 = inv * inv;
        g = fma(lh, inv2, gh);
        u = fma(0.5 * g, inv2, hd * v_in);  // hd (v + hd forceLC(rho))
        u_prev = u;
    }
    __device__ inline void step()
    {
        const double rn = fma(rho, u, rho);
        inv = fast_rcp(rn);
        inv2 = inv * inv;
        g = fma(lh, inv2, gh);
        const double s = rho + rn;
        T = fma(s, s, T);
        u_prev = u;
        u = fma(g, inv2, u);
        rho = rn;
    }
    // step with 1/rn from the reciprocal of the previous rho instead of the
    // hardware reciprocal: rn = rho (1 + u), so y0 = inv (1 - u) has the
    // relative error u^2 and one Newton-Raphson step y0 (2 - rn y0) leaves
    // u^4 (below the double rounding for |u| < 1e-4, the relative change of
    // rho in one micro-step).  Two FMAs in place of the reciprocal, which
    // issues at a quarter of the FMA rate.
    __device__ inline void step_seeded()
    {
        const double rn = fma(rho, u, rho);
        const double y0 = fma(-inv, u, inv);
        inv = fma(y0, fma(-rn, y0, 1.0), y0);
        inv2 = inv * inv;
        g = fma(lh, inv2, gh);
        const double s = rho + rn;
        T = fma(s, s, T);
        u_prev = u;
        u = fma(g, inv2, u);
        rho = rn;
    }
    // step with 1/rn = inv (1 - u + u^2), the series of 1/(1 + u) to second
    // order: relative error u^3 on top of the error inv brings, corrected by
    // the Newton step of the next step_seeded (one FMA fewer than it).
    __device__ inline void step_series()
    {
        const double rn = fma(rho, u, rho);
        inv = fma(inv, fma(u, u, -u), inv);
        inv2 = inv * inv;
        g = fma(lh, inv2, gh);
        const double s = rho + rn;
        T = fma(s, s, T);
        u_prev = u;
        u = fma(g, inv2, u);
        rho = rn;
    }
    // velocity after the steps taken: the closing half kick on the u held
    // before the last full kick (u is v scaled by hd)
    __device__ inline double v_end() const { return fma(0.5 * g, inv2, u_prev) * ihd; }
    __device__ inline double t_end(double t_in) const { return fma(0.5 * hd, T, t_in); }
};

// doMicroLeviCivita, nsphere.c:1990-2036 (force at rho_next reused).
template <bool FAST>
__device__ inline void micro_lc(const Phys &ph, double rho_in, double v_in, double t_in, int64_t sub,
                                double h_tau, double &rho_out, double &v_out, double &t_out)
{
    const double dtau = h_tau / (double)sub;
    double rho = rho_in, v = v_in, t = t_in;
    if constexpr (!FAST) {
        double f = force_lc_strict(ph, rho);
        for (int64_t ss = 0; ss < sub; ss++) {
            const double v_half = v + 0.5 * dtau * f;
            const double rho_next = rho + dtau * (0.5 * rho * v_half);
            const double f2 = force_lc_strict(ph, rho_next);
            const double v_next = v_half + 0.5 * dtau * f2;
            const double rho_mid = 0.5 * (rho + rho_next);
            const double dt_phys = dtau * (rho_mid * rho_mid);
            t += dt_phys;
            rho = rho_next;
            v = v_next;
            f = f2;
        }
    } else if (sub > 0) {
        LcFast lf;
        lf.init(ph, dtau, rho_in, v_in);
#pragma unroll 4
        for (int64_t ss = 0; ss < sub; ss++)
            lf.step();
        rho = lf.rho;
        v = lf.v_end();
        t = lf.t_end(t_in);
    }
    rho_out = rho;
    v_out = v;
    t_out = t;
}

// The reference's Levi-Civita convergence test (nsphere.c:2125-2130): the
// change of rho relative to r = rho_F^2 and of v relative to v_F below the
// tolerances.  The fast form multiplies instead of dividing, which differs
// from it only where a quotient rounds across the tolerance.
template <bool FAST>
__device__ inline bool lc_converged(const StepArgs &a, double rhoC, double vC, double rhoF, double vF)
{
    const double rF = rhoF * rhoF;
    if (FAST)
        return fabs(rhoF - rhoC) < a.radius_tol * (fabs(rF) + 1.0e-30) &&
               fabs(vF - vC) < a.velocity_tol * (fabs(vF) + 1.0e-30);
    const double rhodif = fabs(rhoF - rhoC) / (fabs(rF) + 1.0e-30);
    const double vdif = fabs(vF - vC) / (fabs(vF) + 1.0e-30);
    return (rhodif < a.radius_tol) && (vdif < a.velocity_tol);
}

__device__ inline void lc_output(int out_type, double rhoC, double vC, double tC, double rhoF, double vF,
                                 double tF, double &rho_out, double &v_out, double &t_out)
{
    if (out_type == 0) {
        rho_out = rhoC;
        v_out = vC;
        t_out = tC;
    } else if (out_type == 1) {
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
}

// doSingleTauStepAdaptiveLeviCivita, nsphere.c:2077-2186.  With cap_N > 0
// the loop stops before a level whose N exceeds cap_N and returns kEscalate.
template <bool FAST>
__device__ int single_tau_lc(const Phys &ph, const StepArgs &a, double rho_in, double v_in, double t_in,
                             double h, double &rho_out, double &v_out, double &t_out, Budget &bud,
                             int64_t cap_N)
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
        micro_lc<FAST>(ph, rho_in, v_in, t_in, 2 * N + 1, h, rhoC, vC, tC);
    }
    while (N <= a.max_subdiv) {
        if (cap_N > 0 && N > cap_N)
            return kEscalate;
        if (!bud.take(4 * N + 2)) {
            rho_out = rhoC;
            v_out = vC;
            t_out = tC;
            return last | NS_LVL_NOCONV;
        }
        micro_lc<FAST>(ph, rho_in, v_in, t_in, 4 * N + 1, h, rhoF, vF, tF);
        last = level_of(N);
        if (lc_converged<FAST>(a, rhoC, vC, rhoF, vF)) {
            lc_output(a.out_type, rhoC, vC, tC, rhoF, vF, tF, rho_out, v_out, t_out);
            return last;
        }
        N *= 2;
        if (N <= a.max_subdiv) {
            rhoC = rhoF;
            vC = vF;
            tC = tF;
        }
    }
    // not converged: the fine pass at the overflowed N (nsphere.c:2168-2185)
    if (!bud.take(4 * N + 2)) {
        rho_out = rhoF;
        v_out = vF;
        t_out = tF;
        return last | NS_LVL_NOCONV;
    }
    micro_lc<FAST>(ph, rho_in, v_in, t_in, 4 * N + 1, h, rho_out, v_out, t_out);
    return level_of(N) | NS_LVL_NOCONV;
}

// doSingleTauStepAdaptiveLeviCivita, wave-cooperative form (all 64 lanes,
// same particle).  Lane k runs the fine pass of level k (4*2^k+1 micro-steps,
// k = 0..kmax+1 -- the last one is the non-convergence fallback), lane 63 the
// coarse pass of level 0 (3 micro-steps).  All lanes step in lockstep; after
// 4*2^k+1 iterations lanes k-1 (or 63) and k hold coarse and fine of level k
// and the reference's test is evaluated.  Same results as the lane form, after
// 4*2^K+1 iterations instead of about 8*2^K.
// The fast form's wave passes from level kSeedLevel on take the reciprocal
// from the previous micro-step's (LcFast::step_seeded), and from level
// kSeriesLevel on three micro-steps of four take the second-order series
// instead (LcFast::step_series) between Newton-refined ones.
constexpr int kSeedLevel = 5;
constexpr int kSeriesLevel = 6;
static_assert(kSeedLevel >= 1 && kSeriesLevel >= kSeedLevel,
              "level 0 runs the hardware reciprocal; a level k >= 1 adds 2^(k+1) iterations, whole groups of 4");

template <bool FAST>
__device__ int single_tau_lc_wave(const Phys &ph, const StepArgs &a, double rho_in, double v_in, double t_in,
                                  double h, double &rho_out, double &v_out, double &t_out, Budget &bud)
{
    if (a.max_subdiv < 1)  // degenerate: no level loop, only the fallback pass
        return single_tau_lc<FAST>(ph, a, rho_in, v_in, t_in, h, rho_out, v_out, t_out, bud, 0);
    const int lane = threadIdx.x & 63;
    int kmax = 63 - __clzll((unsigned long long)a.max_subdiv);
    if (kmax > 61)
        kmax = 61;
    const bool coarse_lane = lane == 63;
    const bool active = coarse_lane || lane <= kmax + 1;
    const int64_t sub = coarse_lane ? 3 : (active ? 4 * ((int64_t)1 << lane) + 1 : 1);
    const double dtau = h / (double)sub;
    // strict: the reference's variables; fast: the pass state of micro_lc<true>
    double rho = rho_in, v = v_in, t = t_in, f = 0.0;
    LcFast lf;
    if constexpr (!FAST)
        f = force_lc_strict(ph, rho);
    else
        lf.init(ph, dtau, rho_in, v_in);
    if (!bud.take(4)) {
        rho_out = rho_in;
        v_out = v_in;
        t_out = t_in;
        return NS_LVL_NOCONV | NS_LVL_MASK;
    }
    auto micro_step = [&]() {
        if constexpr (!FAST) {
            const double v_half = v + 0.5 * dtau * f;
            const double rho_next = rho + dtau * (0.5 * rho * v_half);
            const double f2 = force_lc_strict(ph, rho_next);
            const double v_next = v_half + 0.5 * dtau * f2;
            const double rho_mid = 0.5 * (rho + rho_next);
            const double dt_phys = dtau * (rho_mid * rho_mid);
            t += dt_phys;
            rho = rho_next;
            v = v_next;
            f = f2;
        } else {
            lf.step();
        }
    };
    // state at the end of the micro-steps this lane has run
    auto rho_now = [&]() -> double { return FAST ? lf.rho : rho; };
    auto v_now = [&]() -> double {
        if constexpr (FAST)
            return lf.v_end();
        else
            return v;
    };
    auto t_now = [&]() -> double {
        if constexpr (FAST)
            return lf.t_end(t_in);
        else
            return t;
    };
    // Iterations [4*2^(k-1)+1, 4*2^k+1) belong to lanes k..kmax+1, whose
    // passes all run through them, so the lane mask is set once per level.
    // Level 0 (iterations [0, 5)) also carries lane 63's 3-step coarse pass.
    // The fast form's iterations from level kSeedLevel on take the seeded
    // reciprocal (LcFast::step_seeded, and from kSeriesLevel on the series
    // one three times in four): the passes running there have at least
    // 4*2^kSeedLevel+1 micro-steps, each moving rho by a small part of its
    // change over the tau step.
    int64_t it = 0;
    auto run_level = [&](int k) {
        const int64_t end = 4 * ((int64_t)1 << k) + 1;
        if (k == 0) {
            if (active)
                for (int s = 0; s < 3; s++)
                    micro_step();
            if (active && !coarse_lane)
                for (int s = 3; s < 5; s++)
                    micro_step();
        } else if (lane >= k && lane <= kmax + 1) {
            if (FAST && k >= kSeriesLevel)
                repeat_uniform((end - it) / 4, [&]() {  // 2^(k+1) steps, a multiple of 4
                    lf.step_seeded();
                    lf.step_series();
                    lf.step_series();
                    lf.step_series();
                });
            else if (FAST && k >= kSeedLevel)
                repeat_uniform(end - it, [&]() { lf.step_seeded(); });  // 2^(k+1) steps
            else
                repeat_uniform(end - it, micro_step);  // 2^(k+1) steps
        }
        it = end;
    };
    for (int k = 0; k <= kmax; k++) {
        const int64_t Nk = (int64_t)1 << k;
        const int src_c = (k == 0) ? 63 : k - 1;
        if (!bud.take(4 * Nk + 2)) {
            // coarse of level k: the previous fine pass (or lane 63)
            const double rn = rho_now(), vn = v_now(), tn = t_now();
            rho_out = __shfl(rn, src_c, 64);
            v_out = __shfl(vn, src_c, 64);
            t_out = __shfl(tn, src_c, 64);
            bud.chain += it;
            return (k == 0 ? 0 : level_of(Nk >> 1)) | NS_LVL_NOCONV;
        }
        run_level(k);
        const double rn = rho_now(), vn = v_now(), tn = t_now();
        const double rhoC = __shfl(rn, src_c, 64), vC = __shfl(vn, src_c, 64), tC = __shfl(tn, src_c, 64);
        const double rhoF = __shfl(rn, k, 64), vF = __shfl(vn, k, 64), tF = __shfl(tn, k, 64);
        if (lc_converged<FAST>(a, rhoC, vC, rhoF, vF)) {
            lc_output(a.out_type, rhoC, vC, tC, rhoF, vF, tF, rho_out, v_out, t_out);
            bud.chain += it;
            return level_of(Nk);
        }
    }
    // not converged: the fine pass at the overflowed N = 2^(kmax+1)
    const int64_t Nf = (int64_t)2 << kmax;
    if (!bud.take(4 * Nf + 2)) {
        const double rn = rho_now(), vn = v_now(), tn = t_now();
        rho_out = __shfl(rn, kmax, 64);
        v_out = __shfl(vn, kmax, 64);
        t_out = __shfl(tn, kmax, 64);
        bud.chain += it;
        return level_of(Nf >> 1) | NS_LVL_NOCONV;
    }
    run_level(kmax + 1);
    const double rn = rho_now(), vn = v_now(), tn = t_now();
    rho_out = __shfl(rn, kmax + 1, 64);
    v_out = __shfl(vn, kmax + 1, 64);
    t_out = __shfl(tn, kmax + 1, 64);
    bud.chain += it;
    return level_of(Nf) | NS_LVL_NOCONV;
}

// doAdaptiveFullLeviCivita, nsphere.c:2231-2334.  WAVE: every tau step on the
// whole wave (all lanes call with the same particle).  cap_N (lane form
// only): escalate with kEscalate when a tau step needs N > cap_N.
template <bool FAST, bool WAVE = false>
__device__ int adaptive_lc(const Phys &ph, const StepArgs &a, double r_in, double v_in, double &r_out,
                           double &v_out, Budget &bud, int64_t cap_N = 0)
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
        int lv;
        if constexpr (WAVE)
            lv = single_tau_lc_wave<FAST>(ph, a, rho_current, v_current, t_cur, delta_tau, rho_next, v_next,
                                          t_next, bud);
        else
            lv = single_tau_lc<FAST>(ph, a, rho_current, v_current, t_cur, delta_tau, rho_next, v_next, t_next,
                                     bud, cap_N);
        if (lv == kEscalate)
            return kEscalate;
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

// The constants LcFast::init derives from dtau = h / sub for a pass of sub
// micro-steps over h (the same operations, done once per particle-step
// instead of once per tau step).
struct LcPassConst {
    double hd, lh, gh, ihd;
    __device__ inline void set(const Phys &ph, double h, int64_t sub)
    {
        const double dtau = h / (double)sub;
        hd = 0.5 * dtau;
        ihd = 1.0 / hd;
        const double sc = hd * dtau;
        lh = ph.l2 * sc;
        gh = (ph.idg ? 0.0 : ph.gm) * sc;
    }
};



// single_tau_lc<true> with cap_N = 2 (levels 0 and 1), for max_subdiv >= 4,
// written out: the coarse and fine passes of level 0 (3 and 5 micro-steps)
// start from the same state, so they run interleaved, and the level-1 pass
// (9 micro-steps) only when level 0 fails.  The same passes, budget charges,
// tests and outputs as the general form, hence the same results.
__device__ inline int lc_tau_lane_fast(const StepArgs &a, const LcPassConst K[3], double rho_in, double v_in,
                                       double t_in, double &rho_out, double &v_out, double &t_out, Budget &bud)
{
    if (!bud.take(4)) {
        rho_out = rho_in;
        v_out = v_in;
        t_out = t_in;
        return NS_LVL_NOCONV | NS_LVL_MASK;
    }
    const double inv = fast_rcp(rho_in);
    const double inv2 = inv * inv;
    LcFast c, f;
    c.init_from(K[0].hd, K[0].lh, K[0].gh, K[0].ihd, rho_in, inv2, v_in);
    f.init_from(K[1].hd, K[1].lh, K[1].gh, K[1].ihd, rho_in, inv2, v_in);
    c.step();
    f.step();
    c.step();
    f.step();
    c.step();
    f.step();
    f.step();
    f.step();
    const double rhoC = c.rho, vC = c.v_end(), tC = c.t_end(t_in);
    if (!bud.take(6)) {
        rho_out = rhoC;
        v_out = vC;
        t_out = tC;
        return NS_LVL_NOCONV;
    }
    const double rhoF = f.rho, vF = f.v_end(), tF = f.t_end(t_in);
    if (lc_converged<true>(a, rhoC, vC, rhoF, vF)) {
        lc_output(a.out_type, rhoC, vC, tC, rhoF, vF, tF, rho_out, v_out, t_out);
        return 0;
    }
    if (!bud.take(10)) {  // the coarse pass of level 1 is the fine pass of level 0
        rho_out = rhoF;
        v_out = vF;
        t_out = tF;
        return NS_LVL_NOCONV;
    }
    LcFast d;
    d.init_from(K[2].hd, K[2].lh, K[2].gh, K[2].ihd, rho_in, inv2, v_in);
#pragma unroll
    for (int s = 0; s < 9; s++)
        d.step();
    const double rhoD = d.rho, vD = d.v_end(), tD = d.t_end(t_in);
    if (lc_converged<true>(a, rhoF, vF, rhoD, vD)) {
        lc_output(a.out_type, rhoF, vF, tF, rhoD, vD, tD, rho_out, v_out, t_out);
        return 1;
    }
    return kEscalate;  // level 2 is past the lane levels
}

// adaptive_lc<true> (lane form) with cap_N = 2 and max_subdiv >= 4: the
// tau loop of doAdaptiveFullLeviCivita (nsphere.c:2231-2334) around
// lc_tau_lane_fast, the pass constants computed once for the particle-step.
__device__ inline int adaptive_lc_lane_fast(const Phys &ph, const StepArgs &a, double r_in, double v_in,
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
    LcPassConst K[3];
    K[0].set(ph, delta_tau, 3);
    K[1].set(ph, delta_tau, 5);
    K[2].set(ph, delta_tau, 9);
    int64_t step_count = 0;
    const int64_t step_max = 100000000;
    while (true) {
        if (t_cur >= dt)
            break;
        double rho_next, v_next, t_next;
        const int lv = lc_tau_lane_fast(a, K, rho_current, v_current, t_cur, rho_next, v_next, t_next, bud);
        if (lv == kEscalate)
            return kEscalate;
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

} // namespace ns
