// ns_ic.cpp -- seeded initial conditions for the evolution loop (host side).
//
// The reference samples its initial conditions from a GSL spline of its own
// cored profile, seeded from time+pid (nsphere.c:4656-4955, 4699-4700).  The
// workloads here are instead closed-form Plummer and Hernquist spheres in
// G = M = a = 1 units, drawn from a counter-based Philox-4x32-10 stream
// (Salmon et al., SC'11) so that every particle's draw depends only on
// (seed, particle index, attempt) and is reproducible at any N.

#include <cmath>
#include <cstdint>
#include <cstring>

#include "nsphere_gpu.h"
#include "ns_internal.h"

namespace {

inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

// Streams of the per-particle counter (word 3).
enum : uint32_t { S_RADIUS = 1, S_SPEED = 2, S_DIRECTION = 3, S_LOWL = 4, S_GAUSS = 5 };

struct Draw {
    uint64_t seed;
    uint64_t idx;
    // Four 32-bit words for (stream, attempt).
    void block(uint32_t stream, uint32_t attempt, uint32_t out[4]) const
    {
        uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), attempt, stream};
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        ns_philox4x32_10(ctr, key, out);
    }
};

// Uniform on the open interval (0, 1) from 53 random bits.
inline double u01(uint32_t hi, uint32_t lo)
{
    uint64_t b = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)b + 0.5) * 0x1.0p-53;
}

// Plummer, G = M = a = 1: M(r) = r^3/(1+r^2)^{3/2} => r = (U^{-2/3} - 1)^{-1/2}.
double plummer_radius(const Draw &d, double rmax)
{
    for (uint32_t a = 0;; a++) {
        uint32_t w[4];
        d.block(S_RADIUS, a, w);
        double u = u01(w[0], w[1]);
        double r = 1.0 / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        if (r <= rmax && std::isfinite(r))
            return r;
    }
}

// Aarseth, Henon & Wielen (1974): q = v/v_esc with density q^2 (1-q^2)^{7/2},
// accepted under the bound 0.1.
double plummer_speed(const Draw &d, double r)
{
    double vesc = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    for (uint32_t a = 0;; a++) {
        uint32_t w[4];
        d.block(S_SPEED, a, w);
        double q = u01(w[0], w[1]);
        double y = u01(w[2], w[3]);
        if (0.1 * y < q * q * std::pow(1.0 - q * q, 3.5))
            return q * vesc;
    }
}

// Hernquist, G = M = a = 1: M(r) = r^2/(1+r)^2 => r = sqrt(U) / (1 - sqrt(U)).
double hernquist_radius(const Draw &d, double rmax)
{
    for (uint32_t a = 0;; a++) {
        uint32_t w[4];
        d.block(S_RADIUS, a, w);
        double s = std::sqrt(u01(w[0], w[1]));
        double r = s / (1.0 - s);
        if (r <= rmax && std::isfinite(r))
            return r;
    }
}

// Isotropic Jeans dispersion of the Hernquist sphere (Hernquist 1990, eq. 10),
// G = M = a = 1.
double hernquist_sigma2(double r)
{
    double t1 = 12.0 * r * std::pow(1.0 + r, 3) * std::log1p(1.0 / r);
    double t2 = r / (1.0 + r) * (25.0 + 52.0 * r + 42.0 * r * r + 12.0 * r * r * r);
    double s2 = (t1 - t2) / 12.0;
    return s2 > 0.0 ? s2 : 0.0;
}

// Box-Muller pair from one block.
void gauss2(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, double *g0, double *g1)
{
    double u = u01(w0, w1), v = u01(w2, w3);
    double rad = std::sqrt(-2.0 * std::log(u));
    *g0 = rad * std::cos(2.0 * M_PI * v);
    *g1 = rad * std::sin(2.0 * M_PI * v);
}

} // namespace

extern "C" void ns_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c0, &hi0, &lo0);
        mulhilo32(M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

extern "C" int ns_ic_generate(int kind, int64_t n, uint64_t seed, double rmax, double lowl_frac,
                              double lowl_scale, double *r, double *v, double *ell, uint32_t *id)
{
    if (n < 0 || n > (int64_t)UINT32_MAX || kind < 0 || kind > 2 || !(rmax > 0.0))
        return ns::set_error(NS_ERR_ARG, "ns_ic_generate: bad kind/n/rmax");
    if (n > 0 && (!r || !v || !ell || !id))
        return ns::set_error(NS_ERR_ARG, "ns_ic_generate: null output array");
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        Draw d{seed, (uint64_t)i};
        bool hern = (kind == 1) || (kind == 2 && (i & 1));
        double rr, vr, vt;
        if (!hern) {
            rr = plummer_radius(d, rmax);
            double sp = plummer_speed(d, rr);
            uint32_t w[4];
            d.block(S_DIRECTION, 0, w);
            double mu = 2.0 * u01(w[0], w[1]) - 1.0; // isotropic cos(theta)
            vr = sp * mu;
            vt = sp * std::sqrt(1.0 - mu * mu);
        } else {
            rr = hernquist_radius(d, rmax);
            double sig = std::sqrt(hernquist_sigma2(rr));
            uint32_t w[4], x[4];
            d.block(S_GAUSS, 0, w);
            d.block(S_GAUSS, 1, x);
            double g0, g1, g2, g3;
            gauss2(w[0], w[1], w[2], w[3], &g0, &g1);
            gauss2(x[0], x[1], x[2], x[3], &g2, &g3);
            (void)g3;
            vr = sig * g0;
            vt = sig * std::sqrt(g1 * g1 + g2 * g2);
        }
        double l = rr * vt;
        uint32_t w[4];
        d.block(S_LOWL, 0, w);
        if (u01(w[0], w[1]) < lowl_frac)
            l *= lowl_scale;
        r[i] = rr;
        v[i] = vr;
        ell[i] = l;
        id[i] = (uint32_t)i;
    }
    return NS_OK;
}
