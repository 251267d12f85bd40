// ns_ic_sampler.h -- per-particle draw of the seeded initial conditions,
// shared by the host generator (ns_ic.cpp, OpenMP) and the device generator
// (ns_io.hip): every function here is compiled for both sides, so the two
// take the same draws; they differ only where the device math library's
// pow / log / sin / cos round differently from the host's.
//
// Closed-form Plummer and Hernquist spheres in G = M = a = 1 units, drawn
// from a counter-based Philox-4x32-10 stream (Salmon et al., SC'11): every
// particle's draw depends only on (seed, particle index, stream, attempt).
#pragma once

#include <cmath>
#include <cstdint>

#pragma clang force_cuda_host_device begin

namespace ns {
namespace ic {

inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
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

// Rejection loops give up (NaN) after this many attempts, which only a
// degenerate recipe reaches (e.g. rmax so small that almost no draw fits).
constexpr uint32_t kMaxAttempts = 1u << 16;

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
        philox4x32_10(ctr, key, out);
    }
};

// Uniform on the open interval (0, 1) from 53 random bits.
inline double u01(uint32_t hi, uint32_t lo)
{
    uint64_t b = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)b + 0.5) * 0x1.0p-53;
}

// Plummer, G = M = a = 1: M(r) = r^3/(1+r^2)^{3/2} => r = (U^{-2/3} - 1)^{-1/2}.
inline double plummer_radius(const Draw &d, double rmax)
{
    for (uint32_t a = 0; a < kMaxAttempts; a++) {
        uint32_t w[4];
        d.block(S_RADIUS, a, w);
        double u = u01(w[0], w[1]);
        double r = 1.0 / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        if (r <= rmax && std::isfinite(r))
            return r;
    }
    return std::nan("");
}

// Aarseth, Henon & Wielen (1974): q = v/v_esc with density q^2 (1-q^2)^{7/2},
// accepted under the bound 0.1.
inline double plummer_speed(const Draw &d, double r)
{
    double vesc = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    for (uint32_t a = 0; a < kMaxAttempts; a++) {
        uint32_t w[4];
        d.block(S_SPEED, a, w);
        double q = u01(w[0], w[1]);
        double y = u01(w[2], w[3]);
        if (0.1 * y < q * q * std::pow(1.0 - q * q, 3.5))
            return q * vesc;
    }
    return std::nan("");
}

// Hernquist, G = M = a = 1: M(r) = r^2/(1+r)^2 => r = sqrt(U) / (1 - sqrt(U)).
inline double hernquist_radius(const Draw &d, double rmax)
{
    for (uint32_t a = 0; a < kMaxAttempts; a++) {
        uint32_t w[4];
        d.block(S_RADIUS, a, w);
        double s = std::sqrt(u01(w[0], w[1]));
        double r = s / (1.0 - s);
        if (r <= rmax && std::isfinite(r))
            return r;
    }
    return std::nan("");
}

// Isotropic Jeans dispersion of the Hernquist sphere (Hernquist 1990, eq. 10),
// G = M = a = 1.
inline double hernquist_sigma2(double r)
{
    double t1 = 12.0 * r * std::pow(1.0 + r, 3) * std::log1p(1.0 / r);
    double t2 = r / (1.0 + r) * (25.0 + 52.0 * r + 42.0 * r * r + 12.0 * r * r * r);
    double s2 = (t1 - t2) / 12.0;
    return s2 > 0.0 ? s2 : 0.0;
}

// Box-Muller pair from one block.
inline void gauss2(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, double *g0, double *g1)
{
    double u = u01(w0, w1), v = u01(w2, w3);
    double rad = std::sqrt(-2.0 * std::log(u));
    *g0 = rad * std::cos(2.0 * 3.14159265358979323846 * v);
    *g1 = rad * std::sin(2.0 * 3.14159265358979323846 * v);
}

// One particle of recipe `kind` (0 Plummer, 1 Hernquist, 2 alternating by
// index parity): radius, radial velocity and ell, with the low-ell tail.
inline void sample_particle(int kind, uint64_t seed, uint64_t i, double rmax, double lowl_frac, double lowl_scale,
                            double &r, double &v, double &ell)
{
    Draw d{seed, i};
    const bool hern = (kind == 1) || (kind == 2 && (i & 1));
    double rr, vr, vt;
    if (!hern) {
        rr = plummer_radius(d, rmax);
        const double sp = plummer_speed(d, rr);
        uint32_t w[4];
        d.block(S_DIRECTION, 0, w);
        const double mu = 2.0 * u01(w[0], w[1]) - 1.0;  // isotropic cos(theta)
        vr = sp * mu;
        vt = sp * std::sqrt(1.0 - mu * mu);
    } else {
        rr = hernquist_radius(d, rmax);
        const double sig = std::sqrt(hernquist_sigma2(rr));
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
    r = rr;
    v = vr;
    ell = l;
}

} // namespace ic
} // namespace ns

#pragma clang force_cuda_host_device end
