"""CPU oracle of the snapshot density estimate -- TEST INFRASTRUCTURE ONLY.

Restates nsphere.c:7236-7686 (with fft_gaussian_convolution, 2901-3060) in
numpy, step by step:
  1. enclosed mass (i + 1) dM of the sorted radii (nsphere.c:7224), dM =
     M / npts (5179); where radii repeat, the strictly increasing
     subsequence with the first mass of each run (7286-7326);
  2. min(ceil(n / 100), 30000) points at (int) round(i step) (7331-7361),
     non-increasing ones nudged by 1e-12 (7446-7454);
  3. a natural cubic spline (GSL "cspline": interior second-derivative
     system, diagonal 2 (h_i + h_{i+1}), off-diagonal h_{i+1}, solved by
     LDL^T) of M(r) on 131072 points with log10 r uniform from the first
     decimated radius to 1.04 x the last (7377-7425), mass constant outside
     the nodes (7476-7493);
  4. central differences floored at 1e-10, ends copied (7495-7520);
  5. the one-sided normalised Gaussian of width 0.08 (n / 1e4)^-0.4,
     convolved through a zero-padded real FFT of length 2 x 131072 and
     floored at 1e-10 (2901-3060) -- numpy's FFT stands in for FFTW here;
  6. a natural cubic spline of the smoothed grid at every radius, end
     values outside (7565-7660).
The direct (truncated) form of step 5, conv_direct(), is what the device
computes; tests/test_oracle_density.py pins the FFT form against it, the
spline against scipy's natural CubicSpline, and the whole estimate against
the analytic dM/dr of a Plummer sphere.
"""
from __future__ import annotations

import math

import numpy as np

GRID = 131072


def spline_coeffs(x, y) -> np.ndarray:
    """Second-derivative coefficients c (c[0] = c[-1] = 0) of the natural spline."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    m = x.size
    c = np.zeros(m)
    if m < 3:
        return c
    h_i, h_ip1 = x[1:-1] - x[:-2], x[2:] - x[1:-1]
    yd_i, yd_ip1 = y[1:-1] - y[:-2], y[2:] - y[1:-1]
    g_i = np.where(h_i != 0.0, 1.0 / np.where(h_i != 0.0, h_i, 1.0), 0.0)
    g_ip1 = np.where(h_ip1 != 0.0, 1.0 / np.where(h_ip1 != 0.0, h_ip1, 1.0), 0.0)
    off = h_ip1
    diag = 2.0 * (h_ip1 + h_i)
    g = 3.0 * (yd_ip1 * g_ip1 - yd_i * g_i)
    k = m - 2
    if k == 1:
        c[1] = g[0] / diag[0]
        return c
    alpha = np.empty(k)
    gamma = np.empty(k)
    alpha[0] = diag[0]
    gamma[0] = off[0] / alpha[0]
    for i in range(1, k - 1):
        alpha[i] = diag[i] - off[i - 1] * gamma[i - 1]
        gamma[i] = off[i] / alpha[i]
    alpha[k - 1] = diag[k - 1] - off[k - 2] * gamma[k - 2]
    z = np.empty(k)
    z[0] = g[0]
    for i in range(1, k):
        z[i] = g[i] - gamma[i - 1] * z[i - 1]
    xs = np.empty(k)
    xs[k - 1] = z[k - 1] / alpha[k - 1]
    for i in range(k - 2, -1, -1):
        xs[i] = z[i] / alpha[i] - gamma[i] * xs[i + 1]
    c[1:-1] = xs
    return c


def spline_eval(x, y, c, t) -> np.ndarray:
    """Natural spline at t (inside [x[0], x[-1]]); interval x[lo] <= t < x[lo+1]."""
    x = np.asarray(x)
    lo = np.clip(np.searchsorted(x, t, side="right") - 1, 0, x.size - 2)
    h = x[lo + 1] - x[lo]
    dy = y[lo + 1] - y[lo]
    delx = t - x[lo]
    ci, cn = c[lo], c[lo + 1]
    b = dy / h - h * (cn + 2.0 * ci) / 3.0
    d = (cn - ci) / (3.0 * h)
    return y[lo] + delx * (b + delx * (ci + delx * d))


def kernel(dlog: float, sigma: float) -> np.ndarray:
    x = np.arange(GRID, dtype=np.float64) * dlog
    ker = (1.0 / (sigma * math.sqrt(2.0 * math.pi))) * np.exp(-0.5 * (x / sigma) * (x / sigma))
    return ker / np.sum(ker)


def conv_fft(den, ker) -> np.ndarray:
    n2 = 2 * GRID
    a = np.zeros(n2)
    a[:GRID] = den
    k = np.zeros(n2)
    k[:GRID] = ker
    out = np.fft.irfft(np.fft.rfft(a) * np.fft.rfft(k), n=n2)[:GRID]
    return np.maximum(out, 1e-10)


def conv_direct(den, ker, taps: int) -> np.ndarray:
    out = np.zeros(GRID)
    for k in range(taps):
        out[k:] += ker[k] * den[:GRID - k]
    return np.maximum(out, 1e-10)


def grid_of(r_sorted, halo_mass: float):
    """Steps 1-4: (r_grid, density_grid, info)."""
    r = np.asarray(r_sorted, np.float64)
    n = r.size
    dm = np.float64(halo_mass) / np.float64(n)
    mass = np.arange(1, n + 1, dtype=np.float64) * dm
    keep = np.ones(n, bool)
    keep[1:] = r[1:] > r[:-1]
    rf, mf = r[keep], mass[keep]
    count = rf.size
    dec = min(int(math.ceil(n / 100.0)), 30000) or 1
    if count <= dec:
        rd, md = rf.copy(), mf.copy()
    else:
        step = (count - 1) / (dec - 1)
        x = np.arange(dec) * step
        t = np.trunc(x)
        idx = np.minimum((t + (x - t >= 0.5)).astype(np.int64), count - 1)  # C round() of x >= 0
        rd, md = rf[idx].copy(), mf[idx].copy()
    for i in range(1, rd.size):
        if rd[i] <= rd[i - 1]:
            rd[i] = rd[i - 1] + 1e-12
    min_r, max_r = rd[0], rd[-1] * 1.04
    lmin, lmax = math.log10(min_r), math.log10(max_r)
    dlog = (lmax - lmin) / (GRID - 1)
    rg = np.array([math.pow(10.0, lmin + i * dlog) for i in range(GRID)])
    for i in range(1, GRID):
        if rg[i] <= rg[i - 1]:
            rg[i] = rg[i - 1] + 1e-12
    c = spline_coeffs(rd, md)
    mg = np.where(rg < rd[0], md[0], np.where(rg > rd[-1], md[-1], 0.0))
    inside = (rg >= rd[0]) & (rg <= rd[-1])
    mg[inside] = spline_eval(rd, md, c, rg[inside])
    den = np.empty(GRID)
    dr = rg[2:] - rg[:-2]
    den[1:-1] = np.maximum((mg[2:] - mg[:-2]) / dr, 1e-10)
    den[0], den[-1] = den[1], den[-2]
    sigma = 0.08 * math.pow(n / 10000.0, -0.40)
    return rg, den, {"kept": count, "decimated": rd.size, "sigma": sigma, "dlog": dlog, "r_min": min_r,
                     "r_max": max_r}


def density(r_sorted, halo_mass: float, smooth=conv_fft):
    """The estimate at every radius (the reference's FFT smoothing by default)."""
    r = np.asarray(r_sorted, np.float64)
    rg, den, info = grid_of(r, halo_mass)
    ker = kernel(info["dlog"], info["sigma"])
    sm = smooth(den, ker)
    c = spline_coeffs(rg, sm)
    out = np.where(r <= rg[0], sm[0], np.where(r >= rg[-1], sm[-1], 0.0))
    inside = (r > rg[0]) & (r < rg[-1])
    out[inside] = spline_eval(rg, sm, c, r[inside])
    return out, info
