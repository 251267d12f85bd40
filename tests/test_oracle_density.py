"""Pins of the density oracle (oracle/density.py) on the CPU."""
import math

import numpy as np
import pytest

from oracle import density as od


def test_natural_spline_matches_scipy():
    interpolate = pytest.importorskip("scipy.interpolate")
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.uniform(0.1, 1.0, 200))
    y = np.sin(x) + 0.1 * x
    c = od.spline_coeffs(x, y)
    t = np.sort(rng.uniform(x[0], x[-1], 5000))
    want = interpolate.CubicSpline(x, y, bc_type="natural")(t)
    np.testing.assert_allclose(od.spline_eval(x, y, c, t), want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(od.spline_eval(x, y, c, x[:-1]), y[:-1], rtol=0, atol=1e-13)  # through the nodes
    # linear data: exactly linear (all second derivatives zero)
    assert np.all(np.abs(od.spline_coeffs(x, 3.0 * x - 1.0)) < 1e-11)


def test_fft_convolution_equals_direct():
    rng = np.random.default_rng(5)
    den = np.exp(-np.linspace(0, 30, od.GRID)) * 1e6 + rng.uniform(0, 1, od.GRID)
    ker = od.kernel(3.8e-5, 0.005)
    taps = int(math.ceil(37.2 * 0.005 / 3.8e-5)) + 1
    fft, direct = od.conv_fft(den, ker), od.conv_direct(den, ker, taps)
    assert np.max(np.abs(fft - direct)) <= 1e-12 * np.max(direct)
    assert abs(np.sum(ker) - 1.0) < 1e-12


def test_estimate_follows_plummer_density():
    # sorted Plummer radii, G = M = a = 1: dM/dr = 3 r^2 (1 + r^2)^(-5/2); the
    # estimate must equal that density put through the same one-sided
    # smoothing (which lags it by about sigma in log10 r)
    n = 200_000
    u = (np.arange(n) + 0.5) / n
    r = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    est, info = od.density(r, 1.0)
    assert info["decimated"] == 2000 and info["kept"] == n
    rg, den, _ = od.grid_of(r, 1.0)
    exact = 3 * rg ** 2 * (1 + rg ** 2) ** -2.5
    lagged = od.conv_fft(exact, od.kernel(info["dlog"], info["sigma"]))
    mid = (r > 0.2) & (r < 5.0)
    want = np.interp(r[mid], rg, lagged)
    assert np.max(np.abs(est[mid] / want - 1.0)) < 0.01
    # and not the unsmoothed density: the lag is visible
    assert np.max(np.abs(est[mid] / (3 * r[mid] ** 2 * (1 + r[mid] ** 2) ** -2.5) - 1.0)) > 0.05
