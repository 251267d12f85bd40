"""Host-side checks of the native library (no GPU needed).

* the C-ABI library loads and exports every function include/nsphere_gpu.h
  declares; without a GPU, context creation fails with a code, not a crash;
* the default parameters are the reference's constants (nsphere.c:364-373,
  5777-5782);
* Philox-4x32-10 matches the Random123 known-answer vectors;
* the seeded samplers reproduce the Plummer and Hernquist distributions
  (radii against the closed-form M(r), speeds against the Aarseth-Henon-Wielen
  density, the Hernquist dispersion against a numerical Jeans integral).
"""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy import integrate, stats

import nsphere_amd as ns
from nsphere_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "nsphere_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ns_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_context_without_gpu_is_an_error_not_a_crash():
    h = ctypes.c_void_p()
    rc = _lib.lib.ns_gpu_ctx_create(1000, 0, ctypes.byref(h))
    if rc == _lib.NS_OK:
        _lib.lib.ns_gpu_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc in (_lib.NS_ERR_NODEV, _lib.NS_ERR_HIP)
    assert _lib.lib.ns_gpu_last_error()
    assert _lib.lib.ns_gpu_ctx_create(-5, 0, ctypes.byref(h)) == _lib.NS_ERR_ARG


def test_null_context_is_rejected():
    assert _lib.lib.ns_gpu_sort_particles(None) == _lib.NS_ERR_ARG
    assert _lib.lib.ns_gpu_step(None, None) == _lib.NS_ERR_ARG
    assert b"null context" in _lib.lib.ns_gpu_last_error()


def test_default_params_are_reference_constants():
    p = ns.default_params()
    assert p.G == 4.3e-6 and p.halo_mass == 1.0e12
    assert p.vel_conv_sq == 1.02271e-3 * 1.02271e-3
    assert (p.radius_tol, p.velocity_tol) == (1e-7, 1e-7)
    assert p.max_subdiv == 4096 * 4096 * 16
    assert (p.out_type, p.N_taumin, p.alpha) == (2, 10, 0.05)
    u = ns.unit_params()
    assert (u.G, u.halo_mass, u.vel_conv_sq, u.dt) == (1.0, 1.0, 1.0, 1e-3)


def test_philox_known_answers():
    assert ns.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert ns.philox4x32_10((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                       0x6D5451FD)
    assert ns.philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
                            (0xA4093822, 0x299F31D0)) == (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_ic_deterministic_and_prefix_stable():
    a = ns.generate_ic(0, 5000, 42, 100.0, 0.05, 1e-3)
    b = ns.generate_ic(0, 5000, 42, 100.0, 0.05, 1e-3)
    c = ns.generate_ic(0, 2000, 42, 100.0, 0.05, 1e-3)
    for x, y, z in zip(a, b, c):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x[:2000], z)
    d = ns.generate_ic(0, 5000, 43, 100.0, 0.05, 1e-3)
    assert not np.array_equal(a[0], d[0])


def test_plummer_sampler():
    n = 200_000
    r, v, ell, ids = ns.generate_ic(0, n, 7, 100.0, 0.0, 1e-3)
    rmax = 100.0
    mass = lambda x: x ** 3 / (1 + x * x) ** 1.5
    cdf = lambda x: mass(np.minimum(x, rmax)) / mass(rmax)
    # fixed seeds make these deterministic; 1e-4 is the rejection level
    assert stats.kstest(r, cdf).pvalue > 1e-4
    vt = ell / r
    speed = np.sqrt(v * v + vt * vt)
    q = speed / (np.sqrt(2.0) * (1 + r * r) ** -0.25)
    # q has density proportional to q^2 (1 - q^2)^{7/2} on (0, 1)
    norm = integrate.quad(lambda x: x * x * (1 - x * x) ** 3.5, 0, 1)[0]
    qcdf = np.vectorize(lambda x: integrate.quad(lambda y: y * y * (1 - y * y) ** 3.5, 0, x)[0] / norm)
    qs = np.sort(q[:20000])
    assert stats.kstest(qs, qcdf).pvalue > 1e-4
    mu = v / speed
    assert abs(mu.mean()) < 0.01 and abs((mu * mu).mean() - 1 / 3) < 0.01
    np.testing.assert_array_equal(ids, np.arange(n, dtype=np.uint32))


def test_plummer_virial_ratio():
    import oracle
    n = 200_000
    r, v, ell, _ = ns.generate_ic(0, n, 11, 100.0, 0.0, 1e-3)
    q = np.argsort(r, kind="stable")
    K, W = oracle.energy(r[q], v[q], ell[q], oracle.params())
    assert 2 * K / abs(W) == pytest.approx(1.0, abs=0.02)


def test_hernquist_sampler():
    n = 200_000
    r, v, ell, _ = ns.generate_ic(1, n, 8, 200.0, 0.0, 1e-3)
    rmax = 200.0
    mass = lambda x: x * x / (1 + x) ** 2
    assert stats.kstest(r, lambda x: mass(np.minimum(x, rmax)) / mass(rmax)).pvalue > 1e-4
    # isotropic Gaussian: v_r variance in shells against the isotropic Jeans
    # dispersion sigma^2(r) = (1/rho) int_r^inf rho M / r'^2 dr' (G = M = a = 1)
    rho = lambda x: 1.0 / (2 * np.pi) / (x * (1 + x) ** 3)
    for lo, hi in [(0.3, 0.35), (1.0, 1.1), (3.0, 3.4)]:
        sel = (r > lo) & (r < hi)
        mid = 0.5 * (lo + hi)
        s2 = integrate.quad(lambda x: rho(x) * mass(x) / (x * x), mid, np.inf)[0] / rho(mid)
        assert np.var(v[sel]) == pytest.approx(s2, rel=0.08)
        vt2 = (ell[sel] / r[sel]) ** 2
        assert vt2.mean() == pytest.approx(2 * s2, rel=0.08)


def test_low_l_tail_fraction():
    a = ns.generate_ic(0, 100_000, 5, 100.0, 0.0, 1e-3)
    b = ns.generate_ic(0, 100_000, 5, 100.0, 0.05, 1e-3)
    scaled = b[2] != a[2]
    assert scaled.mean() == pytest.approx(0.05, abs=0.003)
    np.testing.assert_allclose(b[2][scaled], a[2][scaled] * 1e-3, rtol=1e-15)
    np.testing.assert_array_equal(a[0], b[0])


def test_mixed_ic_alternates_recipes():
    mix = ns.generate_ic(2, 10_000, 3, 100.0, 0.0, 1e-3)
    pl = ns.generate_ic(0, 10_000, 3, 100.0, 0.0, 1e-3)
    he = ns.generate_ic(1, 10_000, 3, 100.0, 0.0, 1e-3)
    np.testing.assert_array_equal(mix[0][0::2], pl[0][0::2])
    np.testing.assert_array_equal(mix[0][1::2], he[0][1::2])


def test_chain_floor_rejects_bad_arguments():
    """Refused on the host before any device call."""
    import ctypes
    from nsphere_amd._lib import NS_ERR_ARG, lib
    out = ctypes.c_double()
    assert lib.ns_gpu_chain_floor(0, 10, ctypes.byref(out)) == NS_ERR_ARG
    assert lib.ns_gpu_chain_floor(0, 1 << 40, ctypes.byref(out)) == NS_ERR_ARG
    assert lib.ns_gpu_chain_floor(0, 1 << 20, None) == NS_ERR_ARG
