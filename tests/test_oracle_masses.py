"""Oracle for per-particle masses (an extension: the reference's particles all
weigh M/npts and M(r) comes from the rank, nsphere.c:513).

Pins: the enclosed-mass prefix against exact integer sums and the closed form
i*m; the force with an enclosed-mass array against the rank formula; the
energy with masses against the brute-force pairwise shell potential and
against the equal-mass energy.
"""
import numpy as np

import oracle


def test_prefix_exact_on_integer_masses():
    rng = np.random.default_rng(3)
    m = rng.integers(1, 1 << 20, 100_003).astype(np.float64)  # sums stay below 2^53: exact
    out = oracle.mass_prefix(m)
    want = np.concatenate([[0], np.cumsum(m.astype(np.int64))[:-1]]).astype(np.float64)
    np.testing.assert_array_equal(out[:-2], want)
    assert out[-2] == float(m.astype(np.int64).sum())
    assert out[-1] == m[0]


def test_prefix_closed_form_equal_masses():
    n = 1_000_000
    m = np.full(n, 1.0 / n)
    out = oracle.mass_prefix(m)
    i = np.arange(n, dtype=np.float64)
    # i * fl(1/n), summed to ~2^-64 and rounded once: |fl(1/n) - 1/n| and the
    # final rounding each contribute at most 2^-53 relative
    np.testing.assert_allclose(out[:n], i / n, rtol=2.3e-16, atol=0)
    assert abs(out[n] - 1.0) < 1e-15


def test_force_with_menc_matches_rank_formula():
    n = 1000
    p = oracle.params()
    menc = np.concatenate([np.arange(n) / n, [1.0, 1.0 / n]])
    q = oracle.with_menc(p, menc)
    L = oracle.lib()
    for rank in (0, 1, 17, 999):
        for r in (1e-3, 0.7, 35.0):
            a = L.orc_gravitational_force(r, rank, n, oracle.ctypes.byref(p))
            b = L.orc_gravitational_force(r, rank, n, oracle.ctypes.byref(q))
            assert abs(a - b) <= 2e-16 * abs(a) + 1e-300


def test_energy_masses_equals_pairwise_shells():
    rng = np.random.default_rng(5)
    n = 1500
    r = np.sort(rng.uniform(0.01, 10.0, n))
    v = rng.normal(0, 0.5, n)
    ell = rng.uniform(0, 1, n)
    m = rng.uniform(0.2, 2.0, n) / n
    p = oracle.params(G=1.3)
    K, W = oracle.energy_masses(r, v, ell, m, p)
    k_want = np.sum(m * 0.5 * (v * v + ell * ell / (r * r)))
    # every pair interacts through the outer radius (shell theorem)
    outer = np.maximum.outer(r, r)
    w_want = -1.3 * np.sum(np.triu(np.outer(m, m) / outer, 1))
    assert abs(K - k_want) <= 1e-13 * abs(k_want)
    assert abs(W - w_want) <= 1e-12 * abs(w_want)


def test_energy_masses_equal_mass_limit():
    rng = np.random.default_rng(6)
    n = 2000
    r = np.sort(rng.uniform(0.01, 10.0, n))
    v = rng.normal(0, 0.5, n)
    ell = rng.uniform(0, 1, n)
    p = oracle.params()
    k0, w0 = oracle.energy(r, v, ell, p)
    k1, w1 = oracle.energy_masses(r, v, ell, np.full(n, 1.0 / n), p)
    assert abs(k1 - k0) <= 1e-14 * abs(k0)
    assert abs(w1 - w0) <= 1e-14 * abs(w0)
