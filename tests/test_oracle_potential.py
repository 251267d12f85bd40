"""Pins of the dynamic-potential oracle (oracle/potential.py)."""
import numpy as np

from oracle import potential


def test_closed_form_with_exact_sums():
    r = np.array([0.25, 0.5, 1.0, 2.0, 4.0])  # 1/r = 4, 2, 1, 0.5, 0.25: exact sums
    psi, psi0 = potential.psi_a(r, 2.0, 0.5, 10.0)  # dM = 2
    tail = np.array([3.75, 1.75, 0.75, 0.25, 0.0])  # sum of 1/r beyond each rank
    want = 2.0 * ((np.arange(1, 6) * 2.0) / r + 2.0 * tail) * 0.5
    np.testing.assert_array_equal(psi, want)
    assert psi0 == 2.0 * 2.0 * 7.75 * 0.5


def test_outermost_and_monotone():
    rng = np.random.default_rng(1)
    r = np.sort(rng.uniform(0.01, 50, 10_000))
    psi, psi0 = potential.psi_a(r, 4.3e-6, 1.0459e-6, 1e12)
    assert psi[-1] == 4.3e-6 * (1e12 / r[-1]) * 1.0459e-6  # all the mass inside, nothing outside
    assert np.all(np.diff(psi) < 0) and psi0 > psi[0]       # a potential well, deepest at r = 0
