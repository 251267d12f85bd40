"""CPU oracle of the dynamic potential, method A -- TEST INFRASTRUCTURE ONLY.

Restates nsphere.c:7698-7738 literally (np.cumsum adds sequentially, so the
suffix recurrence suffix[j] = suffix[j + 1] + 1/r_j and the forward sum of
PsiA(0) are the reference's own roundings):
  psiA[i] = G ((i + 1) dM / r_i + dM suffix[i + 1]) vel_conv_sq  (0 outer term
  for the last rank), PsiA0 = G dM (sum 1/r) vel_conv_sq, dM = M / npts
  (nsphere.c:5179), M_r = (rank + 1) dM (nsphere.c:7224).

Pinned by tests/test_oracle_potential.py: radii whose reciprocals and sums
are exact in binary (so every value has a closed form), and the outermost
value G M / r_max.
"""
from __future__ import annotations

import numpy as np


def psi_a(r_sorted, G: float, vel_conv_sq: float, halo_mass: float):
    r = np.asarray(r_sorted, np.float64)
    n = r.size
    dm = np.float64(halo_mass) / np.float64(n)
    inv = 1.0 / r
    suffix = np.cumsum(inv[::-1])[::-1]
    outer = np.zeros(n)
    outer[:-1] = dm * suffix[1:]
    m_r = np.arange(1, n + 1, dtype=np.float64) * dm
    psi = np.float64(G) * (m_r / r + outer)
    psi = psi * np.float64(vel_conv_sq)
    psi0 = np.float64(G) * dm * np.cumsum(inv)[-1] * np.float64(vel_conv_sq)
    return psi, psi0
