"""CPU oracle of the Rank_Mass_Rad_VRad files -- TEST INFRASTRUCTURE ONLY.

Restates nsphere.c:7043-7231 (unsorted and sorted arrays of a snapshot),
7698-7850 (PsiA, its linear spline with strictly increasing nodes, E) and
7911-7989 (the rows, written by fprintf_bin: int32 for %d, float32 for %f),
from a snapshot's records in id order (oracle/snapshot.py's APD_RECORD):
the sort by float32 radius keeps ties in id order, as glibc's merge-sorting
qsort leaves the id-ordered array it sorts.  psi comes from
oracle/potential.py and the density column from oracle/density.py.

Pins (tests/test_oracle_rankfiles.py): the node recurrence and the linear
lookup by hand and against numpy's interp, the row layout by struct.pack.
"""
from __future__ import annotations

import numpy as np

from . import density as od
from . import potential as op

UNSORTED = np.dtype([("rank", "<i4"), ("mass", "<f4"), ("r", "<f4"), ("v", "<f4"), ("psi", "<f4"),
                     ("E", "<f4"), ("ell", "<f4")])
SORTED = np.dtype(UNSORTED.descr + [("density", "<f4")])


def nodes(r_sorted, psi, psi0):
    """Spline nodes (0, psi0), (r_i, psi_i), forced strictly increasing (7760-7769)."""
    x = np.empty(r_sorted.size + 1)
    x[0] = 0.0
    x[1:] = r_sorted
    for i in range(1, x.size):
        if x[i] <= x[i - 1]:
            x[i] = x[i - 1] + 1e-5
    y = np.empty_like(x)
    y[0] = psi0
    y[1:] = psi
    return x, y


def linear(x, y, t):
    """The linear spline at t: x[lo] <= t < x[lo + 1] (last node: last interval)."""
    lo = np.clip(np.searchsorted(x, t, side="right") - 1, 0, x.size - 2)
    dx = x[lo + 1] - x[lo]
    return y[lo] + (t - x[lo]) / dx * (y[lo + 1] - y[lo])


def rank_files(records, G: float, vel_conv_sq: float, halo_mass: float):
    rec = np.asarray(records)
    n = rec.size
    dm = np.float64(halo_mass) / np.float64(n)
    rank = rec["rank"].astype(np.int64)
    r_un = rec["r"].astype(np.float64)
    v_un = rec["v"].astype(np.float64)
    l_un = rec["ell"].astype(np.float64)
    order = np.argsort(rec["r"], kind="stable")
    r_s, v_s, l_s = r_un[order], v_un[order], l_un[order]
    psi, psi0 = op.psi_a(r_s, G, vel_conv_sq, halo_mass)
    x, y = nodes(r_s, psi, psi0)
    un = np.empty(n, UNSORTED)
    un["rank"] = rank
    un["mass"] = (rank + 1).astype(np.float64) * dm
    un["r"], un["v"], un["ell"] = r_un, v_un, l_un
    p_un = linear(x, y, r_un)
    un["psi"] = p_un
    un["E"] = p_un - (l_un * l_un / (2.0 * r_un * r_un)) - 0.5 * v_un * v_un
    so = np.empty(n, SORTED)
    i = np.arange(n)
    so["rank"] = i
    so["mass"] = (i + 1).astype(np.float64) * dm
    so["r"], so["v"], so["ell"] = r_s, v_s, l_s
    p_s = linear(x, y, r_s)
    so["psi"] = p_s
    so["E"] = p_s - (l_s * l_s / (2.0 * r_s * r_s)) - 0.5 * v_s * v_s
    so["density"] = od.density(r_s, halo_mass)[0]
    return un, so
