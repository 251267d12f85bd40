"""The reference's own C signatures (include/nsphere_dropin.h) from C and Python.

tests/c/dropin_driver is a plain C program that calls sort_particles_with_alg
and nsphere_method5_step on a caller-owned particles[5][npts] array exactly as
the reference's main() would (nsphere.c:5339-5342, 5775-5838), in the
reference's kpc / Myr / Msun units.  Its output must equal the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c", "dropin_driver")


def scaled_ic(n, seed):
    from nsphere_amd import generate_ic
    r, v, ell, ids = generate_ic(0, n, seed, 100.0, 0.05, 1e-3)
    gm = 1.02271e-3 ** 2 * 4.3e-6 * 1e12   # VEL_CONV_SQ * G_CONST * HALO_MASS
    L = 20.0                                # kpc
    vs = np.sqrt(gm / L)
    return r * L, v * vs, ell * L * vs, ids.astype(np.float64)


def test_c_driver_matches_oracle(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/c/dropin_driver is not built (run __graft_entry__.build())")
    n, steps, dt = 50_000, 3, 0.04
    r, v, ell, idf = scaled_ic(n, 5)
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([n], np.int64).tofile(f)
        for a in (r, v, ell, idf):
            a.astype(np.float64).tofile(f)
    proc = subprocess.run([DRIVER, str(inp), str(out), repr(dt), str(steps)], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stderr
    raw = np.fromfile(out, dtype=np.uint8)
    assert int(raw[:8].view(np.int64)[0]) == n
    cols = raw[8:8 + 32 * n].view(np.float64).reshape(4, n)
    inv = raw[8 + 32 * n:].view(np.int32)

    ids = idf.astype(np.uint32)
    q = oracle.sort_perm(r)
    r, v, ell, ids = r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy()
    p = oracle.reference_units(dt=dt)
    for _ in range(steps):
        _, inv_o = oracle.step(r, v, ell, ids, p)
    np.testing.assert_array_equal(cols[3], ids.astype(np.float64))
    np.testing.assert_array_equal(cols[0].view(np.uint64), r.view(np.uint64))
    np.testing.assert_array_equal(cols[1].view(np.uint64), v.view(np.uint64))
    np.testing.assert_array_equal(cols[2].view(np.uint64), ell.view(np.uint64))
    np.testing.assert_array_equal(inv, inv_o)


def test_python_mirror_sort_particles():
    import nsphere_amd as ns
    n = 30_000
    r, v, ell, idf = scaled_ic(n, 9)
    particles = np.stack([r, v, ell, idf, np.zeros(n)])
    mu_before = particles[4].copy()
    ns.sort_particles(particles, n)
    q = oracle.sort_perm(r)
    np.testing.assert_array_equal(particles[0], r[q])
    np.testing.assert_array_equal(particles[3], idf[q])
    np.testing.assert_array_equal(particles[4], mu_before)  # mu is not touched


def test_python_mirror_method5_step():
    import nsphere_amd as ns
    n = 30_000
    r, v, ell, idf = scaled_ic(n, 10)
    q = oracle.sort_perm(r)
    particles = np.stack([r[q], v[q], ell[q], idf[q], np.zeros(n)])
    p = ns.default_params(dt=0.04)
    inv = np.empty(n, np.int32)
    lv = ns.method5_step(particles, n, p, inv)
    rr, vv, ll, ii = r[q].copy(), v[q].copy(), ell[q].copy(), idf[q].astype(np.uint32)
    lv_o, inv_o = oracle.step(rr, vv, ll, ii, oracle.from_step_params(p))
    np.testing.assert_array_equal(particles[0], rr)
    np.testing.assert_array_equal(particles[1], vv)
    np.testing.assert_array_equal(lv, lv_o)
    np.testing.assert_array_equal(inv, inv_o)
