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


def run_driver(tmp_path, cols_in, dt, steps, env_extra=None):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/c/dropin_driver is not built (run __graft_entry__.build())")
    n = cols_in[0].size
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([n], np.int64).tofile(f)
        for a in cols_in:
            a.astype(np.float64).tofile(f)
    env = dict(os.environ)
    env.pop("NSPHERE_GPU_BUDGET", None)
    env.update(env_extra or {})
    proc = subprocess.run([DRIVER, str(inp), str(out), repr(dt), str(steps)], capture_output=True, text=True,
                          timeout=120, env=env)
    assert proc.returncode == 0, proc.stderr
    raw = np.fromfile(out, dtype=np.uint8)
    assert int(raw[:8].view(np.int64)[0]) == n
    cols = raw[8:8 + 40 * n].view(np.float64).reshape(5, n)
    inv = raw[8 + 40 * n:].view(np.int32)
    return cols, inv, proc.stderr


def oracle_run(r, v, ell, idf, mu, dt, steps, budget=0):
    """The reference's sequence on the oracle: exact initial sort, then steps;
    mu moves with every reordering (nsphere.c:8613-8647)."""
    ids = idf.astype(np.uint32)
    q = oracle.sort_perm(r)
    r, v, ell, ids, mu = r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy(), mu[q].copy()
    p = oracle.reference_units(dt=dt, budget=budget)
    inv_o = None
    for _ in range(steps):
        pos = np.argsort(ids, kind="stable")  # position of each id before the step
        _, inv_o = oracle.step(r, v, ell, ids, p)
        mu = mu[pos[ids]]                    # the step's own reordering of every column
    return r, v, ell, ids, mu, inv_o


def test_c_driver_matches_oracle(tmp_path):
    n, steps, dt = 50_000, 3, 0.04
    r, v, ell, idf = scaled_ic(n, 5)
    mu = np.random.default_rng(5).uniform(-1.0, 1.0, n)
    cols, inv, err = run_driver(tmp_path, (r, v, ell, idf, mu), dt, steps)
    assert "NSPHERE_GPU_BUDGET" not in err
    r, v, ell, ids, mu, inv_o = oracle_run(r, v, ell, idf, mu, dt, steps)
    np.testing.assert_array_equal(cols[3], ids.astype(np.float64))
    np.testing.assert_array_equal(cols[0].view(np.uint64), r.view(np.uint64))
    np.testing.assert_array_equal(cols[1].view(np.uint64), v.view(np.uint64))
    np.testing.assert_array_equal(cols[2].view(np.uint64), ell.view(np.uint64))
    np.testing.assert_array_equal(cols[4].view(np.uint64), mu.view(np.uint64))
    np.testing.assert_array_equal(inv, inv_o)


def test_c_driver_budget_reports_capped_steps(tmp_path):
    """NSPHERE_GPU_BUDGET caps the work as the library's NS_OPT_BUDGET does:
    results equal the capped oracle and every capping step is reported on
    stderr (the reference has no cap, nsphere.c:1718-1734)."""
    n, steps, dt, budget = 20_000, 2, 0.04, 64
    r, v, ell, idf = scaled_ic(n, 6)
    mu = np.random.default_rng(6).uniform(-1.0, 1.0, n)
    cols, inv, err = run_driver(tmp_path, (r, v, ell, idf, mu), dt, steps, {"NSPHERE_GPU_BUDGET": str(budget)})
    r, v, ell, ids, mu, inv_o = oracle_run(r, v, ell, idf, mu, dt, steps, budget=budget)
    np.testing.assert_array_equal(cols[0].view(np.uint64), r.view(np.uint64))
    np.testing.assert_array_equal(cols[1].view(np.uint64), v.view(np.uint64))
    np.testing.assert_array_equal(cols[4].view(np.uint64), mu.view(np.uint64))
    np.testing.assert_array_equal(inv, inv_o)
    lines = [x for x in err.splitlines() if "NSPHERE_GPU_BUDGET=64" in x]
    assert len(lines) == steps, err


def test_python_mirror_sort_particles():
    import nsphere_amd as ns
    n = 30_000
    r, v, ell, idf = scaled_ic(n, 9)
    mu = np.random.default_rng(9).uniform(-1.0, 1.0, n)
    particles = np.stack([r, v, ell, idf, mu])
    ns.sort_particles(particles, n)
    q = oracle.sort_perm(r)
    np.testing.assert_array_equal(particles[0], r[q])
    np.testing.assert_array_equal(particles[3], idf[q])
    np.testing.assert_array_equal(particles[4], mu[q])  # all five columns move (nsphere.c:8613-8647)


def test_python_mirror_method5_step():
    import nsphere_amd as ns
    n = 30_000
    r, v, ell, idf = scaled_ic(n, 10)
    q = oracle.sort_perm(r)
    mu = np.random.default_rng(10).uniform(-1.0, 1.0, n)
    particles = np.stack([r[q], v[q], ell[q], idf[q], mu])
    p = ns.default_params(dt=0.04)
    inv = np.empty(n, np.int32)
    lv = ns.method5_step(particles, n, p, inv)
    rr, vv, ll, ii = r[q].copy(), v[q].copy(), ell[q].copy(), idf[q].astype(np.uint32)
    pos = np.argsort(ii, kind="stable")
    lv_o, inv_o = oracle.step(rr, vv, ll, ii, oracle.from_step_params(p))
    np.testing.assert_array_equal(particles[0], rr)
    np.testing.assert_array_equal(particles[1], vv)
    np.testing.assert_array_equal(particles[4], mu[pos[ii]])
    np.testing.assert_array_equal(lv, lv_o)
    np.testing.assert_array_equal(inv, inv_o)


def test_permutation_entry_points():
    """ns_gpu_permutation / ns_gpu_permute_column after a sort, a step and an
    overlapped run; the energy's own sort leaves the permutation alone."""
    import nsphere_amd as ns
    n = 40_000
    r, v, ell, ids = ns.generate_ic(0, n, 31, 100.0, 0.05, 1e-3)
    col = np.random.default_rng(31).standard_normal(n)
    p = ns.unit_params()
    with ns.GpuContext(n) as ctx:
        ctx.upload(r, v, ell, ids)
        with pytest.raises(ns.NsError):
            ctx.permutation()
        ctx.sort_particles()
        perm = ctx.permutation()
        np.testing.assert_array_equal(perm, oracle.sort_perm(r))
        np.testing.assert_array_equal(ctx.permute_column(col), col[perm])
        ctx.set_option(ns.NS_OPT_INTEGRATOR, ns.FORM_BINNED_FAST)
        _, _, _, id0 = ctx.download()
        ctx.step(p)
        ctx.sync()
        _, _, _, id1 = ctx.download()
        perm = ctx.permutation()
        np.testing.assert_array_equal(id0[perm], id1)
        ctx.energy(p)
        np.testing.assert_array_equal(ctx.permutation(), perm)
        state = ctx.download()
        # an overlapped run of three: the permutation of its last step, i.e.
        # from the state after two steps (the same in a run of two)
        with ns.GpuContext(n) as two:
            two.upload(*state)
            two.set_option(ns.NS_OPT_INTEGRATOR, ns.FORM_BINNED_FAST)
            two.run(p, 2)
            two.sync()
            _, _, _, id_before = two.download()
        ctx.run(p, 3)
        ctx.sync()
        _, _, _, id_after = ctx.download()
        np.testing.assert_array_equal(id_before[ctx.permutation()], id_after)
