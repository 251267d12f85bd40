"""Per-particle masses: the enclosed-mass scan M(r) on the GPU (decoupled
lookback; float64 for float64 state, float32 two-sum pairs for float32 state)
and the step that uses it.

The reference has no per-particle masses (every particle weighs M/npts and
M_enc comes from the rank, nsphere.c:513); the extension's oracle is
oracle.mass_prefix (compensated, rounded once) for M(r) and the strict oracle
step run with the GPU's own M(r) array for the integration, which must then
agree bit for bit.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TILE = 4096  # masses per scan tile


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def sorted_ic(ns, kind, n, seed, lowl=0.05):
    r, v, ell, ids = ns.generate_ic(kind, n, seed, 100.0 if kind != 1 else 200.0, lowl, 1e-3)
    q = oracle.sort_perm(r)
    return r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy()


def assert_bits(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:
        bad = np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))
    else:
        bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{what}: {bad.size} of {a.size} differ, first at {bad[0]}: {a[bad[0]]!r} vs {b[bad[0]]!r}"


def scan_on_gpu(ns, m, precision):
    """M(r) of rank-ordered masses m (the state is sorted, so integrate scans in place)."""
    n = m.size
    r = np.linspace(1.0, 2.0, n)
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.upload(r, np.zeros(n), np.full(n, 0.5), np.arange(n, dtype=np.uint32))
        ctx.set_masses(m)
        p = ns.unit_params()
        ctx.set_option(ns.NS_OPT_BUDGET, 64)
        ctx.integrate(p)
        return ctx.enclosed_mass()


# bound on |M_gpu - M_exact| / M_total: float64 -- one rounding per tile
# step of the within-tile tree (~24 additions) plus one fixed-point
# truncation (2^-61 of the total) per preceding tile; float32 pairs -- the
# same with the pair's 2^-46 relative error per addition.
def bound(n, precision):
    tiles = (n + TILE - 1) // TILE
    per_add = 2.0 ** -46 if precision else 2.0 ** -53
    return 32 * per_add + tiles * 2.0 ** -61 + 2.0 ** -52


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [1, 63, 4096, 4097, 100_003, 1_200_000])
def test_scan_against_exact_prefix(ns, precision, n):
    rng = np.random.default_rng(n + precision)
    m = rng.uniform(0.5, 1.5, n) / n
    if precision:
        m = m.astype(np.float32).astype(np.float64)  # what the float32 state holds
    got = scan_on_gpu(ns, m, precision)
    want = oracle.mass_prefix(m)
    total = want[n]
    err = np.max(np.abs(got[: n + 1] - want[: n + 1])) / total
    assert err <= bound(n, precision), f"max error {err:.3e} of the total"
    assert got[n + 1] == m[0]
    assert got[0] == 0.0


def test_float32_pairs_beat_plain_float32(ns):
    n = 1_200_000
    m = (np.random.default_rng(9).uniform(0.5, 1.5, n) / n).astype(np.float32)
    got = scan_on_gpu(ns, m.astype(np.float64), 1)
    want = oracle.mass_prefix(m.astype(np.float64))
    naive = np.concatenate([[0], np.cumsum(m, dtype=np.float32)[:-1]]).astype(np.float64)
    e_pair = np.max(np.abs(got[:n] - want[:n]))
    e_naive = np.max(np.abs(naive - want[:n]))
    assert e_pair < 1e-12 and e_naive > 1e-6 and e_pair < 1e-5 * e_naive, (e_pair, e_naive)


def test_equal_masses_match_rank_formula(ns):
    n = 300_000
    got = scan_on_gpu(ns, np.full(n, 1.0 / n), 0)
    np.testing.assert_allclose(got[:n], np.arange(n) / n, rtol=0, atol=bound(n, 0))


@pytest.mark.parametrize("precision", [0, 1])
def test_scan_reproducible(ns, precision):
    n = 2_000_003
    m = np.random.default_rng(4).uniform(0.1, 10.0, n)
    a = scan_on_gpu(ns, m, precision)
    b = scan_on_gpu(ns, m, precision)
    assert_bits(a, b, "M(r) run to run")


@pytest.mark.parametrize("form", [0, 1])
def test_steps_with_masses_equal_oracle(ns, form):
    """Strict forms with per-particle masses: every step's positions, velocities
    and levels equal the oracle's step run with the GPU's M(r); the GPU's M(r)
    is the exact prefix of the rank-ordered masses within the bound; the masses
    travel with their particles."""
    n = 120_000
    r, v, ell, ids = sorted_ic(ns, 2, n, 71)
    mass_by_id = np.random.default_rng(72).uniform(0.2, 3.0, n) / n
    p = ns.unit_params()
    po = oracle.from_step_params(p)
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
        ctx.upload(r, v, ell, ids)
        ctx.set_masses(mass_by_id[ids])
        for s in range(3):
            ctx.step(p)
            gr, gv, gl, gid = ctx.download()
            menc = ctx.enclosed_mass()
            perm = oracle.sort_perm(r)
            want_m = oracle.mass_prefix(mass_by_id[ids[perm]])
            assert np.max(np.abs(menc - want_m)) <= bound(n, 0) * want_m[n], f"step {s}: M(r)"
            lv, _ = oracle.step(r, v, ell, ids, oracle.with_menc(po, menc))
            assert_bits(gid, ids, f"step {s}: id")
            assert_bits(gr, r, f"step {s}: r")
            assert_bits(gv, v, f"step {s}: v")
            assert_bits(ctx.levels(), lv, f"step {s}: levels")


def test_energy_with_masses(ns):
    n = 150_000
    r, v, ell, ids = sorted_ic(ns, 0, n, 81)
    m = np.random.default_rng(82).uniform(0.5, 2.0, n) / n
    p = ns.unit_params()
    with ns.GpuContext(n) as ctx:
        ctx.upload(r, v, ell, ids)
        ctx.set_masses(m)
        K, W = ctx.energy(p)
        ctx.set_masses(None)
        K0, W0 = ctx.energy(p)
    k_want, w_want = oracle.energy_masses(r, v, ell, m, oracle.from_step_params(p))
    k0, w0 = oracle.energy(r, v, ell, oracle.from_step_params(p))
    assert abs(K - k_want) <= 1e-12 * abs(k_want)
    assert abs(W - w_want) <= 1e-12 * abs(w_want)
    assert abs(K0 - k0) <= 1e-12 * abs(k0) and abs(W0 - w0) <= 1e-12 * abs(w0)
    assert abs(W - W0) > 1e-6 * abs(W0)  # the masses do change the potential


def test_masses_rejected(ns):
    n = 10
    with ns.GpuContext(n) as ctx:
        ctx.upload(np.linspace(1, 2, n), np.zeros(n), np.ones(n), np.arange(n, dtype=np.uint32))
        for bad in (np.full(n, -1.0), np.full(n, np.nan), np.zeros(n)):
            with pytest.raises(ns.NsError):
                ctx.set_masses(bad)
        with pytest.raises(ns.NsError):
            ctx.enclosed_mass()  # none set


@pytest.mark.parametrize("precision", [0, 1])
def test_upload_after_odd_steps_keeps_masses_by_position(ns, precision):
    """After set_masses and an odd number of steps the masses sit in the
    second state buffer; an upload must carry them to the first (masses stay
    by position), so the next step equals a fresh context given the same
    masses in that order."""
    n = 60_000
    rng = np.random.default_rng(77)
    m = rng.uniform(0.5, 1.5, n) / n
    r, v, ell, ids = ns.generate_ic(0, n, 71, 100.0, 0.05, 1e-3)
    state2 = ns.generate_ic(0, n, 72, 100.0, 0.05, 1e-3)
    p = ns.unit_params()
    with ns.GpuContext(n) as a:
        a.set_option(ns.NS_OPT_PRECISION, precision)
        a.set_option(ns.NS_OPT_INTEGRATOR, ns.FORM_BINNED)
        a.upload(r, v, ell, ids)
        a.set_masses(m)
        a.step(p)
        a.sync()
        id1 = a.download()[3]
        a.upload(*state2)
        a.step(p)
        a.sync()
        got = a.download()
        ka, wa = a.energy(p)
    with ns.GpuContext(n) as b:
        b.set_option(ns.NS_OPT_PRECISION, precision)
        b.set_option(ns.NS_OPT_INTEGRATOR, ns.FORM_BINNED)
        b.upload(*state2)
        b.set_masses(m[id1])  # ids were 0..n-1 at set_masses: position k then held mass m[id1[k]]
        b.step(p)
        b.sync()
        want = b.download()
        kb, wb = b.energy(p)
    for x, y, what in zip(got, want, ("r", "v", "ell", "id")):
        assert_bits(x, y, what)
    assert (ka, wa) == (kb, wb)
