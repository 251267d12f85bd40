"""Overlapped runs (NS_OPT_OVERLAP=1): the particles whose previous step was
slow (not converged, a leapfrog level of 4 or deeper or a Levi-Civita tau
step at level 2 or deeper; else among a halo's 64 innermost ranks) are
integrated on a second stream while the rest of the step runs and
the next step's radii are sorted, then merged back at their ranks.  The merge must reproduce the
stable sort of the true keys, so every result is bit-identical to the run
with every kernel on one stream -- including exact radius ties, more deferred
particles than the 1024 slots, float32 state, and the work counters.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def assert_bits(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:
        bad = np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))
    else:
        bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{what}: {bad.size} differ, first at {bad[0]}: {a[bad[0]]!r} vs {b[bad[0]]!r}"


def plunging_ic(ns, n, seed, ties=False, ell_range=(1e-6, 1e-4)):
    """Plummer with a 20 % low-ell tail and particles on radial plunges near
    the centre (deep chains that hit a small work cap).  With ell in
    [5e-3, 2e-2] the plunging particles sit inside r_crit and take the
    Levi-Civita branch."""
    r, v, ell, ids = ns.generate_ic(0, n, seed, 100.0, 0.2, 1e-3)
    rng = np.random.default_rng(seed)
    k = n // 100
    pick = rng.choice(n, k, replace=False)
    r[pick] = rng.uniform(1e-4, 3e-3, k)
    v[pick] = rng.uniform(-0.6, 0.6, k)
    ell[pick] = rng.uniform(*ell_range, k)
    if ties:
        r = np.round(r, 3) + 1e-3  # a 1e-3 grid: many exact ties, all positive
    return r, v, ell, ids


def run(ns, state, overlap, form, precision, budget, steps, profile=False):
    r, v, ell, ids = state
    n = r.size
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_BUDGET, budget)
        ctx.set_option(ns.NS_OPT_OVERLAP, overlap)
        ctx.set_option(ns.NS_OPT_PROFILE, int(profile))
        ctx.upload(r, v, ell, ids)
        ctx.run(ns.unit_params(), steps)
        ctx.sync()
        out = dict(state=ctx.download(), levels=ctx.levels(), hits=ctx.cap_hits())
        if profile:
            out["evals"] = ctx.force_evals()
            out["stats"] = ctx.kernel_stats()
            out["chain"] = ctx.chain_steps()
            ctx.reset_kernel_stats()
            out["chain_after_reset"] = ctx.chain_steps()
    return out


def assert_same(a, b):
    for x, y, what in zip(a["state"], b["state"], ("r", "v", "ell", "id")):
        assert_bits(x, y, what)
    assert_bits(a["levels"], b["levels"], "levels")
    assert a["hits"] == b["hits"]


@pytest.mark.parametrize("form,precision", [(1, 0), (2, 0), (2, 1)])
@pytest.mark.parametrize("budget", [64, 4096])
def test_overlap_is_bit_identical(ns, form, precision, budget):
    state = plunging_ic(ns, 30_000, 5)
    one = run(ns, state, 0, form, precision, budget, 6)
    two = run(ns, state, 1, form, precision, budget, 6)
    assert one["hits"] > 0
    assert_same(one, two)


def test_overlap_more_deferred_than_slots(ns):
    # a budget of 2 stops every particle at its first level: all are flagged
    # not converged, far more than the 1024 slots of the slow list
    state = plunging_ic(ns, 50_000, 9)
    one = run(ns, state, 0, 2, 0, 2, 4)
    two = run(ns, state, 1, 2, 0, 2, 4)
    assert one["hits"] >= 4 * 50_000 - 4096
    assert_same(one, two)


@pytest.mark.parametrize("precision", [0, 1])
def test_overlap_with_radius_ties(ns, precision):
    state = plunging_ic(ns, 40_000, 21, ties=True)
    one = run(ns, state, 0, 2, precision, 64, 6)
    two = run(ns, state, 1, 2, precision, 64, 6)
    assert_same(one, two)


def test_overlap_counters_and_timing(ns):
    state = plunging_ic(ns, 30_000, 13)
    one = run(ns, state, 0, 2, 0, 4096, 5, profile=True)
    two = run(ns, state, 1, 2, 0, 4096, 5, profile=True)
    assert_same(one, two)
    # a deferred leapfrog particle's lane-per-particle attempt is counted with
    # the late kernel, so only the total is the same
    assert sum(one["evals"].values()) == sum(two["evals"].values())
    assert two["stats"]["integrate_late"][0] == 5
    assert two["stats"]["overlap_plan"][0] == 5
    assert two["stats"]["overlap_place"][0] == 5
    assert "integrate_late" not in one["stats"]


def test_overlap_chain_counter(ns):
    """The sum over late launches of each launch's longest Levi-Civita chain.
    A wave's tau step that stops at level K runs 4 * 2^K + 1 micro-steps one
    after another and is charged at least that many force evaluations, so the
    sum is bounded by the late kernel's force evaluations and, per launch, by
    one particle's work cap."""
    # 157 Levi-Civita particles, all stopped by the cap (checked on the oracle)
    state = plunging_ic(ns, 30_000, 13, ell_range=(5e-3, 2e-2))
    budget, steps = 4096, 5
    one = run(ns, state, 0, 2, 0, budget, steps, profile=True)
    two = run(ns, state, 1, 2, 0, budget, steps, profile=True)
    assert one["chain"] == 0
    assert 0 < two["chain"] <= two["evals"]["late"]
    assert two["chain"] <= steps * budget
    assert two["chain_after_reset"] == 0


def test_overlap_strict_matches_capped_oracle(ns):
    n = 20_000
    r, v, ell, ids = plunging_ic(ns, n, 17)
    p = ns.unit_params()
    budget = 4096
    got = run(ns, (r, v, ell, ids), 1, 1, 0, budget, 4)
    po = oracle.from_step_params(p, budget)
    capped = 0
    for _ in range(4):
        lv, _ = oracle.step(r, v, ell, ids, po)
        capped += int(((lv & oracle.LVL_NOCONV) != 0).sum())
    assert capped > 0
    for x, y, what in zip(got["state"], (r, v, ell, ids), ("r", "v", "ell", "id")):
        assert_bits(x, y, what)
    assert_bits(got["levels"], lv, "levels")


def test_steps_after_an_overlapped_run(ns):
    # the last step's late ranks get their true radius keys back, so a plain
    # step after the run sorts as if the run had not overlapped
    state = plunging_ic(ns, 30_000, 3)
    r, v, ell, ids = state
    p = ns.unit_params()
    outs = []
    for overlap in (0, 1):
        with ns.GpuContext(r.size) as ctx:
            ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
            ctx.set_option(ns.NS_OPT_BUDGET, 64)
            ctx.set_option(ns.NS_OPT_OVERLAP, overlap)
            ctx.upload(r, v, ell, ids)
            ctx.run(p, 4)
            ctx.step(p)
            ctx.run(p, 3)
            outs.append(ctx.download())
    for x, y, what in zip(outs[0], outs[1], ("r", "v", "ell", "id")):
        assert_bits(x, y, what)


@pytest.mark.parametrize("n", [1, 2, 3, 1025])
def test_overlap_tiny(ns, n):
    state = plunging_ic(ns, max(n, 200), 7)
    state = tuple(x[:n].copy() for x in state)
    state = (state[0], state[1], state[2], np.arange(n, dtype=np.uint32))
    one = run(ns, state, 0, 2, 0, 2, 3)
    two = run(ns, state, 1, 2, 0, 2, 3)
    assert_same(one, two)


def test_chunked_runs_seed_from_levels(ns):
    # a run of short runs: each one seeds its slow list from the level records
    # the previous call left
    state = plunging_ic(ns, 30_000, 31)
    r, v, ell, ids = state
    p = ns.unit_params()
    outs = []
    for overlap in (0, 1):
        with ns.GpuContext(r.size) as ctx:
            ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
            ctx.set_option(ns.NS_OPT_BUDGET, 256)
            ctx.set_option(ns.NS_OPT_OVERLAP, overlap)
            ctx.upload(r, v, ell, ids)
            for _ in range(5):
                ctx.run(p, 2)
            outs.append((ctx.download(), ctx.levels(), ctx.cap_hits()))
    for x, y, what in zip(outs[0][0], outs[1][0], ("r", "v", "ell", "id")):
        assert_bits(x, y, what)
    assert_bits(outs[0][1], outs[1][1], "levels")
    assert outs[0][2] == outs[1][2] > 0


def test_profile_kernel_mask(ns):
    state = plunging_ic(ns, 20_000, 41)
    with ns.GpuContext(20_000) as ctx:
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_option(ns.NS_OPT_BUDGET, 4096)
        ctx.upload(*state)
        fam = ctx.kernel_families()
        assert fam.index("integrate_late") >= 0 and len(fam) <= 32
        ctx.set_option(ns.NS_OPT_PROFILE, 1)
        ctx.profile_kernels(["sort_scatter", "integrate_late"])
        ctx.run(ns.unit_params(), 3)
        st = ctx.kernel_stats()
        assert set(st) == {"sort_scatter", "integrate_late"}
        # float64 keys: four radix passes on the high halves per sort
        assert st["integrate_late"][0] == 3 and st["sort_scatter"][0] == 3 * 4
        with pytest.raises(ValueError):
            ctx.profile_kernels(["no_such_kernel"])
        ctx.profile_kernels(None)
        ctx.reset_kernel_stats()
        ctx.run(ns.unit_params(), 2)
        assert {"integrate_bulk", "integrate_deep", "overlap_plan", "overlap_place"} <= set(ctx.kernel_stats())


def ensemble_state(ns, sizes, seed, ties=False):
    parts = [plunging_ic(ns, s, seed + h, ties) if s >= 100 else
             tuple(x[:s].copy() for x in plunging_ic(ns, 100, seed + h, ties)) for h, s in enumerate(sizes)]
    base = np.cumsum([0] + list(sizes[:-1]))
    r = np.concatenate([p[0] for p in parts])
    v = np.concatenate([p[1] for p in parts])
    ell = np.concatenate([p[2] for p in parts])
    ids = np.arange(r.size, dtype=np.uint32)
    return r, v, ell, ids


@pytest.mark.parametrize("budget,ties,form,precision",
                         [(64, False, 2, 0), (64, True, 2, 0), (2, False, 2, 0), (64, False, 2, 1), (64, False, 1, 0)])
def test_overlap_ensemble_is_bit_identical(ns, budget, ties, form, precision):
    # halos sort as segments of their own: the late particles of each halo
    # come back inside their halo, and a budget of 2 makes every particle slow
    sizes = [20_000, 1, 5_000, 777, 15_000, 3]
    state = ensemble_state(ns, sizes, 90, ties)
    outs = []
    for overlap in (0, 1):
        with ns.GpuContext(state[0].size) as ctx:
            ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
            ctx.set_option(ns.NS_OPT_PRECISION, precision)
            ctx.set_option(ns.NS_OPT_BUDGET, budget)
            ctx.set_option(ns.NS_OPT_OVERLAP, overlap)
            ctx.set_halos(sizes)
            ctx.upload(*state)
            ctx.run(ns.unit_params(), 3)
            ctx.run(ns.unit_params(), 3)
            outs.append(dict(state=ctx.download(), levels=ctx.levels(), hits=ctx.cap_hits()))
    assert outs[0]["hits"] > 0
    assert_same(outs[0], outs[1])


def test_chain_floor_measures_one_wave(ns):
    """ns_gpu_chain_floor: one wave alone running the late kernel's micro-step;
    a few tens of ns per micro-step on MI355X, and stable between calls."""
    a = ns.chain_floor(0, 1 << 20)
    b = ns.chain_floor(0, 1 << 21)
    assert 5.0 < a < 200.0 and 5.0 < b < 200.0
    assert abs(a - b) < 0.2 * max(a, b)
