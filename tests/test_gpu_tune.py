"""Launch-configuration knobs and the autotuner (ns_gpu_autotune).

Every knob (sort pass algorithm and tile, leapfrog levels per lane, the wave
kernel's grid) changes only how the work is laid out, never the result; the
autotuner must leave the resident state exactly as it found it and only the
chosen options behind.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def assert_state(a, b, what):
    for k, name in enumerate(("r", "v", "ell", "id")):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        bad = np.flatnonzero(x.view(np.uint64) != y.view(np.uint64)) if x.dtype == np.float64 else \
            np.flatnonzero(x != y)
        assert bad.size == 0, f"{what} {name}: {bad.size} differ"


def sorted_ic(ns, kind, n, seed):
    r, v, ell, ids = ns.generate_ic(kind, n, seed, 100.0, 0.05, 1e-3)
    q = oracle.sort_perm(r)
    return r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy()


@pytest.mark.parametrize("form", [1, 2])
def test_knobs_do_not_change_results(ns, form):
    from nsphere_amd import _lib
    state = sorted_ic(ns, 2, 120_000, 61)
    p = ns.unit_params()
    ref = None
    for lanes, blocks in [(4, 1024), (1, 64), (2, 4096), (6, 300)]:
        with ns.GpuContext(state[0].size) as ctx:
            ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
            ctx.set_option(ns.NS_OPT_BUDGET, 1 << 16)
            ctx.set_option(_lib.NS_OPT_LANE_LEVELS, lanes)
            ctx.set_option(_lib.NS_OPT_DEEP_BLOCKS, blocks)
            ctx.upload(*state)
            ctx.run(p, 4)
            got = (ctx.download(), ctx.levels())
        if ref is None:
            ref = got
        else:
            assert_state(got[0], ref[0], f"lanes {lanes} blocks {blocks}")
            np.testing.assert_array_equal(got[1], ref[1])


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_restores_state_and_keeps_results(ns, precision):
    state = sorted_ic(ns, 0, 200_000, 62)
    p = ns.unit_params()
    n = state[0].size
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_option(ns.NS_OPT_BUDGET, 1 << 16)
        ctx.upload(*state)
        ctx.run(p, 2)
        before = ctx.download()
        hits = ctx.cap_hits()
        res = ctx.autotune(p, steps=2)
        assert_state(ctx.download(), before, "after autotune")
        assert ctx.cap_hits() == hits
        ctx.run(p, 3)
        tuned = ctx.download()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_option(ns.NS_OPT_BUDGET, 1 << 16)
        ctx.upload(*state)
        ctx.run(p, 5)
        plain = ctx.download()
    assert_state(tuned, plain, "tuned vs default options")
    # float64: 3 forms of the sort between steps; 2 pass algorithms x 6 tile shapes, 4 lane-level
    # settings, 4 wave-kernel grids
    assert len(res["tried"]) == (20 if precision else 23) and all(t["ms_per_step"] > 0 for t in res["tried"])
    assert {t["sort_keys"] for t in res["tried"]} == ({1} if precision else {0, 1, 2})
    assert res["best"]["ms_per_step"] == min(t["ms_per_step"] for t in res["tried"])


def test_autotune_ensemble(ns):
    sizes = [40_000, 1_000, 25_000]
    parts = [sorted_ic(ns, h % 2, s, 70 + h) for h, s in enumerate(sizes)]
    base = np.cumsum([0] + sizes[:-1])
    state = [np.concatenate([pt[k] + (b if k == 3 else 0) for pt, b in zip(parts, base)]) for k in range(4)]
    state[3] = state[3].astype(np.uint32)
    p = ns.unit_params()
    with ns.GpuContext(state[0].size) as ctx:
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_halos(sizes)
        ctx.upload(*state)
        res = ctx.autotune(p, steps=2)
        assert all(t["sort_alg"] == 0 for t in res["tried"])  # an ensemble sorts by segments
        assert_state(ctx.download(), state, "ensemble after autotune")
