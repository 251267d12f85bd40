"""NS_OPT_GRAPH: ns_gpu_run replays captured steps.

Without the second stream each graph launch replays two steps; an overlapped
run (binned forms, NS_OPT_OVERLAP) takes one step eagerly and then replays 16
overlapped steps per launch, the second stream's plan and late kernels as a
branch of the graph.  Every replay must equal the eagerly launched steps bit
for bit: several runs in a row with downloads, energies and option changes
between them (each forces a new capture or a replay of the existing one), odd
and even step counts, both precisions, an ensemble, and per-particle masses
(whose scan keeps its lookback epoch in device memory so that replays stay
valid).
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
        if x.dtype == np.float64:
            bad = np.flatnonzero(x.view(np.uint64) != y.view(np.uint64))
        else:
            bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{what} {name}: {bad.size} differ"


def sorted_ic(ns, kind, n, seed):
    r, v, ell, ids = ns.generate_ic(kind, n, seed, 100.0, 0.05, 1e-3)
    q = oracle.sort_perm(r)
    return r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy()


def run_both(ns, state, setup, schedule):
    """Run the schedule [(params, nsteps), ...] eagerly and with graphs."""
    out = []
    for graph in (0, 1):
        n = state[0].size
        with ns.GpuContext(n) as ctx:
            setup(ctx)
            ctx.set_option(ns.NS_OPT_GRAPH, graph)
            ctx.upload(*state)
            trace = []
            for p, k in schedule:
                ctx.run(p, k)
                trace.append((ctx.download(), ctx.energy(p), ctx.cap_hits()))
            out.append(trace)
    for i, (a, b) in enumerate(zip(*out)):
        assert_state(a[0], b[0], f"after run {i}")
        assert a[1] == b[1] and a[2] == b[2], f"after run {i}: energy / cap hits"


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("form", [1, 2])
def test_graph_replays_equal_eager(ns, precision, form):
    state = sorted_ic(ns, 0, 150_000, 31)
    p = ns.unit_params()
    p2 = p.replace(dt=5e-4)

    def setup(ctx):
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
        ctx.set_option(ns.NS_OPT_BUDGET, 1 << 16)
        ctx.set_option(ns.NS_OPT_OVERLAP, 0)

    run_both(ns, state, setup, [(p, 7), (p, 4), (p2, 5), (p, 3), (p, 1), (p, 6)])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("form", [1, 2])
def test_graph_overlapped_equal_eager(ns, precision, form):
    """Runs of 17 steps or more replay captured overlapped steps: 33 = 1 + 2 x 16
    (two launches, none left over), 20 and 40 (launches then eager steps), 17
    with another dt (a new capture), 16 and 2 (eager only, after replays)."""
    state = sorted_ic(ns, 0, 200_000, 33)
    p = ns.unit_params()
    p2 = p.replace(dt=5e-4)

    def setup(ctx):
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
        ctx.set_option(ns.NS_OPT_OVERLAP, 1)

    run_both(ns, state, setup, [(p, 33), (p, 20), (p2, 17), (p, 40), (p, 16), (p, 2)])


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_whole_run_prepared(ns, precision):
    """An even overlapped run is replayed as one graph, captured ahead by
    ns_gpu_graph_prepare (which runs nothing) or by the run itself; the state
    equals eager runs after every run, including runs whose keys were
    invalidated by an energy evaluation in between (a separate capture)."""
    state = sorted_ic(ns, 0, 200_000, 34)
    p = ns.unit_params()
    out = []
    for graph in (0, 1):
        with ns.GpuContext(state[0].size) as ctx:
            ctx.set_option(ns.NS_OPT_PRECISION, precision)
            ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
            ctx.set_option(ns.NS_OPT_GRAPH, graph)
            ctx.upload(*state)
            ctx.run(p, 5)
            ctx.graph_prepare(p, 20)
            before = ctx.download()  # the prepare ran nothing
            trace = [before]
            for k in (20, 20, 6, 20):
                ctx.run(p, k)
                trace.append(ctx.download())
            trace.append(ctx.energy(p))
            ctx.run(p, 20)
            trace.append(ctx.download())
            out.append(trace)
    for i, (a, b) in enumerate(zip(out[0][:-2], out[1][:-2])):
        assert_state(a, b, f"after run {i}")
    assert out[0][-2] == out[1][-2]
    assert_state(out[0][-1], out[1][-1], "after the energy and another run")


def test_graph_overlapped_ensemble(ns):
    sizes = [30_000, 5_000, 60_000]
    parts = [sorted_ic(ns, h % 2, s, 50 + h) for h, s in enumerate(sizes)]
    base = np.cumsum([0] + sizes[:-1])
    state = [np.concatenate([pt[k] + (b if k == 3 else 0) for pt, b in zip(parts, base)]) for k in range(4)]
    state[3] = state[3].astype(np.uint32)
    p = ns.unit_params()

    def ens(ctx):
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_halos(sizes, [1.0, 2.0, 0.5])

    run_both(ns, state, ens, [(p, 18), (p, 35)])


def test_graph_ensemble_and_masses(ns):
    sizes = [20_000, 3_000, 45_000, 7_777]
    parts = [sorted_ic(ns, h % 2, s, 40 + h) for h, s in enumerate(sizes)]
    base = np.cumsum([0] + sizes[:-1])
    state = [np.concatenate([pt[k] + (b if k == 3 else 0) for pt, b in zip(parts, base)]) for k in range(4)]
    state[3] = state[3].astype(np.uint32)
    p = ns.unit_params()

    def ens(ctx):
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_option(ns.NS_OPT_OVERLAP, 0)
        ctx.set_halos(sizes, [1.0, 2.0, 0.5, 1.0])

    run_both(ns, state, ens, [(p, 5), (p, 4)])

    n = state[0].size
    mass = np.random.default_rng(5).uniform(0.5, 1.5, n) / n

    def masses(ctx):
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_option(ns.NS_OPT_PRECISION, 1)
        ctx.upload(*state)
        ctx.set_masses(mass)

    run_both(ns, state, masses, [(p, 6), (p, 3)])
