"""NS_OPT_TIEBREAK=1: exact radius ties ordered by particle id.

The sort becomes a stable sort on (radius key, id) -- the ids first, then the
radius keys gathered through the id order -- so its permutation must equal
numpy's lexsort on (id, radius) and the oracle's id-tiebreak sort, for every
pass algorithm and tile, in both precisions and per halo of an ensemble; a
method-5 step with it must equal the oracle's step with the same tiebreak bit
for bit, eagerly and by graph replay.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def tie_heavy(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.integers(1, 400, n) * 0.25                    # many exact ties
    r[rng.integers(0, n, n // 10)] = rng.uniform(0.1, 100, n // 10)
    ids = rng.permutation(n).astype(np.uint32)
    return r, ids


@pytest.mark.parametrize("alg", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [1, 1000, 77_777])
def test_tiebreak_sort_is_lexsort(ns, alg, precision, n):
    from nsphere_amd import _lib
    r, ids = tie_heavy(n, n + 3 * alg + precision)
    if precision:
        r = r.astype(np.float32).astype(np.float64)
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(_lib.NS_OPT_SORT_ALG, alg)
        ctx.set_option(ns.NS_OPT_TIEBREAK, 1)
        ctx.upload(r, np.zeros(n), np.zeros(n), ids)
        ctx.sort_particles()
        gr, _, _, gid = ctx.download()
    want = np.lexsort((ids, r))
    np.testing.assert_array_equal(gid, ids[want])
    np.testing.assert_array_equal(gr, r[want])
    np.testing.assert_array_equal(oracle.sort_perm(r, ids, tiebreak_id=True), want.astype(np.uint32))


def test_tiebreak_per_halo(ns):
    sizes = [5000, 1, 20_000, 777]
    parts = [tie_heavy(s, 40 + h) for h, s in enumerate(sizes)]
    base = np.cumsum([0] + sizes[:-1])
    r = np.concatenate([p[0] for p in parts])
    ids = np.concatenate([p[1] + b for p, b in zip(parts, base)]).astype(np.uint32)
    n = r.size
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_TIEBREAK, 1)
        ctx.set_halos(sizes)
        ctx.upload(r, np.zeros(n), np.zeros(n), ids)
        ctx.sort_particles()
        _, _, _, gid = ctx.download()
    for (pr, pid), b, s in zip(parts, base, sizes):
        np.testing.assert_array_equal(gid[b:b + s], (pid + b)[np.lexsort((pid, pr))])


@pytest.mark.parametrize("graph", [0, 1])
def test_tiebreak_steps_equal_oracle(ns, graph):
    n = 30_000
    r, v, ell, ids = ns.generate_ic(0, n, 8, 100.0, 0.05, 1e-3)
    r = np.round(r, 3) + 1e-3    # radii on a 1e-3 grid: many ties, all positive
    p = ns.unit_params()
    po = oracle.from_step_params(p)
    ro, vo, lo, io = r.copy(), v.copy(), ell.copy(), ids.copy()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_TIEBREAK, 1)
        ctx.set_option(ns.NS_OPT_GRAPH, graph)
        ctx.upload(r, v, ell, ids)
        ctx.run(p, 4)
        gr, gv, gl, gid = ctx.download()
        glv = ctx.levels()
    for _ in range(4):
        lv, _ = oracle.step(ro, vo, lo, io, po, tiebreak_id=True)
    for x, y, what in ((gid, io, "id"), (gr, ro, "r"), (gv, vo, "v"), (gl, lo, "ell"), (glv, lv, "levels")):
        x, y = np.asarray(x), np.asarray(y)
        bad = np.flatnonzero(x.view(np.uint64) != y.view(np.uint64)) if x.dtype == np.float64 else \
            np.flatnonzero(x != y)
        assert bad.size == 0, f"{what}: {bad.size} differ"
