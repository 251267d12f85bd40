"""NS_OPT_SORT_KEYS: between steps the float64 radii are sorted on 24-bit
quantile keys instead of the two-plane 64-bit keys: by bucket (1: one
scatter into per-bucket regions, then one workgroup per bucket ordering it by
sub-bin, radius and position; buckets too large for a workgroup, or with a
run of equal sub-bins past its limit, go to a fallback kernel) or by three
radix passes and a fix-up of each run of equal keys by the full radius key (2).

The reference sorts the radii at the start of every method-5 step with a
stable sort (nsphere.c:5784-5787, quadsort at nsphere.c:8356-8359); the
oracle restates it (orc_sort_perm).  The quantile keys only arise after an
integrate at N >= 65536, so these tests run 200k-300k particles: each step's
permutation after a run must equal the oracle's stable sort of the radii that
step started from, and whole runs with either key form must agree bit for
bit (eager and overlapped, the strict plain and the fast binned forms).
Blocks of identical particles give runs of equal keys far longer than a
wave's look-ahead, which reach the long-run kernel, and buckets past a
workgroup's capacity or with runs past its limit, which reach the bucket
sort's fallback.
"""
import ctypes

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
        assert bad.size == 0, f"{what} {name}: {bad.size} differ, first at {bad[:5]}"


def ic(ns, n, seed):
    return ns.generate_ic(0, n, seed, 100.0, 0.05, 1e-3)


def run_then_step(ns, state, keys, form, steps, overlap):
    """`steps` steps in one run, then one more step: the final state, the
    radii that last step started from and its permutation."""
    from nsphere_amd._lib import check, lib
    n = state[0].size
    p = ns.unit_params()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_SORT_KEYS, keys)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
        ctx.set_option(ns.NS_OPT_OVERLAP, overlap)
        ctx.set_option(ns.NS_OPT_BUDGET, 1 << 18)
        ctx.upload(*state)
        ctx.run(p, steps)
        r = np.empty(n, np.float64)
        check(lib.ns_gpu_download_soa(ctx._h, r.ctypes.data, None, None, None))
        ctx.step(p)
        return ctx.download(), r, ctx.permutation()


@pytest.mark.parametrize("keys", [1, 2])
@pytest.mark.parametrize("form,overlap", [(0, 0), (2, 0), (2, 1)])
def test_quantile_keys_match_64bit_keys_and_oracle(ns, form, overlap, keys):
    state = ic(ns, 300_000, 71)
    fin_q, r_q, perm_q = run_then_step(ns, state, keys, form, 6, overlap)
    fin_w, r_w, perm_w = run_then_step(ns, state, 0, form, 6, overlap)
    assert_state(fin_q, fin_w, "quantile vs 64-bit keys")
    np.testing.assert_array_equal(r_q.view(np.uint64), r_w.view(np.uint64))
    np.testing.assert_array_equal(perm_q, perm_w)
    np.testing.assert_array_equal(perm_q, oracle.sort_perm(r_q))
    # the window moves: the step reordered most particles
    assert np.count_nonzero(perm_q != np.arange(perm_q.size, dtype=np.uint32)) > perm_q.size // 2


@pytest.mark.parametrize("keys", [1, 2])
def test_quantile_keys_long_runs_of_exact_ties(ns, keys):
    """Blocks of identical particles: the first step's quantile table holds
    about 14 equal quantiles per block, so the block's new radii (split only
    by the rank-dependent enclosed mass, by about 1e-7) land in one or two
    sub-bins of one bucket, and the second step's sort meets runs of tens of
    thousands of equal keys (buckets far past a workgroup's capacity), and,
    from the blocks of 100 and 200, runs below and above the bucket kernel's
    limit of 128 in buckets it can hold."""
    r, v, ell, ids = ic(ns, 200_000, 72)
    # 40 000 copies of one particle and 30 000 of another, 200 and 100 of
    # two more, 2 000 twin pairs
    r, v, ell = r.copy(), v.copy(), ell.copy()
    for lo, hi, src in ((1_000, 41_000, 0), (100_000, 130_000, 150_000), (160_000, 160_200, 170_000),
                        (180_000, 180_100, 190_000)):
        r[lo:hi], v[lo:hi], ell[lo:hi] = r[src], v[src], ell[src]
    twins = np.arange(150_001, 154_001, 2)
    r[twins + 1], v[twins + 1], ell[twins + 1] = r[twins], v[twins], ell[twins]
    state = (r, v, ell, ids)
    for form, overlap in ((0, 0), (2, 1)):
        fin_q, r_q, perm_q = run_then_step(ns, state, keys, form, 4, overlap)
        fin_w, _, perm_w = run_then_step(ns, state, 0, form, 4, overlap)
        assert_state(fin_q, fin_w, f"form {form}: quantile vs 64-bit keys")
        np.testing.assert_array_equal(perm_q, perm_w)
        np.testing.assert_array_equal(perm_q, oracle.sort_perm(r_q))


def test_quantile_keys_graph_replay(ns):
    """Captured overlapped steps sort quantile keys too: the same state as eager."""
    state = ic(ns, 250_000, 73)
    p = ns.unit_params()
    out = []
    for graph in (0, 1):
        with ns.GpuContext(state[0].size) as ctx:
            ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
            ctx.set_option(ns.NS_OPT_GRAPH, graph)
            ctx.upload(*state)
            ctx.run(p, 35)
            out.append(ctx.download())
    assert_state(out[0], out[1], "graph vs eager")


def test_bucket_sort_large_regions(ns):
    """At N = 2e7 a bucket holds about 4900 start radii and the last one
    about 8200, past the smaller bucket kernel's 6144: the sort takes the
    16384-entry regions and kernel.  Steps of the fast binned form, overlapped,
    agree bit for bit with the 64-bit key sort, and the last permutation is
    the oracle's stable sort."""
    state = ic(ns, 20_000_000, 74)
    fin_q, r_q, perm_q = run_then_step(ns, state, 1, 2, 3, 1)
    fin_w, _, perm_w = run_then_step(ns, state, 0, 2, 3, 1)
    assert_state(fin_q, fin_w, "bucket sort vs 64-bit keys at N = 2e7")
    np.testing.assert_array_equal(perm_q, perm_w)
    np.testing.assert_array_equal(perm_q, oracle.sort_perm(r_q))


@pytest.mark.parametrize("n", [65_536, 65_537, 70_001])
def test_bucket_sort_smallest_sizes(ns, n):
    """The smallest N that takes quantile keys: 16 or 17 start radii per bucket,
    so that a 4096-position tile spans about 250 buckets of the scatter's
    window; at 65 537 the stride falls back to floor(N / 4096) because
    ceil(N / 4096) = 17 would put the last bucket's first rank past N.
    Steps of the fast binned form, eager and overlapped, agree bit for bit
    with the 64-bit key sort."""
    state = ic(ns, n, 75)
    for overlap in (0, 1):
        fin_q, r_q, perm_q = run_then_step(ns, state, 1, 2, 5, overlap)
        fin_w, _, perm_w = run_then_step(ns, state, 0, 2, 5, overlap)
        assert_state(fin_q, fin_w, f"N={n} overlap={overlap}: bucket sort vs 64-bit keys")
        np.testing.assert_array_equal(perm_q, perm_w)
        np.testing.assert_array_equal(perm_q, oracle.sort_perm(r_q))
