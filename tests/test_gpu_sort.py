"""GPU radial sort vs the oracle's stable sort (bit-exact permutations).

ns_gpu_sort_particles replaces sort_particles (nsphere.c:8662-8665) with the
exact stable semantics of quadsort (nsphere.c:8356-8359) under the comparator
of nsphere.c:8297-8306: ties (including -0 vs +0) keep their input order.
NaN, which that comparator leaves unordered, is placed after +inf.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def gpu_sort_perm(ns, r):
    n = r.size
    with ns.GpuContext(n) as ctx:
        ids = np.arange(n, dtype=np.uint32)
        ctx.upload(r, np.arange(n, dtype=np.float64), np.zeros(n), ids)
        ctx.sort_particles()
        rs, vs, ls, ids_out = ctx.download()
    # every column moved with its particle
    np.testing.assert_array_equal(vs, ids_out.astype(np.float64))
    np.testing.assert_array_equal(rs.view(np.uint64), r[ids_out].view(np.uint64))
    return ids_out


def radii(kind, n, rng):
    if kind == "uniform":
        return rng.uniform(0, 100, n)
    if kind == "plummer":
        from nsphere_amd import generate_ic
        return generate_ic(0, n, int(rng.integers(1 << 62)), 100.0)[0]
    if kind == "nearly_sorted":
        x = np.sort(rng.uniform(0, 10, n))
        return x + rng.normal(0, 1e-4, n)
    if kind == "reversed":
        return np.sort(rng.uniform(0, 10, n))[::-1].copy()
    if kind == "all_equal":
        return np.full(n, 1.25)
    if kind == "few_values":
        return rng.choice(np.array([0.5, 1.0, 2.0, 3.5]), n)
    if kind == "signed_zero":
        x = rng.choice(np.array([0.0, -0.0, 1e-300, -1e-300, 5e-324, -5e-324]), n)
        return x
    if kind == "negatives":
        return rng.normal(0, 3, n)
    if kind == "specials":
        x = rng.normal(0, 1, n)
        pick = rng.integers(0, 8, n)
        x[pick == 0] = np.inf
        x[pick == 1] = -np.inf
        x[pick == 2] = np.nan
        x[pick == 3] = -0.0
        x[pick == 4] = 0.0
        x[pick == 5] = -np.nan
        return x
    raise ValueError(kind)


SIZES = [1, 2, 63, 64, 65, 4095, 4096, 4097, 12289, 100_003]
KINDS = ["uniform", "plummer", "nearly_sorted", "reversed", "all_equal", "few_values",
         "signed_zero", "negatives", "specials"]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", KINDS)
def test_sort_matches_oracle(ns, n, kind):
    rng = np.random.default_rng(n * 131 + KINDS.index(kind))
    r = radii(kind, n, rng)
    want = oracle.sort_perm(r)
    got = gpu_sort_perm(ns, r)
    np.testing.assert_array_equal(got, want)


def test_sort_empty(ns):
    with ns.GpuContext(0) as ctx:
        e = np.empty(0)
        ctx.upload(e, e, e, np.empty(0, np.uint32))
        ctx.sort_particles()
        out = ctx.download()
    assert all(a.size == 0 for a in out)


@pytest.mark.parametrize("n", [12_000_000])
def test_sort_full_size_properties(ns, n):
    """At the benchmark size: the permutation equals numpy's stable argsort of
    the order-preserving keys (an independent stable sort)."""
    r = ns.generate_ic(0, n, 99, 100.0, 0.05, 1e-3)[0]
    u = r.view(np.uint64)
    key = np.where(u >> np.uint64(63) == 1, ~u, u | np.uint64(1 << 63))
    want = np.argsort(key, kind="stable").astype(np.uint32)
    got = gpu_sort_perm(ns, r)
    np.testing.assert_array_equal(got, want)


def test_sort_repeated_steps_stable_on_ties(ns):
    """Sorting an already-sorted array with ties is the identity (stability)."""
    rng = np.random.default_rng(5)
    r = np.sort(rng.choice(np.arange(50, dtype=np.float64), 20000))
    got = gpu_sort_perm(ns, r)
    np.testing.assert_array_equal(got, np.arange(r.size, dtype=np.uint32))


def grouped_radii(rng, lengths, shuffle=True):
    """Groups of radii sharing the high 32 bits of their key (1 + g 2^-20 plus
    less than 2^-20), each of the given length, with exact ties inside."""
    parts = []
    for g, m in enumerate(lengths):
        low = rng.integers(0, 1 << 32, m)
        low[rng.random(m) < 0.1] = low[0]            # exact ties
        parts.append(1.0 + g * 2.0 ** -20 + low * 2.0 ** -52)
    r = np.concatenate(parts)
    if shuffle:
        r = r[rng.permutation(r.size)]
    return r


@pytest.mark.parametrize("case", ["short_runs", "window_edges", "one_long_run", "mixed"])
def test_sort_runs_of_equal_high_halves(ns, case):
    """The float64 sort orders the high 32 bits of the keys by radix passes and
    then each run of equal high halves by the low halves: runs inside one
    block's window, runs that cross tiles, and runs longer than any window
    (sorted on one workgroup) must all give the oracle's permutation."""
    rng = np.random.default_rng(["short_runs", "window_edges", "one_long_run", "mixed"].index(case) + 100)
    if case == "short_runs":
        lengths = rng.integers(1, 9, 60_000)
    elif case == "window_edges":
        lengths = [2047, 2048, 2049, 2559, 2560, 2561, 2562, 1, 4095, 4096, 4097, 5000, 3]
    elif case == "one_long_run":
        lengths = [300_000]
    else:
        lengths = np.concatenate([rng.integers(1, 40, 3000), [70_000], rng.integers(100, 3000, 40)])
    r = grouped_radii(rng, lengths)
    np.testing.assert_array_equal(gpu_sort_perm(ns, r), oracle.sort_perm(r))


def test_sort_runs_in_input_order_keep_it(ns):
    """A long run already in low-half order (sorted input) stays as it is."""
    rng = np.random.default_rng(7)
    r = np.sort(grouped_radii(rng, [200_000, 5, 100_000]))
    np.testing.assert_array_equal(gpu_sort_perm(ns, r), np.arange(r.size, dtype=np.uint32))
