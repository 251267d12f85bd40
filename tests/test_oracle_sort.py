"""Pins the oracle's radial sort.

* against the reference's own quadsort (golden permutations written by
  tests/golden/make_golden.py from oracle/_ref, and live when _ref is built);
* against numpy's stable argsort of the order-preserving keys (an independent
  stable sort) on inputs with NaN and infinities;
* the key transform against IEEE-754 total-order known answers.
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "quadsort_perm.npz")


def golden_cases():
    with np.load(GOLDEN) as z:
        names = sorted({k.split("__")[0] for k in z.files})
        return [(n, z[f"{n}__r"], z[f"{n}__perm"]) for n in names]


@pytest.mark.parametrize("name,r,perm", golden_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_matches_reference_quadsort_golden(name, r, perm):
    np.testing.assert_array_equal(oracle.sort_perm(r), perm)


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref (reference quadsort) not built here")
@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_reference_quadsort_live(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 200_000))
    kind = seed % 3
    if kind == 0:
        r = rng.uniform(0, 100, n)
    elif kind == 1:
        r = rng.choice(np.linspace(0, 1, 37), n)          # heavy ties
    else:
        r = np.sort(rng.uniform(0, 5, n)) + rng.normal(0, 1e-3, n)
    np.testing.assert_array_equal(oracle.sort_perm(r), oracle.ref_quadsort_perm(r))


def np_key(r):
    r = np.where(np.isnan(r), np.nan, r).astype(np.float64)
    r = np.where(r == 0.0, 0.0, r)
    u = r.view(np.uint64).copy()
    u[np.isnan(r)] = np.uint64(0x7FF8000000000000)
    return np.where(u >> np.uint64(63) == 1, ~u, u | np.uint64(1 << 63))


@pytest.mark.parametrize("n", [0, 1, 5, 4097, 100_000])
def test_oracle_sort_equals_numpy_stable_argsort(n):
    rng = np.random.default_rng(n)
    r = rng.normal(0, 2, n)
    pick = rng.integers(0, 10, n)
    r[pick == 0] = np.nan
    r[pick == 1] = np.inf
    r[pick == 2] = -np.inf
    r[pick == 3] = -0.0
    r[pick == 4] = 0.0
    want = np.argsort(np_key(r), kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(oracle.sort_perm(r), want)


def test_key_total_order_known_answers():
    seq = [-np.inf, -1e308, -1.0, -5e-324, -0.0, 0.0, 5e-324, 2.2250738585072014e-308, 1.0,
           1.0000000000000002, 1e308, np.inf, np.nan]
    keys = [oracle.radius_key(x) for x in seq]
    assert keys[4] == keys[5], "-0 and +0 compare equal"
    strictly = keys[:4] + keys[5:]
    assert all(a < b for a, b in zip(strictly, strictly[1:]))
    assert oracle.radius_key(float("nan")) == oracle.radius_key(-float("nan"))
    assert oracle.radius_key(1.0) == 0xBFF0000000000000
    assert oracle.radius_key(-1.0) == 0x400FFFFFFFFFFFFF


def test_tiebreak_id_orders_ties_by_id():
    r = np.array([2.0, 1.0, 2.0, 1.0, 2.0])
    ids = np.array([9, 7, 3, 8, 1], np.uint32)
    assert oracle.sort_perm(r).tolist() == [1, 3, 0, 2, 4]            # stable on position
    assert oracle.sort_perm(r, ids, True).tolist() == [1, 3, 4, 2, 0]  # ties by id


def test_chunked_sort_exact_on_nearly_sorted_data():
    """The reference default (24 chunks + seam windows, nsphere.c:8475-8542)
    is exact when no particle crosses more than a seam window."""
    rng = np.random.default_rng(3)
    n = 240_000
    r = np.sort(rng.uniform(0, 10, n)) + rng.normal(0, 1e-6, n)
    np.testing.assert_array_equal(oracle.sort_perm_chunked(r), oracle.sort_perm(r))


def test_chunked_sort_is_approximate_on_random_data():
    """...and mis-orders data whose elements move further than the windows."""
    rng = np.random.default_rng(4)
    r = rng.uniform(0, 10, 100_000)
    p = oracle.sort_perm_chunked(r)
    assert np.any(np.diff(r[p]) < 0)
    assert np.array_equal(np.sort(p), np.arange(r.size))
