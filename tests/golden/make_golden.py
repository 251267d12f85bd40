"""Writes tests/golden/quadsort_perm.npz: radial-sort permutations produced by
the reference's own vendored quadsort 1.2.1.3 (compiled from the reference
tree into oracle/_ref by `make -C oracle ref`), called the way
sort_particles_with_alg(..., "quadsort") calls it (nsphere.c:8605-8652,
8356-8359).  Inputs are small seeded arrays chosen to exercise ties, signed
zeros, negative radii and nearly-sorted data (no NaN: the reference's
comparator leaves NaN unordered).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def cases():
    rng = np.random.default_rng(20250401)
    yield "uniform_1000", rng.uniform(0, 100, 1000)
    yield "plummer_like_4096", 1.0 / np.sqrt(rng.uniform(0, 1, 4096) ** (-2.0 / 3.0) - 1.0)
    x = np.sort(rng.uniform(0, 5, 3000))
    yield "nearly_sorted_3000", x + rng.normal(0, 2e-3, x.size)
    yield "ties_2000", rng.choice(np.arange(20, dtype=np.float64) * 0.5, 2000)
    yield "signed_zero_500", rng.choice(np.array([0.0, -0.0, 1.0, -1.0, 5e-324, -5e-324]), 500)
    yield "negatives_1500", rng.normal(0, 2, 1500)
    yield "reversed_777", np.sort(rng.uniform(0, 9, 777))[::-1].copy()
    yield "all_equal_300", np.full(300, 2.5)
    yield "infinities_400", rng.choice(np.array([np.inf, -np.inf, 1.0, 3.0, -2.0]), 400)


def main():
    if oracle.ref_lib() is None and not oracle.build_ref():
        raise SystemExit("needs the reference tree to build oracle/_ref")
    out = {}
    for name, r in cases():
        out[f"{name}__r"] = r
        out[f"{name}__perm"] = oracle.ref_quadsort_perm(r)
    np.savez_compressed(os.path.join(HERE, "quadsort_perm.npz"), **out)
    print(f"wrote {len(out) // 2} cases")


if __name__ == "__main__":
    main()
