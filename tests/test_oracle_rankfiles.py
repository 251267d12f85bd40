"""Pins of the Rank-file oracle (oracle/rankfiles.py) on the CPU."""
import struct

import numpy as np

from oracle import rankfiles as orf


def test_nodes_by_hand():
    x, y = orf.nodes(np.array([1.0, 1.0, 1.0, 1.000001, 2.0]), np.arange(5.0), -1.0)
    assert x.tolist() == [0.0, 1.0, 1.0 + 1e-5, 1.0 + 1e-5 + 1e-5, 1.0 + 1e-5 + 1e-5 + 1e-5, 2.0]
    assert y.tolist() == [-1.0, 0.0, 1.0, 2.0, 3.0, 4.0]


def test_linear_matches_interp_on_increasing_nodes():
    rng = np.random.default_rng(2)
    x = np.concatenate([[0.0], np.cumsum(rng.uniform(0.01, 1.0, 500))])
    y = rng.normal(size=x.size)
    t = rng.uniform(0, x[-1], 10_000)
    np.testing.assert_allclose(orf.linear(x, y, t), np.interp(t, x, y), rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(orf.linear(x, y, x[1:-1]), y[1:-1])  # exact at the nodes


def test_row_layout():
    row = np.zeros(1, orf.SORTED)
    row["rank"], row["mass"], row["r"], row["v"] = 7, 0.5, 2.0, -1.0
    row["psi"], row["E"], row["ell"], row["density"] = 3.0, 4.0, 0.25, 1e-10
    assert row.tobytes() == struct.pack("<ifffffff", 7, 0.5, 2.0, -1.0, 3.0, 4.0, 0.25, np.float32(1e-10))
    assert orf.UNSORTED.itemsize == 28 and orf.SORTED.itemsize == 32
