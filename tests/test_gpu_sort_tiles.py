"""Every onesweep tile shape (NS_OPT_SORT_TILE) and both key widths give the
oracle's stable permutation, bit for bit."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TILES = [(256, 16), (256, 24), (512, 12), (512, 16)]


@pytest.mark.parametrize("threads,items", TILES)
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [1, 4097, 12289, 333_333])
def test_tile_shapes_match_oracle(threads, items, precision, n):
    import nsphere_amd as ns
    from nsphere_amd import _lib
    rng = np.random.default_rng(n + 7 * threads + items + precision)
    r = np.sort(rng.uniform(0, 10, n)) + rng.normal(0, 1e-3, n)   # nearly sorted, as between steps
    r[rng.integers(0, n, n // 50)] = rng.choice(np.array([0.0, -0.0, 2.5]), n // 50)  # ties
    if precision:
        r = r.astype(np.float32).astype(np.float64)
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(_lib.NS_OPT_SORT_TILE, (threads << 16) | items)
        ctx.upload(r, np.zeros(n), np.zeros(n), np.arange(n, dtype=np.uint32))
        ctx.sort_particles()
        _, _, _, gid = ctx.download()
    np.testing.assert_array_equal(gid, oracle.sort_perm(r))


def test_unsupported_tile_rejected():
    import nsphere_amd as ns
    from nsphere_amd import _lib
    with pytest.raises(ns.NsError):
        with ns.GpuContext(10) as ctx:
            ctx.set_option(_lib.NS_OPT_SORT_TILE, (128 << 16) | 16)
