"""Every radix pass algorithm (NS_OPT_SORT_ALG), tile shape (NS_OPT_SORT_TILE)
and key width gives the oracle's stable permutation, bit for bit."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TILES = [(256, 16), (256, 24), (512, 8), (512, 12), (512, 16), (1024, 6)]


@pytest.mark.parametrize("alg", [0, 1])
@pytest.mark.parametrize("threads,items", TILES)
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [1, 4097, 12289, 333_333])
def test_tile_shapes_match_oracle(alg, threads, items, precision, n):
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
        ctx.set_option(_lib.NS_OPT_SORT_ALG, alg)
        ctx.upload(r, np.zeros(n), np.zeros(n), np.arange(n, dtype=np.uint32))
        ctx.sort_particles()
        _, _, _, gid = ctx.download()
    np.testing.assert_array_equal(gid, oracle.sort_perm(r))


def test_switching_algorithms_between_sorts():
    """Reduce-then-scan keeps its tile counts where onesweep keeps its lookback
    words; alternating the two on one context must not let either read the
    other's leftovers."""
    import nsphere_amd as ns
    from nsphere_amd import _lib
    n = 200_003
    rng = np.random.default_rng(11)
    with ns.GpuContext(n) as ctx:
        for k, alg in enumerate([1, 0, 1, 0, 0, 1, 1]):
            r = rng.uniform(0, 1, n)
            ctx.set_option(_lib.NS_OPT_SORT_ALG, alg)
            ctx.upload(r, np.zeros(n), np.zeros(n), np.arange(n, dtype=np.uint32))
            ctx.sort_particles()
            _, _, _, gid = ctx.download()
            np.testing.assert_array_equal(gid, np.argsort(r, kind="stable"), err_msg=f"sort {k} alg {alg}")


def test_unsupported_tile_rejected():
    import nsphere_amd as ns
    from nsphere_amd import _lib
    with pytest.raises(ns.NsError):
        with ns.GpuContext(10) as ctx:
            ctx.set_option(_lib.NS_OPT_SORT_TILE, (128 << 16) | 16)
    with pytest.raises(ns.NsError):
        with ns.GpuContext(10) as ctx:
            ctx.set_option(_lib.NS_OPT_SORT_ALG, 2)
