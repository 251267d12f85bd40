"""Snapshot path: all_particle_data records, the file writer, tracked ids.

The records kernel must give, for every original id, exactly the record the
reference's per-dtwrite block fills (oracle/snapshot.py restates it), bit for
bit, in both state precisions, after steps have reordered the particles; the
writer's file must equal the bytes the reference appends for the same
snapshots (whole blocks only), with snapshots written between steps while
the state keeps evolving.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import snapshot as osn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def sorted_ic(ns, kind, n, seed, lowl=0.05):
    r, v, ell, ids = ns.generate_ic(kind, n, seed, 100.0, lowl, 1e-3)
    q = oracle.sort_perm(r)
    return r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy()


def expected(ctx):
    r, v, ell, ids = ctx.download()
    return osn.records(r, v, ell, ids)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [1, 1000, 123_457])
def test_records_match_oracle(ns, precision, n):
    state = sorted_ic(ns, 2, n, 5 + n)
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.upload(*state)
        for _ in range(3):
            ctx.step(ns.unit_params())
        got = ctx.snapshot_records()
        want = expected(ctx)
    assert got.tobytes() == want.tobytes()


def test_gather_ids(ns):
    n = 50_000
    state = sorted_ic(ns, 0, n, 9)
    with ns.GpuContext(n) as ctx:
        ctx.upload(*state)
        ctx.run(ns.unit_params(), 2)
        r, v, ell, ids = ctx.download()
        pick = np.array([0, 7, n - 1, 12_345, 3, 49_000], np.uint32)
        gr, gv, gl, grank = ctx.gather_ids(pick)
        bad = ctx.gather_ids(np.array([n + 5], np.uint32))
    pos = np.empty(n, np.int64)
    pos[ids] = np.arange(n)
    np.testing.assert_array_equal(grank, pos[pick])
    np.testing.assert_array_equal(gr, r[pos[pick]])
    np.testing.assert_array_equal(gv, v[pos[pick]])
    np.testing.assert_array_equal(gl, ell[pos[pick]])
    assert bad[3][0] == -1


@pytest.mark.parametrize("precision", [0, 1])
def test_writer_file_equals_reference_bytes(ns, precision, tmp_path):
    n = 40_000
    state = sorted_ic(ns, 1, n, 17)
    path = tmp_path / "all_particle_data.dat"
    prefix = b"earlier content "
    path.write_bytes(prefix)  # the reference appends ("ab")
    snaps = []
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.upload(*state)
        w = ns.ApdWriter(ctx, path, block_size=3)
        for s in range(8):
            ctx.run(ns.unit_params(), 1 + s % 2)
            w.write()
            snaps.append(expected(ctx))
        kept = w.close()
    data = path.read_bytes()
    assert kept == 6
    assert data[:len(prefix)] == prefix
    assert data[len(prefix):] == osn.file_bytes(snaps, 3)
    body = data[len(prefix):]
    for s in range(6):
        np.testing.assert_array_equal(osn.read_snapshot(body, n, 3, s), snaps[s])
    # the package's reader at the reference's offsets (no prefix here)
    clean = tmp_path / "clean.dat"
    clean.write_bytes(body)
    np.testing.assert_array_equal(ns.read_all_particle_data(str(clean), n, 4), snaps[4])


def test_writer_rejections(ns, tmp_path):
    with ns.GpuContext(100) as ctx:
        with pytest.raises(ns.NsError):
            ns.ApdWriter(ctx, tmp_path / "x.dat", 3)  # no state uploaded
        r, v, ell, ids = sorted_ic(ns, 0, 100, 1)
        ctx.upload(r, v, ell, ids)
        with pytest.raises(ns.NsError):
            ns.ApdWriter(ctx, tmp_path / "x.dat", 0)
        with pytest.raises(ns.NsError):
            ns.ApdWriter(ctx, tmp_path / "no_such_dir" / "x.dat", 3)
        with ns.ApdWriter(ctx, tmp_path / "y.dat", 2) as w:
            w.write()
        assert os.path.getsize(tmp_path / "y.dat") == 0  # one snapshot: no whole block
