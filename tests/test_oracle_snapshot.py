"""Pins of the snapshot oracle (oracle/snapshot.py) on the CPU."""
import struct

import numpy as np

from oracle import snapshot as osn


def test_records_by_hand():
    # rank order: position 0 holds id 2, position 1 id 0, position 2 id 1
    r = np.array([0.5, 1.25, 3.0])
    v = np.array([-1.0, 0.1, 2.0])
    ell = np.array([0.3, 0.0, 1e-40])
    ids = np.array([2, 0, 1], np.uint32)
    rec = osn.records(r, v, ell, ids)
    assert rec["rank"].tolist() == [1, 2, 0]
    assert rec["r"].tolist() == [np.float32(1.25), np.float32(3.0), np.float32(0.5)]
    assert rec["v"].tolist() == [np.float32(0.1), np.float32(2.0), np.float32(-1.0)]
    assert rec["ell"][1] == np.float32(1e-40)  # float32 subnormal, as the C cast gives


def test_file_bytes_by_hand():
    snaps = []
    for s in range(5):
        rec = np.zeros(2, osn.APD_RECORD)
        rec["rank"] = [1, 0]
        rec["r"] = [s + 0.5, s + 1.5]
        rec["v"] = [-s, s]
        rec["ell"] = [0.25, 0.75]
        snaps.append(rec)
    want = b"".join(struct.pack("<ifff", int(snaps[s]["rank"][i]), snaps[s]["r"][i], snaps[s]["v"][i],
                                snaps[s]["ell"][i]) for s in range(4) for i in range(2))
    assert osn.file_bytes(snaps, 2) == want          # 2 whole blocks of 2; the fifth is dropped
    assert osn.file_bytes(snaps, 5) == want + struct.pack("<ifff", 1, 4.5, -4.0, 0.25) + \
        struct.pack("<ifff", 0, 5.5, 4.0, 0.75)
    assert osn.file_bytes(snaps[:1], 2) == b""


def test_reader_offsets():
    rng = np.random.default_rng(3)
    snaps = []
    for _ in range(9):
        rec = np.zeros(7, osn.APD_RECORD)
        rec["rank"] = rng.permutation(7)
        rec["r"] = rng.random(7)
        snaps.append(rec)
    data = osn.file_bytes(snaps, 3)
    for s in range(9):
        np.testing.assert_array_equal(osn.read_snapshot(data, 7, 3, s), snaps[s])
