"""Initial-conditions file: the oracle by hand, and the library's host reader
and writer against it (no GPU needed)."""
import struct

import numpy as np
import pytest

from oracle import icfile


def test_oracle_bytes_and_conversion_by_hand():
    rec = np.array([[1.5, 200.0, 30.0, 1.0, -0.5], [0.25, 10.0, 2.0, 0.0, 0.75]])
    want = struct.pack("<i", 2) + struct.pack("<5d", *rec[0]) + struct.pack("<5d", *rec[1])
    assert icfile.file_bytes(rec) == want
    k = 1.02271e-3
    r, v, ell, ids = icfile.convert(rec, k)
    assert v[0] == (200.0 * -0.5) * k and ell[1] == 2.0 * k and ids.tolist() == [1, 0]


def test_library_file_io(tmp_path):
    ns = pytest.importorskip("nsphere_amd")
    rng = np.random.default_rng(2)
    rec = rng.normal(size=(1001, 5))
    rec[:, 3] = rng.permutation(1001)
    path = tmp_path / "ic.dat"
    ns.write_ic_file(path, rec)
    assert path.read_bytes() == icfile.file_bytes(rec)
    back = ns.read_ic_file(path)
    assert back.tobytes() == rec.tobytes()
    # a count that does not match is refused, as the reference refuses it
    import ctypes
    from nsphere_amd import _lib
    n = ctypes.c_int32(1000)
    buf = np.empty((1000, 5))
    assert _lib.lib.ns_ic_file_read(str(path).encode(), ctypes.byref(n), buf.ctypes.data) == 6
    with pytest.raises(ns.NsError):
        ns.read_ic_file(tmp_path / "missing.dat")
