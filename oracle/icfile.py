"""CPU oracle of the initial-conditions file -- TEST INFRASTRUCTURE ONLY.

* file_bytes(): write_initial_conditions (nsphere.c:2386-2416): int32 npts,
  then for each particle its five doubles r, |v|, ell, id, mu in turn;
* convert(): the conversion after reading (nsphere.c:4908-4920):
  particles[1] *= particles[4]; particles[1] *= kmsec_to_kpcmyr;
  particles[2] *= kmsec_to_kpcmyr -- two roundings for v_r, in that order.

Pinned by tests/test_oracle_icfile.py (struct.pack of a hand case).
"""
from __future__ import annotations

import numpy as np


def file_bytes(records) -> bytes:
    rec = np.ascontiguousarray(records, "<f8")
    return np.int32(rec.shape[0]).astype("<i4").tobytes() + rec.tobytes()


def convert(records, conv: float):
    rec = np.asarray(records, np.float64)
    v = rec[:, 1] * rec[:, 4]
    v = v * np.float64(conv)
    return rec[:, 0].copy(), v, rec[:, 2] * np.float64(conv), rec[:, 3].astype(np.uint32)
