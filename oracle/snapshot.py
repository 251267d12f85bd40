"""CPU oracle of the snapshot path -- TEST INFRASTRUCTURE ONLY.

Restates, in numpy, what the reference does every dtwrite steps:

* records(): the per-dtwrite block (nsphere.c:6456-6471) -- for every
  original id, rank = inverse_map[id] (the particle's position in the
  rank-ordered arrays) and (float) r, v_r, ell of that position, stored at
  index id;
* file_bytes(): append_all_particle_data_chunk_to_file (nsphere.c:1295-1324)
  called whenever a block of block_size snapshots is full
  (nsphere.c:6515-6527): each record as int32 rank, float32 r, v, ell, the
  particles of one snapshot contiguous, snapshots in order; a partial last
  block is never written;
* read_snapshot(): the reader's offset (nsphere.c:1380-1385),
  block_size * npts * 16 * (snap / block_size) + npts * 16 * (snap % block_size).

Pins (tests/test_oracle_snapshot.py): records of a hand-checked state, the
file of a hand-packed case byte for byte (struct.pack), and the reader's
offset formula against slicing the concatenated snapshots.
"""
from __future__ import annotations

import numpy as np

APD_RECORD = np.dtype([("rank", "<i4"), ("r", "<f4"), ("v", "<f4"), ("ell", "<f4")])


def records(r, v, ell, ids) -> np.ndarray:
    """Records of a state held in rank order (position i = rank i)."""
    n = len(r)
    ids = np.asarray(ids, np.int64)
    inverse_map = np.empty(n, np.int64)
    inverse_map[ids] = np.arange(n)              # nsphere.c:5789-5794
    out = np.empty(n, APD_RECORD)
    for name, col in (("r", r), ("v", v), ("ell", ell)):
        out[name] = np.asarray(col, np.float64)[inverse_map].astype(np.float32)
    out["rank"] = inverse_map.astype(np.int32)
    return out


def file_bytes(snapshots, block_size: int) -> bytes:
    """The bytes the reference appends for these snapshots (whole blocks only)."""
    keep = (len(snapshots) // block_size) * block_size
    parts = []
    for b in range(0, keep, block_size):        # one append per full block
        for step in range(block_size):
            parts.append(np.ascontiguousarray(snapshots[b + step], APD_RECORD).tobytes())
    return b"".join(parts)


def read_snapshot(data: bytes, npts: int, block_size: int, snap: int) -> np.ndarray:
    block_number, index_in_block = snap // block_size, snap % block_size
    step_data_size = npts * 16
    offset = block_size * step_data_size * block_number + step_data_size * index_in_block
    return np.frombuffer(data, APD_RECORD, count=npts, offset=offset)
