#!/usr/bin/env python3
"""Writes tests/golden/c2_reference.npz: the 4096 x 4096 rank-3000 RREF configuration mod 2^31-1
(RankedMatrix(4096, 4096, 3000) with 64 rows overwritten by combinations of other rows, salt 0x31).
The RREF and its pivot columns come from the reference's own engine::elim::rref (oracle/_ref) and
must equal the oracle's smallest-row-rule elimination, which also supplies pivot_rows; the fixture
keeps rank, pivots, pivot_rows and the FNV-1a-64 checksum of the RREF (rank rows)."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

p, m, n, r, nover = 2147483647, 4096, 4096, 3000, 64
t0 = time.time()
A = O.overwrite_rows(O.ranked_matrix(m, n, r, p, 8), p, nover)
rank, piv, prow, R = O.rref_rule(p, A, True, 8)
t1 = time.time()
rank_ref, piv_ref, R_ref = O.ref_rref(p, A)
t2 = time.time()
assert rank == rank_ref and np.array_equal(piv, piv_ref) and np.array_equal(R, R_ref[:rank]) and not R_ref[rank:].any()
out = dict(p=np.array([p], np.uint64), shape=np.array([m, n, r, nover], np.uint64), rank=np.array([rank], np.uint64),
           pivots=piv, pivot_rows=prow, rref_fnv=np.array([O.fnv_matrix(R)], np.uint64),
           input_fnv=np.array([O.fnv_matrix(A)], np.uint64))
np.savez_compressed(os.path.join(HERE, "c2_reference.npz"), **out)
print(f"rank {rank}; oracle {t1 - t0:.1f} s, reference elim::rref {t2 - t1:.1f} s; fnv {out['rref_fnv'][0]:#x}")
