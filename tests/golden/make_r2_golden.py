#!/usr/bin/env python3
"""Writes tests/golden/r2_reference.npz: inputs and outputs of the reference's own blocked LU over
Z/2^k (ring2k::echelon_blocked, MinValuation; Binary/Ring2k.h:538-694, compiled into oracle/_ref) on
the native u32 and u64 rungs -- the LU the device session's host driver must reproduce entrywise
(Binary/Ring2kDeviceLU.h).  Cases: k = 30 / 32 on u32, k = 61 / 40 / 64 on u64, panels 8..64, a column
limit below n, rows scaled by powers of two (valuation structure), duplicated and zero rows."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

if not O.ref_available():
    raise SystemExit("oracle/_ref/libhacsref.so is not built (make -C oracle ref)")
rng = np.random.default_rng(0x2C)
cases = [(np.uint32, 30, 48, 60, 8, 60), (np.uint32, 32, 100, 90, 16, 90), (np.uint32, 32, 130, 170, 32, 150),
         (np.uint64, 61, 70, 50, 16, 50), (np.uint64, 40, 120, 140, 32, 140), (np.uint64, 64, 90, 90, 64, 70)]
out = {}
for i, (dt, k, m, n, panel, climit) in enumerate(cases):
    A = rng.integers(0, np.iinfo(dt).max, (m, n), dtype=np.uint64, endpoint=True).astype(dt)
    for r0 in range(0, m, 4):                                     # rows scaled by 2^s: valuation structure
        A[r0] = (A[r0] << dt(rng.integers(0, 8))).astype(dt)
    A[7] = A[3]                                                   # a duplicated row
    A[11] = 0                                                     # a zero row
    R, r, piv, pv = O.ref_r2_echelon(k, A, panel, climit)
    out[f"case{i}_meta"] = np.array([np.dtype(dt).itemsize, k, m, n, panel, climit, r], np.uint64)
    out[f"case{i}_A"], out[f"case{i}_R"] = A, R
    out[f"case{i}_piv"], out[f"case{i}_pv"] = piv, pv
out["count"] = np.array([len(cases)], np.uint64)
np.savez_compressed(os.path.join(HERE, "r2_reference.npz"), **out)
print("wrote", len(cases), "cases")
