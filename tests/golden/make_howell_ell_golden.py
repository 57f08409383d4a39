#!/usr/bin/env python3
"""Writes tests/golden/howell_ell_reference.npz: inputs and outputs of the reference's own Howell form over
Z/ell^k (elim::howell with ModCtx::prime_power(ell, k), DenseElim.cpp:1450 -> howell_rows :1130-1300, compiled
into oracle/_ref): H, pivots and pivot_v.  Cases: ell = 3, 5, 7, 11, 251, 65521 and the widest powers that fit
32 bits (3^20, 5^13, 7^11), random entries, rows scaled by ell and by ell^(k/2) (valuation structure),
duplicated and zero rows, tall and wide shapes, [[ell, 1]] over Z/ell^2."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

if not O.ref_available():
    raise SystemExit("oracle/_ref/libhacsref.so is not built (make -C oracle ref)")
rng = np.random.default_rng(0x4E11)


def job(ell, k, m, n):
    q = ell ** k
    A = rng.integers(0, q, (m, n), dtype=np.uint64)
    A[::3] = A[::3] * ell % q
    A[1::4] = A[1::4] * (ell ** (k // 2)) % q
    if m > 6:
        A[5] = A[2]
        A[6] = 0
    return A.astype(np.uint32)


cases = [(3, 2, job(3, 2, 12, 9)), (3, 5, job(3, 5, 40, 30)), (5, 3, job(5, 3, 25, 50)), (7, 4, job(7, 4, 60, 40)),
         (11, 9, job(11, 9, 33, 33)), (251, 2, job(251, 2, 50, 70)), (65521, 2, job(65521, 2, 20, 16)),
         (3, 20, job(3, 20, 48, 48)), (5, 13, job(5, 13, 30, 25)), (7, 11, job(7, 11, 20, 35)),
         (3, 2, np.array([[3, 1]], np.uint32)), (5, 2, np.array([[5, 1]], np.uint32)),
         (3, 4, np.zeros((4, 6), np.uint32))]
out = {}
for i, (ell, k, A) in enumerate(cases):
    H, piv, pv = O.ref_howell_ell(ell, k, A)
    out[f"case{i}_meta"] = np.array([ell, k], np.uint64)
    out[f"case{i}_A"], out[f"case{i}_H"] = A, H
    out[f"case{i}_piv"], out[f"case{i}_pv"] = piv, pv
out["count"] = np.array([len(cases)], np.uint64)
np.savez_compressed(os.path.join(HERE, "howell_ell_reference.npz"), **out)
print("wrote", len(cases), "cases")
