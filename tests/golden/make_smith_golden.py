#!/usr/bin/env python3
"""Writes tests/golden/smith_reference.npz: inputs and outputs of the reference's own Smith valuations over
Z/2^k by levels (ring2k::smith, Binary/Ring2k.h:701-841, compiled into oracle/_ref) on the native u32 and u64
rungs: sigma (ascending) and level_ranks.  Each case is also run through the reference's smith_naive (global
minimal-valuation pivoting, Ring2k.h:1004-1035), whose sigma must agree.  Cases: random words at k = 32 / 5 / 1
/ 16 / 61 / 64 / 40, planted valuations (a diagonal 2^v u under row and column operations), rows scaled by
powers of two, duplicated and zero rows, tall and wide shapes, a zero matrix, 1 x 1 and 1 x n."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

if not O.ref_available():
    raise SystemExit("oracle/_ref/libhacsref.so is not built (make -C oracle ref)")
rng = np.random.default_rng(0x5A17)


def words(dt, shape):
    return rng.integers(0, np.iinfo(dt).max, shape, dtype=np.uint64, endpoint=True).astype(dt)


def planted(dt, m, n, vals):
    """diag(2^v u) (u odd) under unit-triangular row and column mixing and permutations (all unimodular)."""
    D = np.zeros((m, n), dt)
    for i, v in enumerate(vals):
        D[i, i] = dt((int(words(dt, 1)[0]) | 1) << v & int(np.iinfo(dt).max))
    A = D[rng.permutation(m)][:, rng.permutation(n)]
    with np.errstate(over="ignore"):
        for d in (1, 2, 3, 5, 8, 13):
            if d < m:
                A[d:] += words(dt, (m - d, 1)) * A[:-d]
            if d < n:
                A[:, d:] += A[:, :-d] * words(dt, (1, n - d))
    return A


cases = []
cases.append((32, words(np.uint32, (40, 50))))
cases.append((5, words(np.uint32, (60, 45))))
cases.append((32, planted(np.uint32, 64, 64, [0] * 20 + [1] * 10 + [3] * 9 + [7] * 8 + [31] * 5 + [30] * 4)))
cases.append((1, words(np.uint32, (30, 30))))
cases.append((61, planted(np.uint64, 50, 70, [0] * 10 + [2] * 10 + [33] * 10 + [60] * 10 + [12] * 5)))
A = words(np.uint64, (80, 60))
for r0 in range(0, 80, 3):
    A[r0] = A[r0] << np.uint64(rng.integers(0, 40))
cases.append((64, A))
A = words(np.uint32, (70, 90))
A[1::4] = A[0::4][:len(A[1::4])]
A[5::9] = 0
cases.append((16, A))
cases.append((40, np.zeros((10, 12), np.uint64)))
cases.append((32, np.array([[7]], np.uint32)))
cases.append((32, np.array([[4, 8, 0, 12, 1 << 31]], np.uint32)))
cases.append((3, planted(np.uint32, 24, 18, [0, 0, 1, 1, 1, 2, 2, 2, 2])))

out = {}
for i, (k, A) in enumerate(cases):
    sig, lev = O.ref_smith(k, A)
    sig_naive, _ = O.ref_smith(k, A, naive=True)
    assert np.array_equal(sig, sig_naive), f"case {i}: smith != smith_naive"
    out[f"case{i}_k"] = np.array([k], np.uint64)
    out[f"case{i}_A"] = A
    out[f"case{i}_sigma"] = sig
    out[f"case{i}_levels"] = lev
out["count"] = np.array([len(cases)], np.uint64)
np.savez_compressed(os.path.join(HERE, "smith_reference.npz"), **out)
print("wrote", len(cases), "cases")
