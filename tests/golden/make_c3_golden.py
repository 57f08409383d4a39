#!/usr/bin/env python3
"""Writes tests/golden/c3_reference.npz: the 8192^3 GEMM configuration mod 2147483629
(A = EngineMatrix(8192, 8192, p, 0x01), B = salt 0x02, C = A B mod p).  C comes from the oracle's
restatement of kernels::gemm<uint32_t> (hxo_gemm_fast) and must equal the reference's own
kernels::gemm (oracle/_ref) byte for byte; the fixture keeps the FNV-1a-64 checksums of A, B and C
(rows, cols, entries row-major) and three full rows of C for diagnosis."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

p, n = 2147483629, 8192
rows = np.array([0, 4095, 8191], np.uint64)
t0 = time.time()
A = O.engine_matrix(n, n, p, 0x01)
B = O.engine_matrix(n, n, p, 0x02)
C = O.gemm_fast(p, A, B, None, 0, 8)
t1 = time.time()
C_ref = O.ref_gemm(p, A, B, None, 0, 8)
t2 = time.time()
assert np.array_equal(C, C_ref), "oracle and reference kernels::gemm differ"
out = dict(p=np.array([p], np.uint64), n=np.array([n], np.uint64),
           a_fnv=np.array([O.fnv_matrix(A)], np.uint64), b_fnv=np.array([O.fnv_matrix(B)], np.uint64),
           c_fnv=np.array([O.fnv_matrix(C)], np.uint64), rows=rows, c_rows=C[rows.astype(np.int64)].copy())
np.savez_compressed(os.path.join(HERE, "c3_reference.npz"), **out)
print(f"oracle {t1 - t0:.1f} s, reference kernels::gemm {t2 - t1:.1f} s; C fnv {out['c_fnv'][0]:#x}")
