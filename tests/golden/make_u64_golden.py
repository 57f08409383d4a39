#!/usr/bin/env python3
"""Writes tests/golden/u64_reference.npz: C = A B mod p (Set) and C0 - A B (Sub) on 64-bit residues
from the reference's own kernels::gemm<uint64_t> (ModArithmetic/Kernels.h:184, compiled into
oracle/_ref), at the moduli of its own u64 gates (2^61 - 1, KernelBitIdentityValidation.cpp) and at
2^64 - 59, 2^32 + 15 and 2^63 + 29."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

if not O.ref_available():
    raise SystemExit("oracle/_ref/libhacsref.so is not built (make -C oracle ref)")
rng = np.random.default_rng(0x64)
cases = [(2305843009213693951, 33, 29, 130), (18446744073709551557, 20, 40, 300), (4294967311, 48, 17, 257),
         (9223372036854775837, 16, 16, 1000)]
out = {}
for i, (p, m, n, k) in enumerate(cases):
    def draw(shape):
        hi = rng.integers(0, 1 << 32, shape, dtype=np.uint64)
        lo = rng.integers(0, 1 << 32, shape, dtype=np.uint64)
        return ((hi << np.uint64(32)) | lo) % np.uint64(p)
    A, B, C0 = draw((m, k)), draw((k, n)), draw((m, n))
    out[f"case{i}_p"] = np.array([p], np.uint64)
    out[f"case{i}_A"], out[f"case{i}_B"], out[f"case{i}_C0"] = A, B, C0
    out[f"case{i}_set"] = O.ref_gemm_u64(p, A, B, None, 0)
    out[f"case{i}_sub"] = O.ref_gemm_u64(p, A, B, C0, 2)
out["count"] = np.array([len(cases)], np.uint64)
np.savez_compressed(os.path.join(HERE, "u64_reference.npz"), **out)
print("wrote", len(cases), "cases")
