#!/usr/bin/env python3
"""Writes tests/golden/reference_vectors.npz from the reference's OWN CPU engine, compiled from its
sources by oracle/Makefile (oracle/_ref/libhacsref.so, a C-ABI shim over engine::kernels::gemm,
engine::elim::rref / column_rank_profile, engine::ref::rref, engine::ring2k::howell (u16 and u32
rungs), engine::ref::howell and engine::elim::howell).  Run in the build container only
(`python tests/golden/make_golden.py`); the fixtures are data: inputs and the reference's outputs.

Inputs use the reference micro-benchmark generators (RandomEntries / SeedFor / RankedMatrix,
PerformanceSmokeValidation.cpp:64-170) so they are reproducible from their seeds; they are stored
anyway so that a fixture is self-contained."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libhacsref.so is not built (make -C oracle ref)")
    out = {}
    # GEMM: C = A B mod p (Set) and C -= A B (Sub) through engine::kernels::gemm<uint32_t>
    gemm_cases = [(65521, 37, 29, 300), (1000003, 20, 31, 257), (2147483629, 33, 17, 515), (2147483647, 16, 16, 64),
                  (4294967291, 9, 13, 100), (251, 40, 40, 1000), (134217757, 24, 20, 1100)]
    for i, (p, m, n, k) in enumerate(gemm_cases):
        A = O.engine_matrix(m, k, p, 0x01)
        B = O.engine_matrix(k, n, p, 0x02)
        C0 = O.engine_matrix(m, n, p, 0x03)
        out[f"gemm{i}_p"] = np.array([p], np.uint64)
        out[f"gemm{i}_A"], out[f"gemm{i}_B"], out[f"gemm{i}_C0"] = A, B, C0
        out[f"gemm{i}_set"] = O.ref_gemm(p, A, B, None, 0, 4)
        out[f"gemm{i}_sub"] = O.ref_gemm(p, A, B, C0, 2, 4)
    # the CPU-runnable config: C = A B mod 65521, 1024^2 (checksum only: FNV-1a over (rows, cols, entries))
    A = O.engine_matrix(1024, 1024, 65521, 0x01)
    B = O.engine_matrix(1024, 1024, 65521, 0x02)
    out["c1_fnv"] = np.array([O.fnv_matrix(O.ref_gemm(65521, A, B, None, 0, 8))], np.uint64)
    # RREF through engine::elim::rref (blocked PLU) and engine::ref::rref (naive Gauss-Jordan);
    # pivot-row sets through engine::elim::column_rank_profile of the transpose
    rref_cases = [(2, 40, 50, 0), (3, 33, 33, 20), (13, 64, 48, 30), (65521, 70, 90, 45), (1000003, 100, 80, 0),
                  (2147483647, 96, 96, 60), (2147483629, 50, 120, 0)]
    for i, (p, m, n, r) in enumerate(rref_cases):
        A = O.ranked_matrix(m, n, r, p)
        if r:
            A = O.overwrite_rows(A, p, 4)
        rank, piv, R = O.ref_rref(p, A)
        rank2, piv2, R2 = O.ref_rref(p, A, naive=True)
        assert rank == rank2 and np.array_equal(piv, piv2) and np.array_equal(R, R2)
        out[f"rref{i}_p"] = np.array([p], np.uint64)
        out[f"rref{i}_A"], out[f"rref{i}_R"], out[f"rref{i}_piv"] = A, R, piv
        out[f"rref{i}_rowprofile"] = O.ref_column_rank_profile(p, A.T.copy())
    # Howell over Z/2^k through ring2k::howell (u16 rung for k <= 16, u32 rung), ref::howell, elim::howell
    rng = np.random.default_rng(2024)
    howell_cases = [(16, 30, 40), (16, 64, 64), (2, 20, 25), (8, 33, 17), (31, 25, 30), (32, 20, 28), (1, 30, 30)]
    for i, (k, m, n) in enumerate(howell_cases):
        A = rng.integers(0, 1 << k, (m, n), dtype=np.uint64).astype(np.uint32)
        if i % 2 == 1:    # valuation structure: rows scaled by 2^s
            A = ((A.astype(np.uint64) << rng.integers(0, k, (m, 1)).astype(np.uint64)) % (1 << k)).astype(np.uint32)
        H, piv, pv = O.ref_howell(k, A, "ring2k", 32)
        if k <= 16:
            H16, piv16, pv16 = O.ref_howell(k, A, "ring2k", 16)
            assert np.array_equal(H, H16) and np.array_equal(piv, piv16) and np.array_equal(pv, pv16)
        if k <= 31:
            Hn, pn, vn = O.ref_howell(k, A, "naive")
            assert np.array_equal(H, Hn) and np.array_equal(piv, pn) and np.array_equal(pv, vn)
        if 2 <= k <= 31:
            He, pe, ve = O.ref_howell(k, A, "elim")
            assert np.array_equal(H, He) and np.array_equal(piv, pe) and np.array_equal(pv, ve)
        out[f"howell{i}_k"] = np.array([k], np.uint64)
        out[f"howell{i}_A"], out[f"howell{i}_H"] = A, H
        out[f"howell{i}_piv"], out[f"howell{i}_pv"] = piv, pv
    out["counts"] = np.array([len(gemm_cases), len(rref_cases), len(howell_cases)], np.uint64)
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
