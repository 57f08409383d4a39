"""The oracle and the device against vectors written by the reference's own CPU engine
(tests/golden/reference_vectors.npz, made by tests/golden/make_golden.py from oracle/_ref)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def G():
    return np.load(GOLD, allow_pickle=False)


def _cases(G, prefix, idx):
    return range(int(G["counts"][idx]))


def test_oracle_gemm_vs_reference(O, G):
    for i in _cases(G, "gemm", 0):
        p = int(G[f"gemm{i}_p"][0])
        A, B, C0 = G[f"gemm{i}_A"], G[f"gemm{i}_B"], G[f"gemm{i}_C0"]
        assert np.array_equal(O.gemm_fast(p, A, B), G[f"gemm{i}_set"]), f"gemm{i} Set p={p}"
        assert np.array_equal(O.gemm_naive(p, A, B, C0, 2), G[f"gemm{i}_sub"]), f"gemm{i} Sub p={p}"
    A = O.engine_matrix(1024, 1024, 65521, 0x01)
    B = O.engine_matrix(1024, 1024, 65521, 0x02)
    assert O.fnv_matrix(O.gemm_fast(65521, A, B)) == int(G["c1_fnv"][0])


def test_oracle_rref_vs_reference(O, G):
    for i in _cases(G, "rref", 1):
        p = int(G[f"rref{i}_p"][0])
        A = G[f"rref{i}_A"]
        rank, piv, prow, R = O.rref_rule(p, A)
        assert rank == len(G[f"rref{i}_piv"]) and np.array_equal(piv, G[f"rref{i}_piv"])
        assert np.array_equal(R, G[f"rref{i}_R"][:rank]) and not G[f"rref{i}_R"][rank:].any()
        assert np.array_equal(np.sort(prow), G[f"rref{i}_rowprofile"])


def test_oracle_howell_vs_reference(O, G):
    for i in _cases(G, "howell", 2):
        k = int(G[f"howell{i}_k"][0])
        A = G[f"howell{i}_A"]
        for fn in (O.howell_fast, O.howell_naive) if k <= 31 else (O.howell_fast,):
            H, piv, pv = fn(k, A)
            assert np.array_equal(H, G[f"howell{i}_H"]), f"howell{i} k={k}"
            assert np.array_equal(piv, G[f"howell{i}_piv"]) and np.array_equal(pv, G[f"howell{i}_pv"])


@pytest.mark.gpu
def test_device_vs_reference_vectors(hx_gpu, G):
    import torch
    from hacs_amd import engine as E
    for i in _cases(G, "gemm", 0):
        p = int(G[f"gemm{i}_p"][0])
        ctx = E.ModCtx(p)
        for v in [E.HX_GEMM_PLAIN] + ([E.HX_GEMM_LIMB] if ctx.limbs else []):
            dC = E.to_device(G[f"gemm{i}_C0"])
            E.gemm(ctx, E.to_device(G[f"gemm{i}_A"]), E.to_device(G[f"gemm{i}_B"]), dC, E.GemmOp.Sub, variant=v)
            torch.cuda.synchronize()
            assert np.array_equal(E.from_device(dC), G[f"gemm{i}_sub"]), f"gemm{i} p={p} variant={v}"
    for i in _cases(G, "rref", 1):
        p = int(G[f"rref{i}_p"][0])
        A = G[f"rref{i}_A"]
        R = torch.zeros(A.shape, dtype=torch.int32, device="cuda")
        res = E.rref(E.ModCtx(p), E.to_device(A), R)
        torch.cuda.synchronize()
        assert np.array_equal(E.from_device(R), G[f"rref{i}_R"]), f"rref{i} p={p}"
        assert np.array_equal(res.pivots, G[f"rref{i}_piv"])
        assert np.array_equal(np.sort(res.pivot_rows), G[f"rref{i}_rowprofile"])
    ks = {}
    for i in _cases(G, "howell", 2):
        ks.setdefault(int(G[f"howell{i}_k"][0]), []).append(i)
    for k, idx in ks.items():
        b = E.HowellBatch(k, [E.to_device(G[f"howell{i}_A"]) for i in idx])
        b.run()
        torch.cuda.synchronize()
        for i, res in zip(idx, b.results()):
            assert np.array_equal(res.H, G[f"howell{i}_H"]), f"howell{i} k={k}"
            assert np.array_equal(res.pivots, G[f"howell{i}_piv"]) and np.array_equal(res.pivot_v, G[f"howell{i}_pv"])
