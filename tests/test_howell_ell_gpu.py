"""Howell form over Z/ell^k for odd primes ell on the device (hx_howell_batch_create_ell; the reference's
elim::howell with a prime-power context, DenseElim.cpp:1450 -> howell_rows :1130-1300): H, pivots and
pivot_v equal the reference's own on the committed fixture, and the oracle's restatement on batches of
uneven random and valuation-structured jobs over ell = 3 ... 65521 up to the widest powers below 2^32;
k = 1 gives the RREF over F_ell; ell = 2 through the same entry equals the Z/2^k batch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _structured(rng, ell, k, m, n):
    """Random residues with rows scaled by ell and by ell^(k/2): many non-unit pivots and closure rows."""
    q = ell ** k
    A = rng.integers(0, q, (m, n), dtype=np.uint64)
    A[::3] = A[::3] * ell % q
    A[1::4] = A[1::4] * (ell ** (k // 2)) % q
    return A.astype(np.uint32)


def _run(E, ell, k, cases):
    mats = [E.to_device(A) for A in cases]
    b = E.HowellBatch(k, mats, ell=ell)
    b.run()
    torch.cuda.synchronize()
    return b.results()


def test_howell_ell_matches_reference_fixture(hx_gpu):
    import os
    from hacs_amd import engine as E
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "howell_ell_reference.npz"),
                allow_pickle=False)
    for i in range(int(G["count"][0])):
        ell, k = (int(x) for x in G[f"case{i}_meta"])
        print(f"Howell fixture case {i} ell={ell} k={k} {G[f'case{i}_A'].shape}", flush=True)
        res = _run(E, ell, k, [G[f"case{i}_A"]])[0]
        assert np.array_equal(res.H, G[f"case{i}_H"]), f"case {i}: H"
        assert np.array_equal(res.pivots, G[f"case{i}_piv"]) and np.array_equal(res.pivot_v, G[f"case{i}_pv"])


@pytest.mark.parametrize("ell,k", [(3, 2), (3, 7), (3, 20), (5, 13), (7, 4), (11, 9), (251, 2), (65521, 2)])
def test_howell_ell_batch_matches_oracle(hx_gpu, O, ell, k):
    from hacs_amd import engine as E
    rng = np.random.default_rng(ell * 100 + k)
    q = ell ** k
    cases = [np.array([[ell, 1]], np.uint32), np.zeros((3, 5), np.uint32)]
    for m, n in [(1, 1), (5, 9), (9, 5), (17, 17), (40, 33), (33, 70), (64, 64), (100, 130), (200, 150), (600, 500)]:
        cases.append(rng.integers(0, q, (m, n), dtype=np.uint64).astype(np.uint32))
        cases.append(_structured(rng, ell, k, m, n))
    print(f"Howell batch ell={ell} k={k}: {len(cases)} jobs", flush=True)
    for A, res in zip(cases, _run(E, ell, k, cases)):
        H, piv, pv = O.howell_ell(ell, k, A)
        assert np.array_equal(res.pivots, piv), f"{A.shape}: pivots"
        assert np.array_equal(res.pivot_v, pv), f"{A.shape}: pivot_v"
        assert np.array_equal(res.H, H), f"{A.shape}: H"
    r0 = _run(E, ell, 2, [np.array([[ell, 1]], np.uint32)])[0] if k != 2 else None
    if r0 is not None:
        assert np.array_equal(r0.H, np.array([[ell, 1], [0, ell]], np.uint32))


def test_howell_ell_field_and_two(hx_gpu, O):
    """k = 1: the RREF over F_ell (pivots, zero valuations); ell = 2: the Z/2^k batch's forms."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(0x4E15)
    for ell in (3, 65521):
        cases = [rng.integers(0, ell, (m, n)).astype(np.uint32) for (m, n) in [(30, 40), (50, 20), (64, 64)]]
        for A, res in zip(cases, _run(E, ell, 1, cases)):
            r, piv, _, R = O.rref_rule(ell, A)
            assert np.array_equal(res.pivots, piv) and np.array_equal(res.H, R) and not res.pivot_v.any()
    k = 9
    cases = [rng.integers(0, 1 << k, (m, n)).astype(np.uint32) for (m, n) in [(20, 30), (70, 50), (300, 200)]]
    a = _run(E, 2, k, cases)
    mats = [E.to_device(A) for A in cases]
    b = E.HowellBatch(k, mats)
    b.run()
    torch.cuda.synchronize()
    for x, y in zip(a, b.results()):
        assert np.array_equal(x.H, y.H) and np.array_equal(x.pivots, y.pivots) and np.array_equal(x.pivot_v, y.pivot_v)


def test_howell_ell_limits(hx_gpu):
    from hacs_amd import engine as E
    from hacs_amd import _lib
    M = [E.to_device(np.zeros((2, 2), np.uint32))]
    for ell, k in ((9, 2), (3, 21), (1, 3), (7, 0)):          # not prime, 3^21 >= 2^32, ell < 2, k < 1
        with pytest.raises(_lib.HxError):
            E.HowellBatch(k, M, ell=ell)
