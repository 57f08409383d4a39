"""F_3 on bit-sliced rows (hx_f3.hip; the reference's binary::TritMatrix, Binary/TritMatrix.h): the in-place
RREF, its rank and pivot columns against the oracle's smallest-row rule at p = 3 (elim::rref's definition,
pinned by the reference's own vectors in tests/test_golden.py; over a field the RREF is unique, so
TritMatrix's rref gives the same entries, TritMatrix.h:26-29) on random, planted low-rank, duplicated-row,
sparse, zero and permuted-identity inputs, padding columns and padding words included; at 8192 x 8192
(too large for the oracle in seconds) by what the RREF must satisfy: the planted rank, identity on the
pivot columns, input rows rebuilt from their pivot entries (A = A[:, pivots] R over F_3), idempotence.
The product C = A B (multiply) against the oracle's GEMM at p = 3, k not a multiple of 64, strided rows,
accumulation."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _planted(rng, m, n, r):
    L = rng.integers(0, 3, (m, r)).astype(np.float64)           # sums of products of 0..2 stay exact
    R = rng.integers(0, 3, (r, n)).astype(np.float64)
    return ((L @ R).astype(np.int64) % 3).astype(np.uint32)


def _planted_rows(rng, m, n, r):
    """r random rows (full rank w.h.p.), every other row a combination of two of them, shuffled."""
    A = np.empty((m, n), np.uint32)
    A[:r] = rng.integers(0, 3, (r, n))
    for i in range(r, m):
        a, b = rng.choice(r, 2, replace=False)
        A[i] = (A[a] + 2 * A[b]) % 3
    return A[rng.permutation(m)]


def _device_rref(E, A, ldw=None):
    m, n = A.shape
    W = E.pack_trits(A, ldw)
    dW = torch.from_numpy(W.view(np.int64).copy()).cuda()
    rank, piv = E.f3_rref(dW, n)
    torch.cuda.synchronize()
    return rank, piv, dW.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("shape,kind", [
    ((1, 1), "random"), ((5, 70), "random"), ((64, 64), "random"), ((100, 300), "random"),
    ((700, 129), "planted"), ((129, 700), "planted"), ((1500, 1000), "random"), ((1000, 1500), "planted"),
    ((2048, 512), "dup"), ((300, 3000), "sparse"), ((3000, 300), "sparse"), ((257, 257), "zero"),
    ((1024, 1024), "identity"), ((5000, 400), "planted"),
])
def test_f3_rref_matches_oracle(hx_gpu, O, shape, kind):
    from hacs_amd import engine as E
    m, n = shape
    rng = np.random.default_rng(m * 7907 + n)
    if kind == "random":
        A = rng.integers(0, 3, (m, n)).astype(np.uint32)
    elif kind == "planted":
        A = _planted(rng, m, n, min(m, n) // 3)
    elif kind == "dup":
        A = rng.integers(0, 3, (m, n)).astype(np.uint32)
        A[1::3] = (2 * A[0::3][:len(A[1::3])]) % 3
        A[7::11] = 0
    elif kind == "sparse":
        A = ((rng.random((m, n)) < 0.004) * rng.integers(1, 3, (m, n))).astype(np.uint32)
    elif kind == "zero":
        A = np.zeros((m, n), np.uint32)
    else:
        A = (np.eye(m, n, dtype=np.uint32) * 2)[rng.permutation(m)]
    print(f"F3 RREF {m}x{n} {kind}", flush=True)
    ldw = (n + 63) // 64 + (1 if kind == "sparse" else 0)
    rank, piv, out = _device_rref(E, A, ldw)
    r, opiv, _, R = O.rref_rule(3, A, threads=16)
    assert rank == r, (rank, r)
    assert np.array_equal(piv, opiv)
    got = E.unpack_trits(out, n)
    assert np.array_equal(got[:r], R) and not got[r:].any()
    nw = (n + 63) // 64
    assert not out[:, nw:ldw].any() and not out[:, ldw + nw:].any()      # padding words stay zero
    if n % 64:
        for pl in (0, ldw):
            assert not (out[:, pl + (n - 1) // 64] >> np.uint64(n % 64)).any()   # padding bits stay zero


def test_f3_rref_8192_properties(hx_gpu):
    """8192 x 8192 of planted rank 6000: the rank, identity on the pivot columns, A = A[:, pivots] R over
    F_3 on 256 sampled rows, and the RREF of the RREF is itself."""
    from hacs_amd import engine as E
    m = n = 8192
    rng = np.random.default_rng(83)
    A = _planted_rows(rng, m, n, 6000)
    rank, piv, out = _device_rref(E, A)
    assert rank == 6000
    R = E.unpack_trits(out[:rank], n)
    assert np.array_equal(R[:, piv], np.eye(rank, dtype=np.uint32))
    assert not out[rank:].any()
    rows = rng.choice(m, 256, replace=False)
    rebuilt = (A[rows][:, piv].astype(np.float64) @ R.astype(np.float64)).astype(np.int64) % 3
    assert np.array_equal(rebuilt, A[rows].astype(np.int64))
    rank2, piv2, out2 = _device_rref(E, E.unpack_trits(out, n))
    assert rank2 == rank and np.array_equal(piv2, piv) and np.array_equal(out2, out)


def test_f3_rref_limits(hx_gpu):
    """Past 8192 rows the elimination runs in stacked passes, which need the rank to stay below 8192: a
    full-rank 8300 x 8300 matrix is refused; so is ldw < ceil(n / 64)."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(8300)
    A = rng.integers(0, 3, (8300, 8300)).astype(np.uint32)
    dW = torch.from_numpy(E.pack_trits(A).view(np.int64).copy()).cuda()
    with pytest.raises(E.HxError):
        E.f3_rref(dW, 8300)
    with pytest.raises(E.HxError):
        E.f3_rref(torch.zeros((8, 2), dtype=torch.int64, device="cuda"), 65)   # ldw < ceil(n / 64)


@pytest.mark.parametrize("m,n,r", [(8193, 64, 64), (20000, 600, 400), (30000, 200, 199)])
def test_f3_rref_tall_stacked(hx_gpu, m, n, r):
    """More than 8192 rows (stacked passes of the one-pass elimination): A = L R mod 3 with R a random RREF
    (formed on the device in float32, exact), so RREF(A) = R: its pivots and rank, zero rows below."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(m + n)
    piv = np.sort(rng.choice(n, r, replace=False))
    R = rng.integers(0, 3, (r, n)).astype(np.float32)
    for i, c in enumerate(piv):
        R[i, :c] = 0
    R[:, piv] = np.eye(r, dtype=np.float32)
    L = torch.from_numpy(rng.integers(0, 3, (m, r)).astype(np.float32)).cuda()
    A = torch.remainder(L @ torch.from_numpy(R).cuda(), 3).to(torch.int64).cpu().numpy().astype(np.uint32)
    print(f"F3 tall RREF {m}x{n} rank {r}", flush=True)
    rank, gp, out = _device_rref(E, A)
    assert rank == r and np.array_equal(gp, piv)
    assert np.array_equal(E.unpack_trits(out[:r], n), R.astype(np.uint32)) and not out[r:].any()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 70, 65), (300, 129, 1000), (1024, 1024, 1024), (129, 4096, 200)])
def test_f3_gemm_matches_oracle(hx_gpu, O, m, n, k):
    from hacs_amd import engine as E
    rng = np.random.default_rng(m + 5 * n + 11 * k)
    A = rng.integers(0, 3, (m, k)).astype(np.uint32)
    B = rng.integers(0, 3, (k, n)).astype(np.uint32)
    C0 = rng.integers(0, 3, (m, n)).astype(np.uint32)
    dA = torch.from_numpy(E.pack_trits(A, (k + 63) // 64 + 1).view(np.int64).copy()).cuda()   # strided rows
    dB = torch.from_numpy(E.pack_trits(B).view(np.int64).copy()).cuda()
    want = O.gemm_fast(3, A, B)
    for acc in (False, True):
        dC = torch.from_numpy(E.pack_trits(C0).view(np.int64).copy()).cuda()
        E.f3_gemm(dA, dB, dC, k, n, accumulate=acc)
        torch.cuda.synchronize()
        got = E.unpack_trits(dC.cpu().numpy().view(np.uint64), n)
        assert np.array_equal(got, (want + C0) % 3 if acc else want), f"accumulate={acc}"
