"""F_2 on bit-packed rows (hx_f2.hip; the reference's binary::BitMatrix, Binary/BitMatrix.h): the in-place
RREF, its rank and pivot columns against the oracle's smallest-row rule at p = 2 (which is elim::rref's
definition, pinned by the reference's own vectors in tests/test_golden.py; over F_2 the RREF is unique, so
BitMatrix's rref gives the same bytes, BitMatrix.h:24-28) on random, low-rank, duplicated-row, sparse and
edge shapes, padding columns included; at 8192 x 8192 (too large for the oracle in seconds) by what the
RREF must satisfy: the planted rank, identity on the pivot columns, every input row rebuilt from its pivot
entries (A = A[:, pivots] R over F_2), and idempotence.  The product C = A B (multiply_m4rm) against the
oracle's GEMM at p = 2, with k not a multiple of 64, strided rows and accumulation."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _planted(rng, m, n, r):
    L = rng.integers(0, 2, (m, r)).astype(np.float64)           # 0 / 1 sums are exact in float64
    R = rng.integers(0, 2, (r, n)).astype(np.float64)
    return ((L @ R).astype(np.int64) & 1).astype(np.uint32)


def _planted_rows(rng, m, n, r):
    """r random rows (full rank w.h.p.), the other m - r rows each the sum of two of them, shuffled."""
    A = np.empty((m, n), np.uint32)
    A[:r] = rng.integers(0, 2, (r, n))
    for i in range(r, m):
        a, b = rng.choice(r, 2, replace=False)
        A[i] = A[a] ^ A[b]
    return A[rng.permutation(m)]


def _device_rref(E, A, ldw=None):
    m, n = A.shape
    W = E.pack_bits(A, ldw)
    dW = torch.from_numpy(W.view(np.int64).copy()).cuda()
    rank, piv = E.f2_rref(dW, n)
    torch.cuda.synchronize()
    out = dW.cpu().numpy().view(np.uint64)
    return rank, piv, out


@pytest.mark.parametrize("shape,kind", [
    ((1, 1), "random"), ((5, 70), "random"), ((64, 64), "random"), ((100, 300), "random"),
    ((700, 129), "planted"), ((129, 700), "planted"), ((1500, 1000), "random"), ((1000, 1500), "planted"),
    ((2048, 512), "dup"), ((300, 3000), "sparse"), ((3000, 300), "sparse"), ((257, 257), "zero"),
    ((1024, 1024), "identity"),
])
def test_f2_rref_matches_oracle(hx_gpu, O, shape, kind):
    from hacs_amd import engine as E
    m, n = shape
    rng = np.random.default_rng(m * 7919 + n)
    if kind == "random":
        A = rng.integers(0, 2, (m, n)).astype(np.uint32)
    elif kind == "planted":
        A = _planted(rng, m, n, min(m, n) // 3)
    elif kind == "dup":
        A = rng.integers(0, 2, (m, n)).astype(np.uint32)
        A[1::3] = A[0::3][:len(A[1::3])]
        A[7::11] = 0
    elif kind == "sparse":
        A = (rng.random((m, n)) < 0.004).astype(np.uint32)
    elif kind == "zero":
        A = np.zeros((m, n), np.uint32)
    else:
        A = np.eye(m, n, dtype=np.uint32)[rng.permutation(m)]
    print(f"F2 RREF {m}x{n} {kind}", flush=True)
    rank, piv, out = _device_rref(E, A, ldw=(n + 63) // 64 + (1 if kind == "sparse" else 0))
    r, opiv, _, R = O.rref_rule(2, A, threads=16)
    assert rank == r, (rank, r)
    assert np.array_equal(piv, opiv)
    got = E.unpack_bits(out, n)
    assert np.array_equal(got[:r], R) and not got[r:].any()
    assert not (out[:, (n + 63) // 64:]).any()                       # padding words stay zero
    if n % 64:
        assert not (out[:, (n - 1) // 64] >> np.uint64(n % 64)).any()   # padding bits stay zero


def test_f2_rref_8192_properties(hx_gpu):
    """8192 x 8192 of planted rank 6000: the rank, identity on the pivot columns, A = A[:, pivots] R over
    F_2 on 512 sampled rows, and the RREF of the RREF is itself."""
    from hacs_amd import engine as E
    m = n = 8192
    rng = np.random.default_rng(81)
    A = _planted_rows(rng, m, n, 6000)
    rank, piv, out = _device_rref(E, A)
    assert rank == 6000
    R = E.unpack_bits(out[:rank], n)
    assert np.array_equal(R[:, piv], np.eye(rank, dtype=np.uint32))
    assert not out[rank:].any()
    rows = rng.choice(m, 512, replace=False)
    rebuilt = (A[rows][:, piv].astype(np.float64) @ R.astype(np.float64)).astype(np.int64) & 1
    assert np.array_equal(rebuilt, A[rows].astype(np.int64))
    rank2, piv2, out2 = _device_rref(E, E.unpack_bits(out, n))
    assert rank2 == rank and np.array_equal(piv2, piv) and np.array_equal(out2, out)


def test_f2_rref_limits(hx_gpu):
    """Past 16384 rows the elimination runs in stacked passes, which need the rank to stay below 16384: a
    full-rank 16400 x 16400 matrix is refused; so is ldw < ceil(n / 64)."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(16400)
    W = torch.from_numpy(rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, (16400, 257), dtype=np.int64,
                                      endpoint=True)).cuda()                  # every bit of every word random
    W[:, 256] &= (1 << 16) - 1                                            # 16400 columns: 256 words + 16 bits
    with pytest.raises(E.HxError):
        E.f2_rref(W, 16400)
    with pytest.raises(E.HxError):
        E.f2_rref(torch.zeros((8, 1), dtype=torch.int64, device="cuda"), 65)   # ldw < ceil(n / 64)


def _planted_rref(rng, p, m, n, r):
    """A = L R mod p with R (r x n) a random matrix in reduced row echelon form (pivots `piv`) and L (m x r)
    random (full column rank with overwhelming probability), formed on the device in float32 (exact: every
    sum stays below 2^24): RREF(A) = R, since the RREF is a function of the row space alone."""
    piv = np.sort(rng.choice(n, r, replace=False))
    R = rng.integers(0, p, (r, n)).astype(np.float32)
    for i, c in enumerate(piv):
        R[i, :c] = 0
    R[:, piv] = np.eye(r, dtype=np.float32)
    L = torch.from_numpy(rng.integers(0, p, (m, r)).astype(np.float32)).cuda()
    A = torch.remainder(L @ torch.from_numpy(R).cuda(), p).to(torch.uint8).cpu().numpy()
    return A, R.astype(np.uint32), piv


@pytest.mark.parametrize("m,n,r", [(16385, 64, 64), (40000, 1000, 700), (70000, 300, 299)])
def test_f2_rref_tall_stacked(hx_gpu, m, n, r):
    """More than 16384 rows (stacked passes of the one-pass elimination): the planted RREF, its pivots and
    rank, zero rows below."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(m + n)
    A, R, piv = _planted_rref(rng, 2, m, n, r)
    print(f"F2 tall RREF {m}x{n} rank {r}", flush=True)
    rank, gp, out = _device_rref(E, A)
    assert rank == r and np.array_equal(gp, piv)
    assert np.array_equal(E.unpack_bits(out[:r], n), R) and not out[r:].any()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 70, 65), (300, 129, 1000), (1024, 1024, 1024), (129, 4096, 200)])
def test_f2_gemm_matches_oracle(hx_gpu, O, m, n, k):
    from hacs_amd import engine as E
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = rng.integers(0, 2, (m, k)).astype(np.uint32)
    B = rng.integers(0, 2, (k, n)).astype(np.uint32)
    C0 = rng.integers(0, 2, (m, n)).astype(np.uint32)
    dA = torch.from_numpy(E.pack_bits(A, (k + 63) // 64 + 1).view(np.int64).copy()).cuda()   # strided rows
    dB = torch.from_numpy(E.pack_bits(B).view(np.int64).copy()).cuda()
    want = O.gemm_fast(2, A, B)
    for acc in (False, True):
        dC = torch.from_numpy(E.pack_bits(C0).view(np.int64).copy()).cuda()
        E.f2_gemm(dA, dB, dC, k, n, accumulate=acc)
        torch.cuda.synchronize()
        got = E.unpack_bits(dC.cpu().numpy().view(np.uint64), n)
        assert np.array_equal(got, want ^ C0 if acc else want), f"accumulate={acc}"
