"""GEMM mod p on 64-bit residues (hx_gemm_u64) against the oracle's exact unsigned __int128 product
(itself equal to the reference's kernels::gemm<uint64_t>, tests/test_oracle.py): the moduli of the
reference's u64 gates (2^61 - 1, KernelBitIdentityValidation.cpp:520-604) and 2^64 - 59, moduli just
above 2^32 and 2^63, an even modulus (the long-division epilogue), every op, ragged and strided
views, k = 0, and the all-(p-1) worst case over a K that fills the 160-bit accumulator's top word."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PRIMES = [2305843009213693951, 18446744073709551557, 4294967311, 9223372036854775837, (1 << 64) - 2, 7]


def _draw(rng, p, shape):
    return ((rng.integers(0, 1 << 32, shape, dtype=np.uint64) << np.uint64(32)) |
            rng.integers(0, 1 << 32, shape, dtype=np.uint64)) % np.uint64(p)


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 17, 33), (64, 64, 64), (65, 129, 300), (130, 70, 17), (3, 5, 0)])
def test_gemm_u64_matches_oracle(hx_gpu, O, p, shape):
    from hacs_amd import engine as E
    ctx = E.ModCtx64(p)
    m, n, k = shape
    rng = np.random.default_rng((p ^ (m * 1000003 + n * 1009 + k)) & 0xFFFFFFFF)
    A, B, C0 = _draw(rng, p, (m, k)), _draw(rng, p, (k, n)), _draw(rng, p, (m, n))
    dA, dB = E.to_device64(A), E.to_device64(B)
    for op in (0, 1, 2):
        dC = E.to_device64(C0)
        E.gemm_u64(ctx, dA, dB, dC, op)
        torch.cuda.synchronize()
        want = O.gemm_u64(p, A, B, C0, op, 16) if k else (np.zeros_like(C0) if op == 0 else C0)
        assert np.array_equal(E.from_device64(dC), want), f"op {op}"


def test_gemm_u64_strided_views(hx_gpu, O):
    """A transposed (column-major) A, a strided window of B and C."""
    from hacs_amd import engine as E
    p = 18446744073709551557
    ctx = E.ModCtx64(p)
    rng = np.random.default_rng(5)
    m, n, k = 90, 75, 130
    At = _draw(rng, p, (k, m))
    Bbig = _draw(rng, p, (k + 3, n + 9))
    Cbig = _draw(rng, p, (m + 4, n + 6))
    dAt, dBbig, dCbig = E.to_device64(At), E.to_device64(Bbig), E.to_device64(Cbig)
    E.gemm_u64(ctx, dAt.t(), dBbig[2:2 + k, 5:5 + n], dCbig[1:1 + m, 3:3 + n], 2)
    torch.cuda.synchronize()
    want = Cbig.copy()
    want[1:1 + m, 3:3 + n] = O.gemm_u64(p, At.T.copy(), Bbig[2:2 + k, 5:5 + n].copy(), Cbig[1:1 + m, 3:3 + n].copy(), 2)
    assert np.array_equal(E.from_device64(dCbig), want)


@pytest.mark.parametrize("p", [18446744073709551557, (1 << 64) - 2, 2305843009213693951])
def test_gemm_u64_worst_case_carries(hx_gpu, O, p):
    """Every entry p - 1 over K = 70000: each product is nearly 2^128, so the sum's top word counts up to
    K; the result is the closed form K (p - 1)^2 = K mod p."""
    from hacs_amd import engine as E
    ctx = E.ModCtx64(p)
    K = 70000
    dA = E.to_device64(np.full((8, K), p - 1, np.uint64))
    dB = E.to_device64(np.full((K, 72), p - 1, np.uint64))
    dC = torch.zeros((8, 72), dtype=torch.int64, device="cuda")
    E.gemm_u64(ctx, dA, dB, dC, 0)
    torch.cuda.synchronize()
    assert (E.from_device64(dC) == np.uint64(K % p)).all()
