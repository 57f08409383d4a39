"""GEMM mod p on the GPU against the CPU oracle: bit-exact (memcmp), both device kernels.

Cases follow the reference's device gate (Cuda/CudaValidation.cpp:102-195): every op, strided /
transposed operands, ragged sizes, k = 0, the "rank-1 x all-(p-1)" worst case, and the moduli of
the configs (65521, 1000003, 2147483629, 2^31-1) plus 2^32-5 (plain only: no limb split fits it)
and small moduli exercising 1-, 2- and 3-limb splits.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PRIMES = [251, 65521, 1000003, 16777213, 2147483629, 2147483647, 4294967291]


def _run(E, ctx, A, B, C0, op, variant):
    dA, dB = E.to_device(A), E.to_device(B)
    dC = E.to_device(C0)
    E.gemm(ctx, dA, dB, dC, op, variant=variant)
    torch.cuda.synchronize()
    return E.from_device(dC)


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 17, 33), (64, 128, 64), (65, 129, 65), (130, 70, 300),
                                   (257, 191, 1000), (3, 5, 0)])
def test_gemm_matches_oracle(hx_gpu, O, p, shape):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    m, n, k = shape
    rng = np.random.default_rng(p ^ (m * 1000003 + n * 1009 + k))
    A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
    B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
    C0 = rng.integers(0, p, (m, n), dtype=np.uint64).astype(np.uint32)
    variants = [E.HX_GEMM_PLAIN] + ([E.HX_GEMM_LIMB] if ctx.limbs > 0 else [])
    for op in (E.GemmOp.Set, E.GemmOp.Add, E.GemmOp.Sub):
        want = O.gemm_fast(p, A, B, C0, op) if k else (np.zeros_like(C0) if op == 0 else C0.copy())
        for v in variants:
            got = _run(E, ctx, A, B, C0, op, v)
            assert np.array_equal(got, want), f"p={p} shape={shape} op={op} variant={v}"


@pytest.mark.parametrize("p", [65521, 2147483629, 4294967291])
def test_gemm_worst_case_and_extremes(hx_gpu, O, p):
    """rank-1 A x all-(p-1) B (CudaValidation.cpp:176-194) and the balanced extremes (p +- 1) / 2."""
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    m, n, k = 96, 160, 2048
    rng = np.random.default_rng(7)
    u = rng.integers(0, p, (m, 1), dtype=np.uint64)
    v = rng.integers(0, p, (1, k), dtype=np.uint64)
    A = ((u * v) % p).astype(np.uint32)
    B = np.full((k, n), p - 1, np.uint32)
    ext = np.array([(p - 1) // 2, (p + 1) // 2, p - 1, 0, 1], np.uint32)
    A2 = ext[rng.integers(0, len(ext), (m, k))]
    B2 = ext[rng.integers(0, len(ext), (k, n))]
    variants = [E.HX_GEMM_PLAIN] + ([E.HX_GEMM_LIMB] if ctx.limbs > 0 else [])
    for (a, b) in ((A, B), (A2, B2), (np.full((m, k), p - 1, np.uint32), B)):
        exact = (a.astype(object).dot(b.astype(object)) % p).astype(np.uint32)
        for var in variants:
            got = _run(E, ctx, a, b, np.zeros((m, n), np.uint32), E.GemmOp.Set, var)
            assert np.array_equal(got, exact), f"p={p} variant={var}"


@pytest.mark.parametrize("variant", [1, 2])
def test_gemm_strided_views(hx_gpu, O, variant):
    """Transposed and windowed operands (every layout x transpose combination is one view)."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(3)
    big = rng.integers(0, p, (300, 400), dtype=np.uint64).astype(np.uint32)
    dbig = E.to_device(big)
    A = dbig[10:110, 5:205]            # 100 x 200 window
    Bt = dbig[50:180, 200:400]         # 130 x 200 -> B = Bt^T is 200 x 130
    B = Bt.t()
    C = torch.zeros((130, 100), dtype=torch.int32, device="cuda").t()   # 100 x 130 column-major
    E.gemm(ctx, A, B, C, E.GemmOp.Set, variant=variant)
    torch.cuda.synchronize()
    want = O.gemm_fast(p, big[10:110, 5:205], big[50:180, 200:400].T.copy())
    assert np.array_equal(E.from_device(C.contiguous()), want)


def test_gemm_rejects_bad_input(hx_gpu):
    from hacs_amd import engine as E
    ctx = E.ModCtx(4294967291)
    a = torch.zeros((4, 4), dtype=torch.int32, device="cuda")
    b = torch.zeros((5, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(E.HxError):
        E.gemm(ctx, a, b, torch.zeros((4, 4), dtype=torch.int32, device="cuda"))      # shape
    with pytest.raises(E.HxError):
        E.gemm(ctx, a, a, a)                                                          # aliasing
    with pytest.raises(E.HxError):
        E.gemm(ctx, a, a, torch.zeros((4, 4), dtype=torch.int32, device="cuda"), variant=E.HX_GEMM_LIMB)


def test_config1_gemm_1024_mod_65521(hx_gpu, O):
    """The CPU-runnable config: C = A B mod 65521, 1024^2, A/B = EngineMatrix(1024, 1024, 65521, 0x01/0x02)."""
    from hacs_amd import engine as E
    p, n = 65521, 1024
    ctx = E.ModCtx(p)
    A = O.engine_matrix(n, n, p, 0x01)
    B = O.engine_matrix(n, n, p, 0x02)
    want = O.gemm_fast(p, A, B)
    dA = E.random_entries(n, n, p, O.seed_for(n, n, p, 0x01))
    dB = E.random_entries(n, n, p, O.seed_for(n, n, p, 0x02))
    assert np.array_equal(E.from_device(dA), A) and np.array_equal(E.from_device(dB), B)
    for v in (E.HX_GEMM_PLAIN, E.HX_GEMM_LIMB):
        dC = torch.empty((n, n), dtype=torch.int32, device="cuda")
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=v)
        torch.cuda.synchronize()
        got = E.from_device(dC)
        assert np.array_equal(got, want), f"variant {v}"
        assert O.fnv_matrix(got) == O.fnv_matrix(want)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_gemm_batch_uneven(hx_gpu, O, variant):
    """Uneven shapes, mixed ops and per-job variants in one batch, each job against the oracle."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(40 + variant)
    host, jobs = [], []
    for t in range(37):
        m, n, k = (int(x) for x in rng.integers(1, 300, 3))
        if t % 9 == 0:
            k = 0
        A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
        C0 = rng.integers(0, p, (m, n), dtype=np.uint64).astype(np.uint32)
        op = t % 3
        jv = (0, 1, 2)[t % 3] if variant == 0 else 0
        host.append((A, B, C0, op))
        jobs.append((E.to_device(A), E.to_device(B), E.to_device(C0), op, jv))
    b = E.GemmBatch(ctx, jobs, variant)
    for rep in range(2):     # run twice: the second run must not depend on the first
        for (A, B, C0, op), jb in zip(host, jobs):
            jb[2].copy_(E.to_device(C0))
        b.run()
        torch.cuda.synchronize()
        for (A, B, C0, op), jb in zip(host, jobs):
            if A.shape[1]:
                want = O.gemm_fast(p, A, B, C0, op)
            else:
                want = np.zeros_like(C0) if op == 0 else C0
            assert np.array_equal(E.from_device(jb[2]), want), f"rep {rep} {A.shape} x {B.shape} op {op}"


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("p", [251, 65521, 16777213, 2147483629])
def test_limb_kernels_all_copy_paths(hx_gpu, O, p, tile):
    """Every limb kernel (tile 1: register-staged LDS copy; tile 2: direct-to-LDS three-stage
    ring; tile 3: the ring with the LDS reads pipelined against the MFMAs; tile 4: tile 3's
    pipeline on eight waves, one 32 x 32 block each; tile 5: 96 x 128 tiles on twelve waves, whose
    last row tile reads past the plane's rows when m is not a multiple of 96) and the residue route
    (tile 6: one int8 product per small modulus on 256 x 256 tiles, CRT into C) for every limb count,
    with k covering 1..5 LDS stages (the rings' unroll by three), ragged m / n, every op, and a k
    above one limb chunk (two accumulation passes for the limb kernels, one for the residues)."""
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    assert ctx.limbs in (1, 2, 3, 4)
    rng = np.random.default_rng(p + tile)
    shapes = [(64, 128, 64), (65, 129, 128), (200, 130, 192), (31, 300, 256), (129, 257, 320), (96, 96, 1000),
              (193, 64, 448), (288, 200, 64)]
    for (m, n, k) in shapes:
        A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
        C0 = rng.integers(0, p, (m, n), dtype=np.uint64).astype(np.uint32)
        dA, dB = E.to_device(A), E.to_device(B)
        for op in (E.GemmOp.Set, E.GemmOp.Add, E.GemmOp.Sub):
            dC = E.to_device(C0)
            E.gemm(ctx, dA, dB, dC, op, variant=E.HX_GEMM_LIMB, tile=tile)
            torch.cuda.synchronize()
            assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B, C0, op)), f"{(m, n, k)} op {op}"
    if ctx.limbs == 4:
        m, n, k = 70, 130, 32704 + 300          # limb_k_chunk(4) = 32704
        A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
        dC = torch.empty((m, n), dtype=torch.int32, device="cuda")
        E.gemm(ctx, E.to_device(A), E.to_device(B), dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=tile)
        torch.cuda.synchronize()
        assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B))


def test_mfma_rate_calibration(hx_gpu):
    """hx_mfma_rate: a finite positive rate below the published 5000 TOP/s dense INT8 figure for both operand
    kinds, and out-of-range sizes rejected."""
    from hacs_amd import engine as E
    for rnd in (True, False):
        r = E.mfma_rate(512, 200, rnd)
        assert 0.0 < r < 5000.0 * 1.05, (rnd, r)
    for blocks, iters in ((0, 10), (65537, 10), (8, 0)):
        with pytest.raises(E.HxError):
            E.mfma_rate(blocks, iters)


def test_residue_route_bounds(hx_gpu, O):
    """The residue route (tile 6) at its edges: operands all equal to v with v = 127 mod 255, so the first
    modulus' int32 sums reach k * 127^2 = 2147479576 at k = 133144 (the largest k it takes; 10 moduli for
    p = 2147483629), and all equal to -(p-1)/2, the largest |x| = k (p-1)^2/4 the CRT must place; the closed
    form k v^2 mod p is the answer.  Past k = 133144 it runs the limb kernels (same bytes).  A caller workspace
    of hx_gemm_workspace_bytes_tile bytes serves it; one byte less fails."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    half = (p - 1) // 2
    v127 = 127 + 255 * 4000000
    assert v127 % 255 == 127 and v127 < p
    m, n = 64, 128
    dC = torch.empty((m, n), dtype=torch.int32, device="cuda")
    for k in (133144, 133144 + 64):
        for v in (v127, (p - half) % p):
            print(f"residue route k={k} v={v}", flush=True)
            A = torch.full((m, k), v, dtype=torch.int32, device="cuda")
            B = torch.full((k, n), v, dtype=torch.int32, device="cuda")
            E.gemm(ctx, A, B, dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=6)
            torch.cuda.synchronize()
            want = (k * (v * v % p)) % p
            assert (E.from_device(dC) == want).all(), (k, v)
            del A, B
    rng = np.random.default_rng(6)
    m, n, k = 300, 260, 700
    A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
    B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
    need = E.gemm_workspace_bytes(ctx, m, n, k, E.HX_GEMM_LIMB, tile=6)
    assert E.residue_moduli(p, k) == 9                       # 20 M_9 > 41 k (p/2)^2 at k = 700
    assert need == 9 * (512 * 704 + 512 * 704 + 512 * 512)
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    dC = torch.empty((m, n), dtype=torch.int32, device="cuda")
    E.gemm(ctx, E.to_device(A), E.to_device(B), dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=6, workspace=ws)
    torch.cuda.synchronize()
    assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B))
    with pytest.raises(E.HxError):
        E.gemm(ctx, E.to_device(A), E.to_device(B), dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=6,
               workspace=ws[:need - 1])


@pytest.mark.parametrize("k", [701, 1030, 2048])
def test_residue_route_operand_layouts(hx_gpu, O, k):
    """The residue route's operand splits on every layout: row-major A with 16-byte aligned rows (the
    direct split; k = 1030 ends in a partial four-entry group), rows of an odd length (k = 701), a row-major
    B and a transposed A view (the column-strided split; k = 701 and 1030 end in a partial 128-k tile), a
    transposed B view, an A view with no unit stride (the LDS-staged split), and every op, against the
    oracle."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(k)
    m, n = 300, 260
    A = rng.integers(0, p, (m, k), dtype=np.uint64).astype(np.uint32)
    B = rng.integers(0, p, (k, n), dtype=np.uint64).astype(np.uint32)
    C0 = rng.integers(0, p, (m, n), dtype=np.uint64).astype(np.uint32)
    dA, dB = E.to_device(A), E.to_device(B)
    dAt = E.to_device(np.ascontiguousarray(A.T)).t()          # m x k view with unit row stride
    dBt = E.to_device(np.ascontiguousarray(B.T)).t()
    big = np.zeros((2 * m, 2 * k), np.uint32)
    big[::2, ::2] = A
    dAs = E.to_device(big)[::2, ::2]                             # neither stride unit: the LDS-staged split
    for (a, b) in ((dA, dB), (dAt, dB), (dA, dBt), (dAs, dB)):
        for op in (E.GemmOp.Set, E.GemmOp.Add, E.GemmOp.Sub):
            dC = E.to_device(C0)
            E.gemm(ctx, a, b, dC, op, variant=E.HX_GEMM_LIMB, tile=6)
            torch.cuda.synchronize()
            assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B, C0, op)), (a.stride(), b.stride(), op)
