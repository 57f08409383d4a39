"""The limb route's workspace contract (GemmOpts.imma_workspace, Cuda/CudaKernels.h:72-75, 100-110):
a caller workspace smaller than one K piece runs K in chunks (same bytes), also inside a hipGraph
capture; a NULL workspace uses planes private to the stream, so two streams never share them; the
limb digits at their extremes over a full int32-exact K chunk; the device arena; and RREF shapes that
one workgroup cannot hold (more than 4096 rows, rows wider than its LDS panel)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rand(rng, p, shape):
    return rng.integers(0, p, shape, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("p", [65521, 16777213, 2147483629])
def test_limb_chunked_in_small_workspace(hx_gpu, O, p):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(p)
    m, n, k = 300, 260, 1000
    A, B, C0 = _rand(rng, p, (m, k)), _rand(rng, p, (k, n)), _rand(rng, p, (m, n))
    dA, dB = E.to_device(A), E.to_device(B)
    for chunk in (64, 128, 448):
        nbytes = E.gemm_workspace_bytes(ctx, m, n, chunk)
        assert E.limb_k_chunk(ctx, m, n, k, nbytes) == chunk
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        for op in (E.GemmOp.Set, E.GemmOp.Add, E.GemmOp.Sub):
            dC = E.to_device(C0)
            E.gemm(ctx, dA, dB, dC, op, variant=E.HX_GEMM_LIMB, workspace=ws)
            torch.cuda.synchronize()
            assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B, C0, op)), f"chunk {chunk} op {op}"
    # below one 64-deep chunk: LIMB refuses, AUTO runs the plain kernel
    tiny = torch.empty(E.gemm_workspace_bytes(ctx, m, n, 64) - 1, dtype=torch.uint8, device="cuda")
    dC = torch.empty((m, n), dtype=torch.int32, device="cuda")
    with pytest.raises(E.HxError):
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, workspace=tiny)
    E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_AUTO, workspace=tiny)
    torch.cuda.synchronize()
    assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B))


def test_limb_small_workspace_inside_graph_capture(hx_gpu, O):
    """A limb GEMM with an undersized caller workspace (K in 128-deep chunks) captured in a hipGraph:
    every replay equals the oracle, with eager work and synchronisations between the replays."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(7)
    m, n, k = 192, 320, 700
    A, B = _rand(rng, p, (m, k)), _rand(rng, p, (k, n))
    dA, dB = E.to_device(A), E.to_device(B)
    ws = torch.empty(E.gemm_workspace_bytes(ctx, m, n, 128), dtype=torch.uint8, device="cuda")
    want = O.gemm_fast(p, A, B)
    dC = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    L = _lib.lib()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        g = C.c_void_p()
        _lib.check(L.hx_graph_begin(h))
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, stream=h, variant=E.HX_GEMM_LIMB, workspace=ws)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            dC.fill_(12345)
            s.synchronize()
            _lib.check(L.hx_graph_launch(g, h))
            s.synchronize()
            assert np.array_equal(E.from_device(dC), want), f"replay {rep}"
            other = torch.empty((m, n), dtype=torch.int32, device="cuda")     # eager work between replays
            E.gemm(ctx, dA, dB, other, E.GemmOp.Set, stream=h, variant=E.HX_GEMM_PLAIN)
            s.synchronize()
            assert np.array_equal(E.from_device(other), want)
        _lib.check(L.hx_graph_destroy(g))


def test_limb_null_workspace_inside_capture(hx_gpu, O):
    """NULL workspace inside a capture: planes already grown on this stream are used; planes that would
    have to grow make LIMB fail with HX_E_CAPTURE (nothing launched) and AUTO run the plain kernel."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(8)
    L = _lib.lib()
    s = torch.cuda.Stream()
    small, big = (128, 128, 256), (256, 384, 512)
    ops = {}
    for shp in (small, big):
        m, n, k = shp
        A, B = _rand(rng, p, (m, k)), _rand(rng, p, (k, n))
        ops[shp] = (E.to_device(A), E.to_device(B), torch.zeros((m, n), dtype=torch.int32, device="cuda"),
                    O.gemm_fast(p, A, B))
    with torch.cuda.stream(s):
        h = s.cuda_stream
        dA, dB, dC, want = ops[small]
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, stream=h, variant=E.HX_GEMM_LIMB)   # grows this stream's planes
        s.synchronize()
        g = C.c_void_p()
        _lib.check(L.hx_graph_begin(h))
        dC.fill_(0)
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, stream=h, variant=E.HX_GEMM_LIMB)
        bA, bB, bC, bwant = ops[big]
        with pytest.raises(E.HxError) as ei:
            E.gemm(ctx, bA, bB, bC, E.GemmOp.Set, stream=h, variant=E.HX_GEMM_LIMB)
        assert ei.value.code == _lib.HX_E_CAPTURE
        E.gemm(ctx, bA, bB, bC, E.GemmOp.Set, stream=h, variant=E.HX_GEMM_AUTO)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        for rep in range(2):
            _lib.check(L.hx_graph_launch(g, h))
            s.synchronize()
            assert np.array_equal(E.from_device(dC), want) and np.array_equal(E.from_device(bC), bwant), rep
        _lib.check(L.hx_graph_destroy(g))


def test_two_streams_null_workspace(hx_gpu, O):
    """Limb GEMMs with NULL workspaces on two streams at once (different shapes, each stream's planes
    growing while the other runs) all equal the oracle: the planes are private to each stream."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(9)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    cases = []
    for i, (m, n, k) in enumerate([(256, 256, 512), (512, 384, 1024), (384, 512, 2048), (640, 640, 768)]):
        A, B = _rand(rng, p, (m, k)), _rand(rng, p, (k, n))
        cases.append((i % 2, E.to_device(A), E.to_device(B), torch.zeros((m, n), dtype=torch.int32, device="cuda"),
                      O.gemm_fast(p, A, B)))
    torch.cuda.synchronize()
    for rnd in range(3):
        for si, dA, dB, dC, _ in cases:
            with torch.cuda.stream(streams[si]):
                E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, stream=streams[si].cuda_stream, variant=E.HX_GEMM_LIMB)
        torch.cuda.synchronize()
        for si, _, _, dC, want in cases:
            assert np.array_equal(E.from_device(dC), want), f"round {rnd} stream {si} {tuple(dC.shape)}"
            dC.zero_()
        torch.cuda.synchronize()


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6])
def test_limb_extreme_digits_full_chunk(hx_gpu, O, tile):
    """Operands whose balanced limbs sit at the ends of the digit range, over K = 32704 (one full
    int32-exact chunk at four limbs, the bound |D_d| <= pairs(d) K 2^14 nearly met on every diagonal)
    and K = 32704 + 64 (two chunks), for every limb kernel."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    assert ctx.limbs == 4
    half = (p - 1) // 2
    # balanced values whose four digits are all -128 or all +127 where the range allows: -(p-1)/2 ..
    ext = [(p - half) % p, half, (-(128 + 128 * 256 + 128 * 65536)) % p, 127 + 127 * 256 + 127 * 65536, p - 1, 1]
    rng = np.random.default_rng(tile)
    for k in (32704, 32704 + 64):
        m, n = 64, 128
        A = np.array(ext, np.uint32)[rng.integers(0, 2, (m, k))]          # the two extreme halves
        B = np.array(ext, np.uint32)[rng.integers(0, 2, (k, n))]
        A[:, :16] = np.array(ext, np.uint32)[rng.integers(0, len(ext), (m, 16))]
        dC = torch.empty((m, n), dtype=torch.int32, device="cuda")
        E.gemm(ctx, E.to_device(A), E.to_device(B), dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=tile)
        torch.cuda.synchronize()
        assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B)), f"K = {k}"
        # all-equal operands: every product on a diagonal has the same sign, the bound's worst case
        A1 = np.full((m, k), (p - half) % p, np.uint32)
        B1 = np.full((k, n), (p - half) % p, np.uint32)
        E.gemm(ctx, E.to_device(A1), E.to_device(B1), dC, E.GemmOp.Set, variant=E.HX_GEMM_LIMB, tile=tile)
        torch.cuda.synchronize()
        want = np.full((m, n), (k * ((half * half) % p)) % p, np.uint32)    # closed form: k (p-1)^2/4
        assert np.array_equal(E.from_device(dC), want), f"all-equal K = {k}"


def test_device_arena_reuse_and_trim(hx_gpu):
    from hacs_amd import engine as E
    from hacs_amd import _lib
    L = _lib.lib()
    st0 = E.arena_stats()
    ptrs = []
    for nbytes in (1000, 5 << 20, 5 << 20, 3000):
        p_ = C.c_void_p()
        _lib.check(L.hx_malloc(C.byref(p_), nbytes))
        assert p_.value
        ptrs.append((p_, nbytes))
    st1 = E.arena_stats()
    assert st1["live_bytes"] >= st0["live_bytes"] + 1000 + (10 << 20) + 3000
    for p_, _ in ptrs:
        _lib.check(L.hx_free(p_))
    p_ = C.c_void_p()
    _lib.check(L.hx_malloc(C.byref(p_), 5 << 20))        # served from the free list
    st2 = E.arena_stats()
    assert st2["reuses"] >= st1["reuses"] + 1
    _lib.check(L.hx_free(p_))
    assert L.hx_free(p_) == _lib.HX_E_ARGUMENT          # freed twice
    freed = E.arena_trim()
    st3 = E.arena_stats()
    assert freed >= (10 << 20) and st3["reserved_bytes"] == st3["live_bytes"]


@pytest.mark.parametrize("p,m,n,rank", [(1000003, 5000, 400, 300), (2147483647, 16, 40000, 12),
                                         (65521, 600, 37000, 450), (13, 4500, 900, 0)])
def test_rref_shapes_beyond_one_workgroup(hx_gpu, O, p, m, n, rank):
    """Tall (m > 4096) and wide (rows beyond the one-workgroup LDS panel) shapes take the whole-GPU
    path and equal the oracle."""
    from hacs_amd import engine as E
    from tests.matrix_kinds import make
    rng = np.random.default_rng(m * 7 + n)
    A = make(rng, m, n, p, "planted" if rank else "random", rank=rank)
    R = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    res = E.rref(E.ModCtx(p), E.to_device(A), R)
    torch.cuda.synchronize()
    r, piv, prow, want = O.rref_rule(p, A, threads=16)
    assert res.rank == r and np.array_equal(res.pivots, piv) and np.array_equal(res.pivot_rows, prow)
    got = E.from_device(R)
    assert np.array_equal(got[:r], want) and not got[r:].any()


def test_residue_route_inside_graph_capture(hx_gpu, O):
    """The residue route (tile 6) with a caller workspace captured in a hipGraph (its splits, product and CRT
    are kernels only, no memset): three replays each equal the oracle, with eager work and synchronisations
    between them; without a workspace inside the capture it is refused (the library's own cannot grow there)."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(66)
    m, n, k = 300, 260, 1030
    A, B = _rand(rng, p, (m, k)), _rand(rng, p, (k, n))
    C0 = _rand(rng, p, (m, n))
    dA, dB = E.to_device(A), E.to_device(B)
    ws = torch.empty(E.gemm_workspace_bytes(ctx, m, n, k, E.HX_GEMM_LIMB, tile=6), dtype=torch.uint8, device="cuda")
    want = O.gemm_fast(p, A, B, C0, 1)
    dC = E.to_device(C0)
    s = torch.cuda.Stream()
    L = _lib.lib()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        g = C.c_void_p()
        _lib.check(L.hx_graph_begin(h))
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Add, stream=h, variant=E.HX_GEMM_LIMB, tile=6, workspace=ws)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            dC.copy_(E.to_device(C0))
            s.synchronize()
            _lib.check(L.hx_graph_launch(g, h))
            s.synchronize()
            assert np.array_equal(E.from_device(dC), want), f"replay {rep}"
            other = E.to_device(C0)                                           # eager work between replays
            E.gemm(ctx, dA, dB, other, E.GemmOp.Add, stream=h, variant=E.HX_GEMM_PLAIN)
            s.synchronize()
            assert np.array_equal(E.from_device(other), want)
        _lib.check(L.hx_graph_destroy(g))
        fresh = torch.cuda.Stream()
    with torch.cuda.stream(fresh):
        h2 = fresh.cuda_stream
        g2 = C.c_void_p()
        _lib.check(L.hx_graph_begin(h2))
        try:
            with pytest.raises(E.HxError):
                E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, stream=h2, variant=E.HX_GEMM_LIMB, tile=6)
        finally:
            rc = L.hx_graph_end(h2, C.byref(g2))
            if rc == 0 and g2.value:
                L.hx_graph_destroy(g2)
