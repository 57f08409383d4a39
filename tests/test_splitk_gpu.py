"""Split-K of the plain GEMM kernel (GemmOpts::splits / workspace, Cuda/CudaKernels.h:64-66; the reference
combines the partials in a second kernel, cuda_kernels.cu:302-317, here the last slice of a tile to finish
combines them): every split count gives the oracle's bytes for every op, ragged and strided shapes and
moduli up to 2^32; the automatic policy splits thin deep products; a caller workspace that is too small is
refused for an explicit count and skipped by the automatic one; a split product captured in a hipGraph
equals the oracle on every replay."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rand(rng, p, shape):
    return rng.integers(0, p, shape, dtype=np.uint64).astype(np.uint32)


def _stats():
    from hacs_amd import _lib
    st = _lib.hx_stats()
    _lib.check(_lib.lib().hx_stats_get(C.byref(st)))
    return st


@pytest.mark.parametrize("p", [2147483629, 65521, 4294967291, 3])
@pytest.mark.parametrize("shape", [(64, 64, 5000), (100, 70, 3333), (1, 1, 40000), (130, 257, 2048), (7, 9, 40)])
def test_split_k_matches_oracle(hx_gpu, O, p, shape):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    m, n, k = shape
    rng = np.random.default_rng(p ^ (m * 131 + n * 7 + k))
    A, B, C0 = _rand(rng, p, (m, k)), _rand(rng, p, (k, n)), _rand(rng, p, (m, n))
    dA, dB = E.to_device(A), E.to_device(B)
    for splits in (2, 7, 64):
        ws = torch.empty(E.split_workspace_bytes(m, n, k, splits), dtype=torch.uint8, device="cuda")
        for op in (E.GemmOp.Set, E.GemmOp.Add, E.GemmOp.Sub):
            print(f"split-K p={p} {m}x{n}x{k} splits={splits} op={op}", flush=True)
            dC = E.to_device(C0)
            before = _stats().splitk_launches
            E.gemm(ctx, dA, dB, dC, op, variant=E.HX_GEMM_PLAIN, splits=splits, split_workspace=ws)
            torch.cuda.synchronize()
            assert _stats().splitk_launches == before + 1
            assert np.array_equal(E.from_device(dC), O.gemm_fast(p, A, B, C0, op)), f"splits {splits} op {op}"


def test_split_k_strided_views(hx_gpu, O):
    """A transposed A and windows of larger B and C, split nine ways with the stream's own partials."""
    from hacs_amd import engine as E
    p = 2147483647
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(3)
    m, n, k = 90, 75, 4100
    At = _rand(rng, p, (k, m))
    Bbig = _rand(rng, p, (k + 3, n + 9))
    Cbig = _rand(rng, p, (m + 4, n + 6))
    dAt, dBbig, dCbig = E.to_device(At), E.to_device(Bbig), E.to_device(Cbig)
    E.gemm(ctx, dAt.t(), dBbig[2:2 + k, 5:5 + n], dCbig[1:1 + m, 3:3 + n], E.GemmOp.Sub, variant=E.HX_GEMM_PLAIN,
           splits=9)
    torch.cuda.synchronize()
    want = Cbig.copy()
    want[1:1 + m, 3:3 + n] = O.gemm_fast(p, At.T.copy(), Bbig[2:2 + k, 5:5 + n].copy(),
                                         Cbig[1:1 + m, 3:3 + n].copy(), E.GemmOp.Sub)
    assert np.array_equal(E.from_device(dCbig), want)


def test_split_k_automatic_policy(hx_gpu, O):
    """A 256 x 256 x 16384 plain product: the policy splits it (16 tiles, 64 slices of 256), the bytes are
    the oracle's and the unsplit kernel's; an 8192-row product is not split; AUTO takes the split plain
    kernel for this shape rather than the limb route."""
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(4)
    m, n, k = 256, 256, 16384
    assert E.suggest_splits(m, n, k) == 64
    A, B = _rand(rng, p, (m, k)), _rand(rng, p, (k, n))
    dA, dB = E.to_device(A), E.to_device(B)
    want = O.gemm_fast(p, A, B)
    dC = torch.empty((m, n), dtype=torch.int32, device="cuda")
    before = _stats().splitk_launches
    E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_PLAIN)              # splits = 0: automatic
    torch.cuda.synchronize()
    assert _stats().splitk_launches == before + 1
    assert np.array_equal(E.from_device(dC), want)
    dC.fill_(0)
    E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_PLAIN, splits=1)    # never split
    torch.cuda.synchronize()
    assert _stats().splitk_launches == before + 1
    assert np.array_equal(E.from_device(dC), want)
    assert E.suggest_splits(8192, 8192, 8192) == 1
    # AUTO: the limb route's 8 tiles cannot fill the GPU, so the product runs on the split plain kernel
    st0 = _stats()
    dC.fill_(0)
    E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_AUTO)
    torch.cuda.synchronize()
    st1 = _stats()
    assert st1.splitk_launches == st0.splitk_launches + 1 and st1.limb_launches == st0.limb_launches
    assert np.array_equal(E.from_device(dC), want)


def test_split_k_workspace_too_small(hx_gpu, O):
    from hacs_amd import engine as E
    p = 65521
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(5)
    m, n, k = 128, 128, 8192
    A, B = _rand(rng, p, (m, k)), _rand(rng, p, (k, n))
    dA, dB = E.to_device(A), E.to_device(B)
    want = O.gemm_fast(p, A, B)
    tiny = torch.empty(E.split_workspace_bytes(m, n, k, 8) - 4, dtype=torch.uint8, device="cuda")
    dC = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    with pytest.raises(E.HxError):
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_PLAIN, splits=8, split_workspace=tiny)
    before = _stats().splitk_launches
    E.gemm(ctx, dA, dB, dC, E.GemmOp.Set, variant=E.HX_GEMM_PLAIN, splits=0, split_workspace=tiny)
    torch.cuda.synchronize()
    assert _stats().splitk_launches == before                                  # automatic: not split
    assert np.array_equal(E.from_device(dC), want)


def test_split_k_inside_graph_capture(hx_gpu, O):
    """A split product with a caller workspace captured in a hipGraph: the counters are cleared by a kernel
    (no memset node), every replay equals the oracle, with eager work and synchronisations between."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(6)
    m, n, k = 192, 130, 6000
    A, B, C0 = _rand(rng, p, (m, k)), _rand(rng, p, (k, n)), _rand(rng, p, (m, n))
    dA, dB = E.to_device(A), E.to_device(B)
    ws = torch.empty(E.split_workspace_bytes(m, n, k, 12), dtype=torch.uint8, device="cuda")
    want = O.gemm_fast(p, A, B, C0, E.GemmOp.Add)
    dC = E.to_device(C0)
    s = torch.cuda.Stream()
    L = _lib.lib()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        g = C.c_void_p()
        _lib.check(L.hx_graph_begin(h))
        E.gemm(ctx, dA, dB, dC, E.GemmOp.Add, stream=h, variant=E.HX_GEMM_PLAIN, splits=12, split_workspace=ws)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            dC.copy_(E.to_device(C0))
            s.synchronize()
            _lib.check(L.hx_graph_launch(g, h))
            s.synchronize()
            assert np.array_equal(E.from_device(dC), want), f"replay {rep}"
            other = E.to_device(C0)                                            # eager work between replays
            E.gemm(ctx, dA, dB, other, E.GemmOp.Add, stream=h, variant=E.HX_GEMM_PLAIN, splits=5)
            s.synchronize()
            assert np.array_equal(E.from_device(other), want)
        _lib.check(L.hx_graph_destroy(g))
