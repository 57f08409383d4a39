"""The level-1/2 and tape-support kernels (gemv, axpy, scal, copy, transpose, fill, axpy_identity,
gather, canonicalise, gemm_nt) on the GPU against the CPU oracle, bit-exact, through the C-ABI.

Cases follow the reference's kernel gate (Cuda/CudaValidation.cpp): every layout a DeviceView can
take (row-major, column-major, strided sub-windows), ragged and empty shapes, both gemv kernels,
the row-list / column-list / identity-source forms of gather, the argument checks, graph capture,
and the launch counters.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PRIMES = [65521, 2147483647, 4294967291]


def _dev(E, M, layout):
    """A device view of M (uint32) in the given layout: row-major, column-major, or a strided
    window (every second row and every third column of a larger buffer)."""
    m, n = M.shape
    if layout == "row":
        return E.to_device(M)
    if layout == "col":
        return E.to_device(np.ascontiguousarray(M.T)).t()
    big = np.zeros((2 * m + 1, 3 * n + 2), np.uint32)
    big[1::2, 2::3][:m, :n] = M
    t = E.to_device(big)
    return t[1::2, 2::3][:m, :n]


def _host(E, t):
    return E.from_device(t)


def _rand(rng, shape, hi):
    return rng.integers(0, hi, shape, dtype=np.uint64).astype(np.uint32)


LAYOUTS = ["row", "col", "strided"]
SHAPES = [(1, 1), (37, 300), (300, 37), (257, 1031)]


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("layout", LAYOUTS)
def test_elementwise_match_oracle(hx_gpu, O, p, layout):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    for shape in SHAPES:
        rng = np.random.default_rng([p, LAYOUTS.index(layout), shape[0], shape[1]])
        x, y = _rand(rng, shape, p), _rand(rng, shape, p)
        for alpha in (0, 1, p - 1, int(rng.integers(0, p))):
            for sub in (False, True):
                dx, dy = _dev(E, x, layout), _dev(E, y, "row" if layout == "col" else layout)
                E.axpy(ctx, alpha, dx, dy, subtract=sub)
                torch.cuda.synchronize()
                assert np.array_equal(_host(E, dy), O.axpy(p, alpha, x, y, sub)), (shape, alpha, sub)
            dx = _dev(E, x, layout)
            E.scal(ctx, alpha, dx)
            torch.cuda.synchronize()
            assert np.array_equal(_host(E, dx), O.scal(p, alpha, x)), (shape, alpha)
        lazy = _rand(rng, shape, min(2 * p, 1 << 32))
        dv = _dev(E, lazy, layout)
        E.canonicalise(ctx, dv)
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, dv), O.canonicalise(p, lazy))
        # copy into another layout, fill, axpy_identity
        dd = _dev(E, np.zeros(shape, np.uint32), "strided" if layout == "row" else "row")
        E.copy(_dev(E, x, layout), dd)
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, dd), x)
        E.fill(dd, 123457)
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, dd), np.full(shape, 123457, np.uint32))
        for sub in (False, True):
            dy = _dev(E, y, layout)
            a = int(rng.integers(0, p))
            E.axpy_identity(ctx, a, dy, subtract=sub)
            torch.cuda.synchronize()
            assert np.array_equal(_host(E, dy), O.axpy_identity(p, a, y, sub))


@pytest.mark.parametrize("src_layout", LAYOUTS)
@pytest.mark.parametrize("dst_layout", LAYOUTS)
def test_transpose_and_copy_layouts(hx_gpu, src_layout, dst_layout):
    from hacs_amd import engine as E
    rng = np.random.default_rng(5)
    for m, n in [(1, 1), (1, 200), (64, 64), (65, 130), (333, 71)]:
        S = _rand(rng, (m, n), 1 << 32)
        ds = _dev(E, S, src_layout)
        dt = _dev(E, np.zeros((n, m), np.uint32), dst_layout)
        E.transpose(ds, dt)
        dc = _dev(E, np.zeros((m, n), np.uint32), dst_layout)
        E.copy(ds, dc)
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, dt), S.T), (m, n)
        assert np.array_equal(_host(E, dc), S), (m, n)
        E.copy(ds, ds)           # identical views: a no-op
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, ds), S)


@pytest.mark.parametrize("p", PRIMES)
def test_canonicalise_acc64(hx_gpu, O, p):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(p)
    for m, n in [(1, 1), (100, 77), (513, 300)]:
        acc = rng.integers(0, (1 << 64) - 1, (m, n), dtype=np.uint64, endpoint=True)
        special = np.array([0, (1 << 64) - 1, p * (p - 1)], np.uint64)[:m * n]
        acc.ravel()[:len(special)] = special
        for layout in ("row", "col"):
            src = acc if layout == "row" else np.ascontiguousarray(acc.T)
            dacc = torch.from_numpy(src.view(np.int64)).cuda()
            if layout == "col":
                dacc = dacc.t()
            dd = E.to_device(np.zeros((m, n), np.uint32))
            E.canonicalise_acc64(ctx, dacc, dd)
            torch.cuda.synchronize()
            assert np.array_equal(_host(E, dd), O.canonicalise_acc64(p, acc)), (m, n, layout)


@pytest.mark.parametrize("p", PRIMES)
def test_gather_forms(hx_gpu, O, p):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(p + 1)
    ma, na = 300, 517
    a = _rand(rng, (ma, na), p)
    for m, n in [(1, 1), (40, 250), (260, 301)]:
        rows = rng.integers(0, ma, m).astype(np.uint32)
        cols = rng.integers(0, na, n).astype(np.uint32)
        c = _rand(rng, (m, n), p)
        drows, dcols = E.to_device(rows.reshape(-1, 1)), E.to_device(cols.reshape(-1, 1))
        alpha, beta = int(rng.integers(0, p)), int(rng.integers(1, p))
        for layout in ("row", "col"):
            for acc in (False, True):
                for off in (0, 3):
                    dc = _dev(E, c, layout)
                    E.gather(ctx, alpha, beta, _dev(E, a, layout), off, dcols, dc, acc, rows=drows)
                    torch.cuda.synchronize()
                    want = O.gather(p, alpha, beta, a, off, cols, c, acc, rows)
                    assert np.array_equal(_host(E, dc), want), (m, n, layout, acc, off)
        # no lists (a the same shape as c), and the identity source
        a2 = _rand(rng, (m, n), p)
        dc = _dev(E, c, "row")
        E.gather(ctx, alpha, beta, _dev(E, a2, "strided"), 1, None, dc, True)
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, dc), O.gather(p, alpha, beta, a2, 1, None, c, True))
        dc = _dev(E, c, "row")
        E.gather(ctx, alpha, beta, None, 0, dcols, dc, False)
        torch.cuda.synchronize()
        assert np.array_equal(_host(E, dc), O.gather(p, alpha, beta, None, 0, cols, c, False))


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("a_layout", LAYOUTS)
def test_gemv_both_kernels(hx_gpu, O, p, a_layout):
    """a_layout "col" takes the thread-per-row kernel, the others the wave-per-row kernel."""
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(p ^ 0x77)
    for m, k in [(1, 1), (5, 0), (100, 1), (77, 1000), (1000, 4097)]:
        A = _rand(rng, (m, k), p)
        x = _rand(rng, (k, 1), p)
        y0 = _rand(rng, (m, 1), p)
        for op in (E.GemmOp.Set, E.GemmOp.Add, E.GemmOp.Sub):
            # x as a column of a wider matrix (row stride 3), y contiguous
            xb = np.zeros((k, 3), np.uint32)
            xb[:, 1] = x[:, 0]
            dx = E.to_device(xb)[:, 1:2]
            dy = E.to_device(y0)
            E.gemv(ctx, _dev(E, A, a_layout), dx, dy, op)
            torch.cuda.synchronize()
            want = O.gemm_naive(p, A, x, y0, op) if k else (np.zeros_like(y0) if op == 0 else y0.copy())
            assert np.array_equal(_host(E, dy), want), (m, k, op)
    # the worst case: all-(p-1) row times all-(p-1) vector, long k
    k = 20000
    A = np.full((3, k), p - 1, np.uint32)
    x = np.full((k, 1), p - 1, np.uint32)
    dy = E.to_device(np.zeros((3, 1), np.uint32))
    E.gemv(ctx, _dev(E, A, a_layout), E.to_device(x), dy)
    torch.cuda.synchronize()
    assert np.array_equal(_host(E, dy).ravel(), np.full(3, (k * (p - 1) ** 2) % p, np.uint32))


@pytest.mark.parametrize("p", [65521, 2147483629])
def test_gemm_nt(hx_gpu, O, p):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(9)
    for m, n, k in [(7, 9, 11), (130, 70, 300)]:
        A, Bt, C0 = _rand(rng, (m, k), p), _rand(rng, (n, k), p), _rand(rng, (m, n), p)
        for v in (E.HX_GEMM_PLAIN, E.HX_GEMM_LIMB):
            dC = E.to_device(C0)
            E.gemm_nt(ctx, E.to_device(A), E.to_device(Bt), dC, E.GemmOp.Sub, variant=v)
            torch.cuda.synchronize()
            assert np.array_equal(_host(E, dC), O.gemm_naive(p, A, np.ascontiguousarray(Bt.T), C0, 2))


def test_argument_checks(hx_gpu):
    from hacs_amd import engine as E
    from hacs_amd._lib import HxError, HX_E_ARGUMENT
    ctx = E.ModCtx(65521)
    X = E.to_device(np.arange(64 * 64, dtype=np.uint32).reshape(64, 64) % 65521)
    Y = E.to_device(np.zeros((64, 64), np.uint32))
    bad = [
        lambda: E.axpy(ctx, 65521, X, Y),                        # alpha not canonical
        lambda: E.axpy(ctx, 1, X, Y[:, :10]),                    # shape mismatch
        lambda: E.scal(ctx, 70000, X),
        lambda: E.copy(X[:, :32], X[:, 16:48]),                  # overlapping windows
        lambda: E.transpose(X, X),                               # aliases
        lambda: E.transpose(X[:10, :20], Y[:10, :20]),           # shape mismatch
        lambda: E.gemv(ctx, X, X[:, 0:1], X[:, 1:2]),            # y aliases A
        lambda: E.gemv(ctx, X, Y[:, 0:2], Y[:, 2:3]),            # x not a column
        lambda: E.gather(ctx, 1, 1, X, 0, None, X),              # output aliases the source
        lambda: E.gather(ctx, 1, 1, X[:10, :10], 0, None, Y),    # count mismatch without lists
        lambda: E.axpy_identity(ctx, 65521, Y),
    ]
    for i, f in enumerate(bad):
        with pytest.raises(HxError) as ei:
            f()
        assert ei.value.code == HX_E_ARGUMENT, i
    # nothing was written by the refused calls
    torch.cuda.synchronize()
    assert int(Y.abs().sum()) == 0


def test_capture_and_replay(hx_gpu, O):
    """axpy, scal, gather and gemv captured in one hipGraph; every replay equals the eager result."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    p = 2147483647
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(3)
    x, y0, a = _rand(rng, (200, 300), p), _rand(rng, (200, 300), p), _rand(rng, (300, 200), p)
    cols = rng.integers(0, 200, 300).astype(np.uint32)
    dx, dy, da, dcols = E.to_device(x), E.to_device(y0), E.to_device(a), E.to_device(cols.reshape(-1, 1))
    dg = E.to_device(np.zeros((300, 300), np.uint32))
    dv = E.to_device(np.zeros((200, 1), np.uint32))
    want_y = O.scal(p, 5, O.axpy(p, 7, x, y0))
    want_g = O.gather(p, 3, 11, a, 0, cols, np.zeros((300, 300), np.uint32))
    want_v = O.gemm_naive(p, want_y, x[0:1, :].T.copy())
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        g = C.c_void_p()
        _lib.check(_lib.lib().hx_graph_begin(h))
        E.axpy(ctx, 7, dx, dy, stream=h)
        E.scal(ctx, 5, dy, stream=h)
        E.gather(ctx, 3, 11, da, 0, dcols, dg, stream=h)
        E.gemv(ctx, dy, dx[0:1, :].t(), dv, stream=h)
        _lib.check(_lib.lib().hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            dy.copy_(E.to_device(y0))
            dg.zero_()
            dv.zero_()
            s.synchronize()
            _lib.check(_lib.lib().hx_graph_launch(g, h))
            s.synchronize()
            assert np.array_equal(_host(E, dy), want_y), rep
            assert np.array_equal(_host(E, dg), want_g), rep
            assert np.array_equal(_host(E, dv), want_v), rep
        _lib.check(_lib.lib().hx_graph_destroy(g))


def test_launch_stats(hx_gpu):
    from hacs_amd import engine as E
    ctx = E.ModCtx(2147483629)
    A = E.to_device(np.ones((128, 128), np.uint32))
    Cm = E.to_device(np.zeros((128, 128), np.uint32))
    torch.cuda.synchronize()
    E.reset_launch_stats()
    E.gemm(ctx, A, A, Cm, variant=E.HX_GEMM_LIMB)
    E.gemm(ctx, A, A, Cm, variant=E.HX_GEMM_PLAIN)
    E.copy(A, Cm)
    torch.cuda.synchronize()
    st = E.launch_stats()
    assert st["gemm_launches"] == 2 and st["limb_launches"] == 1
    assert st["kernel_launches"] >= 5          # limb: two preps + the product; plain; copy
    E.reset_launch_stats()
    assert E.launch_stats()["kernel_launches"] == 0
