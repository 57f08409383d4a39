"""The whole-GPU RREF path (tournament pivot search, limb-MFMA trailing updates, TRSM back
substitution) against the oracle, and the full 4096 x 4096 rank-3000 configuration mod 2^31-1
against the fixture written by the reference's own elim::rref (tests/golden/c2_reference.npz)."""
import os

import numpy as np
import pytest
import torch

from tests.matrix_kinds import make

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_reference.npz")


@pytest.mark.parametrize("p,m,n,kind,rank", [
    (2147483647, 2048, 1536, "planted", 1100),
    (1000003, 1100, 2048, "planted", 900),
    (65521, 1536, 1536, "random", 0),
    (3, 1280, 1700, "random", 0),
    (2, 1024, 2100, "random", 0),
    (2147483629, 1300, 1700, "duprows", 0),
    (13, 1024, 2048, "boundaries", 0),
])
def test_rref_large_matches_oracle(hx_gpu, O, p, m, n, kind, rank):
    from hacs_amd import engine as E
    rng = np.random.default_rng(m + n + p)
    A = make(rng, m, n, p, kind, rank=rank)
    R = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    res = E.rref(E.ModCtx(p), E.to_device(A), R)
    torch.cuda.synchronize()
    r, piv, prow, want = O.rref_rule(p, A, threads=16)
    assert res.rank == r, (res.rank, r)
    assert np.array_equal(res.pivots, piv)
    assert np.array_equal(res.pivot_rows, prow)
    got = E.from_device(R)
    assert np.array_equal(got[:r], want) and not got[r:].any()


def test_rref_c2_full_size(hx_gpu, O):
    from hacs_amd import engine as E
    from hacs_amd import workloads as W
    G = np.load(GOLD, allow_pickle=False)
    p = int(G["p"][0])
    A = W.c2_device()
    assert O.fnv_matrix(E.from_device(A)) == int(G["input_fnv"][0]), "generator differs from the fixture's input"
    R = torch.empty_like(A)
    res = E.rref(E.ModCtx(p), A, R)
    torch.cuda.synchronize()
    assert res.rank == int(G["rank"][0]) == 3000
    assert np.array_equal(res.pivots, G["pivots"])
    assert np.array_equal(res.pivot_rows, G["pivot_rows"])
    Rh = E.from_device(R)
    assert O.fnv_matrix(Rh[:res.rank]) == int(G["rref_fnv"][0])
    assert not Rh[res.rank:].any()


def test_elimination_counters(hx_gpu, O):
    """The elimination counters (ImmaElimStats, Cuda/CudaKernels.h:289-298) over one whole-GPU RREF: one
    call, ceil(n / 32) panels over n columns, at least one update product per panel but the last, the
    call's workspace as the peak footprint and a positive wall time; reset clears them."""
    import ctypes as C
    from hacs_amd import _lib
    from hacs_amd import engine as E
    L = _lib.lib()
    p, m, n = 2147483647, 1500, 1400
    A = make(np.random.default_rng(9), m, n, p, "planted", rank=1000)
    _lib.check(L.hx_elim_stats_reset())
    res = E.rref(E.ModCtx(p), E.to_device(A))
    st = _lib.hx_elim_stats()
    _lib.check(L.hx_elim_stats_get(C.byref(st)))
    assert res.rank == O.rref_rule(p, A, threads=16)[0]
    assert st.calls == 1 and st.panels == -(-n // 32) and st.panel_columns == n
    assert st.gemms >= st.panels - 1 and st.kernel_launches >= st.panels + st.gemms
    assert st.peak_device_bytes >= L.hx_rref_workspace_bytes(m, n) and st.seconds > 0.0
    assert st.h2d_bytes == 0 and st.d2h_bytes == 0
    _lib.check(L.hx_elim_stats_reset())
    _lib.check(L.hx_elim_stats_get(C.byref(st)))
    assert st.calls == 0 and st.peak_device_bytes == 0 and st.seconds == 0.0


def test_trsm_unit_matches_oracle(hx_gpu, O):
    from hacs_amd import engine as E
    p = 2147483629
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(8)
    for n, q, upper in [(70, 33, True), (70, 33, False), (300, 129, True), (257, 64, False)]:
        T = rng.integers(0, p, (n, n), dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, p, (n, q), dtype=np.uint64).astype(np.uint32)
        dB = E.to_device(B)
        E.trsm_unit(ctx, E.to_device(T), dB, upper)
        torch.cuda.synchronize()
        X = E.from_device(dB)
        Tu = np.triu(T, 1) if upper else np.tril(T, -1)
        Tu = Tu + np.eye(n, dtype=np.uint32)
        assert np.array_equal(O.gemm_fast(p, Tu, X), B), f"T X != B for n={n} upper={upper}"


def test_rref_device_entry_captured_and_replayed(hx_gpu, O):
    """hx_rref_u32_device (no allocation, no read-back) captured in a hipGraph: three replays with eager
    work and synchronisations between them, each equal to the oracle; and the synchronous entry on
    the same matrix allocates nothing from the driver after its first call (the arena reuses its block)."""
    import ctypes as C
    from hacs_amd import engine as E
    from hacs_amd import _lib
    p = 2147483647
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(77)
    m, n = 1200, 1400
    A = make(rng, m, n, p, "planted", rank=900)
    want = O.rref_rule(p, A, threads=16)
    dA = E.to_device(A)
    R = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    out = torch.zeros(1 + 2 * min(m, n), dtype=torch.int32, device="cuda")
    ws = torch.empty(E.rref_workspace_bytes(m, n), dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    L = _lib.lib()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        E.rref_device(ctx, dA, R, out, ws, stream=h)          # eager once (creates the side stream)
        s.synchronize()
        res = E.rref_device_result(out, m, n)
        assert res.rank == want[0]
        g = C.c_void_p()
        _lib.check(L.hx_graph_begin(h))
        E.rref_device(ctx, dA, R, out, ws, stream=h)
        _lib.check(L.hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            R.fill_(7)
            out.fill_(0)
            s.synchronize()
            _lib.check(L.hx_graph_launch(g, h))
            s.synchronize()
            res = E.rref_device_result(out, m, n)
            assert res.rank == want[0] and np.array_equal(res.pivots, want[1]) and np.array_equal(res.pivot_rows, want[2])
            got = E.from_device(R)
            assert np.array_equal(got[:res.rank], want[3]) and not got[res.rank:].any(), f"replay {rep}"
            other = torch.empty((m, n), dtype=torch.int32, device="cuda")
            E.copy(dA, other, stream=h)                       # eager work between replays
            s.synchronize()
        _lib.check(L.hx_graph_destroy(g))
    res1 = E.rref(ctx, dA, R)
    torch.cuda.synchronize()
    st0 = E.arena_stats()
    res2 = E.rref(ctx, dA, R)
    torch.cuda.synchronize()
    st1 = E.arena_stats()
    assert res1.rank == res2.rank == want[0]
    assert st1["driver_mallocs"] == st0["driver_mallocs"], "the second call must reuse the arena block"


def test_staged_transfers(hx_gpu):
    """Pageable uploads / downloads through the pinned chunks (several chunks, a ragged tail) and a pinned
    host buffer straight to the DMA give the bytes back unchanged; the staging counters move."""
    import ctypes as C
    from hacs_amd import _lib
    L = _lib.lib()
    nbytes = (40 << 20) + 12345
    host = np.random.default_rng(1).integers(0, 256, nbytes, dtype=np.uint8)
    dev = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    st0 = _lib.hx_staging_stats()
    _lib.check(L.hx_staging_stats_get(C.byref(st0)))
    _lib.check(L.hx_upload(C.c_void_p(dev.data_ptr()), host.ctypes.data, nbytes, None))
    back = np.zeros_like(host)
    _lib.check(L.hx_download(back.ctypes.data, C.c_void_p(dev.data_ptr()), nbytes, None))
    assert np.array_equal(back, host)
    st1 = _lib.hx_staging_stats()
    _lib.check(L.hx_staging_stats_get(C.byref(st1)))
    assert st1.acquisitions >= st0.acquisitions + 6 and st1.chunk_bytes == 16 << 20
    pin = C.c_void_p()
    _lib.check(L.hx_host_alloc(C.byref(pin), nbytes))
    try:
        C.memmove(pin, host.ctypes.data, nbytes)
        dev.zero_()
        _lib.check(L.hx_upload(C.c_void_p(dev.data_ptr()), pin, nbytes, None))
        assert np.array_equal(dev.cpu().numpy(), host)
    finally:
        _lib.check(L.hx_host_free(pin))
