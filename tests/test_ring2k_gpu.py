"""The blocked LU over Z/2^32 and Z/2^64 on the device session (the device route of
ring2k::echelon_blocked, Binary/Ring2kDeviceLU.h) against the oracle's restatement of the host code,
entrywise: the matrix left behind (native words, not only mod 2^k), the rank, the pivots and their
valuations -- at u32 k = 30 / 32 and u64 k = 61 / 40 (the reference's gate,
DescentSmithGpuValidation.cpp:143-145, 243-246), resident and tiled (phases, panel replays, the left
columns' host permutation, an early phase end when the L record is full), with a column limit below n.
The oracle itself equals the reference's own echelon_blocked on the committed fixture
(tests/test_oracle.py)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _input(rng, dt, m, n, scaled=True):
    A = rng.integers(0, np.iinfo(dt).max, (m, n), dtype=np.uint64, endpoint=True).astype(dt)
    if scaled:                                   # rows scaled by 2^s: pivots of every valuation
        for r0 in range(0, m, 3):
            A[r0] = (A[r0] << dt(rng.integers(0, 12))).astype(dt)
        A[5:m:17] = 0
        if m > 40:
            A[40] = A[7]
    return A


def _device(k, A, panel, climit=None, budget=0, store=0):
    from hacs_amd import _lib
    L = _lib.lib()
    A = np.ascontiguousarray(A).copy()
    m, n = A.shape
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    pv = np.zeros(max(1, min(m, n)), np.uint32)
    rank = C.c_uint32(0)
    _lib.check(L.hx_ring2k_echelon(A.dtype.itemsize, k, m, n, A.ctypes.data_as(C.c_void_p), n, panel,
                                   n if climit is None else climit, budget, store, C.byref(rank),
                                   piv.ctypes.data_as(_lib._U32P), pv.ctypes.data_as(_lib._U32P)))
    r = rank.value
    return A, r, piv[:r].copy(), pv[:r].copy()


def _stats():
    from hacs_amd import _lib
    st = _lib.hx_ring2k_stats()
    _lib.check(_lib.lib().hx_ring2k_stats_get(C.byref(st)))
    return {n: int(getattr(st, n)) for n, _ in st._fields_}


def _check(O, k, A, panel, climit=None, **kw):
    got = _device(k, A, panel, climit, **kw)
    want = O.r2_echelon(k, A, panel, climit)
    assert got[1] == want[1], f"rank {got[1]} vs {want[1]}"
    assert np.array_equal(got[2], want[2]) and np.array_equal(got[3], want[3]), "pivots / valuations"
    bad = np.argwhere(got[0] != want[0])
    assert bad.size == 0, f"{len(bad)} entries differ, first at {bad[:3].tolist()}"
    return got


@pytest.mark.parametrize("dt,k", [(np.uint32, 30), (np.uint32, 32), (np.uint64, 61), (np.uint64, 40)])
def test_echelon_resident(hx_gpu, O, dt, k):
    from hacs_amd import _lib
    rng = np.random.default_rng(k * 7 + np.dtype(dt).itemsize)
    _lib.lib().hx_ring2k_stats_reset()
    for (m, n, panel, climit) in [(300, 260, 32, None), (257, 400, 64, None), (190, 230, 16, 150), (70, 500, 8, None)]:
        A = _input(rng, dt, m, n)
        _check(O, k, A, panel, climit)
    st = _stats()
    assert st["device"] == 4 and st["resident"] == 4 and st["tiled"] == 0 and st["lookaheads"] > 0, st


@pytest.mark.parametrize("dt,k", [(np.uint32, 32), (np.uint64, 61)])
def test_echelon_tiled(hx_gpu, O, dt, k):
    """A memory budget for a window of two panels: phases, replays onto the slabs right of the window,
    the host permutation of the columns left of it; then a small L record that ends phases early."""
    from hacs_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(k + 5)
    e = np.dtype(dt).itemsize
    m, n, panel = 240, 330, 32
    fixed = L.hx_ring2k_session_bytes(e, m, 0, panel)
    budget = fixed + m * e * 2 * panel
    assert L.hx_ring2k_plan(e, m, n, panel, budget) == 2 * panel
    assert L.hx_ring2k_plan(e, m, n, panel, 0) == n
    L.hx_ring2k_stats_reset()
    A = _input(rng, dt, m, n)
    _check(O, k, A, panel, None, budget=budget)
    st = _stats()
    assert st["tiled"] == 1 and st["replays"] > 0 and st["slabs"] > 1, st
    _check(O, k, A, panel, 300, budget=budget, store=int(m * panel * e * 1.5))
    st2 = _stats()
    assert st2["slabs"] > st["slabs"] + 1, st2


def test_echelon_square_large(hx_gpu, O):
    """2048 x 2048 over Z/2^32 with panel 64 (the default-sized route) entrywise."""
    rng = np.random.default_rng(11)
    A = _input(rng, np.uint32, 2048, 2048)
    _check(O, 32, A, 64)


def test_echelon_declines_without_room(hx_gpu):
    from hacs_amd import _lib
    L = _lib.lib()
    A = np.ones((64, 64), np.uint32)
    piv = np.zeros(64, np.uint32)
    pv = np.zeros(64, np.uint32)
    rank = C.c_uint32(0)
    rc = L.hx_ring2k_echelon(4, 32, 64, 64, A.ctypes.data_as(C.c_void_p), 64, 32, 64, 1024, 0, C.byref(rank),
                             piv.ctypes.data_as(_lib._U32P), pv.ctypes.data_as(_lib._U32P))
    assert rc == _lib.HX_E_NOMEM and np.all(A == 1), "a declined call must leave A untouched"


@pytest.mark.parametrize("elem", [4, 8])
def test_gemm_sub_host(hx_gpu, elem):
    """C -= L U in wraparound words against exact integer products reduced mod 2^(8 elem), K above one
    64-deep slice."""
    from hacs_amd import _lib
    L_ = _lib.lib()
    dt = np.uint32 if elem == 4 else np.uint64
    rng = np.random.default_rng(elem)
    m, n, k = 150, 170, 150
    Cm = rng.integers(0, np.iinfo(dt).max, (m, n), dtype=np.uint64, endpoint=True).astype(dt)
    Lm = rng.integers(0, np.iinfo(dt).max, (m, k), dtype=np.uint64, endpoint=True).astype(dt)
    U = rng.integers(0, np.iinfo(dt).max, (k, n), dtype=np.uint64, endpoint=True).astype(dt)
    mod = 1 << (8 * elem)
    want = np.array((Cm.astype(object) - Lm.astype(object).dot(U.astype(object))) % mod, dtype=dt)
    got = Cm.copy()
    _lib.check(L_.hx_ring2k_gemm_sub_host(elem, got.ctypes.data, n, Lm.ctypes.data, k, U.ctypes.data, n, m, n, k))
    assert np.array_equal(got, want)


def test_session_protocol_errors(hx_gpu):
    """Steps outside the session's contract are refused before any device work."""
    from hacs_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.hx_ring2k_open(4, 100, 80, 16, C.byref(h)))
    try:
        bad = _lib.hx_ring2k_step_args(m=99, r=0, np=0, w=0)
        assert L.hx_ring2k_step(h, C.byref(bad)) == _lib.HX_E_ARGUMENT
        bad = _lib.hx_ring2k_step_args(m=100, r=95, np=16, w=0)
        assert L.hx_ring2k_step(h, C.byref(bad)) == _lib.HX_E_ARGUMENT
        out = np.zeros((4, 4), np.uint32)
        assert L.hx_ring2k_fetch(h, out.ctypes.data, 4) == _lib.HX_E_ARGUMENT     # no lookahead pending
        assert L.hx_ring2k_get(h, out.ctypes.data, 4, 98, 4, 0, 4) == _lib.HX_E_ARGUMENT
    finally:
        _lib.check(L.hx_ring2k_close(h))
    assert L.hx_ring2k_open(4, 100, 80, 65, C.byref(h)) == _lib.HX_E_ARGUMENT     # panel above 64
