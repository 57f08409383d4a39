"""The Z/2^32 / Z/2^64 LU session (open / put / step / fetch / get / close) driven by a host panel
factorisation like the reference's (Binary/Ring2kDeviceLU.h: minimal-valuation pivots on the host,
swaps + TRSM + trailing GEMM on the device).  The rows it leaves span the input's row space, so
their Howell form must equal the oracle's Howell form of the input (unique, ENGINE_API.md Sec. 13.5);
gemm_sub_host is checked against exact wraparound products."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _wrap(x, elem):
    return x & np.uint64((1 << 32) - 1) if elem == 4 else x


def device_echelon(L_, A, k, panel, elem=4, lookahead=True):
    from hacs_amd import _lib
    dt = np.uint32 if elem == 4 else np.uint64
    m, n = A.shape
    h = C.c_void_p()
    _lib.check(L_.hx_ring2k_open(elem, m, n, panel, C.byref(h)))
    try:
        A = np.ascontiguousarray(A.astype(dt))
        _lib.check(L_.hx_ring2k_put(h, A.ctypes.data, n, 0, m, 0, n))
        mask = (1 << k) - 1
        r, c = 0, 0
        pend = None
        while c < n and r < m:
            c1 = min(c + panel, n)
            pw = c1 - c
            Pn = np.zeros((m - r, pw), dt)
            if pend is not None:
                Pn[:] = pend
            else:
                _lib.check(L_.hx_ring2k_get(h, Pn.ctypes.data, pw, r, m - r, c, pw))
            Lh = np.zeros((m - r, panel), dt)
            swaps = []
            npv = 0
            rows = m - r
            for j in range(pw):
                best, bv = None, k
                for i in range(npv, rows):
                    x = int(Pn[i, j]) & mask
                    if x:
                        v = (x & -x).bit_length() - 1
                        if v < bv:
                            best, bv = i, v
                            if v == 0:
                                break
                if best is None:
                    continue
                if best != npv:
                    Pn[[best, npv]] = Pn[[npv, best]]
                    Lh[[best, npv]] = Lh[[npv, best]]
                    swaps += [r + best, r + npv]
                P = Pn[npv].astype(object)
                u = (int(P[j]) >> bv) & ((1 << (8 * elem)) - 1)
                ui = pow(u, -1, 1 << (8 * elem))
                for i in range(npv + 1, rows):
                    x = int(Pn[i, j])
                    if x & mask == 0:
                        Lh[i, npv] = 0
                        continue
                    f = ((x >> bv) * ui) % (1 << (8 * elem))
                    Lh[i, npv] = f
                    Pn[i, j:] = ((Pn[i, j:].astype(object) - f * P[j:]) % (1 << (8 * elem))).astype(dt)
                npv += 1
            la = lookahead and r + npv < m and c1 < n
            la1 = min(c1 + panel, n)
            sw = np.array(swaps, np.uint32) if swaps else np.zeros(2, np.uint32)
            st = _lib.hx_ring2k_step_args(m=m, r=r, np=npv, pc0=c, pc1=c1, panel=Pn.ctypes.data, ldp=pw,
                                          swaps=sw.ctypes.data, nswaps=len(swaps) // 2, L=Lh.ctypes.data, ldl=panel,
                                          u0=c1, la0=c1 if la else 0, la1=la1 if la else 0, w=0)
            _lib.check(L_.hx_ring2k_step(h, C.byref(st)))
            r += npv
            if la:
                pend = np.zeros((m - r, la1 - c1), dt)
                _lib.check(L_.hx_ring2k_fetch(h, pend.ctypes.data, la1 - c1))
            else:
                pend = None
            c = c1
        out = np.zeros((m, n), dt)
        _lib.check(L_.hx_ring2k_get(h, out.ctypes.data, n, 0, m, 0, n))
        return out[:r], r
    finally:
        L_.hx_ring2k_close(h)


@pytest.mark.parametrize("k,elem,panel,shape", [(16, 4, 8, (60, 50)), (32, 4, 16, (90, 70)), (5, 4, 4, (40, 45)),
                                                (30, 8, 8, (50, 40))])
def test_ring2k_session_lu(hx_gpu, O, k, elem, panel, shape):
    rng = np.random.default_rng(k + panel)
    m, n = shape
    A = rng.integers(0, 1 << min(k, 32), (m, n), dtype=np.uint64)
    A = ((A << rng.integers(0, k, (m, 1)).astype(np.uint64)) % np.uint64(1 << k)).astype(np.uint64)
    E, r = device_echelon(hx_gpu, A, k, panel, elem)
    mask = (1 << k) - 1
    Ech = (E.astype(np.uint64) & np.uint64(mask)).astype(np.uint32)
    H1 = O.howell_fast(k, Ech)
    H2 = O.howell_fast(k, A.astype(np.uint32))
    for a, b in zip(H1, H2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("elem", [4, 8])
def test_ring2k_gemm_sub_host(hx_gpu, elem):
    dt = np.uint32 if elem == 4 else np.uint64
    rng = np.random.default_rng(elem)
    m, n, k = 70, 90, 33
    top = 1 << 32 if elem == 4 else 1 << 63
    Cm = rng.integers(0, top, (m, n), dtype=np.uint64).astype(dt)
    Lm = rng.integers(0, top, (m, k), dtype=np.uint64).astype(dt)
    Um = rng.integers(0, top, (k, n), dtype=np.uint64).astype(dt)
    want = ((Cm.astype(object) - Lm.astype(object).dot(Um.astype(object))) % (1 << (8 * elem))).astype(dt)
    got = Cm.copy()
    from hacs_amd import _lib
    _lib.check(hx_gpu.hx_ring2k_gemm_sub_host(elem, got.ctypes.data, n, Lm.ctypes.data, k, Um.ctypes.data, n, m, n, k))
    assert np.array_equal(got, want)
