"""Howell form over Z/2^k on the GPU against the CPU oracle (and its special cases): H, pivots and
pivot_v bit-exact.  Inputs follow the reference's ladder validation (Ring2kLadderValidation.cpp):
random and planted (valuation-structured) matrices over k in {1..32}, plus the known answer
howell([[2, 1]]) over Z/4 = [[2, 1], [0, 2]] (ENGINE_API.md:1939-1942)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _planted(rng, m, n, k):
    """Rows scaled by powers of two and mixed: many non-unit pivots and closure rows."""
    A = rng.integers(0, 1 << min(k, 32), (m, n), dtype=np.uint64)
    shifts = rng.integers(0, k, (m, 1))
    A = (A << shifts.astype(np.uint64)) & np.uint64((1 << k) - 1 if k < 64 else ~0)
    return A.astype(np.uint32)


@pytest.mark.parametrize("k", [1, 2, 3, 8, 16, 31, 32])
def test_howell_batch(hx_gpu, O, k):
    from hacs_amd import engine as E
    rng = np.random.default_rng(100 + k)
    cases = [np.array([[2, 1]], np.uint32) if k == 2 else np.zeros((1, 1), np.uint32)]
    for m, n in [(1, 1), (5, 9), (9, 5), (17, 17), (40, 33), (33, 70), (64, 64), (100, 130), (200, 150)]:
        top = (1 << k) if k < 32 else (1 << 32)
        cases.append(rng.integers(0, top, (m, n), dtype=np.uint64).astype(np.uint32))
        cases.append(_planted(rng, m, n, k))
    cases.append(np.zeros((6, 4), np.uint32))
    mats = [E.to_device(A) for A in cases]
    b = E.HowellBatch(k, mats)
    b.run()
    torch.cuda.synchronize()
    for A, res in zip(cases, b.results()):
        H, piv, pv = O.howell_fast(k, A)
        assert np.array_equal(res.pivots, piv), f"k={k} {A.shape}: pivots"
        assert np.array_equal(res.pivot_v, pv), f"k={k} {A.shape}: pivot_v"
        assert np.array_equal(res.H, H), f"k={k} {A.shape}: H"
    if k == 2:
        r0 = b.results()[0]
        assert r0.H.tolist() == [[2, 1], [0, 2]] and r0.pivots.tolist() == [0, 1] and r0.pivot_v.tolist() == [1, 1]


def test_howell_k1_is_rref(hx_gpu, O):
    """k = 1: the Howell form over F_2 is the RREF without zero rows (ENGINE_API.md:1988-1995)."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(4)
    A = rng.integers(0, 2, (70, 90), dtype=np.uint64).astype(np.uint32)
    b = E.HowellBatch(1, [E.to_device(A)])
    b.run()
    torch.cuda.synchronize()
    res = b.results()[0]
    r, piv, R = O.rref_naive(2, A)
    assert np.array_equal(res.H, R[:r]) and np.array_equal(res.pivots, piv) and not res.pivot_v.any()


@pytest.mark.parametrize("k", [5, 16, 32])
def test_howell_row_classes_and_panel_edges(hx_gpu, O, k):
    """Every job class of the batch (m + n <= 256, <= 512, <= 1024 on the panel-blocked kernel,
    larger on the column-at-a-time kernel) in ONE batch; widths around the 16-column panel
    (3, 15, 16, 17, 33), zero columns (panel steps without a pivot), zero rows, sparse and
    planted inputs."""
    from hacs_amd import engine as E
    rng = np.random.default_rng(700 + k)
    top = (1 << k) if k < 32 else (1 << 32)
    cases = []
    for m, n in [(3, 16), (16, 3), (20, 17), (31, 32), (15, 15), (40, 33), (300, 200), (600, 400),
                 (700, 400), (40, 600)]:
        cases.append(rng.integers(0, top, (m, n), dtype=np.uint64).astype(np.uint32))
        cases.append(_planted(rng, m, n, k))
    S = rng.integers(0, top, (90, 70), dtype=np.uint64).astype(np.uint32)
    S[rng.random((90, 70)) > 0.05] = 0
    S[:, 10:30] = 0                      # whole panels without a pivot
    S[40:50, :] = 0
    cases.append(S)
    mats = [E.to_device(A) for A in cases]
    b = E.HowellBatch(k, mats)
    for rep in range(2):                 # a second run over the same inputs gives the same forms
        b.run()
        torch.cuda.synchronize()
        for A, res in zip(cases, b.results()):
            H, piv, pv = O.howell_fast(k, A)
            assert np.array_equal(res.pivots, piv), f"rep {rep} k={k} {A.shape}: pivots"
            assert np.array_equal(res.pivot_v, pv), f"rep {rep} k={k} {A.shape}: pivot_v"
            assert np.array_equal(res.H, H), f"rep {rep} k={k} {A.shape}: H"


def test_howell_batch_graph_replay(hx_gpu, O):
    """A batch spanning every job class (the large classes run on the caller's stream, the small
    ones on the batch's second stream, forked and joined by events) captured in one hipGraph on a
    side stream and replayed three times; every replay equals the oracle."""
    import ctypes as C
    from hacs_amd import engine as E
    from hacs_amd import _lib
    k = 16
    rng = np.random.default_rng(4242)
    cases = [rng.integers(0, 1 << k, (m, n), dtype=np.uint64).astype(np.uint32)
             for m, n in [(10, 12), (100, 90), (200, 250), (400, 500), (700, 400), (30, 31)]]
    want = [O.howell_fast(k, A) for A in cases]
    mats = [E.to_device(A) for A in cases]
    b = E.HowellBatch(k, mats)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        g = C.c_void_p()
        _lib.check(_lib.lib().hx_graph_begin(h))
        b.run(stream=h)
        _lib.check(_lib.lib().hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            _lib.check(_lib.lib().hx_graph_launch(g, h))
            s.synchronize()
            for A, (H, piv, pv), res in zip(cases, want, b.results()):
                assert np.array_equal(res.pivots, piv) and np.array_equal(res.pivot_v, pv), (rep, A.shape)
                assert np.array_equal(res.H, H), (rep, A.shape)
        _lib.check(_lib.lib().hx_graph_destroy(g))
