"""Device RREF over F_p against the CPU oracle: rank, pivot columns, pivot_rows (the smallest-
original-row rule) and the RREF rows, all bit-exact (Cuda/CudaImmaElimValidation.cpp arm (A)).
The pivot_rows set is also checked against the row rank profile computed independently from the
transpose (:137-143), so a wrong tie rule cannot pass by agreeing with a wrong oracle."""
import numpy as np
import pytest
import torch

from tests.matrix_kinds import KINDS, make

pytestmark = pytest.mark.gpu

PRIMES = [2, 3, 5, 13, 127, 65521, 1000003, 2147483629, 2147483647]


def _check(O, p, A, res, R, tag):
    rank, piv, prow, want = O.rref_rule(p, A)
    assert res.rank == rank, f"{tag}: rank {res.rank} vs {rank}"
    assert np.array_equal(res.pivots, piv), f"{tag}: pivots"
    assert np.array_equal(res.pivot_rows, prow), f"{tag}: pivot_rows"
    assert np.array_equal(R[:rank], want), f"{tag}: RREF rows"
    assert not R[rank:].any(), f"{tag}: zero rows"
    if rank:
        rrp = O.rref_naive(p, A.T.copy())[1]
        assert np.array_equal(np.sort(res.pivot_rows), rrp), f"{tag}: pivot_rows set != row rank profile"


@pytest.mark.parametrize("p", PRIMES)
def test_rref_batch_kinds(hx_gpu, O, p):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(p)
    shapes = [(1, 1), (1, 7), (7, 1), (16, 16), (33, 17), (17, 33), (64, 64), (65, 97), (130, 70), (100, 300),
              (257, 129)]
    cases = []
    for i, kind in enumerate(KINDS):
        for (m, n) in shapes[i % 3::3] + [shapes[(i * 5) % len(shapes)]]:
            cases.append((kind, make(rng, m, n, p, kind, rank=max(1, min(m, n) // 3))))
    mats = [E.to_device(A) for _, A in cases]
    b = E.RrefBatch(ctx, mats)
    b.run()
    b.status()
    torch.cuda.synchronize()
    for (kind, A), M, res in zip(cases, mats, b.results()):
        _check(O, p, A, res, E.from_device(M), f"p={p} {kind} {A.shape}")


@pytest.mark.parametrize("p", [3, 65521, 2147483647])
@pytest.mark.parametrize("shape", [(200, 200), (513, 300), (300, 700), (1000, 1000)])
def test_rref_single(hx_gpu, O, p, shape):
    from hacs_amd import engine as E
    ctx = E.ModCtx(p)
    m, n = shape
    rng = np.random.default_rng(m * n + p)
    A = make(rng, m, n, p, "planted", rank=min(m, n) * 3 // 4)
    dA = E.to_device(A)
    R = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    res = E.rref(ctx, dA, R)
    torch.cuda.synchronize()
    assert np.array_equal(E.from_device(dA), A), "A must not be modified"
    _check(O, p, A, res, E.from_device(R), f"p={p} {shape}")


def test_rref_rank_only_and_not_invertible(hx_gpu, O):
    from hacs_amd import engine as E
    p = 1000003
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(5)
    A = make(rng, 90, 60, p, "planted", rank=40)
    res = E.rref(ctx, E.to_device(A))
    rank, piv, prow, _ = O.rref_rule(p, A, want_rref=False)
    assert res.rank == rank and np.array_equal(res.pivots, piv) and np.array_equal(res.pivot_rows, prow)
    bad = E.ModCtx(15)   # not a field: 3 has no inverse mod 15
    with pytest.raises(E.HxError):
        E.rref(bad, E.to_device(np.array([[3, 1], [1, 1]], np.uint32)))


def test_rref_batch_graph_replay(hx_gpu, O):
    """The persistent batch captured in a hipGraph and replayed: every replay equals the eager result."""
    from hacs_amd import engine as E
    from hacs_amd import _lib
    import ctypes as C
    p = 1000003
    ctx = E.ModCtx(p)
    rng = np.random.default_rng(11)
    inputs = [make(rng, int(m), int(n), p, "coboundary") for m, n in rng.integers(16, 160, (40, 2))]
    src = [E.to_device(A) for A in inputs]
    work = [s.clone() for s in src]
    b = E.RrefBatch(ctx, work)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        h = s.cuda_stream
        g = C.c_void_p()
        _lib.check(_lib.lib().hx_graph_begin(h))
        b.run(stream=h)
        _lib.check(_lib.lib().hx_graph_end(h, C.byref(g)))
        for rep in range(3):
            for w_, s_ in zip(work, src):
                w_.copy_(s_)
            _lib.check(_lib.lib().hx_graph_launch(g, h))
            s.synchronize()
            for A, M, res in zip(inputs, work, b.results()):
                _check(O, p, A, res, E.from_device(M), f"replay {rep} {A.shape}")
        _lib.check(_lib.lib().hx_graph_destroy(g))
