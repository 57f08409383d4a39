"""Seeded input kinds for the elimination tests, after the reference's device elimination gate
(Cuda/CudaImmaElimValidation.cpp:80-129): Random, Planted rank, Sparse, Extreme, Zero, Identity,
DupRows, Boundaries (zero columns at panel boundaries, duplicated columns, a zero row), plus the
sparse {-2,-1,1,2} 5%-density kind of the batched-echelon workload."""
import numpy as np

KINDS = ["random", "planted", "sparse", "extreme", "zero", "identity", "duprows", "boundaries", "coboundary"]


def make(rng: np.random.Generator, m: int, n: int, p: int, kind: str, rank: int = 0) -> np.ndarray:
    M = np.zeros((m, n), np.uint64)
    if kind in ("random", "boundaries", "duprows"):
        M = rng.integers(0, p, (m, n), dtype=np.uint64)
    elif kind == "planted":
        r = max(1, min(rank or max(1, min(m, n) // 2), m, n))
        L = rng.integers(0, p, (m, r), dtype=np.uint64).astype(np.uint32)
        R = rng.integers(0, p, (r, n), dtype=np.uint64).astype(np.uint32)
        # the exact product mod p by the oracle's GEMM (itself pinned against exact integers and the
        # reference's kernels::gemm); numpy object arithmetic would take hours at the large shapes
        from oracle import oracle as O
        M = O.gemm_fast(p, L, R, None, 0, 16).astype(np.uint64)
    elif kind == "sparse":
        z = rng.integers(0, 1000, (m, n))
        vals = rng.integers(0, p, (m, n), dtype=np.uint64)
        pick = rng.integers(0, 3, (m, n))
        M = np.where(z < 30, np.where(pick == 0, p - 1, np.where(pick == 1, 1, vals)), 0).astype(np.uint64)
    elif kind == "extreme":
        M = np.where(rng.integers(0, 2, (m, n)) == 1, p // 2, p - p // 2).astype(np.uint64)
    elif kind == "identity":
        for i in range(min(m, n)):
            M[i, i] = 1
    elif kind == "coboundary":
        z = rng.integers(0, 20, (m, n))
        v = np.array([p - 2, p - 1, 1, 2], dtype=np.uint64)[rng.integers(0, 4, (m, n))] % p
        M = np.where(z == 0, v, 0).astype(np.uint64)
    if kind == "boundaries":
        for c in (0, 15, 16, 31, 32, 63, 64, 255, 256, 257, 511, 512):
            if c < n:
                M[:, c] = 0
        for _ in range(max(2, n // 40)):
            c, d = sorted(rng.integers(0, n, 2))
            if c != d:
                M[:, d] = M[:, c]
        if m > 3:
            M[m // 2, :] = 0
    if kind == "duprows":
        for _ in range(max(2, m // 3)):
            a, b = rng.integers(0, m, 2)
            if a != b:
                s = int(rng.integers(0, p))
                M[b, :] = (M[a, :].astype(object) * s % p).astype(np.uint64)
    return (M % p).astype(np.uint32)
