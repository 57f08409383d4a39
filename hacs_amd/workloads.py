"""Seeded synthetic inputs of the benchmark configurations.

All draws come from SplitMix64 streams seeded like the reference's micro-benchmark
(SeedFor / RandomEntries, PerformanceSmokeValidation.cpp:64-89).  SplitMix64 is
counter-based -- draw i of the stream seeded s is mix(s + (i + 1) * gamma) -- so the
streams are generated vectorised here (numpy) or on the device (hx_random_entries)
and agree with the sequential generator draw for draw.

Configurations (sizes and moduli as benchmarked):
  C1  GEMM 1024^3 mod 65521, A / B = RandomEntries with salts 0x01 / 0x02
  C2  RREF of a 4096 x 4096 rank-3000 matrix mod 2^31-1: RankedMatrix (U salt 0x21,
      V salt 0x22) with 64 rows overwritten by combinations of other rows (salt 0x31)
  C3  GEMM 8192^3 mod 2147483629, salts 0x01 / 0x02
  C4  4096 sparse echelon jobs mod 1000003, rows and columns log-uniform in [16, 512],
      entries in {-2, -1, 1, 2} with density 5 %
  C5  20000 mixed jobs: GEMM mod 2147483629 (n log-uniform in [64, 4096]), echelon as C4
      with n in [32, 1024], Howell over Z/2^16 of dense [0, 2^16) matrices, n in [32, 512]
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M64 = (1 << 64) - 1

P_C1 = 65521
P_C2 = 2147483647
P_C3 = 2147483629
P_C4 = 1000003


def seed_for(r: int, c: int, ell: int, salt: int) -> int:
    """SeedFor (PerformanceSmokeValidation.cpp:74-80)."""
    s = 0xC0FFEE123456789 ^ salt
    s = (s * 0x100000001B3 + r) & M64
    s = (s * 0x100000001B3 + c) & M64
    s = (s * 0x100000001B3 + ell) & M64
    return s


def splitmix_draws(seed: int, start: int, count: int) -> np.ndarray:
    """Draws start .. start+count-1 of the SplitMix64 stream seeded `seed` (uint64)."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def random_entries(count: int, ell: int, seed: int) -> np.ndarray:
    """RandomEntries (PerformanceSmokeValidation.cpp:82-89)."""
    return (splitmix_draws(seed, 0, count) % np.uint64(ell)).astype(np.uint32)


def engine_matrix(rows: int, cols: int, ell: int, salt: int) -> np.ndarray:
    return random_entries(rows * cols, ell, seed_for(rows, cols, ell, salt)).reshape(rows, cols)


def _log_uniform(u53: np.ndarray, lo: int, hi: int) -> np.ndarray:
    """floor(lo * (hi / lo) ** u), u = draw >> 11 scaled to [0, 1), clamped to [lo, hi]."""
    u = (u53 >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    x = np.floor(lo * np.exp2(u * np.log2(hi / lo))).astype(np.int64)
    return np.clip(x, lo, hi)


def overwrite_rows(A: np.ndarray, ell: int, nover: int) -> np.ndarray:
    """The C2 rule (restated by the oracle's hxo_overwrite_rows): per overwritten row, a target row and
    four (source row, coefficient) pairs from the stream SeedFor(rows, cols, ell, 0x31); rows in order."""
    rows, cols = A.shape
    st = seed_for(rows, cols, ell, 0x31)
    d = splitmix_draws(st, 0, nover * 9)
    A = A.copy()
    for t in range(nover):
        i = int(d[9 * t] % np.uint64(rows))
        acc = np.zeros(cols, np.uint64)
        for s in range(4):
            j = int(d[9 * t + 1 + 2 * s] % np.uint64(rows))
            if j == i:
                j = (j + 1) % rows
            c = np.uint64(int(d[9 * t + 2 + 2 * s] % np.uint64(ell)))
            acc = (acc + (c * A[j].astype(np.uint64)) % np.uint64(ell)) % np.uint64(ell)
        A[i] = acc.astype(np.uint32)
    return A


def c4_shapes(njobs: int = 4096, p: int = P_C4, lo: int = 16, hi: int = 512, salt: int = 0x41) -> np.ndarray:
    """(rows, cols) per job, each log-uniform in [lo, hi] (two draws per job of SeedFor(njobs, 0, p, salt))."""
    d = splitmix_draws(seed_for(njobs, 0, p, salt), 0, 2 * njobs)
    return np.stack([_log_uniform(d[0::2], lo, hi), _log_uniform(d[1::2], lo, hi)], axis=1)


def sparse_job(j: int, m: int, n: int, p: int, salt: int = 0x42) -> np.ndarray:
    """Job j's entries: draw e of SeedFor(m, n, p, salt) ^ j decides entry e (row-major): non-zero iff
    (draw >> 32) % 20 == 0 (5 %), value {-2, -1, 1, 2}[draw & 3] reduced mod p."""
    d = splitmix_draws(seed_for(m, n, p, salt) ^ j, 0, m * n)
    vals = np.array([p - 2, p - 1, 1, 2], dtype=np.uint64)[(d & np.uint64(3)).astype(np.int64)] % np.uint64(p)
    nz = ((d >> np.uint64(32)) % np.uint64(20)) == 0
    return np.where(nz, vals, 0).astype(np.uint32).reshape(m, n)


def dense_job(j: int, m: int, n: int, bound: int, salt: int) -> np.ndarray:
    return (splitmix_draws(seed_for(m, n, bound, salt) ^ j, 0, m * n) % np.uint64(bound)).astype(np.uint32).reshape(m, n)


def c5_stream(njobs: int = 20000, salt: int = 0x51):
    """The mixed stream: (kind, n) per job, kinds gemm / echelon / howell in equal proportion (job j
    has kind j mod 3 after a seeded shuffle), n log-uniform in the kind's range."""
    d = splitmix_draws(seed_for(njobs, 3, 0, salt), 0, 2 * njobs)
    kinds = np.arange(njobs) % 3
    perm = np.argsort(d[0::2], kind="stable")
    kinds = kinds[perm]
    out = []
    for j in range(njobs):
        k = int(kinds[j])
        lo, hi = ((64, 4096), (32, 1024), (32, 512))[k]
        n = int(_log_uniform(d[2 * j + 1:2 * j + 2], lo, hi)[0])
        out.append((("gemm", "echelon", "howell")[k], n))
    return out


def c2_device(m: int = 4096, n: int = 4096, rank: int = 3000, p: int = P_C2, nover: int = 64):
    """The C2 matrix on the device: RankedMatrix's factors (salts 0x21 / 0x22) generated in HBM, their
    product by the library's GEMM (a unique residue, so any exact kernel gives these bytes), then the
    nover row overwrites (salt 0x31) applied on the host."""
    import torch
    from . import engine as E
    U = E.random_entries(m, rank, p, seed_for(m, rank, p, 0x21))
    V = E.random_entries(rank, n, p, seed_for(rank, n, p, 0x22))
    A = torch.empty((m, n), dtype=torch.int32, device="cuda")
    E.gemm(E.ModCtx(p), U, V, A, E.GemmOp.Set)
    torch.cuda.synchronize()
    if nover:
        A = E.to_device(overwrite_rows(E.from_device(A), p, nover))
    return A
