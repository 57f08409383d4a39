"""ctypes wrapper over the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  It loads

* ``oracle/libhxoracle.so``  -- the C restatement of the reference's CPU path
  (``oracle/hacs_oracle.c``, each function citing the reference file:line), and
* ``oracle/_ref/libhacsref.so`` -- the reference's own CPU engine compiled from
  its sources (``oracle/ref_driver.cpp``), when it was built.

All matrices are numpy ``uint32`` arrays, row-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "libhxoracle.so")
_REF = os.path.join(HERE, "_ref", "libhacsref.so")

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u32p_or_null = C.c_void_p


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and, when the reference tree is present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src/ModArithmetic")
    if with_ref:
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


def _load(path):
    lib = C.CDLL(path)
    return lib


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"oracle library missing: {_LIB} (run make -C oracle)")
        L = _load(_LIB)
        L.hxo_splitmix64.restype = C.c_uint64
        L.hxo_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.hxo_seed_for.restype = C.c_uint64
        L.hxo_seed_for.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        L.hxo_random_entries.restype = None
        L.hxo_random_entries.argtypes = [u32p, C.c_size_t, C.c_uint32, C.c_uint64]
        L.hxo_barrett_m.restype = C.c_uint64
        L.hxo_barrett_m.argtypes = [C.c_uint32]
        L.hxo_k_max.restype = C.c_uint64
        L.hxo_k_max.argtypes = [C.c_uint32]
        L.hxo_reduce_u64.restype = C.c_uint32
        L.hxo_reduce_u64.argtypes = [C.c_uint32, C.c_uint64]
        L.hxo_inv.restype = C.c_uint32
        L.hxo_inv.argtypes = [C.c_uint32, C.c_uint32]
        L.hxo_fnv1a_bytes.restype = C.c_uint64
        L.hxo_fnv1a_bytes.argtypes = [C.c_char_p, C.c_size_t]
        L.hxo_fnv_matrix.restype = C.c_uint64
        L.hxo_fnv_matrix.argtypes = [u32p, C.c_uint32, C.c_uint32, C.c_size_t]
        L.hxo_elim_work.restype = C.c_double
        L.hxo_elim_work.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.hxo_ranked_matrix.restype = None
        L.hxo_ranked_matrix.argtypes = [u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
        L.hxo_overwrite_rows.restype = None
        L.hxo_overwrite_rows.argtypes = [u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.hxo_gemm_naive.restype = None
        L.hxo_gemm_naive.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64,
                                     u32p, C.c_int64, u32p, C.c_int64, C.c_int]
        L.hxo_gemm_fast.restype = None
        L.hxo_gemm_fast.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                    C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int]
        L.hxo_rref_naive.restype = C.c_uint32
        L.hxo_rref_naive.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p]
        L.hxo_rref_rule.restype = C.c_uint32
        L.hxo_rref_rule.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p, u32p,
                                    _u32p_or_null, C.c_int64, C.c_int]
        u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
        L.hxo_gemm_u64.restype = None
        L.hxo_gemm_u64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, u64p, u64p, u64p, C.c_int, C.c_int]
        for name, E in (("hxo_r2_echelon32", C.c_uint32), ("hxo_r2_echelon64", C.c_uint64)):
            f = getattr(L, name)
            f.restype = C.c_uint32
            f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int64, C.c_uint32, C.c_uint32,
                          u32p, u32p]
        for name in ("hxo_howell_naive", "hxo_howell_fast"):
            f = getattr(L, name)
            f.restype = C.c_uint32
            f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p, u32p, u32p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF)


def ref():
    """The compiled reference engine (oracle/_ref), or None when it was not built."""
    global _ref
    if _ref is None and ref_available():
        R = _load(_REF)
        R.hxr_gemm_u32.restype = C.c_int
        R.hxr_gemm_u32.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p,
                                   C.c_int64, u32p, C.c_int64, C.c_int, C.c_int]
        for name in ("hxr_rref_u32", "hxr_ref_rref_u32"):
            f = getattr(R, name)
            f.restype = C.c_int64
            f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p]
        R.hxr_column_rank_profile_u32.restype = C.c_int64
        R.hxr_column_rank_profile_u32.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p]
        R.hxr_ring2k_howell.restype = C.c_int64
        R.hxr_ring2k_howell.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, C.c_int,
                                        u32p, u32p, u32p]
        R.hxr_gemm_u64.restype = C.c_int
        R.hxr_gemm_u64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int]
        R.hxr_r2_echelon.restype = C.c_int64
        R.hxr_r2_echelon.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int64, C.c_uint32,
                                     C.c_uint32, u32p, u32p]
        R.hxr_fnv_words.restype = C.c_uint64
        R.hxr_fnv_words.argtypes = [C.c_void_p, C.c_size_t]
        R.hxr_fnv_bytes.restype = C.c_uint64
        R.hxr_fnv_bytes.argtypes = [C.c_char_p, C.c_size_t]
        for name in ("hxr_ref_howell", "hxr_elim_howell"):
            f = getattr(R, name)
            f.restype = C.c_int64
            f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p, u32p, u32p]
        _ref = R
    return _ref


# ---------------------------------------------------------------------------
# generators (PerformanceSmokeValidation.cpp:64-89, 144-170)
# ---------------------------------------------------------------------------
def splitmix64_stream(seed: int, count: int) -> list[int]:
    st = C.c_uint64(seed)
    return [lib().hxo_splitmix64(C.byref(st)) for _ in range(count)]


def seed_for(r: int, c: int, ell: int, salt: int) -> int:
    return lib().hxo_seed_for(r, c, ell, salt)


def random_entries(count: int, ell: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    lib().hxo_random_entries(out, count, ell, seed)
    return out


def engine_matrix(rows: int, cols: int, ell: int, salt: int) -> np.ndarray:
    """RandomEntries(rows*cols, ell, SeedFor(rows, cols, ell, salt)) as a rows x cols matrix."""
    return random_entries(rows * cols, ell, seed_for(rows, cols, ell, salt)).reshape(rows, cols)


def ranked_matrix(rows: int, cols: int, rank: int, ell: int, threads: int = 8) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.uint32)
    lib().hxo_ranked_matrix(out, rows, cols, rank, ell, threads)
    return out


def overwrite_rows(A: np.ndarray, ell: int, nover: int) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.uint32).copy()
    lib().hxo_overwrite_rows(A, A.shape[0], A.shape[1], ell, nover)
    return A


def fnv_matrix(M: np.ndarray) -> int:
    M = np.ascontiguousarray(M, dtype=np.uint32)
    r, c = (M.shape + (1,))[:2] if M.ndim == 2 else (M.shape[0], 1)
    return lib().hxo_fnv_matrix(M, r, c, c)


def gemm_u64(p: int, A: np.ndarray, B: np.ndarray, C0: np.ndarray | None = None, op: int = 0, threads: int = 8):
    """C op= A B mod p on 64-bit residues, exact (unsigned __int128)."""
    A = np.ascontiguousarray(A, dtype=np.uint64)
    B = np.ascontiguousarray(B, dtype=np.uint64)
    m, k = A.shape
    n = B.shape[1]
    Cm = np.zeros((m, n), np.uint64) if C0 is None else np.ascontiguousarray(C0, dtype=np.uint64).copy()
    lib().hxo_gemm_u64(p, m, n, k, A, B, Cm, op, threads)
    return Cm


def ref_gemm_u64(p: int, A: np.ndarray, B: np.ndarray, C0: np.ndarray | None = None, op: int = 0):
    """The reference's own kernels::gemm<uint64_t> (oracle/_ref)."""
    A = np.ascontiguousarray(A, dtype=np.uint64)
    B = np.ascontiguousarray(B, dtype=np.uint64)
    m, k = A.shape
    n = B.shape[1]
    Cm = np.zeros((m, n), np.uint64) if C0 is None else np.ascontiguousarray(C0, dtype=np.uint64).copy()
    if ref().hxr_gemm_u64(p, m, n, k, A.ctypes.data_as(C.c_void_p), B.ctypes.data_as(C.c_void_p),
                          Cm.ctypes.data_as(C.c_void_p), op) != 0:
        raise RuntimeError("reference kernels::gemm<uint64_t> failed")
    return Cm


def r2_echelon(k: int, A: np.ndarray, panel: int, climit: int | None = None):
    """ring2k::echelon_blocked (MinValuation) on the native rung of A's dtype (uint32: Z/2^32 words,
    uint64: Z/2^64 words), entries understood mod 2^k: (A after the LU, rank, pivots, pivot_v)."""
    A = np.ascontiguousarray(A).copy()
    m, n = A.shape
    fn = lib().hxo_r2_echelon32 if A.dtype == np.uint32 else lib().hxo_r2_echelon64
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    pv = np.zeros(max(1, min(m, n)), np.uint32)
    r = fn(k, m, n, A.ctypes.data_as(C.c_void_p), n, panel, n if climit is None else climit, piv, pv)
    return A, int(r), piv[:r].copy(), pv[:r].copy()


def ref_r2_echelon(k: int, A: np.ndarray, panel: int, climit: int | None = None):
    """The reference's own echelon_blocked (oracle/_ref) on the same rung: (A, rank, pivots, pivot_v)."""
    A = np.ascontiguousarray(A).copy()
    m, n = A.shape
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    pv = np.zeros(max(1, min(m, n)), np.uint32)
    r = ref().hxr_r2_echelon(A.dtype.itemsize, k, m, n, A.ctypes.data_as(C.c_void_p), n, panel,
                             n if climit is None else climit, piv, pv)
    if r < 0:
        raise RuntimeError("reference echelon_blocked failed")
    return A, int(r), piv[:r].copy(), pv[:r].copy()


def smith(k: int, A: np.ndarray) -> np.ndarray:
    """The Smith valuations of A over Z/2^k (ascending), by the oracle's restatement of smith_naive
    (global minimal-valuation pivoting); entries read mod 2^k, k <= 64."""
    M = np.ascontiguousarray(A, dtype=np.uint64).copy()
    m, n = M.shape
    sig = np.zeros(max(1, min(m, n)), np.uint32)
    L = lib()
    L.hxo_smith_naive.restype = C.c_int64
    L.hxo_smith_naive.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, u32p]
    r = L.hxo_smith_naive(k, m, n, M.ctypes.data_as(C.c_void_p), sig)
    return sig[:r].copy()


def ref_smith(k: int, A: np.ndarray, naive: bool = False):
    """The reference's own ring2k::smith (levels) or smith_naive (oracle/_ref) on A's rung (uint32 / uint64):
    (sigma, level_ranks)."""
    A = np.ascontiguousarray(A)
    m, n = A.shape
    sig = np.zeros(max(1, min(m, n)), np.uint32)
    lev = np.zeros(max(1, k), np.uint32)
    nl = C.c_uint32(0)
    R = ref()
    R.hxr_smith.restype = C.c_int64
    R.hxr_smith.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int64, C.c_int,
                            u32p, u32p, C.POINTER(C.c_uint32)]
    r = R.hxr_smith(A.dtype.itemsize, k, m, n, A.ctypes.data_as(C.c_void_p), n, 1 if naive else 0, sig, lev,
                    C.byref(nl))
    if r < 0:
        raise RuntimeError("reference smith failed")
    return sig[:r].copy(), lev[:nl.value].copy()


def fnv1a_bytes(b: bytes) -> int:
    return int(lib().hxo_fnv1a_bytes(b, len(b)))


def ref_fnv_words(words: np.ndarray) -> int:
    """engine::FnvHasher over 64-bit words (oracle/_ref)."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    R = ref()
    return int(R.hxr_fnv_words(w.ctypes.data_as(C.c_void_p), len(w)))


def ref_fnv_bytes(b: bytes) -> int:
    """engine::FnvBytes from the offset basis (oracle/_ref)."""
    return int(ref().hxr_fnv_bytes(b, len(b)))


def elim_work(rows: int, cols: int, rank: int) -> float:
    return lib().hxo_elim_work(rows, cols, rank)


def barrett_m(p: int) -> int:
    return lib().hxo_barrett_m(p)


def k_max(p: int) -> int:
    return lib().hxo_k_max(p)


def reduce_u64(p: int, x: int) -> int:
    return lib().hxo_reduce_u64(p, x)


def inv(a: int, p: int) -> int:
    return lib().hxo_inv(a, p)


# ---------------------------------------------------------------------------
# GEMM / RREF / Howell
# ---------------------------------------------------------------------------
def _c(M):
    return np.ascontiguousarray(M, dtype=np.uint32)


def gemm_naive(p, A, B, C=None, op=0):
    A, B = _c(A), _c(B)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), np.uint32) if C is None else _c(C).copy()
    lib().hxo_gemm_naive(p, m, n, k, A, k, B, n, C, n, op)
    return C


def gemm_fast(p, A, B, C=None, op=0, threads=8):
    A, B = _c(A), _c(B)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), np.uint32) if C is None else _c(C).copy()
    lib().hxo_gemm_fast(p, m, n, k, A.ctypes.data, k, 1, B.ctypes.data, n, 1, C.ctypes.data, n, 1, op, threads)
    return C


def rref_naive(p, A):
    M = _c(A).copy()
    m, n = M.shape
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    r = lib().hxo_rref_naive(p, m, n, M, n, piv)
    return r, piv[:r].copy(), M


def rref_rule(p, A, want_rref=True, threads=8):
    """(rank, pivots, pivot_rows, R) by the smallest-original-row rule; R is rank x n."""
    A = _c(A)
    m, n = A.shape
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    prow = np.zeros(max(1, min(m, n)), np.uint32)
    R = np.zeros((max(1, min(m, n)), max(n, 1)), np.uint32)
    r = lib().hxo_rref_rule(p, m, n, A, n, piv, prow, R.ctypes.data if want_rref else None, max(n, 1), threads)
    return r, piv[:r].copy(), prow[:r].copy(), (R[:r, :n].copy() if want_rref else None)


def _howell(fn, k, A):
    A = _c(A)
    m, n = A.shape
    cap = max(1, n)
    H = np.zeros((cap, max(n, 1)), np.uint32)
    piv = np.zeros(cap, np.uint32)
    pv = np.zeros(cap, np.uint32)
    h = fn(k, m, n, A, n, H, piv, pv)
    return H[:h, :n].copy(), piv[:h].copy(), pv[:h].copy()


def howell_ell(ell, k, A):
    """The Howell form over Z/ell^k (elim::howell's pass, DenseElim.cpp:1130-1300): (H, pivots, pivot_v)."""
    A = _c(A)
    m, n = A.shape
    cap = max(1, n)
    H = np.zeros((cap, max(n, 1)), np.uint32)
    piv = np.zeros(cap, np.uint32)
    pv = np.zeros(cap, np.uint32)
    f = lib().hxo_howell_ell
    f.restype = C.c_uint32
    f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p, u32p, u32p]
    h = f(ell, k, m, n, A, n, H, piv, pv)
    return H[:h, :n].copy(), piv[:h].copy(), pv[:h].copy()


def ref_howell_ell(ell, k, A):
    """The reference's own elim::howell over ModCtx::prime_power(ell, k) (oracle/_ref), k >= 2."""
    A = _c(A)
    m, n = A.shape
    cap = max(1, n)
    H = np.zeros((cap, max(n, 1)), np.uint32)
    piv = np.zeros(cap, np.uint32)
    pv = np.zeros(cap, np.uint32)
    f = ref().hxr_elim_howell_ell
    f.restype = C.c_int64
    f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_int64, u32p, u32p, u32p]
    h = f(ell, k, m, n, A, n, H, piv, pv)
    if h < 0:
        raise RuntimeError("reference elim::howell raised")
    return H[:h, :n].copy(), piv[:h].copy(), pv[:h].copy()


def howell_naive(k, A):
    return _howell(lib().hxo_howell_naive, k, A)


def howell_fast(k, A):
    return _howell(lib().hxo_howell_fast, k, A)


# ---------------------------------------------------------------------------
# level-1/2 and tape-support kernels (Cuda/cuda_kernels.cu:590-713, 950-1087): exact numpy
# restatements on uint64 (every product of two residues < 2^32 fits)
# ---------------------------------------------------------------------------
def _u64(M):
    return np.asarray(M, dtype=np.uint64)


def axpy(p, alpha, x, y, subtract=False):
    """y +/- alpha x mod p (axpy_kernel, cuda_kernels.cu:609-623: red32(alpha x), then apply_op)."""
    t = (np.uint64(alpha) * _u64(x)) % np.uint64(p)
    y = _u64(y)
    r = (y + np.uint64(p) - t) if subtract else (y + t)
    return (r % np.uint64(p)).astype(np.uint32)


def scal(p, alpha, x):
    """alpha x mod p (scal_kernel, cuda_kernels.cu:624-633)."""
    return ((np.uint64(alpha) * _u64(x)) % np.uint64(p)).astype(np.uint32)


def canonicalise(p, v):
    """v >= p ? v - p : v on a Lazy2p view (canon_kernel, cuda_kernels.cu:634-643)."""
    v = np.asarray(v, dtype=np.uint32)
    return np.where(v >= np.uint32(p), v - np.uint32(p), v).astype(np.uint32)


def canonicalise_acc64(p, acc):
    """acc mod p of 64-bit accumulators (canon64_kernel, cuda_kernels.cu:644-649)."""
    return (_u64(acc) % np.uint64(p)).astype(np.uint32)


def axpy_identity(p, alpha, y, subtract=False):
    """y(i, i) +/- alpha on the diagonal of the window (axpy_identity_kernel, cuda_kernels.cu:666-672)."""
    y = np.array(y, dtype=np.uint32)
    n = min(y.shape)
    d = _u64(y[np.arange(n), np.arange(n)])
    d = (d + np.uint64(p) - np.uint64(alpha)) if subtract else (d + np.uint64(alpha))
    y[np.arange(n), np.arange(n)] = (d % np.uint64(p)).astype(np.uint32)
    return y


def gather(p, alpha, beta, a, row_off, cols, c, accumulate=False, rows=None):
    """c(i, t) [+]= alpha [row_off + r(i) == col(t)] + beta a(r(i), col(t)), r(i) = rows[i] or i,
    col(t) = cols[t] or t; a None or beta == 0: a is not read (gather_kernel, cuda_kernels.cu:673-695).
    row_off + r(i) is a 32-bit sum, as in the kernel."""
    c = np.array(c, dtype=np.uint32)
    m, n = c.shape
    r = np.arange(m, dtype=np.uint64) if rows is None else _u64(rows)
    col = np.arange(n, dtype=np.uint64) if cols is None else _u64(cols)
    v = np.zeros((m, n), np.uint64)
    if a is not None and beta != 0:
        av = _u64(a)[r.astype(np.int64)[:, None], col.astype(np.int64)[None, :]]
        v = (np.uint64(beta) * av) % np.uint64(p)
    if alpha != 0:
        hit = ((r[:, None] + np.uint64(row_off)) & np.uint64(0xFFFFFFFF)) == col[None, :]
        v = np.where(hit, (v + np.uint64(alpha)) % np.uint64(p), v)
    if accumulate:
        v = (_u64(c) + v) % np.uint64(p)
    return v.astype(np.uint32)


# ---------------------------------------------------------------------------
# the compiled reference (None when oracle/_ref was not built)
# ---------------------------------------------------------------------------
def ref_gemm(p, A, B, C=None, op=0, threads=8):
    R = ref()
    A, B = _c(A), _c(B)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), np.uint32) if C is None else _c(C).copy()
    if R.hxr_gemm_u32(p, m, n, k, A, k, B, n, C, n, op, threads) != 0:
        raise RuntimeError("reference gemm raised")
    return C


def ref_rref(p, A, naive=False):
    R = ref()
    M = _c(A).copy()
    m, n = M.shape
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    f = R.hxr_ref_rref_u32 if naive else R.hxr_rref_u32
    r = f(p, m, n, M, n, piv)
    if r < 0:
        raise RuntimeError("reference rref raised")
    return int(r), piv[:r].copy(), M


def ref_column_rank_profile(p, A):
    R = ref()
    A = _c(A)
    m, n = A.shape
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    r = R.hxr_column_rank_profile_u32(p, m, n, A, n, piv)
    if r < 0:
        raise RuntimeError("reference column_rank_profile raised")
    return piv[:r].copy()


def ref_howell(k, A, which="ring2k", rung=32):
    R = ref()
    A = _c(A)
    m, n = A.shape
    cap = max(1, n)
    H = np.zeros((cap, max(n, 1)), np.uint32)
    piv = np.zeros(cap, np.uint32)
    pv = np.zeros(cap, np.uint32)
    if which == "ring2k":
        h = R.hxr_ring2k_howell(k, m, n, A, n, rung, H, piv, pv)
    elif which == "naive":
        h = R.hxr_ref_howell(k, m, n, A, n, H, piv, pv)
    else:
        h = R.hxr_elim_howell(k, m, n, A, n, H, piv, pv)
    if h < 0:
        raise RuntimeError("reference howell raised")
    return H[:h, :n].copy(), piv[:h].copy(), pv[:h].copy()
