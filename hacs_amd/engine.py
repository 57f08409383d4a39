"""Python mirror of the reference's device-kernel surface (engine::cuda, Cuda/CudaKernels.h)
over the C-ABI of libhacs_mi355x.so.

Matrices are 2-D torch tensors on the GPU holding uint32 residues (dtype int32 or
uint32: the bits are what matter), addressed through their strides exactly like the
reference's DeviceView {base, rs, cs, rows, cols} (Cuda/DeviceBuf.h:50-91).  PyTorch
only provides device memory and the stream; every value is computed by the library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import (HX_GEMM_AUTO, HX_GEMM_LIMB, HX_GEMM_PLAIN, HX_OP_ADD, HX_OP_SET, HX_OP_SUB, HxError,
                   check, lib)

__all__ = ["ModCtx", "GemmOp", "gemm", "gemm_nt", "schur_update", "to_device", "from_device", "HxError",
           "HX_GEMM_AUTO", "HX_GEMM_PLAIN", "HX_GEMM_LIMB", "current_stream", "view"]


class GemmOp:
    """engine::kernels::GemmOp (Set / Add / Sub)."""
    Set = HX_OP_SET
    Add = HX_OP_ADD
    Sub = HX_OP_SUB


class ModCtx:
    """The device-relevant part of engine::ModCtx<uint32_t> (ModCtx.h:202-250)."""

    def __init__(self, p: int):
        self.c = _lib.hx_modctx()
        check(lib().hx_modctx_init(p, C.byref(self.c)))

    @property
    def p(self) -> int:
        return self.c.p

    @property
    def limbs(self) -> int:
        return self.c.limbs

    @property
    def k_max(self) -> int:
        return self.c.k_max

    @property
    def barrett_m(self) -> int:
        return self.c.barrett_m


class ModCtx64:
    """The device-relevant part of engine::ModCtx<uint64_t> (ModCtx.h): p < 2^64, Montgomery constants."""

    def __init__(self, p: int):
        self.c = _lib.hx_modctx64()
        check(lib().hx_modctx64_init(p, C.byref(self.c)))

    @property
    def p(self) -> int:
        return self.c.p


def gemm_u64(ctx: ModCtx64, A: torch.Tensor, B: torch.Tensor, C_: torch.Tensor, op: int = HX_OP_SET,
             stream=None) -> None:
    """C op= A * B mod p on 64-bit residues (int64 / uint64 device tensors; the u64 kernel of
    Cuda/cuda_kernels.cu:104-164, CPU path kernels::gemm<uint64_t>)."""
    check(lib().hx_gemm_u64(C.byref(ctx.c), view64(A), view64(B), view64(C_), op,
                            current_stream() if stream is None else stream))


def to_device64(a: np.ndarray, device: str = "cuda") -> torch.Tensor:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64)).to(device)


def from_device64(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().contiguous().numpy().view(np.uint64)


# ---- F_2 on bit-packed rows (binary::BitMatrix, Binary/BitMatrix.h) -------------------------------------
def pack_bits(a: np.ndarray, ldw: int | None = None) -> np.ndarray:
    """A 0/1 matrix (m x n) as bit-packed rows: ldw uint64 words per row (default ceil(n / 64)), bit j
    of the row is bit j % 64 of word j / 64, padding zero (the BitMatrix layout)."""
    a = np.asarray(a)
    m, n = a.shape
    w = (n + 63) // 64 if ldw is None else ldw
    bits = np.zeros((m, w * 64), np.uint8)
    bits[:, :n] = a & 1
    by = np.packbits(bits.reshape(m, w * 8, 8)[:, :, ::-1], axis=2).reshape(m, w * 8)
    return np.ascontiguousarray(by).view(np.uint64).reshape(m, w)


def unpack_bits(words: np.ndarray, n: int) -> np.ndarray:
    """The inverse of pack_bits: an m x n uint32 matrix of 0 / 1."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    m = w.shape[0]
    by = w.view(np.uint8).reshape(m, -1, 1)
    bits = np.unpackbits(by, axis=2)[:, :, ::-1].reshape(m, -1)
    return bits[:, :n].astype(np.uint32)


def f2_rref(A_words: torch.Tensor, n: int, stream=None) -> tuple[int, np.ndarray]:
    """In-place RREF over F_2 of a bit-packed device matrix (int64 tensor m x ldw); returns (rank, pivot
    columns) -- rref(BitMatrix&), Binary/BitMatrix.h:101-105."""
    m, ldw = A_words.shape
    assert A_words.is_contiguous()
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    rank = C.c_uint32(0)
    check(lib().hx_f2_rref(m, n, C.c_void_p(A_words.data_ptr()), ldw, C.byref(rank),
                           piv.ctypes.data_as(_lib._U32P), current_stream() if stream is None else stream))
    return rank.value, piv[:rank.value].copy()


def f2_gemm(A_words: torch.Tensor, B_words: torch.Tensor, C_words: torch.Tensor, k: int, n: int,
            accumulate: bool = False, stream=None) -> None:
    """C = A B (or C += A B) over F_2 on bit-packed device matrices (multiply_m4rm, BitMatrix.h:129)."""
    m = A_words.shape[0]
    check(lib().hx_f2_gemm(m, n, k, C.c_void_p(A_words.data_ptr()), A_words.stride(0),
                           C.c_void_p(B_words.data_ptr()), B_words.stride(0), C.c_void_p(C_words.data_ptr()),
                           C_words.stride(0), 1 if accumulate else 0, current_stream() if stream is None else stream))


class SmithResult:
    """ring2k::SmithResult (Binary/Ring2k.h:408-413): sigma ascending (the non-zero Smith diagonal entries
    2^sigma_i over Z/2^k), level_ranks[s] = #{sigma == s} for each level run."""

    def __init__(self, k: int, sigma: np.ndarray, level_ranks: np.ndarray):
        self.k, self.sigma, self.level_ranks = k, sigma, level_ranks

    def length(self) -> int:
        return int(sum(self.k - int(x) for x in self.sigma))


def ring2k_smith(A: np.ndarray, k: int, panel: int = 0) -> SmithResult:
    """The Smith valuations of the host matrix A (uint32 or uint64 words, entries mod 2^k) on the device --
    ring2k::smith (Binary/Ring2k.h:701-841).  A is not modified."""
    A = np.ascontiguousarray(A)
    if A.dtype not in (np.uint32, np.uint64):
        raise TypeError("ring2k_smith: A must be uint32 or uint64")
    m, n = A.shape
    sigma = np.zeros(max(1, min(m, n)), np.uint32)
    lev = np.zeros(max(1, k), np.uint32)
    ns, nl = C.c_uint32(0), C.c_uint32(0)
    check(lib().hx_ring2k_smith(A.dtype.itemsize, k, m, n, A.ctypes.data_as(C.c_void_p), n, panel,
                                sigma.ctypes.data_as(_lib._U32P), C.byref(ns), lev.ctypes.data_as(_lib._U32P),
                                C.byref(nl)))
    return SmithResult(k, sigma[:ns.value].copy(), lev[:nl.value].copy())


def pack_trits(a: np.ndarray, ldw: int | None = None) -> np.ndarray:
    """A matrix over F_3 (entries 0 / 1 / 2, m x n) as bit-sliced rows: per row plane `one` ([x == 1]) then
    plane `two` ([x == 2]), ldw uint64 words each (default ceil(n / 64)) -- the TritMatrix layout."""
    a = np.asarray(a)
    return np.hstack([pack_bits((a == 1).astype(np.uint8), ldw), pack_bits((a == 2).astype(np.uint8), ldw)])


def unpack_trits(words: np.ndarray, n: int) -> np.ndarray:
    """The inverse of pack_trits: an m x n uint32 matrix of 0 / 1 / 2."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    ldw = w.shape[1] // 2
    return unpack_bits(w[:, :ldw], n) + 2 * unpack_bits(w[:, ldw:], n)


def f3_rref(A_words: torch.Tensor, n: int, stream=None) -> tuple[int, np.ndarray]:
    """In-place RREF over F_3 of a bit-sliced device matrix (int64 tensor m x 2 ldw); returns (rank, pivot
    columns) -- rref(TritMatrix&), Binary/TritMatrix.h:96-98."""
    m, two_ldw = A_words.shape
    assert A_words.is_contiguous() and two_ldw % 2 == 0
    piv = np.zeros(max(1, min(m, n)), np.uint32)
    rank = C.c_uint32(0)
    check(lib().hx_f3_rref(m, n, C.c_void_p(A_words.data_ptr()), two_ldw // 2, C.byref(rank),
                           piv.ctypes.data_as(_lib._U32P), current_stream() if stream is None else stream))
    return rank.value, piv[:rank.value].copy()


def f3_gemm(A_words: torch.Tensor, B_words: torch.Tensor, C_words: torch.Tensor, k: int, n: int,
            accumulate: bool = False, stream=None) -> None:
    """C = A B (or C += A B) over F_3 on bit-sliced device matrices (multiply, TritMatrix.h:117)."""
    m = A_words.shape[0]
    check(lib().hx_f3_gemm(m, n, k, C.c_void_p(A_words.data_ptr()), A_words.stride(0) // 2,
                           C.c_void_p(B_words.data_ptr()), B_words.stride(0) // 2, C.c_void_p(C_words.data_ptr()),
                           C_words.stride(0) // 2, 1 if accumulate else 0, current_stream() if stream is None else stream))


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _u32_tensor(t: torch.Tensor) -> None:
    if t.dtype not in (torch.int32, torch.uint32):
        raise TypeError(f"expected an int32/uint32 tensor of residues, got {t.dtype}")
    if not t.is_cuda:
        raise TypeError("expected a device tensor")


def view(t: torch.Tensor) -> _lib.hx_view:
    """hx_view of a 2-D device tensor (strides in elements, like DeviceView)."""
    _u32_tensor(t)
    if t.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    return _lib.hx_view(C.c_void_p(t.data_ptr()), t.stride(0), t.stride(1), t.shape[0], t.shape[1])


def to_device(a: np.ndarray, device: str = "cuda") -> torch.Tensor:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return torch.from_numpy(a.view(np.int32)).to(device)


def from_device(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().contiguous().numpy().view(np.uint32)


def _opts(variant: int, tile: int = 0, workspace: torch.Tensor | None = None, splits: int = 0,
          split_workspace: torch.Tensor | None = None) -> _lib.hx_gemm_opts:
    o = _lib.hx_gemm_opts()
    o.variant = variant
    o.tile = tile
    if workspace is not None:
        o.workspace = C.c_void_p(workspace.data_ptr())
        o.workspace_bytes = workspace.numel() * workspace.element_size()
    o.splits = splits
    if split_workspace is not None:
        o.split_workspace = C.c_void_p(split_workspace.data_ptr())
        o.split_workspace_bytes = split_workspace.numel() * split_workspace.element_size()
    return o


def gemm(ctx: ModCtx, A: torch.Tensor, B: torch.Tensor, C_: torch.Tensor, op: int = GemmOp.Set, stream=None,
         variant: int = HX_GEMM_AUTO, tile: int = 0, workspace: torch.Tensor | None = None, splits: int = 0,
         split_workspace: torch.Tensor | None = None) -> None:
    """C op= A * B mod p  (engine::cuda::gemm, Cuda/CudaKernels.h:120-123).  splits: split-K of the plain
    kernel (0 = automatic, 1 = never, more = that many K slices; GemmOpts::splits)."""
    o = _opts(variant, tile, workspace, splits, split_workspace)
    check(lib().hx_gemm_u32(C.byref(ctx.c), view(A), view(B), view(C_), op,
                            current_stream() if stream is None else stream, C.byref(o)))


def gemm_nt(ctx: ModCtx, A, B, C_, op: int = GemmOp.Set, stream=None, variant: int = HX_GEMM_AUTO) -> None:
    """C op= A * B^T with B given n x k (engine::cuda::gemm_nt, CudaKernels.h:125-128)."""
    o = _opts(variant)
    check(lib().hx_gemm_nt_u32(C.byref(ctx.c), view(A), view(B), view(C_), op,
                               current_stream() if stream is None else stream, C.byref(o)))


def suggest_splits(m: int, n: int, k: int) -> int:
    """The automatic split-K count of an m x n x k plain product (suggest_splits, CudaKernels.h:79)."""
    return int(lib().hx_gemm_suggest_splits(m, n, k))


def split_workspace_bytes(m: int, n: int, k: int, splits: int) -> int:
    """Workspace bytes of a split-K product (gemm_workspace_bytes, CudaKernels.h:78)."""
    return int(lib().hx_gemm_split_workspace_bytes(m, n, k, splits))


def _s(stream):
    return current_stream() if stream is None else stream


def _ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def view64(t: torch.Tensor) -> _lib.hx_view:
    """hx_view of a 2-D device tensor of 64-bit accumulators (strides in 8-byte elements)."""
    if t.dtype not in (torch.int64, torch.uint64) or not t.is_cuda or t.dim() != 2:
        raise TypeError("expected a 2-D int64/uint64 device tensor")
    return _lib.hx_view(C.c_void_p(t.data_ptr()), t.stride(0), t.stride(1), t.shape[0], t.shape[1])


def gemv(ctx: ModCtx, A, x, y, op: int = GemmOp.Set, stream=None) -> None:
    """y op= A x with x a k x 1 and y an m x 1 view (engine::cuda::gemv, CudaKernels.h:187-191)."""
    check(lib().hx_gemv_u32(C.byref(ctx.c), view(A), view(x), view(y), op, _s(stream)))


def axpy(ctx: ModCtx, alpha: int, x, y, subtract: bool = False, stream=None) -> None:
    """y = y +/- alpha x (engine::cuda::axpy, CudaKernels.h:193-196)."""
    check(lib().hx_axpy_u32(C.byref(ctx.c), alpha, view(x), view(y), 1 if subtract else 0, _s(stream)))


def scal(ctx: ModCtx, alpha: int, x, stream=None) -> None:
    """x = alpha x (engine::cuda::scal, CudaKernels.h:198-200)."""
    check(lib().hx_scal_u32(C.byref(ctx.c), alpha, view(x), _s(stream)))


def copy(src, dst, stream=None) -> None:
    """dst = src, any strides (engine::cuda::copy, CudaKernels.h:202-204)."""
    check(lib().hx_copy_u32(view(src), view(dst), _s(stream)))


def transpose(src, dst, stream=None) -> None:
    """dst(i, j) = src(j, i) (engine::cuda::transpose, CudaKernels.h:206-208)."""
    check(lib().hx_transpose_u32(view(src), view(dst), _s(stream)))


def fill(x, value: int, stream=None) -> None:
    """x(i, j) = value (engine::cuda::fill, CudaKernels.h:211-213)."""
    check(lib().hx_fill_u32(view(x), value, _s(stream)))


def axpy_identity(ctx: ModCtx, alpha: int, y, subtract: bool = False, stream=None) -> None:
    """y(i, i) = y(i, i) +/- alpha (engine::cuda::axpy_identity, CudaKernels.h:215-219)."""
    check(lib().hx_axpy_identity_u32(C.byref(ctx.c), alpha, view(y), 1 if subtract else 0, _s(stream)))


def gather(ctx: ModCtx, alpha: int, beta: int, a, row_off: int, cols, c, accumulate: bool = False,
           rows=None, stream=None) -> None:
    """c(i, t) [+]= alpha [row_off + r(i) == col(t)] + beta a(r(i), col(t)) (engine::cuda::gather,
    CudaKernels.h:221-230).  a None = the identity source; cols / rows device index tensors or None."""
    av = _lib.hx_view(None, 0, 0, 0, 0) if a is None else view(a)
    check(lib().hx_gather_u32(C.byref(ctx.c), alpha, beta, av, row_off, _ptr(cols), view(c),
                              1 if accumulate else 0, _ptr(rows), _s(stream)))


def canonicalise(ctx: ModCtx, v, stream=None) -> None:
    """Lazy2p -> canonical in place (engine::cuda::canonicalise, CudaKernels.h:232-234)."""
    check(lib().hx_canonicalise_u32(C.byref(ctx.c), view(v), _s(stream)))


def canonicalise_acc64(ctx: ModCtx, acc, dst, stream=None) -> None:
    """dst = acc mod p for 64-bit accumulators (engine::cuda::canonicalise, CudaKernels.h:236-238)."""
    check(lib().hx_canonicalise_acc64_u32(C.byref(ctx.c), view64(acc), view(dst), _s(stream)))


def launch_stats() -> dict:
    """The library's diagnostic counters (launch_stats, CudaKernels.h:299-306; TransferStats,
    CudaCtx.h:94-101)."""
    st = _lib.hx_stats()
    check(lib().hx_stats_get(C.byref(st)))
    return {name: int(getattr(st, name)) for name, _ in st._fields_}


def reset_launch_stats() -> None:
    check(lib().hx_stats_reset())


def arena_stats() -> dict:
    """The device arena's counters (ArenaStats, CudaCtx.h:107-117)."""
    st = _lib.hx_arena_stats()
    check(lib().hx_arena_stats_get(C.byref(st)))
    return {name: int(getattr(st, name)) for name, _ in st._fields_}


def arena_trim() -> int:
    """release_cached (CudaCtx.h:157-168): cached blocks back to the driver; returns the bytes freed."""
    freed = C.c_size_t(0)
    check(lib().hx_arena_trim(C.byref(freed)))
    return int(freed.value)


def limb_k_chunk(ctx: ModCtx, m: int, n: int, k: int, workspace_bytes: int) -> int:
    """The K chunk a limb workspace of this size runs (imma_k_chunk, CudaKernels.h:100-110); 0 = none."""
    return int(lib().hx_gemm_limb_k_chunk(C.byref(ctx.c), m, n, k, workspace_bytes))


def schur_update(ctx: ModCtx, L, U, C_, stream=None, variant: int = HX_GEMM_AUTO) -> None:
    """C -= L * U (engine::cuda::schur_update, CudaKernels.h:130-133)."""
    gemm(ctx, L, U, C_, GemmOp.Sub, stream, variant)


def gemm_workspace_bytes(ctx: ModCtx, m: int, n: int, k: int, variant: int = HX_GEMM_LIMB, tile: int = 0) -> int:
    return lib().hx_gemm_workspace_bytes_tile(C.byref(ctx.c), m, n, k, variant, tile)


def residue_moduli(p: int, k: int) -> int:
    """Moduli count of the residue route (gemm tile 6) at modulus p and depth k; 0 where it does not apply."""
    return int(lib().hx_gemm_residue_moduli(p, k))


def gemm_route(ctx: ModCtx, m: int, n: int, k: int, variant: int = HX_GEMM_LIMB, tile: int = 0) -> int:
    """The kernel gemm runs for these sizes (0 plain, 1..5 limb kernel, 6 residue route), workspace permitting."""
    return int(lib().hx_gemm_route(C.byref(ctx.c), m, n, k, variant, tile))


def mfma_rate(blocks: int = 1024, iters: int = 20000, random_operands: bool = True, stream=None) -> float:
    """Sustained TOP/s of the limb kernel's int8 MFMA instruction issued from registers only (calibration:
    the ceiling the limb GEMM can reach with random operands, and the rate the published peak assumes)."""
    out = C.c_double(0.0)
    check(lib().hx_mfma_rate(blocks, iters, 1 if random_operands else 0,
                             current_stream() if stream is None else stream, C.byref(out)))
    return out.value


def random_entries(rows: int, cols: int, ell: int, seed: int, stream=None) -> torch.Tensor:
    """RandomEntries(rows * cols, ell, seed) generated on the device (PerformanceSmokeValidation.cpp:82-89)."""
    t = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
    check(lib().hx_random_entries(C.c_void_p(t.data_ptr()), rows, cols, cols, ell, seed,
                                  current_stream() if stream is None else stream))
    return t


def random_sparse(rows: int, cols: int, p: int, seed: int, stream=None) -> torch.Tensor:
    """The 5 %-density {-2, -1, 1, 2} entries (the coboundary-like jobs) generated on the device."""
    t = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
    check(lib().hx_random_sparse(C.c_void_p(t.data_ptr()), rows, cols, cols, p, seed,
                                 current_stream() if stream is None else stream))
    return t


@dataclass
class DeviceInfo:
    name: str
    arch: str
    compute_units: int
    clock_khz: int
    total_mem: int


def device_info() -> DeviceInfo:
    d = _lib.hx_device_info()
    check(lib().hx_device_info_get(C.byref(d)))
    return DeviceInfo(d.name.decode(), d.arch.decode(), d.compute_units, d.clock_khz, d.total_mem)


# ---------------------------------------------------------------------------
# elimination
# ---------------------------------------------------------------------------
@dataclass
class RrefResult:
    """rank, pivot columns (ascending) and pivot_rows (ImmaRrefResult, Cuda/CudaKernels.h:269-273)."""
    rank: int
    pivots: np.ndarray
    pivot_rows: np.ndarray


def rref(ctx: ModCtx, A: torch.Tensor, R: torch.Tensor | None = None, stream=None) -> RrefResult:
    """Device RREF of A (m x n, not modified); R (>= m rows) receives RREF(A), zero rows last.

    The device replacement of elim::rref / imma_rref (DenseElim.h:149-151, CudaKernels.h:276-288)."""
    _u32_tensor(A)
    m, n = A.shape
    if A.stride(1) != 1:
        raise ValueError("A must be row-contiguous")
    mn = max(1, min(m, n))
    piv = np.zeros(mn, np.uint32)
    prow = np.zeros(mn, np.uint32)
    rank = C.c_uint32(0)
    rp, ldr = (None, 0)
    if R is not None:
        _u32_tensor(R)
        if R.stride(1) != 1 or R.shape[0] < m or R.shape[1] < n:
            raise ValueError("R must be row-contiguous with at least m rows and n columns")
        rp, ldr = C.c_void_p(R.data_ptr()), R.stride(0)
    check(lib().hx_rref_u32(C.byref(ctx.c), m, n, C.c_void_p(A.data_ptr()), A.stride(0), rp, ldr, C.byref(rank),
                            piv.ctypes.data_as(_lib._U32P), prow.ctypes.data_as(_lib._U32P),
                            current_stream() if stream is None else stream))
    r = rank.value
    return RrefResult(r, piv[:r].copy(), prow[:r].copy())


def rref_workspace_bytes(m: int, n: int) -> int:
    return int(lib().hx_rref_workspace_bytes(m, n))


def rref_device(ctx: ModCtx, A: torch.Tensor, R: torch.Tensor | None, d_out: torch.Tensor, workspace: torch.Tensor,
                stream=None) -> None:
    """The elimination entirely on the device (hx_rref_u32_device): nothing allocated, nothing read back,
    so it can be captured in a graph.  d_out (int32, >= 1 + 2 min(m, n) entries) receives rank, then the
    pivots, then the pivot rows; workspace: >= rref_workspace_bytes(m, n) bytes."""
    _u32_tensor(A)
    m, n = A.shape
    mn = max(1, min(m, n))
    if d_out.numel() < 1 + 2 * mn:
        raise ValueError("d_out too small")
    base = d_out.data_ptr()
    rp, ldr = (None, 0) if R is None else (C.c_void_p(R.data_ptr()), R.stride(0))
    check(lib().hx_rref_u32_device(C.byref(ctx.c), m, n, C.c_void_p(A.data_ptr()), A.stride(0), rp, ldr,
                                   C.c_void_p(base), C.c_void_p(base + 4), C.c_void_p(base + 4 + 4 * mn),
                                   C.c_void_p(workspace.data_ptr()), workspace.numel() * workspace.element_size(),
                                   current_stream() if stream is None else stream))


def rref_device_result(d_out: torch.Tensor, m: int, n: int) -> RrefResult:
    """rank, pivots and pivot rows from rref_device's output array (after a synchronisation)."""
    mn = max(1, min(m, n))
    h = from_device(d_out)
    r = int(h[0])
    if r == 0xFFFFFFFF:
        raise HxError(_lib.HX_E_NOT_INVERTIBLE, "a pivot had no inverse mod p")
    return RrefResult(r, h[1:1 + r].copy(), h[1 + mn:1 + mn + r].copy())


class RrefBatch:
    """Many uneven RREF jobs in one persistent launch; each job's matrix is overwritten by its RREF.

    The job table is uploaded once (not capturable); run() is capturable."""

    def __init__(self, ctx: ModCtx, mats: list[torch.Tensor], panels: list[int] | None = None,
                 mode: int = _lib.HX_RREF_BATCH_AUTO):
        self.ctx = ctx
        self.mats = mats
        n = len(mats)
        self.meta = torch.zeros(max(1, sum(2 * max(1, min(M.shape)) + 1 for M in mats)), dtype=torch.int32,
                                device="cuda")
        jobs = (_lib.hx_rref_job * max(1, n))()
        off = 0
        base = self.meta.data_ptr()
        self.offsets = []
        for i, M in enumerate(mats):
            _u32_tensor(M)
            m_, n_ = M.shape
            mn = max(1, min(m_, n_))
            jobs[i].m, jobs[i].n = m_, n_
            jobs[i].A = C.c_void_p(M.data_ptr())
            jobs[i].lda = M.stride(0)
            jobs[i].pivots = C.c_void_p(base + 4 * off)
            jobs[i].pivot_rows = C.c_void_p(base + 4 * (off + mn))
            jobs[i].rank = C.c_void_p(base + 4 * (off + 2 * mn))
            jobs[i].panel = 0 if panels is None else panels[i]
            self.offsets.append((off, mn))
            off += 2 * mn + 1
        self.h = C.c_void_p()
        check(lib().hx_rref_batch_create(C.byref(ctx.c), jobs, n, mode, C.byref(self.h)))

    def info(self) -> dict:
        """The launch shape create chose (workgroups per CU, grid, LDS, jobs reading X's planes from HBM)."""
        inf = _lib.hx_rref_batch_info()
        check(lib().hx_rref_batch_info_get(self.h, C.byref(inf)))
        return {name: int(getattr(inf, name)) for name, _ in inf._fields_}

    def run(self, stream=None) -> None:
        check(lib().hx_rref_batch_run(self.h, current_stream() if stream is None else stream))

    def status(self, stream=None) -> None:
        check(lib().hx_rref_batch_status(self.h, current_stream() if stream is None else stream))

    def results(self) -> list[RrefResult]:
        meta = from_device(self.meta)
        out = []
        for off, mn in self.offsets:
            r = int(meta[off + 2 * mn])
            out.append(RrefResult(r, meta[off:off + r].copy(), meta[off + mn:off + mn + r].copy()))
        return out

    def __del__(self):
        try:
            if self.h:
                lib().hx_rref_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def trsm_unit(ctx: ModCtx, T: torch.Tensor, B: torch.Tensor, upper: bool, stream=None) -> None:
    """B <- T^{-1} B, T unit triangular (engine::cuda::trsm_unit, Cuda/CudaKernels.h:240-247)."""
    check(lib().hx_trsm_unit_u32(C.byref(ctx.c), view(T), view(B), 1 if upper else 0,
                                 current_stream() if stream is None else stream))


# ---------------------------------------------------------------------------
# Howell form over Z/2^k
# ---------------------------------------------------------------------------
@dataclass
class HowellResult:
    """H (h x n, masked to k bits), strictly increasing pivots, pivot_v (HowellForm, Ring2k.h:415-422)."""
    H: np.ndarray
    pivots: np.ndarray
    pivot_v: np.ndarray


class HowellBatch:
    """Howell forms over Z/2^k (ring2k::howell), or over Z/ell^k for a prime ell (elim::howell with a
    prime-power context), of many uneven matrices in one persistent launch."""

    def __init__(self, k: int, mats: list[torch.Tensor], ell: int = 2):
        self.k, self.ell = k, ell
        self.mats = mats
        n_jobs = len(mats)
        jobs = (_lib.hx_howell_job * max(1, n_jobs))()
        self.outs = []
        for i, M in enumerate(mats):
            _u32_tensor(M)
            m_, n_ = M.shape
            H = torch.zeros((max(1, n_), max(1, n_)), dtype=torch.int32, device="cuda")
            meta = torch.zeros(2 * max(1, n_) + 1, dtype=torch.int32, device="cuda")
            work = torch.empty(max(1, lib().hx_howell_work_words(m_, n_)), dtype=torch.int32, device="cuda")
            jobs[i].m, jobs[i].n = m_, n_
            jobs[i].A = C.c_void_p(M.data_ptr())
            jobs[i].lda = M.stride(0) if m_ else n_
            jobs[i].H = C.c_void_p(H.data_ptr())
            jobs[i].pivots = C.c_void_p(meta.data_ptr())
            jobs[i].pivot_v = C.c_void_p(meta.data_ptr() + 4 * max(1, n_))
            jobs[i].h = C.c_void_p(meta.data_ptr() + 8 * max(1, n_))
            jobs[i].work = C.c_void_p(work.data_ptr())
            self.outs.append((H, meta, work, n_))
        self.h = C.c_void_p()
        if ell == 2:
            check(lib().hx_howell_batch_create(k, jobs, n_jobs, C.byref(self.h)))
        else:
            check(lib().hx_howell_batch_create_ell(ell, k, jobs, n_jobs, C.byref(self.h)))

    def run(self, stream=None) -> None:
        check(lib().hx_howell_batch_run(self.h, current_stream() if stream is None else stream))

    def results(self, jobs: list[int] | None = None) -> list[HowellResult]:
        """The forms of all jobs, or of the listed job indices."""
        res = []
        for i in (range(len(self.outs)) if jobs is None else jobs):
            H, meta, _, n_ = self.outs[i]
            mt = from_device(meta)
            nn = max(1, n_)
            h = int(mt[2 * nn])
            res.append(HowellResult(from_device(H)[:h, :n_].copy(), mt[:h].copy(), mt[nn:nn + h].copy()))
        return res

    def __del__(self):
        try:
            if self.h:
                lib().hx_howell_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


class GemmBatch:
    """Uneven GEMM jobs in one launch per kernel variant (generalises GemmBatch<Elt>,
    Cuda/CudaKernels.h:135-185, which needs one shared shape).  jobs: (A, B, C, op) row-major
    device tensors, or (A, B, C, op, variant), or (A, B, C, op, variant, limb tile 1..4).  Created once (job table and limb workspace
    uploaded: not capturable); run() is capturable."""

    def __init__(self, ctx: ModCtx, jobs, variant: int = HX_GEMM_AUTO):
        self.ctx = ctx
        self.jobs = jobs
        arr = (_lib.hx_gemm_job * max(1, len(jobs)))()
        for i, jb in enumerate(jobs):
            A, B, C_, op = jb[:4]
            for t in (A, B, C_):
                _u32_tensor(t)
                if t.numel() and t.shape[1] > 1 and t.stride(1) != 1:
                    raise ValueError("batched GEMM operands must be row-contiguous")
            arr[i].m, arr[i].k = A.shape
            arr[i].n = B.shape[1]
            arr[i].op = op
            arr[i].variant = jb[4] if len(jb) > 4 else HX_GEMM_AUTO
            arr[i].tile = jb[5] if len(jb) > 5 else 0
            arr[i].A, arr[i].B, arr[i].C = C.c_void_p(A.data_ptr()), C.c_void_p(B.data_ptr()), C.c_void_p(C_.data_ptr())
            ld = [t.stride(0) if t.numel() and t.shape[0] > 1 else t.shape[1] for t in (A, B, C_)]
            arr[i].lda, arr[i].ldb, arr[i].ldc = ld
        self.h = C.c_void_p()
        check(lib().hx_gemm_batch_create(C.byref(ctx.c), arr, len(jobs), variant, C.byref(self.h)))

    def run(self, stream=None) -> None:
        check(lib().hx_gemm_batch_run(self.h, current_stream() if stream is None else stream))

    def __del__(self):
        try:
            if self.h:
                lib().hx_gemm_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass
