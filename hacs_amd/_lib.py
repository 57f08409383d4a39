"""Loader for libhacs_mi355x.so (the C-ABI declared in include/hacs_mi355x.h).

The library is built in-tree (``hacs_amd/csrc/Makefile``) and loaded from this
directory only.  There is no fallback: a missing library or a missing symbol
raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhacs_mi355x.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "hacs_mi355x.h")

HX_OK = 0
HX_E_RUNTIME = -1
HX_E_ARGUMENT = -2
HX_E_UNAVAILABLE = -3
HX_E_NOMEM = -4
HX_E_NOT_INVERTIBLE = -5
HX_E_CAPTURE = -6

HX_RREF_BATCH_AUTO, HX_RREF_BATCH_ONE_PER_CU, HX_RREF_BATCH_TWO_PER_CU = 0, 1, 2

HX_OP_SET, HX_OP_ADD, HX_OP_SUB = 0, 1, 2
HX_GEMM_AUTO, HX_GEMM_PLAIN, HX_GEMM_LIMB = 0, 1, 2


class HxError(RuntimeError):
    """A failed C-ABI call (the engine::cuda::CudaError / ENGINE_CHECK equivalent)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class hx_modctx(C.Structure):
    _fields_ = [("p", C.c_uint32), ("limbs", C.c_uint32), ("barrett_m", C.c_uint64),
                ("k_max", C.c_uint64), ("two64_mod", C.c_uint64)]


class hx_modctx64(C.Structure):
    _fields_ = [("p", C.c_uint64), ("ninv", C.c_uint64), ("r2", C.c_uint64), ("two64", C.c_uint64)]


class hx_view(C.Structure):
    _fields_ = [("base", C.c_void_p), ("rs", C.c_int64), ("cs", C.c_int64),
                ("rows", C.c_uint32), ("cols", C.c_uint32)]


class hx_gemm_opts(C.Structure):
    _fields_ = [("variant", C.c_int), ("tile", C.c_int), ("workspace", C.c_void_p),
                ("workspace_bytes", C.c_size_t), ("main_begin", C.c_void_p), ("main_end", C.c_void_p),
                ("splits", C.c_uint32), ("split_workspace", C.c_void_p), ("split_workspace_bytes", C.c_size_t)]


class hx_device_info(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("arch", C.c_char * 64), ("index", C.c_int),
                ("compute_units", C.c_int), ("clock_khz", C.c_int), ("total_mem", C.c_size_t),
                ("free_mem", C.c_size_t), ("runtime_version", C.c_int), ("driver_version", C.c_int)]


class hx_gemm_job(C.Structure):
    _fields_ = [("m", C.c_uint32), ("n", C.c_uint32), ("k", C.c_uint32), ("op", C.c_int32),
                ("variant", C.c_int32), ("tile", C.c_int32), ("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p),
                ("lda", C.c_int64), ("ldb", C.c_int64), ("ldc", C.c_int64)]


class hx_rref_batch_info(C.Structure):
    _fields_ = [("njobs", C.c_uint32), ("workgroups_per_cu", C.c_uint32), ("grid", C.c_uint32),
                ("region_words", C.c_uint32), ("lds_bytes", C.c_size_t), ("jobs_planes_outside_lds", C.c_uint32)]


class hx_arena_stats(C.Structure):
    _fields_ = [("reserved_bytes", C.c_size_t), ("live_bytes", C.c_size_t), ("peak_live_bytes", C.c_size_t),
                ("peak_reserved_bytes", C.c_size_t), ("driver_mallocs", C.c_uint64), ("reuses", C.c_uint64),
                ("driver_frees", C.c_uint64), ("oom_retries", C.c_uint64)]


class hx_ring2k_stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("calls", "device", "resident", "tiled", "slabs", "replays", "fallbacks",
                                          "panels", "lookaheads", "verified", "ns_open", "ns_fetch", "ns_panel",
                                          "ns_step", "ns_host", "ns_close")]


class hx_staging_stats(C.Structure):
    _fields_ = [("chunk_bytes", C.c_size_t), ("cap_bytes", C.c_size_t), ("allocated_bytes", C.c_size_t),
                ("peak_bytes", C.c_size_t), ("acquisitions", C.c_uint64), ("waits", C.c_uint64)]


class hx_rref_job(C.Structure):
    _fields_ = [("m", C.c_uint32), ("n", C.c_uint32), ("A", C.c_void_p), ("lda", C.c_int64),
                ("pivots", C.c_void_p), ("pivot_rows", C.c_void_p), ("rank", C.c_void_p), ("panel", C.c_uint32)]


class hx_howell_job(C.Structure):
    _fields_ = [("m", C.c_uint32), ("n", C.c_uint32), ("A", C.c_void_p), ("lda", C.c_int64),
                ("H", C.c_void_p), ("pivots", C.c_void_p), ("pivot_v", C.c_void_p), ("h", C.c_void_p),
                ("work", C.c_void_p)]


class hx_stats(C.Structure):
    _fields_ = [("kernel_launches", C.c_uint64), ("gemm_launches", C.c_uint64), ("batched_instances", C.c_uint64),
                ("limb_launches", C.c_uint64), ("h2d_bytes", C.c_uint64), ("d2h_bytes", C.c_uint64),
                ("d2d_bytes", C.c_uint64), ("h2d_calls", C.c_uint64), ("d2h_calls", C.c_uint64),
                ("mallocs", C.c_uint64), ("frees", C.c_uint64), ("splitk_launches", C.c_uint64)]


class hx_elim_stats(C.Structure):
    _fields_ = [("calls", C.c_uint64), ("panels", C.c_uint64), ("panel_columns", C.c_uint64),
                ("kernel_launches", C.c_uint64), ("gemms", C.c_uint64), ("h2d_bytes", C.c_uint64),
                ("d2h_bytes", C.c_uint64), ("peak_device_bytes", C.c_size_t), ("seconds", C.c_double),
                ("spec_panels", C.c_uint64)]


class hx_ring2k_step_args(C.Structure):
    _fields_ = [("m", C.c_uint32), ("r", C.c_uint32), ("np", C.c_uint32), ("pc0", C.c_uint32), ("pc1", C.c_uint32),
                ("panel", C.c_void_p), ("ldp", C.c_size_t), ("swaps", C.c_void_p), ("nswaps", C.c_uint32),
                ("L", C.c_void_p), ("ldl", C.c_size_t), ("u0", C.c_uint32), ("la0", C.c_uint32),
                ("la1", C.c_uint32), ("w", C.c_uint32)]


_V = C.c_void_p
_U32, _U64, _I64, _INT, _SZ, _F = C.c_uint32, C.c_uint64, C.c_int64, C.c_int, C.c_size_t, C.c_float
_U32P = C.POINTER(C.c_uint32)

# name -> (restype, argtypes)
SIGNATURES = {
    "hx_available": (_INT, []),
    "hx_unavailable_reason": (C.c_char_p, []),
    "hx_device_info_get": (_INT, [C.POINTER(hx_device_info)]),
    "hx_last_error": (C.c_char_p, []),
    "hx_version": (C.c_char_p, []),
    "hx_modctx_init": (_INT, [_U32, C.POINTER(hx_modctx)]),
    "hx_stream_create": (_INT, [C.POINTER(_V)]),
    "hx_stream_destroy": (_INT, [_V]),
    "hx_stream_sync": (_INT, [_V]),
    "hx_device_sync": (_INT, []),
    "hx_event_create": (_INT, [C.POINTER(_V), _INT]),
    "hx_event_record": (_INT, [_V, _V]),
    "hx_event_wait": (_INT, [_V, _V]),
    "hx_event_sync": (_INT, [_V]),
    "hx_event_elapsed_ms": (_INT, [_V, _V, C.POINTER(_F)]),
    "hx_event_destroy": (_INT, [_V]),
    "hx_malloc": (_INT, [C.POINTER(_V), _SZ]),
    "hx_free": (_INT, [_V]),
    "hx_upload": (_INT, [_V, _V, _SZ, _V]),
    "hx_download": (_INT, [_V, _V, _SZ, _V]),
    "hx_copy_d2d": (_INT, [_V, _V, _SZ, _V]),
    "hx_host_alloc": (_INT, [C.POINTER(_V), _SZ]),
    "hx_host_free": (_INT, [_V]),
    "hx_staging_stats_get": (_INT, [C.POINTER(hx_staging_stats)]),
    "hx_memset": (_INT, [_V, _INT, _SZ, _V]),
    "hx_graph_begin": (_INT, [_V]),
    "hx_graph_end": (_INT, [_V, C.POINTER(_V)]),
    "hx_graph_launch": (_INT, [_V, _V]),
    "hx_graph_destroy": (_INT, [_V]),
    "hx_gemm_u32": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, hx_view, _INT, _V, C.POINTER(hx_gemm_opts)]),
    "hx_gemm_workspace_bytes": (_SZ, [C.POINTER(hx_modctx), _U32, _U32, _U32, _INT]),
    "hx_gemm_workspace_bytes_tile": (_SZ, [C.POINTER(hx_modctx), _U32, _U32, _U32, _INT, _INT]),
    "hx_gemm_residue_moduli": (_U32, [_U32, _U32]),
    "hx_gemm_route": (_INT, [C.POINTER(hx_modctx), _U32, _U32, _U32, _INT, _INT]),
    "hx_modctx64_init": (_INT, [_U64, C.POINTER(hx_modctx64)]),
    "hx_gemm_u64": (_INT, [C.POINTER(hx_modctx64), hx_view, hx_view, hx_view, _INT, _V]),
    "hx_gemm_limb_k_chunk": (_U32, [C.POINTER(hx_modctx), _U32, _U32, _U32, _SZ]),
    "hx_rref_batch_info_get": (_INT, [_V, C.POINTER(hx_rref_batch_info)]),
    "hx_arena_stats_get": (_INT, [C.POINTER(hx_arena_stats)]),
    "hx_arena_trim": (_INT, [C.POINTER(_SZ)]),
    "hx_f2_rref": (_INT, [_U32, _U32, _V, _U32, _U32P, _U32P, _V]),
    "hx_f2_rref_workspace_bytes": (_SZ, [_U32, _U32, _U32]),
    "hx_f2_gemm": (_INT, [_U32, _U32, _U32, _V, _U32, _V, _U32, _V, _U32, _INT, _V]),
    "hx_f3_rref": (_INT, [_U32, _U32, _V, _U32, _U32P, _U32P, _V]),
    "hx_f3_rref_workspace_bytes": (_SZ, [_U32, _U32, _U32]),
    "hx_f3_gemm": (_INT, [_U32, _U32, _U32, _V, _U32, _V, _U32, _V, _U32, _INT, _V]),
    "hx_elim_stats_get": (_INT, [C.POINTER(hx_elim_stats)]),
    "hx_elim_stats_reset": (_INT, []),
    "hx_gemm_suggest_splits": (_U32, [_U32, _U32, _U32]),
    "hx_gemm_split_workspace_bytes": (_SZ, [_U32, _U32, _U32, _U32]),
    "hx_gemm_u32_noalloc": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, hx_view, _INT, _V, C.POINTER(hx_gemm_opts)]),
    "hx_gemm_noalloc_scratch_bytes": (_SZ, []),
    "hx_mfma_rate": (_INT, [_U32, _U32, _INT, _V, C.POINTER(C.c_double)]),
    "hx_schur_update_u32": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, hx_view, _V, C.POINTER(hx_gemm_opts)]),
    "hx_gemm_batch_create": (_INT, [C.POINTER(hx_modctx), C.POINTER(hx_gemm_job), _U32, _INT, C.POINTER(_V)]),
    "hx_gemm_batch_run": (_INT, [_V, _V]),
    "hx_gemm_batch_destroy": (_INT, [_V]),
    "hx_trsm_unit_u32": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, _INT, _V]),
    "hx_stats_get": (_INT, [C.POINTER(hx_stats)]),
    "hx_stats_reset": (_INT, []),
    "hx_gemm_nt_u32": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, hx_view, _INT, _V, C.POINTER(hx_gemm_opts)]),
    "hx_gemv_u32": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, hx_view, _INT, _V]),
    "hx_axpy_u32": (_INT, [C.POINTER(hx_modctx), _U32, hx_view, hx_view, _INT, _V]),
    "hx_scal_u32": (_INT, [C.POINTER(hx_modctx), _U32, hx_view, _V]),
    "hx_copy_u32": (_INT, [hx_view, hx_view, _V]),
    "hx_transpose_u32": (_INT, [hx_view, hx_view, _V]),
    "hx_fill_u32": (_INT, [hx_view, _U32, _V]),
    "hx_axpy_identity_u32": (_INT, [C.POINTER(hx_modctx), _U32, hx_view, _INT, _V]),
    "hx_gather_u32": (_INT, [C.POINTER(hx_modctx), _U32, _U32, hx_view, _U32, _V, hx_view, _INT, _V, _V]),
    "hx_canonicalise_u32": (_INT, [C.POINTER(hx_modctx), hx_view, _V]),
    "hx_canonicalise_acc64_u32": (_INT, [C.POINTER(hx_modctx), hx_view, hx_view, _V]),
    "hx_rref_u32": (_INT, [C.POINTER(hx_modctx), _U32, _U32, _V, _I64, _V, _I64, _U32P, _U32P, _U32P, _V]),
    "hx_rref_workspace_bytes": (_SZ, [_U32, _U32]),
    "hx_rref_u32_device": (_INT, [C.POINTER(hx_modctx), _U32, _U32, _V, _I64, _V, _I64, _V, _V, _V, _V, _SZ, _V]),
    "hx_rref_u32_host": (_INT, [C.POINTER(hx_modctx), _U32, _U32, _V, _V, _V, _U32, _U32P, _U32P, _U32P]),
    "hx_rref_batch_create": (_INT, [C.POINTER(hx_modctx), C.POINTER(hx_rref_job), _U32, _INT, C.POINTER(_V)]),
    "hx_rref_batch_run": (_INT, [_V, _V]),
    "hx_rref_batch_destroy": (_INT, [_V]),
    "hx_rref_batch_status": (_INT, [_V, _V]),
    "hx_howell_work_words": (_SZ, [_U32, _U32]),
    "hx_howell_batch_create": (_INT, [_U32, C.POINTER(hx_howell_job), _U32, C.POINTER(_V)]),
    "hx_howell_batch_create_ell": (_INT, [_U32, _U32, C.POINTER(hx_howell_job), _U32, C.POINTER(_V)]),
    "hx_howell_batch_run": (_INT, [_V, _V]),
    "hx_howell_batch_destroy": (_INT, [_V]),
    "hx_mlp_forward": (_INT, [_V, _U32, _U32, _V, _V, _U32, _V, _V, _U32, _V, _V, _V]),
    "hx_mlp_backward": (_INT, [_V, _U32, _U32, _V, _U32, _V, _U32, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "hx_adam_step": (_INT, [_V, _V, _V, _V, _U32, _F, _F, _F, _F, _U32, _V]),
    "hx_random_entries": (_INT, [_V, _U32, _U32, _I64, _U32, _U64, _V]),
    "hx_random_sparse": (_INT, [_V, _U32, _U32, _I64, _U32, _U64, _V]),
    "hx_ring2k_session_bytes": (_SZ, [_U32, _U32, _U32, _U32]),
    "hx_ring2k_staging_bytes": (_SZ, [_U32, _U32, _U32]),
    "hx_ring2k_open": (_INT, [_U32, _U32, _U32, _U32, C.POINTER(_V)]),
    "hx_ring2k_put": (_INT, [_V, _V, _SZ, _U32, _U32, _U32, _U32]),
    "hx_ring2k_get": (_INT, [_V, _V, _SZ, _U32, _U32, _U32, _U32]),
    "hx_ring2k_step": (_INT, [_V, C.POINTER(hx_ring2k_step_args)]),
    "hx_ring2k_fetch": (_INT, [_V, _V, _SZ]),
    "hx_ring2k_close": (_INT, [_V]),
    "hx_ring2k_plan": (_U32, [_U32, _U32, _U32, _U32, _U64]),
    "hx_ring2k_store_budget": (_U64, []),
    "hx_ring2k_verify": (_INT, []),
    "hx_ring2k_echelon": (_INT, [_U32, _U32, _U32, _U32, _V, _SZ, _U32, _U32, _U64, _U64, _U32P, _U32P, _U32P]),
    "hx_ring2k_smith": (_INT, [_U32, _U32, _U32, _U32, _V, _SZ, _U32, _U32P, _U32P, _U32P, _U32P]),
    "hx_ring2k_stats_get": (_INT, [C.POINTER(hx_ring2k_stats)]),
    "hx_ring2k_stats_reset": (_INT, []),
    "hx_ring2k_gemm_sub_host": (_INT, [_U32, _V, _SZ, _V, _SZ, _V, _SZ, _U32, _U32, _U32]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"HX_API\s+[\w\s\*]+?\b(hx_\w+)\s*\(", text)))


def lib():
    """The loaded library; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C hacs_amd/csrc` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != HX_OK:
        raise HxError(code, lib().hx_last_error().decode(errors="replace"))


def available() -> bool:
    return bool(lib().hx_available())
