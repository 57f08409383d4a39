// hx_gemm.h -- internal GEMM launch interface shared by the GEMM, elimination and batch sources.
#pragma once

#include "hx_common.h"

namespace hx {

// C_out op= A * B mod p with optional row / column gathers and device-side sizes.
//   A(i, t) = A[ra(i) * ars + ca(t) * acs],  ra(i) = a_rows ? a_rows[i] : i,  ca(t) = a_cols ? a_cols[t] : t
//   B(t, j) = B[t * brs + j * bcs]
//   C_in(i, j) = Cin[rc(i) * cirs + j * cics], rc(i) = c_rows ? c_rows[i] : i   (read when op != SET)
//   C_out(i, j) = Cout[i * cors + j * cocs]
// m, k are upper bounds when m_dev / k_dev are given (the kernels read the true sizes there).
struct GemmArgs {
    uint32_t m = 0, n = 0, k = 0;
    const uint32_t* A = nullptr;
    int64_t ars = 0, acs = 0;
    const uint32_t* a_rows = nullptr;
    const uint32_t* a_cols = nullptr;
    const uint32_t* B = nullptr;
    int64_t brs = 0, bcs = 0;
    const uint32_t* Cin = nullptr;
    int64_t cirs = 0, cics = 0;
    const uint32_t* c_rows = nullptr;
    uint32_t* Cout = nullptr;
    int64_t cors = 0, cocs = 0;
    const uint32_t* m_dev = nullptr;
    const uint32_t* k_dev = nullptr;
    int op = HX_OP_SET;
};

// The plain VALU kernel (any p < 2^32); splits > 1: split-K in ws (plain_split_workspace_bytes).
void gemm_plain(const hx_modctx& ctx, const GemmArgs& g, hipStream_t s, int tile, hipEvent_t ev0 = nullptr,
                hipEvent_t ev1 = nullptr, uint32_t splits = 1, void* ws = nullptr, size_t ws_bytes = 0);
// Split-K of the plain kernel: the split count actually run (no empty split, at most HX_MAX_SPLITS),
// the workspace it needs, and the automatic policy's count for a shape.
uint32_t plain_splits_clamp(uint32_t k, uint32_t splits);
size_t plain_split_workspace_bytes(uint32_t m, uint32_t n, uint32_t splits);
uint32_t plain_suggest_splits(uint32_t m, uint32_t n, uint32_t k);
// The limb-MFMA route (ctx.limbs > 0). ws: the limb planes' workspace (ws_bytes); K runs in the chunks it holds.
void gemm_limb(const hx_modctx& ctx, const GemmArgs& g, hipStream_t s, void* ws, size_t ws_bytes, int tile,
               hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
size_t limb_workspace_bytes(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k);
// Largest K one limb pass can absorb with exact int32 diagonals (multiple of 64).
uint32_t limb_k_chunk(uint32_t limbs);
// The K chunk a limb workspace of `bytes` allows for this shape (0: not one 64-deep chunk).
uint32_t limb_k_chunk_for(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k, size_t bytes);
// The variant the automatic policy picks for this shape.
int gemm_auto_variant(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k);
// splits: 1 = never split K, 0 = the automatic policy, > 1 = that many (plain kernel only); split_ws NULL =
// the stream's own partials.
void gemm_dispatch(const hx_modctx& ctx, const GemmArgs& g, hipStream_t s, int variant, void* ws, size_t ws_bytes,
                   int tile, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, uint32_t splits = 1,
                   void* split_ws = nullptr, size_t split_ws_bytes = 0);

}  // namespace hx
