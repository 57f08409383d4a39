// hx_level1.hip -- the level-1/2 and tape-support kernels of the device layer: gemv, axpy, scal,
// copy, transpose, fill, axpy_identity, gather, canonicalise (u32 and the u64 accumulator form)
// and gemm_nt.  They replace the matching engine::cuda kernels (Cuda/CudaKernels.h:186-238,
// cuda_kernels.cu:590-700 and 950-1087) with the same results and the same argument checks.
//
// Every kernel here moves each byte once and does at most one modular product per element, so all
// are bound by HBM bandwidth.  Layout for MI355X: a 256-thread workgroup covers 256 consecutive
// elements of the fast (unit-stride) dimension times 4 of the slow one; each wave's 64 lanes touch
// 64 consecutive elements per access (one 256-byte burst), and each thread issues its four loads
// before any store so four accesses per lane are in flight.  The fast dimension is the view's
// columns unless its rows are the unit-stride direction (the reference's win_of rule,
// cuda_kernels.cu:597-603).  gemv accumulates exactly (96-bit sums, one Barrett reduction per
// row), the same residue as the reference's blocked u64 sums.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hx_common.h"

namespace hx {
namespace {

constexpr int EF = 4;                  // fast-dimension elements per thread
constexpr int ES = 4;                  // slow-dimension lines per workgroup (one per wave)
constexpr uint32_t ETILE = 64 * EF;    // fast-dimension elements per workgroup

// A view walked as (fast, slow): element (f, s) at base[f * sf + s * ss].
struct Win {
    uint32_t nf, ns;
    int64_t sf, ss;
};

bool rows_fast(const hx_view& v) { return v.cs != 1 && v.rs == 1; }
Win win_of(const hx_view& v, bool rf) { return rf ? Win{v.rows, v.cols, v.rs, v.cs} : Win{v.cols, v.rows, v.cs, v.rs}; }

struct Grid {
    uint32_t tiles_f, blocks;
};
Grid ew_grid(const Win& w) {
    const uint64_t tf = (w.nf + ETILE - 1) / ETILE, ts = (w.ns + ES - 1) / ES;
    HX_CHECK(tf * ts < (1ull << 31), "view too large for one launch");
    return Grid{(uint32_t)tf, (uint32_t)(tf * ts)};
}

// The (fast, slow) coordinates of this thread's EF elements; false when its slow line is outside.
__device__ __forceinline__ bool ew_coords(const Win& w, uint32_t tiles_f, uint32_t& f0, uint32_t& sl) {
    const uint32_t bf = blockIdx.x % tiles_f, bs = blockIdx.x / tiles_f;
    sl = bs * ES + (threadIdx.x >> 6);
    f0 = bf * ETILE + (threadIdx.x & 63);
    return sl < w.ns;
}
__device__ __forceinline__ int64_t at(const Win& w, uint32_t f, uint32_t sl) {
    return (int64_t)f * w.sf + (int64_t)sl * w.ss;
}

// y = y +/- alpha x (axpy_kernel, cuda_kernels.cu:609-623)
__global__ __launch_bounds__(256) void axpy_kernel(uint32_t alpha, const uint32_t* x, Win wx, uint32_t* y, Win wy,
                                                   int subtract, uint32_t tiles_f, DevMod M) {
    uint32_t f0, sl;
    if (!ew_coords(wx, tiles_f, f0, sl)) return;
    uint32_t xv[EF], yv[EF];
#pragma unroll
    for (int u = 0; u < EF; ++u) {
        const uint32_t f = f0 + 64 * u;
        if (f < wx.nf) {
            xv[u] = x[at(wx, f, sl)];
            yv[u] = y[at(wy, f, sl)];
        }
    }
#pragma unroll
    for (int u = 0; u < EF; ++u) {
        const uint32_t f = f0 + 64 * u;
        if (f < wx.nf) y[at(wy, f, sl)] = apply_op(yv[u], mulm(alpha, xv[u], M), subtract ? HX_OP_SUB : HX_OP_ADD, M.p);
    }
}

// x = alpha x (scal_kernel, cuda_kernels.cu:624-633)
__global__ __launch_bounds__(256) void scal_kernel(uint32_t alpha, uint32_t* x, Win wx, uint32_t tiles_f, DevMod M) {
    uint32_t f0, sl;
    if (!ew_coords(wx, tiles_f, f0, sl)) return;
    uint32_t xv[EF];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wx.nf) xv[u] = x[at(wx, f0 + 64 * u, sl)];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wx.nf) x[at(wx, f0 + 64 * u, sl)] = mulm(alpha, xv[u], M);
}

// v = v >= p ? v - p : v (canon_kernel, cuda_kernels.cu:634-643)
__global__ __launch_bounds__(256) void canon_kernel(uint32_t* x, Win wx, uint32_t tiles_f, uint32_t p) {
    uint32_t f0, sl;
    if (!ew_coords(wx, tiles_f, f0, sl)) return;
    uint32_t xv[EF];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wx.nf) xv[u] = x[at(wx, f0 + 64 * u, sl)];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wx.nf) x[at(wx, f0 + 64 * u, sl)] = xv[u] >= p ? xv[u] - p : xv[u];
}

// dst = acc mod p for a u64 accumulator (canon64_kernel, cuda_kernels.cu:644-649)
__global__ __launch_bounds__(256) void canon64_kernel(const uint64_t* acc, Win wa, uint32_t* dst, Win wd,
                                                      uint32_t tiles_f, DevMod M) {
    uint32_t f0, sl;
    if (!ew_coords(wa, tiles_f, f0, sl)) return;
    uint64_t av[EF];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wa.nf) av[u] = acc[at(wa, f0 + 64 * u, sl)];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wa.nf) dst[at(wd, f0 + 64 * u, sl)] = red64(av[u], M);
}

// dst = src (copy_kernel, cuda_kernels.cu:650-656)
__global__ __launch_bounds__(256) void copy_kernel(const uint32_t* src, Win ws, uint32_t* dst, Win wd, uint32_t tiles_f) {
    uint32_t f0, sl;
    if (!ew_coords(ws, tiles_f, f0, sl)) return;
    uint32_t v[EF];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < ws.nf) v[u] = src[at(ws, f0 + 64 * u, sl)];
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < ws.nf) dst[at(wd, f0 + 64 * u, sl)] = v[u];
}

// x = value (fill_kernel, cuda_kernels.cu:659-665)
__global__ __launch_bounds__(256) void fill_kernel(uint32_t* x, Win wx, uint32_t value, uint32_t tiles_f) {
    uint32_t f0, sl;
    if (!ew_coords(wx, tiles_f, f0, sl)) return;
#pragma unroll
    for (int u = 0; u < EF; ++u)
        if (f0 + 64 * u < wx.nf) x[at(wx, f0 + 64 * u, sl)] = value;
}

// y(i, i) +/-= alpha (axpy_identity_kernel, cuda_kernels.cu:666-672)
__global__ __launch_bounds__(256) void axpy_identity_kernel(uint32_t alpha, uint32_t* y, int64_t step, uint32_t n,
                                                            int subtract, uint32_t p) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t* yp = y + (int64_t)i * step;
    *yp = apply_op(*yp, alpha, subtract ? HX_OP_SUB : HX_OP_ADD, p);
}

// c(i, t) [+]= alpha [row_off + r(i) == col(t)] + beta a(r(i), col(t)) (gather_kernel,
// cuda_kernels.cu:673-695); t is the fast index, i the slow one.
struct GatherArgs {
    uint32_t alpha, beta;               // beta == 0 or a == nullptr: a is not read
    const uint32_t* a;
    int64_t a_rs, a_cs;
    uint32_t row_off;
    const uint32_t* cols;               // nullptr: col(t) = t
    const uint32_t* rows;               // nullptr: r(i) = i
    uint32_t* c;
    int64_t c_rs, c_cs;
    uint32_t nrows, ncols;
    int accumulate;
};
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs g, uint32_t tiles_f, DevMod M) {
    const uint32_t bf = blockIdx.x % tiles_f, bs = blockIdx.x / tiles_f;
    const uint32_t i = bs * ES + (threadIdx.x >> 6);
    if (i >= g.nrows) return;
    const uint32_t t0 = bf * ETILE + (threadIdx.x & 63);
    const uint32_t ar = g.rows ? g.rows[i] : i;
    uint32_t col[EF], av[EF], cv[EF];
#pragma unroll
    for (int u = 0; u < EF; ++u) {
        const uint32_t t = t0 + 64 * u;
        if (t >= g.ncols) continue;
        col[u] = g.cols ? g.cols[t] : t;
        av[u] = (g.a != nullptr && g.beta != 0) ? g.a[(int64_t)ar * g.a_rs + (int64_t)col[u] * g.a_cs] : 0u;
        cv[u] = g.accumulate ? g.c[(int64_t)i * g.c_rs + (int64_t)t * g.c_cs] : 0u;
    }
#pragma unroll
    for (int u = 0; u < EF; ++u) {
        const uint32_t t = t0 + 64 * u;
        if (t >= g.ncols) continue;
        uint32_t v = (g.a != nullptr && g.beta != 0) ? mulm(g.beta, av[u], M) : 0u;
        if (g.alpha != 0 && g.row_off + ar == col[u]) v = addm(v, g.alpha, M.p);
        g.c[(int64_t)i * g.c_rs + (int64_t)t * g.c_cs] = g.accumulate ? addm(cv[u], v, M.p) : v;
    }
}

// dst(i, j) = src(j, i) through a 64 x 64 LDS tile (transpose_kernel, cuda_kernels.cu:696-713):
// the loads run along dst's rows = src's columns, the stores along dst's columns.
constexpr int TT = 64;
__global__ __launch_bounds__(256) void transpose_kernel(const uint32_t* src, int64_t srs, int64_t scs, uint32_t* dst,
                                                        int64_t drs, int64_t dcs, uint32_t rows, uint32_t cols,
                                                        uint32_t tiles_j) {
    __shared__ uint32_t tile[TT][TT + 1];
    const uint32_t i0 = (blockIdx.x / tiles_j) * TT, j0 = (blockIdx.x % tiles_j) * TT;
    const uint32_t lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
#pragma unroll 4
    for (int r = 0; r < TT; r += 4) {
        const uint32_t jj = j0 + ly + r, ii = i0 + lx;     // src (jj, ii)
        if (jj < cols && ii < rows) tile[ly + r][lx] = src[(int64_t)jj * srs + (int64_t)ii * scs];
    }
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < TT; r += 4) {
        const uint32_t ii = i0 + ly + r, jj = j0 + lx;     // dst (ii, jj)
        if (ii < rows && jj < cols) dst[(int64_t)ii * drs + (int64_t)jj * dcs] = tile[lx][ly + r];
    }
}

// y op= A x (gemv_warp_kernel / gemv_thread_kernel, cuda_kernels.cu:536-589): exact sums of
// u32 x u32 products in a 64-bit word plus a carry word (< 2^96 for k < 2^32), one reduction.
struct GemvArgs {
    const uint32_t* A;
    int64_t a_rs, a_cs;
    const uint32_t* x;
    int64_t x_rs;
    uint32_t* y;
    int64_t y_rs;
    uint32_t m, k;
    int op;
};
// One wave per row (A's rows unit-stride along k, or no unit stride at all).
__global__ __launch_bounds__(256) void gemv_row_kernel(GemvArgs g, DevMod M) {
    const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= g.m) return;
    uint64_t lo = 0;
    uint32_t hi = 0;
    const uint32_t* arow = g.A + (int64_t)i * g.a_rs;
    for (uint32_t j = lane; j < g.k; j += 64) {
        const uint64_t t = (uint64_t)arow[(int64_t)j * g.a_cs] * g.x[(int64_t)j * g.x_rs];
        lo += t;
        hi += lo < t;
    }
    uint64_t r = red96(hi, lo, M);    // < p < 2^32; 64 of them sum below 2^38
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane == 0) {
        uint32_t* yp = g.y + (int64_t)i * g.y_rs;
        *yp = apply_op(g.op == HX_OP_SET ? 0u : *yp, red64(r, M), g.op, M.p);
    }
}
// One thread per row (A column-major: consecutive rows are consecutive words).
__global__ __launch_bounds__(256) void gemv_col_kernel(GemvArgs g, DevMod M) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.m) return;
    uint64_t lo = 0;
    uint32_t hi = 0;
    for (uint32_t j = 0; j < g.k; ++j) {
        const uint64_t t = (uint64_t)g.A[(int64_t)i * g.a_rs + (int64_t)j * g.a_cs] * g.x[(int64_t)j * g.x_rs];
        lo += t;
        hi += lo < t;
    }
    uint32_t* yp = g.y + (int64_t)i * g.y_rs;
    *yp = apply_op(g.op == HX_OP_SET ? 0u : *yp, red96(hi, lo, M), g.op, M.p);
}

bool empty(const hx_view& v) { return v.rows == 0 || v.cols == 0; }
void check_ctx(const hx_modctx* ctx) { HX_CHECK(ctx != nullptr && ctx->p >= 2, "bad modulus context"); }

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API int hx_gemm_nt_u32(const hx_modctx* ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s,
                          const hx_gemm_opts* opts) {
    // gemm_nt (cuda_kernels.cu:819-823): gemm with B's transposed view
    const hx_view Bt{B.base, B.cs, B.rs, B.cols, B.rows};
    return hx_gemm_u32(ctx, A, Bt, C, op, s, opts);
}

HX_API int hx_gemv_u32(const hx_modctx* ctx, hx_view A, hx_view x, hx_view y, int op, hx_stream s) {
    return guard("hx_gemv_u32", [&] {
        check_ctx(ctx);
        HX_CHECK(x.cols == 1 && y.cols == 1, "gemv: x and y must be column views");
        HX_CHECK(A.cols == x.rows && A.rows == y.rows, "gemv: shape mismatch");
        HX_CHECK(op == HX_OP_SET || op == HX_OP_ADD || op == HX_OP_SUB, "gemv: bad op");
        HX_CHECK(!views_overlap(y, A) && !views_overlap(y, x), "gemv: y aliases an operand");
        if (empty(y)) return;
        GemvArgs g{(const uint32_t*)A.base, A.rs, A.cs, (const uint32_t*)x.base, x.rs, (uint32_t*)y.base, y.rs,
                   A.rows, A.cols, op};
        if (A.cs == 1 || A.rs != 1) {
            gemv_row_kernel<<<(g.m + 3) / 4, 256, 0, (hipStream_t)s>>>(g, devmod(*ctx));
            after_launch("gemv_row_kernel");
        } else {
            gemv_col_kernel<<<(g.m + 255) / 256, 256, 0, (hipStream_t)s>>>(g, devmod(*ctx));
            after_launch("gemv_col_kernel");
        }
    });
}

HX_API int hx_axpy_u32(const hx_modctx* ctx, uint32_t alpha, hx_view x, hx_view y, int subtract, hx_stream s) {
    return guard("hx_axpy_u32", [&] {
        check_ctx(ctx);
        HX_CHECK(x.rows == y.rows && x.cols == y.cols, "axpy: shape mismatch");
        HX_CHECK(alpha < ctx->p, "axpy: alpha must be canonical");
        if (empty(x)) return;
        const bool rf = rows_fast(x);
        const Win wx = win_of(x, rf), wy = win_of(y, rf);
        const Grid gr = ew_grid(wx);
        axpy_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>(alpha, (const uint32_t*)x.base, wx, (uint32_t*)y.base, wy,
                                                            subtract ? 1 : 0, gr.tiles_f, devmod(*ctx));
        after_launch("axpy_kernel");
    });
}

HX_API int hx_scal_u32(const hx_modctx* ctx, uint32_t alpha, hx_view x, hx_stream s) {
    return guard("hx_scal_u32", [&] {
        check_ctx(ctx);
        HX_CHECK(alpha < ctx->p, "scal: alpha must be canonical");
        if (empty(x)) return;
        const Win wx = win_of(x, rows_fast(x));
        const Grid gr = ew_grid(wx);
        scal_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>(alpha, (uint32_t*)x.base, wx, gr.tiles_f, devmod(*ctx));
        after_launch("scal_kernel");
    });
}

HX_API int hx_copy_u32(hx_view src, hx_view dst, hx_stream s) {
    return guard("hx_copy_u32", [&] {
        HX_CHECK(src.rows == dst.rows && src.cols == dst.cols, "copy: shape mismatch");
        if (empty(src)) return;
        if (src.base == dst.base && src.rs == dst.rs && src.cs == dst.cs) return;
        HX_CHECK(!views_overlap(src, dst), "copy: overlapping windows");
        const bool rf = rows_fast(src);
        const Win ws = win_of(src, rf), wd = win_of(dst, rf);
        const Grid gr = ew_grid(ws);
        copy_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>((const uint32_t*)src.base, ws, (uint32_t*)dst.base, wd,
                                                            gr.tiles_f);
        after_launch("copy_kernel");
    });
}

HX_API int hx_transpose_u32(hx_view src, hx_view dst, hx_stream s) {
    return guard("hx_transpose_u32", [&] {
        HX_CHECK(src.rows == dst.cols && src.cols == dst.rows, "transpose: shape mismatch");
        if (empty(src)) return;
        HX_CHECK(!views_overlap(src, dst), "transpose: overlapping windows");
        const uint64_t ti = (dst.rows + TT - 1) / TT, tj = (dst.cols + TT - 1) / TT;
        HX_CHECK(ti * tj < (1ull << 31), "transpose: view too large for one launch");
        transpose_kernel<<<(uint32_t)(ti * tj), 256, 0, (hipStream_t)s>>>(
            (const uint32_t*)src.base, src.rs, src.cs, (uint32_t*)dst.base, dst.rs, dst.cs, dst.rows, dst.cols,
            (uint32_t)tj);
        after_launch("transpose_kernel");
    });
}

HX_API int hx_fill_u32(hx_view x, uint32_t value, hx_stream s) {
    return guard("hx_fill_u32", [&] {
        if (empty(x)) return;
        const Win wx = win_of(x, rows_fast(x));
        const Grid gr = ew_grid(wx);
        fill_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>((uint32_t*)x.base, wx, value, gr.tiles_f);
        after_launch("fill_kernel");
    });
}

HX_API int hx_axpy_identity_u32(const hx_modctx* ctx, uint32_t alpha, hx_view y, int subtract, hx_stream s) {
    return guard("hx_axpy_identity_u32", [&] {
        check_ctx(ctx);
        HX_CHECK(alpha < ctx->p, "axpy_identity: alpha must be canonical");
        const uint32_t n = std::min(y.rows, y.cols);
        if (n == 0) return;
        axpy_identity_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)s>>>(alpha, (uint32_t*)y.base, y.rs + y.cs, n,
                                                                          subtract ? 1 : 0, ctx->p);
        after_launch("axpy_identity_kernel");
    });
}

HX_API int hx_gather_u32(const hx_modctx* ctx, uint32_t alpha, uint32_t beta, hx_view a, uint32_t row_off,
                         const uint32_t* cols, hx_view c, int accumulate, const uint32_t* rows, hx_stream s) {
    return guard("hx_gather_u32", [&] {
        check_ctx(ctx);
        HX_CHECK(alpha < ctx->p && beta < ctx->p, "gather: coefficients must be canonical");
        if (empty(c)) return;
        const bool has_a = a.base != nullptr && beta != 0;
        if (has_a) {
            HX_CHECK(rows != nullptr || a.rows == c.rows, "gather: row count mismatch");
            HX_CHECK(cols != nullptr || a.cols == c.cols, "gather: column count mismatch");
            HX_CHECK(!views_overlap(a, c), "gather: output aliases the source");
        }
        GatherArgs g{alpha, has_a ? beta : 0u, has_a ? (const uint32_t*)a.base : nullptr, a.rs, a.cs, row_off, cols,
                     rows, (uint32_t*)c.base, c.rs, c.cs, c.rows, c.cols, accumulate ? 1 : 0};
        const Grid gr = ew_grid(Win{c.cols, c.rows, c.cs, c.rs});
        gather_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>(g, gr.tiles_f, devmod(*ctx));
        after_launch("gather_kernel");
    });
}

HX_API int hx_canonicalise_u32(const hx_modctx* ctx, hx_view v, hx_stream s) {
    return guard("hx_canonicalise_u32", [&] {
        check_ctx(ctx);
        if (empty(v)) return;
        const Win w = win_of(v, rows_fast(v));
        const Grid gr = ew_grid(w);
        canon_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>((uint32_t*)v.base, w, gr.tiles_f, ctx->p);
        after_launch("canon_kernel");
    });
}

HX_API int hx_canonicalise_acc64_u32(const hx_modctx* ctx, hx_view acc, hx_view dst, hx_stream s) {
    return guard("hx_canonicalise_acc64_u32", [&] {
        check_ctx(ctx);
        HX_CHECK(acc.rows == dst.rows && acc.cols == dst.cols, "canonicalise(acc64): shape mismatch");
        if (empty(acc)) return;
        HX_CHECK(!views_overlap(acc, dst, 8, 4), "canonicalise(acc64): output aliases the accumulator");
        const bool rf = rows_fast(acc);
        const Win wa = win_of(acc, rf), wd = win_of(dst, rf);
        const Grid gr = ew_grid(wa);
        canon64_kernel<<<gr.blocks, 256, 0, (hipStream_t)s>>>((const uint64_t*)acc.base, wa, (uint32_t*)dst.base, wd,
                                                               gr.tiles_f, devmod(*ctx));
        after_launch("canon64_kernel");
    });
}

}  // extern "C"
