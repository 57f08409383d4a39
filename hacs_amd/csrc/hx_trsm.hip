// hx_trsm.hip -- B <- T^{-1} B for a unit-triangular T (Cuda/CudaKernels.h:240-247 trsm_unit).
// Blocked like the reference (cuda_kernels.cu:1089-1131): 32-row diagonal solves (one thread per
// column of B, the 32 x 32 diagonal block in LDS, exact u64+carry dot products) and a GEMM(Sub)
// trailing update per block through the GEMM dispatcher.  Only the strict triangle of T is read.
#include "hx_gemm.h"

namespace hx {
namespace {

constexpr int TB = 32;

__global__ __launch_bounds__(256) void trsm_diag_kernel(const uint32_t* T, int64_t trs, int64_t tcs, uint32_t* B,
                                                        int64_t brs, int64_t bcs, uint32_t nb, uint32_t q,
                                                        int upper, DevMod M) {
    __shared__ uint32_t Ts[TB][TB + 1];
    for (uint32_t idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const uint32_t i = idx / nb, t = idx % nb;
        Ts[i][t] = T[(int64_t)i * trs + (int64_t)t * tcs];
    }
    __syncthreads();
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= q) return;
    uint32_t x[TB];
#pragma unroll
    for (int i = 0; i < TB; ++i) x[i] = (uint32_t)i < nb ? B[(int64_t)i * brs + (int64_t)j * bcs] : 0u;
    if (!upper) {
#pragma unroll
        for (int i = 1; i < TB; ++i) {
            if ((uint32_t)i >= nb) continue;          // (guards, not breaks: the loops unroll fully)
            uint64_t lo = 0;
            uint32_t hi = 0;
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                if (t >= i) continue;
                const uint64_t prod = (uint64_t)Ts[i][t] * x[t];
                lo += prod;
                hi += lo < prod ? 1u : 0u;
            }
            x[i] = subm(x[i], red96(hi, lo, M), M.p);
        }
    } else {
#pragma unroll
        for (int i = TB - 2; i >= 0; --i) {
            if ((uint32_t)i + 1 >= nb) continue;
            uint64_t lo = 0;
            uint32_t hi = 0;
#pragma unroll
            for (int t = TB - 1; t > 0; --t) {
                if (t <= i || (uint32_t)t >= nb) continue;
                const uint64_t prod = (uint64_t)Ts[i][t] * x[t];
                lo += prod;
                hi += lo < prod ? 1u : 0u;
            }
            x[i] = subm(x[i], red96(hi, lo, M), M.p);
        }
    }
#pragma unroll
    for (int i = 0; i < TB; ++i)
        if ((uint32_t)i < nb) B[(int64_t)i * brs + (int64_t)j * bcs] = x[i];
}

}  // namespace

void trsm_unit(const hx_modctx& ctx, hx_view T, hx_view B, int upper, hipStream_t s) {
    const uint32_t n = T.rows, q = B.cols;
    if (n == 0 || q == 0) return;
    const uint32_t* Tp = (const uint32_t*)T.base;
    uint32_t* Bp = (uint32_t*)B.base;
    const uint32_t nblk = (n + TB - 1) / TB;
    for (uint32_t bi = 0; bi < nblk; ++bi) {
        const uint32_t blk = upper ? nblk - 1 - bi : bi;
        const uint32_t r0 = blk * TB, nb = std::min<uint32_t>(TB, n - r0);
        trsm_diag_kernel<<<(q + 255) / 256, 256, 0, s>>>(Tp + (int64_t)r0 * T.rs + (int64_t)r0 * T.cs, T.rs, T.cs,
                                                         Bp + (int64_t)r0 * B.rs, B.rs, B.cs, nb, q, upper,
                                                         devmod(ctx));
        after_launch("trsm_diag_kernel");
        // the rows not yet solved: lower -> below the block, upper -> above it
        const uint32_t o0 = upper ? 0 : r0 + nb, on = upper ? r0 : n - (r0 + nb);
        if (on == 0) continue;
        GemmArgs g;
        g.m = on;
        g.n = q;
        g.k = nb;
        g.A = Tp + (int64_t)o0 * T.rs + (int64_t)r0 * T.cs;
        g.ars = T.rs;
        g.acs = T.cs;
        g.B = Bp + (int64_t)r0 * B.rs;
        g.brs = B.rs;
        g.bcs = B.cs;
        g.Cin = Bp + (int64_t)o0 * B.rs;
        g.cirs = B.rs;
        g.cics = B.cs;
        g.Cout = Bp + (int64_t)o0 * B.rs;
        g.cors = B.rs;
        g.cocs = B.cs;
        g.op = HX_OP_SUB;
        gemm_dispatch(ctx, g, s, HX_GEMM_AUTO, nullptr, 0, 0);
    }
}

}  // namespace hx

using namespace hx;

extern "C" HX_API int hx_trsm_unit_u32(const hx_modctx* ctx, hx_view T, hx_view B, int upper, hx_stream s) {
    return guard("hx_trsm_unit_u32", [&] {
        HX_CHECK(ctx != nullptr && ctx->p >= 2, "bad modulus context");
        HX_CHECK(T.rows == T.cols && T.rows == B.rows, "trsm_unit: T must be square with B.rows rows");
        trsm_unit(*ctx, T, B, upper, (hipStream_t)s);
    });
}
