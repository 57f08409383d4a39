// hx_rref.hip -- the single-matrix RREF entry points (imma_rref's role, Cuda/CudaKernels.h:249-298,
// generalised from p <= 127 to every prime p < 2^31).
#include <chrono>
#include <vector>

#include "hx_common.h"

namespace hx {
int rref_batch_create_impl(const hx_modctx& ctx, const hx_rref_job* jobs, uint32_t njobs, int mode, void** out);
int rref_batch_run_impl(void* b, hipStream_t s);
int rref_batch_status_impl(void* b, hipStream_t s);
void rref_batch_destroy_impl(void* b);
// The multi-kernel path for large matrices (hx_rref_large.hip); false when it does not take the shape.
bool rref_large(const hx_modctx& ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                int64_t ldr, uint32_t* rank, uint32_t* pivots, uint32_t* pivot_rows, hipStream_t s, bool force);
// Whether one job of this shape fits the one-workgroup kernel's LDS (hx_rref_batch.hip).
bool rref_batch_fits(uint32_t m, uint32_t n);
size_t rref_large_workspace_bytes(uint32_t m, uint32_t n);
void rref_large_enqueue(const hx_modctx& ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                        int64_t ldr, uint32_t* d_rank, uint32_t* d_piv, uint32_t* d_prow, void* ws, size_t ws_bytes,
                        hipStream_t s);
}  // namespace hx

using namespace hx;

namespace {

__global__ void rref_zero_rank_kernel(uint32_t* d_rank) {
    if (threadIdx.x == 0) *d_rank = 0;
}
void rref_zero_rank_kernel_launch(uint32_t* d_rank, hipStream_t s) {
    rref_zero_rank_kernel<<<1, 64, 0, s>>>(d_rank);
    after_launch("rref_zero_rank_kernel");
}

// One job through the persistent workgroup kernel: R <- A, then R is eliminated in place.
void rref_via_batch(const hx_modctx& ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                    int64_t ldr, uint32_t* rank, uint32_t* pivots, uint32_t* pivot_rows, hipStream_t s) {
    uint32_t* work = nullptr;
    uint32_t* dst = R;
    int64_t ldd = ldr;
    const uint32_t mn = std::min(m, n);
    const size_t small = (size_t)2 * std::max<uint32_t>(mn, 1) + 1;
    if (dst == nullptr) {
        HX_HIP(hipMalloc(&work, ((size_t)m * n + small) * 4));
        dst = work;
        ldd = n;
    } else {
        HX_HIP(hipMalloc(&work, small * 4));
    }
    void* b = nullptr;
    try {
        HX_HIP(hipMemcpy2DAsync(dst, (size_t)ldd * 4, A, (size_t)lda * 4, (size_t)n * 4, m, hipMemcpyDeviceToDevice, s));
        uint32_t* d_piv = dst == work ? work + (size_t)m * n : work;
        uint32_t* d_prow = d_piv + std::max<uint32_t>(mn, 1);
        uint32_t* d_rank = d_prow + std::max<uint32_t>(mn, 1);
        hx_rref_job job{m, n, dst, ldd, d_piv, d_prow, d_rank, 0};
        rref_batch_create_impl(ctx, &job, 1, HX_RREF_BATCH_AUTO, &b);
        rref_batch_run_impl(b, s);
        rref_batch_status_impl(b, s);
        HX_HIP(hipMemcpyAsync(rank, d_rank, 4, hipMemcpyDeviceToHost, s));
        HX_HIP(hipStreamSynchronize(s));
        if (*rank) {
            if (pivots) HX_HIP(hipMemcpyAsync(pivots, d_piv, (size_t)*rank * 4, hipMemcpyDeviceToHost, s));
            if (pivot_rows) HX_HIP(hipMemcpyAsync(pivot_rows, d_prow, (size_t)*rank * 4, hipMemcpyDeviceToHost, s));
            HX_HIP(hipStreamSynchronize(s));
        }
    } catch (...) {
        if (b) rref_batch_destroy_impl(b);
        (void)hipFree(work);
        throw;
    }
    rref_batch_destroy_impl(b);
    HX_HIP(hipFree(work));
}

}  // namespace

extern "C" {

HX_API int hx_rref_u32(const hx_modctx* ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                       int64_t ldr, uint32_t* rank, uint32_t* pivots, uint32_t* pivot_rows, hx_stream s) {
    return guard("hx_rref_u32", [&] {
        HX_CHECK(ctx != nullptr && rank != nullptr, "NULL argument");
        HX_CHECK(ctx->p >= 2 && ctx->p < 0x80000000u, "rref: needs a prime 2 <= p < 2^31");
        HX_CHECK(lda >= (int64_t)n && (R == nullptr || ldr >= (int64_t)n), "rref: leading dimension < n");
        *rank = 0;
        if (m == 0 || n == 0) return;
        HX_CHECK(A != nullptr, "rref: A is NULL");
        // big matrices, and shapes one workgroup cannot hold well (m > 4096, or rows too wide for a
        // 16-column LDS panel, e.g. 5000 x 400, 16 x 40000, 600 x 37000), take the whole-GPU path
        const bool force = !rref_batch_fits(m, n);
        if (rref_large(*ctx, m, n, A, lda, R, ldr, rank, pivots, pivot_rows, (hipStream_t)s, force)) return;
        rref_via_batch(*ctx, m, n, A, lda, R, ldr, rank, pivots, pivot_rows, (hipStream_t)s);
    });
}

HX_API size_t hx_rref_workspace_bytes(uint32_t m, uint32_t n) {
    return (m == 0 || n == 0) ? 0 : rref_large_workspace_bytes(m, n);
}

HX_API int hx_rref_u32_device(const hx_modctx* ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                              int64_t ldr, uint32_t* d_rank, uint32_t* d_pivots, uint32_t* d_pivot_rows, void* workspace,
                              size_t workspace_bytes, hx_stream s) {
    return guard("hx_rref_u32_device", [&] {
        HX_CHECK(ctx != nullptr && d_rank != nullptr, "NULL argument");
        HX_CHECK(ctx->p >= 2 && ctx->p < 0x80000000u, "rref: needs a prime 2 <= p < 2^31");
        HX_CHECK(lda >= (int64_t)n && (R == nullptr || ldr >= (int64_t)n), "rref: leading dimension < n");
        HX_CHECK(m == 0 || n == 0 || A != nullptr, "rref: A is NULL");
        if (m == 0 || n == 0) {
            rref_zero_rank_kernel_launch(d_rank, (hipStream_t)s);
            return;
        }
        rref_large_enqueue(*ctx, m, n, A, lda, R, ldr, d_rank, d_pivots, d_pivot_rows, workspace, workspace_bytes,
                           (hipStream_t)s);
    });
}

HX_API int hx_rref_u32_host(const hx_modctx* ctx, uint32_t m, uint32_t n, hx_row_source src, hx_row_sink sink,
                            void* user, uint32_t slab_rows, uint32_t* rank, uint32_t* pivots, uint32_t* pivot_rows) {
    return guard("hx_rref_u32_host", [&] {
        HX_CHECK(ctx != nullptr && src != nullptr && rank != nullptr, "NULL argument");
        *rank = 0;
        if (m == 0 || n == 0) return;
        if (slab_rows == 0) slab_rows = (uint32_t)std::max<uint64_t>(1, (64ull << 20) / (4ull * n));
        const auto t0 = std::chrono::steady_clock::now();
        uint32_t* dA = nullptr;
        uint32_t* dR = nullptr;
        hipStream_t s;
        HX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            dA = (uint32_t*)arena_alloc((size_t)m * n * 4);
            if (sink) dR = (uint32_t*)arena_alloc((size_t)m * n * 4);
            elim_peak_bytes((size_t)m * n * 4 * (sink ? 2 : 1) + rref_large_workspace_bytes(m, n));
            for (uint32_t r0 = 0; r0 < m; r0 += slab_rows) {
                const uint32_t nr = std::min(slab_rows, m - r0);
                int64_t ld = n;
                const uint32_t* rows = src(user, r0, nr, &ld);
                HX_CHECK(rows != nullptr && ld >= (int64_t)n, "row source returned no rows");
                // through pinned staging; returns once landed (the source view is valid until the next call)
                staged_copy_2d(dA + (size_t)r0 * n, (size_t)n * 4, rows, (size_t)ld * 4, (size_t)n * 4, nr, true, s);
                elim_stat_add(EL_H2D, (uint64_t)nr * n * 4);
            }
            const int rc = hx_rref_u32(ctx, m, n, dA, n, dR, n, rank, pivots, pivot_rows, (hx_stream)s);
            if (rc != HX_OK) throw Error{rc, hx_last_error()};
            if (sink && *rank > 0) {
                int64_t ld = n;
                uint32_t* out = sink(user, *rank, &ld);
                if (out) {
                    HX_CHECK(ld >= (int64_t)n, "row sink: ld < n");
                    staged_copy_2d(out, (size_t)ld * 4, dR, (size_t)n * 4, (size_t)n * 4, *rank, false, s);
                    elim_stat_add(EL_D2H, (uint64_t)*rank * n * 4);
                }
            }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            arena_free(dA);
            arena_free(dR);
            (void)hipStreamDestroy(s);
            throw;
        }
        arena_free(dA);
        arena_free(dR);
        elim_seconds_add(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        HX_HIP(hipStreamDestroy(s));
    });
}

}  // extern "C"
