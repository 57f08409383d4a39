// hx_ring2k.hip -- the device side of the Z/2^32 / Z/2^64 LU session (Cuda/CudaRing2k.h:28-46): a
// device-resident column window [0, cw) of an m x cw matrix, driven panel by panel by a host that does
// the pivot search (Binary/Ring2kDeviceLU.h).  Per step the device
//   * writes the host-factored panel (rows r .. m-1, columns [pc0, pc1)) into the window,
//   * replays the panel's row swaps (pairs, in order) on every other window column,
//   * runs the unit-lower TRSM on the np pivot rows and the trailing C -= L U over columns [u0, w),
//   * and, on request, copies the next panel's columns (rows r + np .. m-1 of [la0, la1)) back to
//     pinned host memory on a second stream (lookahead; collected by fetch()).
// All arithmetic is native wraparound in u32 / u64 (Z/2^k is the quotient, Ring2k.h:17-22): exact, so
// the values equal the host kernels' bit for bit.
#include <cstring>

#include "hx_common.h"

namespace hx {
namespace {

template <class E>
__global__ void r2_swap_kernel(E* A, size_t lda, uint32_t cw, uint32_t pc0, uint32_t pc1, const uint32_t* sw,
                               uint32_t nswaps) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cw || (j >= pc0 && j < pc1)) return;
    for (uint32_t q = 0; q < nswaps; ++q) {
        E* a = A + (size_t)sw[2 * q] * lda + j;
        E* b = A + (size_t)sw[2 * q + 1] * lda + j;
        const E x = *a;
        *a = *b;
        *b = x;
    }
}

// Unit-lower TRSM on the pivot rows, one column per thread, rows in order.
template <class E>
__global__ void r2_trsm_kernel(E* A, size_t lda, const E* L, size_t ldl, uint32_t np, uint32_t c0, uint32_t c1) {
    const uint32_t j = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= c1) return;
    for (uint32_t t = 1; t < np; ++t) {
        E x = A[(size_t)t * lda + j];
        const E* lt = L + (size_t)t * ldl;
        for (uint32_t t2 = 0; t2 < t; ++t2) x -= lt[t2] * A[(size_t)t2 * lda + j];
        A[(size_t)t * lda + j] = x;
    }
}

// C (M x N) -= L (M x K) U (K x N), wraparound; 64 x 64 tiles, 256 threads (4 x 4 each), K staged by 16.
template <class E>
__global__ __launch_bounds__(256) void r2_gemm_sub_kernel(E* C, size_t ldc, const E* L, size_t ldl, const E* U,
                                                         size_t ldu, uint32_t M, uint32_t N, uint32_t K) {
    __shared__ E Ls[16][65];
    __shared__ E Us[16][65];
    const uint32_t n0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    E acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0;
    for (uint32_t k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = tid + 256 * t;
            const uint32_t r = e / 16, c = e % 16;
            Ls[c][r] = (m0 + r < M && k0 + c < K) ? L[(size_t)(m0 + r) * ldl + k0 + c] : E(0);
            const uint32_t ur = e / 64, uc = e % 64;
            Us[ur][uc] = (k0 + ur < K && n0 + uc < N) ? U[(size_t)(k0 + ur) * ldu + n0 + uc] : E(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            E a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = Ls[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Us[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t r = m0 + ty + 16 * i;
        if (r >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = n0 + tx + 16 * j;
            if (c < N) C[(size_t)r * ldc + c] -= acc[i][j];
        }
    }
}

template <class E>
void gemm_sub(E* C, size_t ldc, const E* L, size_t ldl, const E* U, size_t ldu, uint32_t M, uint32_t N, uint32_t K,
              hipStream_t s) {
    if (M == 0 || N == 0 || K == 0) return;
    for (uint32_t m0 = 0; m0 < M; m0 += 65535u * 64) {
        const uint32_t mm = std::min<uint32_t>(M - m0, 65535u * 64);
        r2_gemm_sub_kernel<E><<<dim3((N + 63) / 64, (mm + 63) / 64), 256, 0, s>>>(C + (size_t)m0 * ldc, ldc,
                                                                                 L + (size_t)m0 * ldl, ldl, U, ldu, mm,
                                                                                 N, K);
        after_launch("r2_gemm_sub_kernel");
    }
}

struct Session {
    uint32_t elem = 0, m = 0, cw = 0, panel = 0;
    char* dA = nullptr;
    char* dL = nullptr;
    uint32_t* dSw = nullptr;
    char* hStage = nullptr;     // panel / L / lookahead staging (pinned), m x panel elements each
    char* hL = nullptr;
    char* hNext = nullptr;
    uint32_t* hSw = nullptr;
    size_t slab = 0;
    hipStream_t s1 = nullptr, s2 = nullptr;
    hipEvent_t ev_la = nullptr, ev_stage = nullptr;
    bool stage_pending = false, la_pending = false;
    uint32_t la_rows = 0, la_cols = 0;
    char* at(uint32_t r, uint32_t c) const { return dA + ((size_t)r * cw + c) * elem; }
    void release() noexcept {
        if (s1) (void)hipStreamSynchronize(s1);
        if (s2) (void)hipStreamSynchronize(s2);
        (void)hipFree(dA);
        (void)hipFree(dL);
        (void)hipFree(dSw);
        (void)hipHostFree(hStage);
        (void)hipHostFree(hL);
        (void)hipHostFree(hNext);
        (void)hipHostFree(hSw);
        if (ev_la) (void)hipEventDestroy(ev_la);
        if (ev_stage) (void)hipEventDestroy(ev_stage);
        if (s1) (void)hipStreamDestroy(s1);
        if (s2) (void)hipStreamDestroy(s2);
        (void)hipGetLastError();
    }
};

void copy_rows(char* dst, size_t ldd_bytes, const char* src, size_t lds_bytes, size_t row_bytes, uint32_t rows) {
    if (ldd_bytes == row_bytes && lds_bytes == row_bytes) {
        std::memcpy(dst, src, row_bytes * rows);
        return;
    }
    for (uint32_t i = 0; i < rows; ++i) std::memcpy(dst + ldd_bytes * i, src + lds_bytes * i, row_bytes);
}

template <class E>
void step_impl(Session& S, const hx_ring2k_step_args& a) {
    const uint32_t rows = a.m - a.r, rn = a.r + a.np, W = a.w ? a.w : S.cw;
    const size_t eb = sizeof(E);
    if (S.stage_pending) {   // the previous step's staging buffers are free once its copies ran
        HX_HIP(hipEventSynchronize(S.ev_stage));
        S.stage_pending = false;
    }
    if (a.panel != nullptr && rows > 0) {
        const uint32_t pw = a.pc1 - a.pc0;
        copy_rows(S.hStage, pw * eb, (const char*)a.panel, a.ldp * eb, pw * eb, rows);
        HX_HIP(hipMemcpy2DAsync(S.at(a.r, a.pc0), S.cw * eb, S.hStage, pw * eb, pw * eb, rows, hipMemcpyHostToDevice, S.s1));
    }
    if (a.nswaps > 0) {
        std::memcpy(S.hSw, a.swaps, sizeof(uint32_t) * 2 * a.nswaps);
        HX_HIP(hipMemcpyAsync(S.dSw, S.hSw, sizeof(uint32_t) * 2 * a.nswaps, hipMemcpyHostToDevice, S.s1));
        r2_swap_kernel<E><<<(W + 255) / 256, 256, 0, S.s1>>>((E*)S.dA, S.cw, W, a.pc0, a.pc1, S.dSw, a.nswaps);
        after_launch("r2_swap_kernel");
    }
    const bool update = a.np > 0 && a.u0 < W;
    if (update) {
        copy_rows(S.hL, a.np * eb, (const char*)a.L, a.ldl * eb, a.np * eb, rows);
        HX_HIP(hipMemcpy2DAsync(S.dL, S.panel * eb, S.hL, a.np * eb, a.np * eb, rows, hipMemcpyHostToDevice, S.s1));
    }
    HX_HIP(hipEventRecord(S.ev_stage, S.s1));
    S.stage_pending = true;
    E* A = (E*)S.dA;
    const E* Ld = (const E*)S.dL;
    const bool la = a.la1 > a.la0 && rn < a.m;
    auto gemm_cols = [&](uint32_t c0, uint32_t c1) {
        if (c1 > c0 && rn < a.m)
            gemm_sub<E>(A + (size_t)rn * S.cw + c0, S.cw, Ld + (size_t)a.np * S.panel, S.panel,
                        A + (size_t)a.r * S.cw + c0, S.cw, a.m - rn, c1 - c0, a.np, S.s1);
    };
    if (update) {
        if (a.np >= 2) {
            r2_trsm_kernel<E><<<(W - a.u0 + 255) / 256, 256, 0, S.s1>>>(A + (size_t)a.r * S.cw, S.cw, Ld, S.panel, a.np,
                                                                        a.u0, W);
            after_launch("r2_trsm_kernel");
        }
        gemm_cols(a.u0, (la && a.la0 == a.u0) ? a.la1 : W);   // the lookahead columns first
    }
    if (la) {
        HX_HIP(hipEventRecord(S.ev_la, S.s1));
        HX_HIP(hipStreamWaitEvent(S.s2, S.ev_la, 0));
        const uint32_t lw = a.la1 - a.la0, lr = a.m - rn;
        HX_HIP(hipMemcpy2DAsync(S.hNext, lw * eb, S.at(rn, a.la0), S.cw * eb, lw * eb, lr, hipMemcpyDeviceToHost, S.s2));
        S.la_pending = true;
        S.la_rows = lr;
        S.la_cols = lw;
        if (update && a.la0 == a.u0) gemm_cols(a.la1, W);
    }
}

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API size_t hx_ring2k_session_bytes(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) {
    return (size_t)m * cw * elem + (size_t)m * panel * elem + sizeof(uint32_t) * 2 * (panel + 1);
}
HX_API size_t hx_ring2k_staging_bytes(uint32_t elem, uint32_t m, uint32_t panel) {
    return 3 * (size_t)m * panel * elem + sizeof(uint32_t) * 2 * (panel + 1);
}

HX_API int hx_ring2k_open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel, hx_ring2k* out) {
    return guard("hx_ring2k_open", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        HX_CHECK(elem == 4 || elem == 8, "ring2k session: elem must be 4 or 8");
        HX_CHECK(m > 0 && cw > 0 && panel > 0, "ring2k session: empty session");
        Session* S = new Session();
        try {
            S->elem = elem;
            S->m = m;
            S->cw = cw;
            S->panel = panel;
            S->slab = (size_t)m * panel * elem;
            HX_HIP(hipStreamCreateWithFlags(&S->s1, hipStreamNonBlocking));
            HX_HIP(hipStreamCreateWithFlags(&S->s2, hipStreamNonBlocking));
            HX_HIP(hipEventCreateWithFlags(&S->ev_la, hipEventDisableTiming));
            HX_HIP(hipEventCreateWithFlags(&S->ev_stage, hipEventDisableTiming));
            HX_HIP(hipMalloc(&S->dA, (size_t)m * cw * elem));
            HX_HIP(hipMalloc(&S->dL, S->slab));
            HX_HIP(hipMalloc(&S->dSw, sizeof(uint32_t) * 2 * (panel + 1)));
            HX_HIP(hipHostMalloc(&S->hStage, S->slab, hipHostMallocDefault));
            HX_HIP(hipHostMalloc(&S->hL, S->slab, hipHostMallocDefault));
            HX_HIP(hipHostMalloc(&S->hNext, S->slab, hipHostMallocDefault));
            HX_HIP(hipHostMalloc(&S->hSw, sizeof(uint32_t) * 2 * (panel + 1), hipHostMallocDefault));
        } catch (...) {
            S->release();
            delete S;
            throw;
        }
        *out = S;
    });
}

HX_API int hx_ring2k_put(hx_ring2k s, const void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0,
                         uint32_t cols) {
    return guard("hx_ring2k_put", [&] {
        HX_CHECK(s != nullptr, "NULL session");
        Session& S = *(Session*)s;
        HX_CHECK((uint64_t)r0 + rows <= S.m && (uint64_t)c0 + cols <= S.cw, "ring2k put: block outside the window");
        if (rows == 0 || cols == 0) return;
        const size_t eb = S.elem;
        HX_HIP(hipMemcpy2DAsync(S.at(r0, c0), S.cw * eb, host, ldh * eb, cols * eb, rows, hipMemcpyHostToDevice, S.s1));
        HX_HIP(hipStreamSynchronize(S.s1));
    });
}

HX_API int hx_ring2k_get(hx_ring2k s, void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols) {
    return guard("hx_ring2k_get", [&] {
        HX_CHECK(s != nullptr, "NULL session");
        Session& S = *(Session*)s;
        HX_CHECK((uint64_t)r0 + rows <= S.m && (uint64_t)c0 + cols <= S.cw, "ring2k get: block outside the window");
        HX_HIP(hipStreamSynchronize(S.s2));
        S.la_pending = false;
        if (rows == 0 || cols == 0) {
            HX_HIP(hipStreamSynchronize(S.s1));
            return;
        }
        const size_t eb = S.elem;
        HX_HIP(hipMemcpy2DAsync(host, ldh * eb, S.at(r0, c0), S.cw * eb, cols * eb, rows, hipMemcpyDeviceToHost, S.s1));
        HX_HIP(hipStreamSynchronize(S.s1));
    });
}

HX_API int hx_ring2k_step(hx_ring2k s, const hx_ring2k_step_args* a) {
    return guard("hx_ring2k_step", [&] {
        HX_CHECK(s != nullptr && a != nullptr, "NULL argument");
        Session& S = *(Session*)s;
        HX_CHECK(!S.la_pending, "ring2k step: the previous lookahead was not fetched");
        HX_CHECK(a->m == S.m && a->r <= a->m && a->np <= S.panel && a->nswaps <= S.panel && a->pc0 <= a->pc1,
                 "ring2k step: arguments outside the session");
        HX_CHECK(a->pc1 - a->pc0 <= S.panel && (a->panel == nullptr || a->pc1 <= S.cw), "ring2k step: bad panel");
        HX_CHECK(a->w <= S.cw && a->la1 <= (a->w ? a->w : S.cw) && (a->la1 <= a->la0 || a->la1 - a->la0 <= S.panel),
                 "ring2k step: bad window / lookahead");
        for (uint32_t q = 0; q < 2 * a->nswaps; ++q) HX_CHECK(a->swaps[q] < S.m, "ring2k step: swap row outside");
        if (S.elem == 4)
            step_impl<uint32_t>(S, *a);
        else
            step_impl<uint64_t>(S, *a);
    });
}

HX_API int hx_ring2k_fetch(hx_ring2k s, void* host, size_t ldh) {
    return guard("hx_ring2k_fetch", [&] {
        HX_CHECK(s != nullptr, "NULL session");
        Session& S = *(Session*)s;
        HX_CHECK(S.la_pending, "ring2k fetch: no lookahead pending");
        HX_HIP(hipStreamSynchronize(S.s2));
        S.la_pending = false;
        copy_rows((char*)host, ldh * S.elem, S.hNext, (size_t)S.la_cols * S.elem, (size_t)S.la_cols * S.elem, S.la_rows);
    });
}

HX_API int hx_ring2k_close(hx_ring2k s) {
    Session* S = (Session*)s;
    if (!S) return HX_OK;
    S->release();
    delete S;
    return HX_OK;
}

HX_API int hx_ring2k_gemm_sub_host(uint32_t elem, void* C, size_t ldc, const void* L, size_t ldl, const void* U,
                                   size_t ldu, uint32_t m, uint32_t n, uint32_t k) {
    return guard("hx_ring2k_gemm_sub_host", [&] {
        HX_CHECK(elem == 4 || elem == 8, "elem must be 4 or 8");
        if (m == 0 || n == 0) return;
        const size_t eb = elem;
        char *dC = nullptr, *dL = nullptr, *dU = nullptr;
        hipStream_t st = nullptr;
        try {
            HX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            HX_HIP(hipMalloc(&dC, (size_t)m * n * eb));
            HX_HIP(hipMalloc(&dL, (size_t)m * std::max<uint32_t>(k, 1) * eb));
            HX_HIP(hipMalloc(&dU, (size_t)std::max<uint32_t>(k, 1) * n * eb));
            HX_HIP(hipMemcpy2DAsync(dC, n * eb, C, ldc * eb, n * eb, m, hipMemcpyHostToDevice, st));
            if (k) {
                HX_HIP(hipMemcpy2DAsync(dL, k * eb, L, ldl * eb, k * eb, m, hipMemcpyHostToDevice, st));
                HX_HIP(hipMemcpy2DAsync(dU, n * eb, U, ldu * eb, n * eb, k, hipMemcpyHostToDevice, st));
            }
            if (elem == 4)
                gemm_sub<uint32_t>((uint32_t*)dC, n, (const uint32_t*)dL, k, (const uint32_t*)dU, n, m, n, k, st);
            else
                gemm_sub<uint64_t>((uint64_t*)dC, n, (const uint64_t*)dL, k, (const uint64_t*)dU, n, m, n, k, st);
            HX_HIP(hipMemcpy2DAsync(C, ldc * eb, dC, n * eb, n * eb, m, hipMemcpyDeviceToHost, st));
            HX_HIP(hipStreamSynchronize(st));
        } catch (...) {
            (void)hipFree(dC);
            (void)hipFree(dL);
            (void)hipFree(dU);
            if (st) (void)hipStreamDestroy(st);
            throw;
        }
        HX_HIP(hipFree(dC));
        HX_HIP(hipFree(dL));
        HX_HIP(hipFree(dU));
        HX_HIP(hipStreamDestroy(st));
    });
}

}  // extern "C"
