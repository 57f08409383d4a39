// hx_f2.hip -- linear algebra over F_2 on bit-packed rows (binary::BitMatrix, Binary/BitMatrix.h:49-140):
// the in-place reduced row echelon form with its rank and pivot columns (rref, BitMatrix.h:101-105),
// and the product C = A B (multiply_m4rm, BitMatrix.h:129).
//
// Layout (the reference's): row i of an m x n bit matrix is `ldw` 64-bit words, bit j of the row is
// bit j % 64 of word j / 64; padding bits (columns >= n) are zero and stay zero.
//
// Over F_2 an exact XOR elimination is exact arithmetic and the RREF, its pivot columns (the column
// rank profile) and hence the rank are unique functions of A, so any correct elimination order gives
// the reference's bytes (BitMatrix.h:24-28) -- this one is designed for the GPU:
//
//  * Column panels of 64 (one word).  f2_panel_kernel (one 1024-thread workgroup) holds the panel word
//    of every row in registers (m <= 16384: sixteen rows per thread) and eliminates the panel:
//    per column the first row (lowest index) that is not yet a pivot and has the bit is the pivot;
//    every other row with the bit -- earlier pivot rows included, so the result is reduced -- takes
//    the pivot's word.  Alongside, every row keeps a 64-bit mask over the panel's pivot slots t:
//    row_i = orig_i + sum_{t in mask_i} orig_{P_t} (orig = the rows at the panel's start), so the
//    rest of each row is one product away.  One barrier per column (two LDS slots).
//  * f2_update_kernel: row_i[w] += sum_{t in mask_i} orig_{P_t}[w] for every other word w, as table
//    lookups: the k <= 64 pivot rows' words (copied before any row changes) give, per group of 8
//    slots, a 256-entry LDS table of all XOR combinations of those 8 rows; a row then needs one
//    lookup per byte of its mask.  A workgroup takes 1024 rows x 4 words; HBM-bound (every word of
//    the matrix read and written once per panel).
//  * f2_finish: the pivot rows in pivot order to the top, zero rows below (the non-pivot rows are
//    zero after the last panel: every row is reduced against every pivot).
//  * The product: f2_gemm_kernel runs the same tables over B's rows, 64 at a time, accumulating the
//    output words in registers (one lookup per byte of A's word).
#include <algorithm>
#include <cstring>
#include <vector>

#include "hx_common.h"

namespace hx {
namespace {

constexpr int F2_PT = 1024;              // threads of the panel kernel
constexpr int F2_RMAX = 16;              // rows per thread held in registers (m <= 16384)
constexpr int F2_UT = 256;               // threads of the table kernels
constexpr int F2_WS = 4;                 // words per table-kernel workgroup
constexpr int F2_ROWS = 1024;            // rows per table-kernel workgroup

// Device-side state of one elimination.
struct F2State {
    uint32_t rank;                       // pivots so far
    uint32_t k;                          // this panel's pivots
};

// One panel (word w) of the elimination.  rows i = tid + F2_PT * r.  done[i] = 1 once row i is a
// pivot row.  Writes the new panel words, every row's slot mask, the panel's pivots (appended to
// prow / pcol), k, and copies the k pivot rows (whole, as they were before this panel) to B.
template <int R>
__global__ __launch_bounds__(F2_PT) void f2_panel_kernel(uint64_t* A, uint32_t ldw, uint32_t m, uint32_t w,
                                                         uint64_t vmask, uint8_t* done, uint64_t* cmask,
                                                         uint32_t* prow, uint32_t* pcol, F2State* st, uint64_t* B) {
    __shared__ uint64_t s_pw[2], s_pc[2];
    __shared__ uint32_t s_min[2][F2_PT / 64];
    __shared__ uint32_t s_slot[64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t x[R], cm[R];
    uint32_t dn = 0;                                   // bit r: row tid + F2_PT r is a pivot row
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = tid + F2_PT * r;
        x[r] = i < m ? A[(size_t)i * ldw + w] : 0u;
        cm[r] = 0;
        if (i >= m || done[i]) dn |= 1u << r;
    }
    const uint32_t rank0 = st->rank;
    uint32_t k = 0;
#pragma unroll 1
    for (uint32_t q = 0; q < 64; ++q) {
        if (!((vmask >> q) & 1u)) continue;            // (uniform) a padding column
        const int slot = q & 1;
        // this thread's first candidate row: the smallest r (hence the smallest row index)
        uint32_t key = 0xffffffffu;
#pragma unroll
        for (int r = R - 1; r >= 0; --r)
            if (!((dn >> r) & 1u) && ((x[r] >> q) & 1u)) key = ((uint32_t)r << 10) | tid;
        // block minimum of the keys (r major, then thread: the row index order)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) key = min(key, (uint32_t)__shfl_xor((int)key, off));
        if (lane == 0) s_min[slot][wave] = key;
        __syncthreads();
        uint32_t best = s_min[slot][0];
#pragma unroll
        for (int v = 1; v < F2_PT / 64; ++v) best = min(best, s_min[slot][v]);
        if (best == 0xffffffffu) continue;             // (uniform) no pivot in this column
        const uint32_t pr = best >> 10, pt = best & 1023u;
        // the pivot row's word and mask, published by its thread (register arrays read by selects)
        if (tid == pt) {
            uint64_t pw = 0, pc = 0;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if ((uint32_t)r == pr) {
                    pw = x[r];
                    pc = cm[r];
                }
            s_pw[slot] = pw;
            s_pc[slot] = pc;
            dn |= 1u << pr;
            s_slot[k] = pt + F2_PT * pr;
        }
        __syncthreads();
        const uint64_t pw = s_pw[slot], pc = s_pc[slot] ^ (1ull << k);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool piv = tid == pt && (uint32_t)r == pr;
            if (!piv && ((x[r] >> q) & 1u)) {
                x[r] ^= pw;
                cm[r] ^= pc;
            }
        }
        if (tid == 0) pcol[rank0 + k] = 64 * w + q;
        ++k;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = tid + F2_PT * r;
        if (i < m) {
            A[(size_t)i * ldw + w] = x[r];
            cmask[i] = cm[r];
            if ((dn >> r) & 1u) done[i] = 1;
        }
    }
    if (tid < k) prow[rank0 + tid] = s_slot[tid];
    if (tid == 0) {
        st->k = k;
        st->rank = rank0 + k;
    }
    // the k pivot rows as they were before this panel (the panel word is not used by the update)
    for (uint32_t idx = tid; idx < k * ldw; idx += F2_PT) {
        const uint32_t t = idx / ldw, c = idx % ldw;
        B[(size_t)t * ldw + c] = A[(size_t)s_slot[t] * ldw + c];
    }
}

// The LDS tables of one workgroup: for each group g of 8 rows of Bsrc (rows 8g .. 8g + 7, fewer than
// `kb` counting as zero) and each of the F2_WS words c0 .. c0 + 3: tab[g][s][c] = XOR of the rows
// whose bit is set in s.
__device__ __forceinline__ void f2_build_tables(uint64_t (*tab)[256][F2_WS], const uint64_t* Bsrc, uint32_t ldb,
                                                uint32_t kb, uint32_t c0, uint32_t nw) {
    __shared__ uint64_t rows[64][F2_WS];
    for (uint32_t idx = threadIdx.x; idx < 64 * F2_WS; idx += F2_UT) {
        const uint32_t t = idx / F2_WS, c = idx % F2_WS;
        rows[t][c] = (t < kb && c0 + c < nw) ? Bsrc[(size_t)t * ldb + c0 + c] : 0u;
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < 8 * 256 * F2_WS; idx += F2_UT) {
        const uint32_t c = idx % F2_WS, s = (idx / F2_WS) % 256, g = idx / (F2_WS * 256);
        uint64_t v = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if ((s >> b) & 1u) v ^= rows[8 * g + b][c];
        tab[g][s][c] = v;
    }
    __syncthreads();
}

// row_i[c] ^= sum over the slots in cmask[i] of B[slot][c], for every word c but the panel word w.
// Workgroup: rows [F2_ROWS * blockIdx.y, +F2_ROWS), words [F2_WS * blockIdx.x, +F2_WS).
__global__ __launch_bounds__(F2_UT) void f2_update_kernel(uint64_t* A, uint32_t ldw, uint32_t m, uint32_t w,
                                                          const uint64_t* cmask, const F2State* st, const uint64_t* B) {
    __shared__ uint64_t tab[8][256][F2_WS];
    const uint32_t k = st->k;
    if (k == 0) return;
    const uint32_t c0 = blockIdx.x * F2_WS;
    f2_build_tables(tab, B, ldw, k, c0, ldw);
    const uint32_t wi = threadIdx.x % F2_WS, c = c0 + wi;
    if (c >= ldw || c == w) return;
    const uint32_t r0 = blockIdx.y * F2_ROWS;
    for (uint32_t u = threadIdx.x / F2_WS; u < F2_ROWS; u += F2_UT / F2_WS) {
        const uint32_t i = r0 + u;
        if (i >= m) break;
        const uint64_t msk = cmask[i];
        if (msk == 0) continue;
        uint64_t acc = 0;
#pragma unroll
        for (int g = 0; g < 8; ++g) acc ^= tab[g][(msk >> (8 * g)) & 255u][wi];
        A[(size_t)i * ldw + c] ^= acc;
    }
}

// The RREF's row order: row t = pivot row prow[t] (t < rank), then zero rows.  T holds the pivot rows.
__global__ void f2_gather_kernel(const uint64_t* A, uint32_t ldw, const uint32_t* prow, const F2State* st, uint64_t* T) {
    const uint32_t rank = st->rank;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)rank * ldw;
         idx += (size_t)gridDim.x * blockDim.x)
        T[idx] = A[(size_t)prow[idx / ldw] * ldw + idx % ldw];
}
__global__ void f2_place_kernel(uint64_t* A, uint32_t ldw, uint32_t m, const F2State* st, const uint64_t* T) {
    const uint32_t rank = st->rank;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)m * ldw;
         idx += (size_t)gridDim.x * blockDim.x)
        A[idx] = idx < (size_t)rank * ldw ? T[idx] : 0u;
}

// C[i][c] (op=) sum over the bits b of A's row i of B[b][c]: per 64 rows of B one set of tables.
__global__ __launch_bounds__(F2_UT) void f2_gemm_kernel(const uint64_t* A, uint32_t lda, const uint64_t* B, uint32_t ldb,
                                                        uint64_t* C, uint32_t ldc, uint32_t m, uint32_t kw, uint32_t kbits,
                                                        uint32_t nw, int accumulate) {
    __shared__ uint64_t tab[8][256][F2_WS];
    const uint32_t c0 = blockIdx.x * F2_WS, wi = threadIdx.x % F2_WS, c = c0 + wi;
    const uint32_t r0 = blockIdx.y * F2_ROWS;
    constexpr int PER = F2_ROWS / (F2_UT / F2_WS);      // rows per thread (16)
    uint64_t acc[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[e] = 0;
    for (uint32_t kb = 0; kb < kw; ++kb) {
        const uint32_t rows_left = kbits - 64 * kb;
        f2_build_tables(tab, B + (size_t)64 * kb * ldb, ldb, min(64u, rows_left), c0, nw);
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t i = r0 + threadIdx.x / F2_WS + (F2_UT / F2_WS) * e;
            const uint64_t a = i < m ? A[(size_t)i * lda + kb] : 0u;
            uint64_t v = 0;
#pragma unroll
            for (int g = 0; g < 8; ++g) v ^= tab[g][(a >> (8 * g)) & 255u][wi];
            acc[e] ^= v;
        }
        __syncthreads();                                // the tables are rebuilt for the next 64 rows
    }
    if (c >= nw) return;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const uint32_t i = r0 + threadIdx.x / F2_WS + (F2_UT / F2_WS) * e;
        if (i < m) C[(size_t)i * ldc + c] = accumulate ? (C[(size_t)i * ldc + c] ^ acc[e]) : acc[e];
    }
}

template <int R>
void launch_panel(uint64_t* A, uint32_t ldw, uint32_t m, uint32_t w, uint64_t vmask, uint8_t* done, uint64_t* cmask,
                  uint32_t* prow, uint32_t* pcol, F2State* st, uint64_t* B, hipStream_t s) {
    f2_panel_kernel<R><<<1, F2_PT, 0, s>>>(A, ldw, m, w, vmask, done, cmask, prow, pcol, st, B);
    after_launch("f2_panel_kernel");
}

__global__ void f2_init_kernel(uint8_t* done, uint32_t m, F2State* st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) done[i] = 0;
    if (i == 0) {
        st->rank = 0;
        st->k = 0;
    }
}

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

}  // extern "C"

namespace {
size_t f2_ws_bytes(uint32_t m, uint32_t n, uint32_t ldw) {
    const size_t mn = std::min(m, n);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    return up(m) + up((size_t)m * 8) + up(mn * 4) * 2 + up(sizeof(F2State)) + up((size_t)64 * ldw * 8) +
           up(mn * ldw * 8);
}
}  // namespace

extern "C" {

HX_API size_t hx_f2_rref_workspace_bytes(uint32_t m, uint32_t n, uint32_t ldw) {
    return (m == 0 || n == 0) ? 0 : f2_ws_bytes(m, n, std::max<uint32_t>(ldw, (n + 63) / 64));
}

}  // extern "C"

namespace {
// One pass of the panel elimination over m <= F2_PT * F2_RMAX rows (the panel kernel holds every row's
// panel word in registers).
void f2_rref_one(uint32_t m, uint32_t n, uint64_t* A, uint32_t ldw, uint32_t* rank, uint32_t* pivots, hx_stream s) {
    const hipStream_t st_ = (hipStream_t)s;
    const uint32_t mn = std::min(m, n), nwords = (n + 63) / 64;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bytes = f2_ws_bytes(m, n, ldw);
    char* ws = (char*)arena_alloc(bytes);
    try {
        char* p = ws;
        uint8_t* done = (uint8_t*)p;
        p += up(m);
        uint64_t* cmask = (uint64_t*)p;
        p += up((size_t)m * 8);
        uint32_t* prow = (uint32_t*)p;
        p += up((size_t)mn * 4);
        uint32_t* pcol = (uint32_t*)p;
        p += up((size_t)mn * 4);
        F2State* state = (F2State*)p;
        p += up(sizeof(F2State));
        uint64_t* B = (uint64_t*)p;
        p += up((size_t)64 * ldw * 8);
        uint64_t* T = (uint64_t*)p;
        f2_init_kernel<<<(m + 255) / 256, 256, 0, st_>>>(done, m, state);
        after_launch("f2_init_kernel");
        const int R = m <= 1024 ? 1 : m <= 2048 ? 2 : m <= 4096 ? 4 : m <= 8192 ? 8 : 16;
        const dim3 ugrid((ldw + F2_WS - 1) / F2_WS, (m + F2_ROWS - 1) / F2_ROWS);
        for (uint32_t w = 0; w < nwords; ++w) {
            const uint32_t bits = std::min<uint32_t>(64, n - 64 * w);
            const uint64_t vmask = bits == 64 ? ~0ull : ((1ull << bits) - 1);
            switch (R) {
                case 1: launch_panel<1>(A, ldw, m, w, vmask, done, cmask, prow, pcol, state, B, st_); break;
                case 2: launch_panel<2>(A, ldw, m, w, vmask, done, cmask, prow, pcol, state, B, st_); break;
                case 4: launch_panel<4>(A, ldw, m, w, vmask, done, cmask, prow, pcol, state, B, st_); break;
                case 8: launch_panel<8>(A, ldw, m, w, vmask, done, cmask, prow, pcol, state, B, st_); break;
                default: launch_panel<16>(A, ldw, m, w, vmask, done, cmask, prow, pcol, state, B, st_); break;
            }
            if (ldw > 1) {
                f2_update_kernel<<<ugrid, F2_UT, 0, st_>>>(A, ldw, m, w, cmask, state, B);
                after_launch("f2_update_kernel");
            }
        }
        const uint32_t g = (uint32_t)std::min<size_t>(4096, ((size_t)m * ldw + 255) / 256);
        f2_gather_kernel<<<g, 256, 0, st_>>>(A, ldw, prow, state, T);
        after_launch("f2_gather_kernel");
        f2_place_kernel<<<g, 256, 0, st_>>>(A, ldw, m, state, T);
        after_launch("f2_place_kernel");
        F2State h{};
        host_status(hx_download(&h, state, sizeof(h), s));
        *rank = h.rank;
        if (pivots && h.rank) host_status(hx_download(pivots, pcol, (size_t)h.rank * 4, s));
    } catch (...) {
        (void)hipStreamSynchronize(st_);
        arena_free(ws);
        throw;
    }
    arena_free(ws);
}
}  // namespace

extern "C" {

HX_API int hx_f2_rref(uint32_t m, uint32_t n, uint64_t* A, uint32_t ldw, uint32_t* rank, uint32_t* pivots, hx_stream s) {
    return guard("hx_f2_rref", [&] {
        HX_CHECK(rank != nullptr, "rank is NULL");
        *rank = 0;
        if (m == 0 || n == 0) return;
        HX_CHECK(A != nullptr, "A is NULL");
        HX_CHECK(ldw >= (n + 63) / 64, "ldw < ceil(n / 64)");
        const uint32_t cap = F2_PT * F2_RMAX;
        if (m <= cap) {
            f2_rref_one(m, n, A, ldw, rank, pivots, s);
            return;
        }
        // taller: the RREF of 16384 rows at a time stacked with the next rows (rank < 16384 needed)
        rref_stacked(m, A, ldw, cap, rank, pivots, (hipStream_t)s,
                     [&](uint32_t rows, uint64_t* W, uint32_t* rk, uint32_t* pv) { f2_rref_one(rows, n, W, ldw, rk, pv, s); },
                     "F_2 RREF: the rank reached 16384 with rows left (more than 16384 rows need a rank below 16384)");
    });
}

HX_API int hx_f2_gemm(uint32_t m, uint32_t n, uint32_t k, const uint64_t* A, uint32_t lda, const uint64_t* B,
                      uint32_t ldb, uint64_t* C, uint32_t ldc, int accumulate, hx_stream s) {
    return guard("hx_f2_gemm", [&] {
        const uint32_t kw = (k + 63) / 64, nw = (n + 63) / 64;
        HX_CHECK(m == 0 || n == 0 || C != nullptr, "C is NULL");
        HX_CHECK(lda >= kw && ldb >= nw && ldc >= nw, "leading dimension below the row's words");
        HX_CHECK(k == 0 || (A != nullptr && B != nullptr), "A or B is NULL");
        if (m == 0 || n == 0) return;
        const dim3 grid((nw + F2_WS - 1) / F2_WS, (m + F2_ROWS - 1) / F2_ROWS);
        f2_gemm_kernel<<<grid, F2_UT, 0, (hipStream_t)s>>>(A, lda, B, ldb, C, ldc, m, kw, k, nw, accumulate);
        after_launch("f2_gemm_kernel");
    });
}

}  // extern "C"
