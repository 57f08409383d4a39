// hx_f3.hip -- linear algebra over F_3 on bit-sliced rows (binary::TritMatrix, Binary/TritMatrix.h:42-125):
// the in-place reduced row echelon form with its rank and pivot columns (rref, TritMatrix.h:96-98), and
// the product C = A B (multiply, TritMatrix.h:117).
//
// Layout (the reference's): row i of an m x n matrix over F_3 is two planes of `ldw` 64-bit words, plane
// `one` then plane `two` (row stride 2 ldw words); bit j of plane one is [x_j == 1], bit j of plane two
// [x_j == 2] (one-hot: no column has both bits); padding bits (columns >= n) are zero and stay zero.
//
// The RREF over F_3 (leading 1 in each pivot column, zero above and below, zero rows last), its pivot
// columns and the rank are unique functions of A, so any exact elimination gives the reference's
// planes (TritMatrix.h:26-29); this one is the F_2 route's design (hx_f2.hip) over F_3:
//
//  * Column panels of 64 (one word per plane).  f3_panel_kernel (one 1024-thread workgroup) holds the
//    panel words of every row in registers (m <= 8192) and eliminates the panel column by column: the
//    first row not yet a pivot with a non-zero entry is the pivot, negated when its entry is 2 (so it
//    leads with 1), and every other row with entry e loses e times it.  Alongside, every row keeps its
//    value as a combination of the rows as they were at the panel's start: own_i orig_i + sum_t c_it
//    orig(P_t) with own_i in {0, 1} (0 once the row is the pivot of slot t: its own row is that slot)
//    and c_i in F_3^64, bit-sliced in two masks -- so the rest of each row is one product away.
//  * f3_update_kernel: row_i[w] = own_i orig_i[w] + sum_t c_it B[t][w] on every other word w, as table
//    lookups: per group of 4 slots an LDS table of the 81 combinations of those rows with coefficients
//    in F_3 (for the workgroup's 4 words, both planes), one lookup per group.
//  * f3_gather / f3_place: the pivot rows in pivot order to the top, zero rows below.
//  * The product: f3_gemm_kernel runs the same tables over B's rows, 64 at a time (one word pair of A).
//
// F_3 arithmetic on 64 entries at once (x = (x1, x2) one-hot, z = x + y): z1 is set where (x, y) is
// (1, 0), (0, 1) or (2, 2), z2 where it is (2, 0), (0, 2) or (1, 1); -x swaps the planes.
#include <algorithm>
#include <cstring>
#include <vector>

#include "hx_common.h"

namespace hx {
namespace {

constexpr int F3_PT = 1024;              // threads of the panel kernel
constexpr int F3_RMAX = 8;               // rows per thread held in registers (m <= 8192: sixteen would spill)
constexpr int F3_UT = 256;               // threads of the table kernels
constexpr int F3_WS = 4;                 // words (per plane) per table-kernel workgroup
constexpr int F3_ROWS = 1024;            // rows per table-kernel workgroup
constexpr int F3_G = 16;                 // groups of 4 slots
constexpr int F3_C = 81;                 // combinations per group

struct F3 {
    uint64_t a, b;                       // plane one, plane two
};

__device__ __forceinline__ F3 f3_add(F3 x, F3 y) {
    const uint64_t x0 = ~(x.a | x.b), y0 = ~(y.a | y.b);
    return F3{(x.a & y0) | (x0 & y.a) | (x.b & y.b), (x.b & y0) | (x0 & y.b) | (x.a & y.a)};
}
__device__ __forceinline__ F3 f3_neg(F3 x) { return F3{x.b, x.a}; }
// e x for a scalar e in {0, 1, 2}
__device__ __forceinline__ F3 f3_scale(F3 x, uint32_t e) { return e == 0 ? F3{0, 0} : (e == 1 ? x : f3_neg(x)); }

struct F3State {
    uint32_t rank;                       // pivots so far
    uint32_t k;                          // this panel's pivots
};

// One panel (word w of both planes).  rows i = tid + F3_PT r.  done[i] = 1 once row i is a pivot row.
// Writes the new panel words, every row's slot coefficients (cm: 2 words) and own flag, the panel's
// pivots (appended to prow / pcol), k, and copies the k pivot rows (whole, as they were before this
// panel) to B.
template <int R>
__global__ __launch_bounds__(F3_PT) void f3_panel_kernel(uint64_t* A, uint32_t ldw, uint32_t m, uint32_t w,
                                                         uint64_t vmask, uint8_t* done, uint64_t* cmask, uint8_t* own,
                                                         uint32_t* prow, uint32_t* pcol, F3State* st, uint64_t* B) {
    __shared__ uint64_t s_p[2][4];                     // the pivot row's panel words and coefficients
    __shared__ uint32_t s_min[2][F3_PT / 64];
    __shared__ uint32_t s_slot[64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t rs = 2 * (size_t)ldw;                 // row stride in words
    F3 x[R], c[R];
    uint32_t dn = 0, ow = 0;                           // bit r: row tid + F3_PT r is a pivot row / has own = 1
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = tid + F3_PT * r;
        x[r] = i < m ? F3{A[(size_t)i * rs + w], A[(size_t)i * rs + ldw + w]} : F3{0, 0};
        c[r] = F3{0, 0};
        if (i >= m || done[i]) dn |= 1u << r;
        ow |= 1u << r;
    }
    const uint32_t rank0 = st->rank;
    uint32_t k = 0;
#pragma unroll 1
    for (uint32_t q = 0; q < 64; ++q) {
        if (!((vmask >> q) & 1u)) continue;            // (uniform) a padding column
        const int slot = q & 1;
        uint32_t key = 0xffffffffu;
#pragma unroll
        for (int r = R - 1; r >= 0; --r)
            if (!((dn >> r) & 1u) && (((x[r].a | x[r].b) >> q) & 1u)) key = ((uint32_t)r << 10) | tid;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) key = min(key, (uint32_t)__shfl_xor((int)key, off));
        if (lane == 0) s_min[slot][wave] = key;
        __syncthreads();
        uint32_t best = s_min[slot][0];
#pragma unroll
        for (int v = 1; v < F3_PT / 64; ++v) best = min(best, s_min[slot][v]);
        if (best == 0xffffffffu) continue;             // (uniform) no pivot in this column
        const uint32_t pr = best >> 10, pt = best & 1023u;
        // the pivot thread: its row becomes slot k (its own row moves into the slot), normalised to lead with 1
        if (tid == pt) {
            F3 px{0, 0}, pc{0, 0};
#pragma unroll
            for (int r = 0; r < R; ++r)
                if ((uint32_t)r == pr) {
                    px = x[r];
                    pc = c[r];
                }
            if ((ow >> pr) & 1u) pc.a |= 1ull << k;    // own coefficient 1 -> slot k
            if ((px.b >> q) & 1u) {                    // entry 2: negate the row
                px = f3_neg(px);
                pc = f3_neg(pc);
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                if ((uint32_t)r == pr) {
                    x[r] = px;
                    c[r] = pc;
                }
            s_p[slot][0] = px.a;
            s_p[slot][1] = px.b;
            s_p[slot][2] = pc.a;
            s_p[slot][3] = pc.b;
            dn |= 1u << pr;
            ow &= ~(1u << pr);
            s_slot[k] = pt + F3_PT * pr;
        }
        __syncthreads();
        const F3 P{s_p[slot][0], s_p[slot][1]}, Pc{s_p[slot][2], s_p[slot][3]};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool piv = tid == pt && (uint32_t)r == pr;
            const uint32_t e = (uint32_t)((x[r].a >> q) & 1u) | ((uint32_t)((x[r].b >> q) & 1u) << 1);
            if (!piv && e != 0) {                      // row -= e P: add (3 - e) P
                x[r] = f3_add(x[r], f3_scale(P, 3 - e));
                c[r] = f3_add(c[r], f3_scale(Pc, 3 - e));
            }
        }
        if (tid == 0) pcol[rank0 + k] = 64 * w + q;
        ++k;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = tid + F3_PT * r;
        if (i < m) {
            A[(size_t)i * rs + w] = x[r].a;
            A[(size_t)i * rs + ldw + w] = x[r].b;
            cmask[2 * (size_t)i] = c[r].a;
            cmask[2 * (size_t)i + 1] = c[r].b;
            own[i] = (uint8_t)((ow >> r) & 1u);
            if ((dn >> r) & 1u) done[i] = 1;
        }
    }
    if (tid < k) prow[rank0 + tid] = s_slot[tid];
    if (tid == 0) {
        st->k = k;
        st->rank = rank0 + k;
    }
    // the k pivot rows as they were before this panel, both planes (the panel word is not used by the update)
    for (uint32_t idx = tid; idx < k * 2 * ldw; idx += F3_PT) {
        const uint32_t t = idx / (2 * ldw), cw = idx % (2 * ldw);
        B[(size_t)t * rs + cw] = A[(size_t)s_slot[t] * rs + cw];
    }
}

// The base-3 index of a group of 4 slots' coefficients (one-hot nibbles d1 = [c == 1], d2 = [c == 2]).
__device__ __forceinline__ uint32_t f3_index(uint32_t d1, uint32_t d2) {
    uint32_t idx = 0, p3 = 1;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        idx += (((d1 >> b) & 1u) + 2u * ((d2 >> b) & 1u)) * p3;
        p3 *= 3;
    }
    return idx;
}

// The LDS tables of one workgroup: for each group g of 4 rows of Bsrc (rows 4g .. 4g + 3, those at or past
// `kb` counting as zero) and each of the F3_WS words c0 .. c0 + 3: tab[g][idx][c] = sum_b d_b Bsrc[4g + b][c],
// idx = sum_b d_b 3^b.  Bsrc rows are two planes of ldb words (plane two at + ldb).
__device__ __forceinline__ void f3_build_tables(F3 (*tab)[F3_C][F3_WS], const uint64_t* Bsrc, uint32_t ldb, size_t rsb,
                                                uint32_t kb, uint32_t c0, uint32_t nw) {
    __shared__ F3 rows[64][F3_WS];
    for (uint32_t idx = threadIdx.x; idx < 64 * F3_WS; idx += F3_UT) {
        const uint32_t t = idx / F3_WS, c = idx % F3_WS;
        const bool ok = t < kb && c0 + c < nw;
        rows[t][c] = ok ? F3{Bsrc[(size_t)t * rsb + c0 + c], Bsrc[(size_t)t * rsb + ldb + c0 + c]} : F3{0, 0};
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < F3_G * F3_C * F3_WS; idx += F3_UT) {
        const uint32_t c = idx % F3_WS, s = (idx / F3_WS) % F3_C, g = idx / (F3_WS * F3_C);
        F3 v{0, 0};
        uint32_t r = s;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            v = f3_add(v, f3_scale(rows[4 * g + b][c], r % 3));
            r /= 3;
        }
        tab[g][s][c] = v;
    }
    __syncthreads();
}

// row_i[c] = own_i row_i[c] + sum over the slots of c_it B[t][c], for every word c but the panel word w.
__global__ __launch_bounds__(F3_UT) void f3_update_kernel(uint64_t* A, uint32_t ldw, uint32_t m, uint32_t w,
                                                          const uint64_t* cmask, const uint8_t* own, const F3State* st,
                                                          const uint64_t* B) {
    __shared__ F3 tab[F3_G][F3_C][F3_WS];
    const uint32_t k = st->k;
    if (k == 0) return;
    const size_t rs = 2 * (size_t)ldw;
    const uint32_t c0 = blockIdx.x * F3_WS;
    f3_build_tables(tab, B, ldw, rs, k, c0, ldw);
    const uint32_t wi = threadIdx.x % F3_WS, c = c0 + wi;
    if (c >= ldw || c == w) return;
    const uint32_t r0 = blockIdx.y * F3_ROWS;
    for (uint32_t u = threadIdx.x / F3_WS; u < F3_ROWS; u += F3_UT / F3_WS) {
        const uint32_t i = r0 + u;
        if (i >= m) break;
        const uint64_t m1 = cmask[2 * (size_t)i], m2 = cmask[2 * (size_t)i + 1];
        const bool keep = own[i] != 0;
        if ((m1 | m2) == 0 && keep) continue;
        F3 acc = keep ? F3{A[(size_t)i * rs + c], A[(size_t)i * rs + ldw + c]} : F3{0, 0};
#pragma unroll
        for (int g = 0; g < F3_G; ++g) {
            const uint32_t d1 = (uint32_t)(m1 >> (4 * g)) & 15u, d2 = (uint32_t)(m2 >> (4 * g)) & 15u;
            if ((d1 | d2) == 0) continue;
            acc = f3_add(acc, tab[g][f3_index(d1, d2)][wi]);
        }
        A[(size_t)i * rs + c] = acc.a;
        A[(size_t)i * rs + ldw + c] = acc.b;
    }
}

// The RREF's row order: row t = pivot row prow[t] (t < rank), then zero rows.  T holds the pivot rows.
__global__ void f3_gather_kernel(const uint64_t* A, uint32_t ldw, const uint32_t* prow, const F3State* st, uint64_t* T) {
    const uint32_t rank = st->rank;
    const size_t rs = 2 * (size_t)ldw;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)rank * rs;
         idx += (size_t)gridDim.x * blockDim.x)
        T[idx] = A[(size_t)prow[idx / rs] * rs + idx % rs];
}
__global__ void f3_place_kernel(uint64_t* A, uint32_t ldw, uint32_t m, const F3State* st, const uint64_t* T) {
    const uint32_t rank = st->rank;
    const size_t rs = 2 * (size_t)ldw;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (size_t)m * rs;
         idx += (size_t)gridDim.x * blockDim.x)
        A[idx] = idx < (size_t)rank * rs ? T[idx] : 0u;
}

// C[i][c] (+)= sum over the entries a of A's row i of a B[t][c]: per 64 rows of B one set of tables.
__global__ __launch_bounds__(F3_UT) void f3_gemm_kernel(const uint64_t* A, uint32_t lda, const uint64_t* B, uint32_t ldb,
                                                        uint64_t* C, uint32_t ldc, uint32_t m, uint32_t kw, uint32_t kbits,
                                                        uint32_t nw, int accumulate) {
    __shared__ F3 tab[F3_G][F3_C][F3_WS];
    const uint32_t c0 = blockIdx.x * F3_WS, wi = threadIdx.x % F3_WS, c = c0 + wi;
    const uint32_t r0 = blockIdx.y * F3_ROWS;
    constexpr int PER = F3_ROWS / (F3_UT / F3_WS);      // rows per thread (16)
    F3 acc[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[e] = F3{0, 0};
    for (uint32_t kb = 0; kb < kw; ++kb) {
        const uint32_t rows_left = kbits - 64 * kb;
        f3_build_tables(tab, B + (size_t)64 * kb * 2 * ldb, ldb, 2 * (size_t)ldb, min(64u, rows_left), c0, nw);
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const uint32_t i = r0 + threadIdx.x / F3_WS + (F3_UT / F3_WS) * e;
            const uint64_t a1 = i < m ? A[(size_t)i * 2 * lda + kb] : 0u;
            const uint64_t a2 = i < m ? A[(size_t)i * 2 * lda + lda + kb] : 0u;
            F3 v = acc[e];
#pragma unroll
            for (int g = 0; g < F3_G; ++g) {
                const uint32_t d1 = (uint32_t)(a1 >> (4 * g)) & 15u, d2 = (uint32_t)(a2 >> (4 * g)) & 15u;
                if ((d1 | d2) != 0) v = f3_add(v, tab[g][f3_index(d1, d2)][wi]);
            }
            acc[e] = v;
        }
        __syncthreads();                                // the tables are rebuilt for the next 64 rows
    }
    if (c >= nw) return;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const uint32_t i = r0 + threadIdx.x / F3_WS + (F3_UT / F3_WS) * e;
        if (i >= m) continue;
        uint64_t* Ci = C + (size_t)i * 2 * ldc;
        F3 v = acc[e];
        if (accumulate) v = f3_add(F3{Ci[c], Ci[ldc + c]}, v);
        Ci[c] = v.a;
        Ci[ldc + c] = v.b;
    }
}

template <int R>
void launch_panel(uint64_t* A, uint32_t ldw, uint32_t m, uint32_t w, uint64_t vmask, uint8_t* done, uint64_t* cmask,
                  uint8_t* own, uint32_t* prow, uint32_t* pcol, F3State* st, uint64_t* B, hipStream_t s) {
    f3_panel_kernel<R><<<1, F3_PT, 0, s>>>(A, ldw, m, w, vmask, done, cmask, own, prow, pcol, st, B);
    after_launch("f3_panel_kernel");
}

__global__ void f3_init_kernel(uint8_t* done, uint32_t m, F3State* st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) done[i] = 0;
    if (i == 0) {
        st->rank = 0;
        st->k = 0;
    }
}

size_t f3_ws_bytes(uint32_t m, uint32_t n, uint32_t ldw) {
    const size_t mn = std::min(m, n);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    return up(m) * 2 + up((size_t)m * 16) + up(mn * 4) * 2 + up(sizeof(F3State)) + up((size_t)64 * 2 * ldw * 8) +
           up(mn * 2 * ldw * 8);
}

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API size_t hx_f3_rref_workspace_bytes(uint32_t m, uint32_t n, uint32_t ldw) {
    return (m == 0 || n == 0) ? 0 : f3_ws_bytes(m, n, std::max<uint32_t>(ldw, (n + 63) / 64));
}

}  // extern "C"

namespace {
// One pass of the panel elimination over m <= F3_PT * F3_RMAX rows (the panel kernel holds every row's
// panel words in registers).
void f3_rref_one(uint32_t m, uint32_t n, uint64_t* A, uint32_t ldw, uint32_t* rank, uint32_t* pivots, hx_stream s) {
    const hipStream_t st_ = (hipStream_t)s;
    const uint32_t mn = std::min(m, n), nwords = (n + 63) / 64;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bytes = f3_ws_bytes(m, n, ldw);
    char* ws = (char*)arena_alloc(bytes);
    try {
        char* p = ws;
        uint8_t* done = (uint8_t*)p;
        p += up(m);
        uint8_t* own = (uint8_t*)p;
        p += up(m);
        uint64_t* cmask = (uint64_t*)p;
        p += up((size_t)m * 16);
        uint32_t* prow = (uint32_t*)p;
        p += up((size_t)mn * 4);
        uint32_t* pcol = (uint32_t*)p;
        p += up((size_t)mn * 4);
        F3State* state = (F3State*)p;
        p += up(sizeof(F3State));
        uint64_t* B = (uint64_t*)p;
        p += up((size_t)64 * 2 * ldw * 8);
        uint64_t* T = (uint64_t*)p;
        f3_init_kernel<<<(m + 255) / 256, 256, 0, st_>>>(done, m, state);
        after_launch("f3_init_kernel");
        const int R = m <= 1024 ? 1 : m <= 2048 ? 2 : m <= 4096 ? 4 : 8;
        const dim3 ugrid((ldw + F3_WS - 1) / F3_WS, (m + F3_ROWS - 1) / F3_ROWS);
        for (uint32_t w = 0; w < nwords; ++w) {
            const uint32_t bits = std::min<uint32_t>(64, n - 64 * w);
            const uint64_t vmask = bits == 64 ? ~0ull : ((1ull << bits) - 1);
            switch (R) {
                case 1: launch_panel<1>(A, ldw, m, w, vmask, done, cmask, own, prow, pcol, state, B, st_); break;
                case 2: launch_panel<2>(A, ldw, m, w, vmask, done, cmask, own, prow, pcol, state, B, st_); break;
                case 4: launch_panel<4>(A, ldw, m, w, vmask, done, cmask, own, prow, pcol, state, B, st_); break;
                default: launch_panel<8>(A, ldw, m, w, vmask, done, cmask, own, prow, pcol, state, B, st_); break;
            }
            if (ldw > 1) {
                f3_update_kernel<<<ugrid, F3_UT, 0, st_>>>(A, ldw, m, w, cmask, own, state, B);
                after_launch("f3_update_kernel");
            }
        }
        const uint32_t g = (uint32_t)std::min<size_t>(4096, ((size_t)m * 2 * ldw + 255) / 256);
        f3_gather_kernel<<<g, 256, 0, st_>>>(A, ldw, prow, state, T);
        after_launch("f3_gather_kernel");
        f3_place_kernel<<<g, 256, 0, st_>>>(A, ldw, m, state, T);
        after_launch("f3_place_kernel");
        F3State h{};
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

HX_API int hx_f3_rref(uint32_t m, uint32_t n, uint64_t* A, uint32_t ldw, uint32_t* rank, uint32_t* pivots, hx_stream s) {
    return guard("hx_f3_rref", [&] {
        HX_CHECK(rank != nullptr, "rank is NULL");
        *rank = 0;
        if (m == 0 || n == 0) return;
        HX_CHECK(A != nullptr, "A is NULL");
        HX_CHECK(ldw >= (n + 63) / 64, "ldw < ceil(n / 64)");
        const uint32_t cap = F3_PT * F3_RMAX;
        if (m <= cap) {
            f3_rref_one(m, n, A, ldw, rank, pivots, s);
            return;
        }
        // taller: the RREF of 8192 rows at a time stacked with the next rows (rank < 8192 needed)
        rref_stacked(m, A, 2 * (size_t)ldw, cap, rank, pivots, (hipStream_t)s,
                     [&](uint32_t rows, uint64_t* W, uint32_t* rk, uint32_t* pv) { f3_rref_one(rows, n, W, ldw, rk, pv, s); },
                     "F_3 RREF: the rank reached 8192 with rows left (more than 8192 rows need a rank below 8192)");
    });
}

HX_API int hx_f3_gemm(uint32_t m, uint32_t n, uint32_t k, const uint64_t* A, uint32_t lda, const uint64_t* B,
                      uint32_t ldb, uint64_t* C, uint32_t ldc, int accumulate, hx_stream s) {
    return guard("hx_f3_gemm", [&] {
        const uint32_t kw = (k + 63) / 64, nw = (n + 63) / 64;
        HX_CHECK(m == 0 || n == 0 || C != nullptr, "C is NULL");
        HX_CHECK(lda >= kw && ldb >= nw && ldc >= nw, "words per plane below the row's words");
        HX_CHECK(k == 0 || (A != nullptr && B != nullptr), "A or B is NULL");
        if (m == 0 || n == 0) return;
        const dim3 grid((nw + F3_WS - 1) / F3_WS, (m + F3_ROWS - 1) / F3_ROWS);
        f3_gemm_kernel<<<grid, F3_UT, 0, (hipStream_t)s>>>(A, lda, B, ldb, C, ldc, m, kw, k, nw, accumulate);
        after_launch("f3_gemm_kernel");
    });
}

}  // extern "C"
