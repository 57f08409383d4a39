// hx_rref_large.hip -- RREF over F_p (p < 2^31 prime) of one large matrix, spread over the whole GPU.
//
// Right-looking elimination in column panels of W = 32 columns over a compacted Schur complement S
// (m_a x n_a, rows in original order: the i-th row of S is original row rows_orig[i]):
//
//  1. Panel pivots by a TOURNAMENT.  The pivot rule -- per column, the unused row of smallest
//     original index with a non-zero Schur entry (Cuda/CudaKernels.h:255-265) -- picks, after column
//     c, the row rank profile RRP(S[:, 0..c]).  A row in RRP of a stack lies in the RRP of its own
//     block (a dependency on the rows above it inside the block is one in the stack), and the span
//     of all rows above a candidate equals the span of the candidates above it, so
//     RRP(S[:, 0..c]) = RRP(candidates[:, 0..c]) for every c: the pivots AND their assignment to
//     columns are those of the candidate rows alone.  Level 0: one workgroup per 256-row chunk
//     finds its chunk's candidates (division-free elimination in LDS, four threads per row); merge
//     levels (by the last workgroup to finish): groups of 16 candidate lists, the rule re-run on
//     their raw rows (two threads per row); the last level writes the panel's pivot rows P (pivot
//     order), columns J and count k.  Before all that, every workgroup runs the rule on the first
//     64 rows alone, which settles most panels without the tournament (see rrp_pivot_kernel).
//  2. X = S[P, J]^{-1} S[P, :]: every workgroup owns a slab of 16 columns and performs the same
//     row reduction on [S[P, J] | S[P, slab]] (division-free Gauss-Jordan, one barrier per pivot,
//     then one inverse per pivot) -- or, when the search over the first rows settled the panel,
//     takes W = S[P, J]^{-1} from that search and computes X[:, slab] = W S[P, slab] -- writing
//     X[:, slab] into the U store (row r + t, original columns), into a contiguous copy and into
//     X's limb planes for the products of step 3.  X has the identity at J.
//  3. The rows R not in P, in order, form the next Schur complement:
//        S'[i, :] = S[R_i, W:] - S[R_i, J] X[:, W:]
//     as k <= 32 products on the INT8 matrix cores (upd32_kernel: A's limbs split on the fly, X's
//     from its limb planes) with a row gather on A and C_in, a column gather on A, and output into
//     the other S buffer (compaction for free).  Lookahead: the next panel's W columns are computed first on
//     the caller's stream, the rest on a second stream, overlapping the next panel's tournament
//     (which reads only those W columns); the next panel's step 2 waits for it.  P, J and the
//     update's sizes are double-buffered by panel parity for the overlap.
//  4. Gauss-Jordan: the pivot rows stored so far (U rows 0 .. r-1) are reduced against the panel's
//     pivots, U[0:r, c:n] -= U[0:r, c + J] X, on the second stream after the rest of the update
//     (their panel columns gathered first, since the product overwrites them); X is exactly the
//     identity at J, so those entries become zero.  When the last panel is done the U rows are the
//     RREF's rows.  The X copy is double-buffered by panel parity, so the next panel's X does not
//     overwrite the one this product reads.
//  All sizes (m_a, k, r) live in device memory; the host only knows upper bounds, so a panel is a
//  fixed sequence of launches.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "hx_elim.h"
#include "hx_gemm.h"
#include "hx_limbs.h"

namespace hx {

namespace {

constexpr int LW = 32;        // panel width
constexpr int CH = 256;       // rows per level-0 chunk
constexpr int GRP = 16;       // candidate lists merged per workgroup
constexpr int LT = 256;       // threads

struct LState {               // device-side counters
    uint32_t m_a, r, k, m_new;
    int status;
    uint32_t snap_k[2], snap_m[2];   // the sizes of panel i's trailing update, slot i & 1
    uint32_t snap_r[2];              // the pivot rows stored before panel i (its Gauss-Jordan rows)
    uint32_t spec;                   // 1: this panel's pivots came from its first SPR rows alone
    uint32_t arrive;                 // blocks of the pivot search done (the last one merges when needed)
    uint32_t spec_hits;              // panels the spec search settled (a diagnostic counter)
};
static_assert(sizeof(LState) <= 16 * 4, "the state fits its 16 workspace words");

// The smallest-row rule on an LDS panel pan[u * ws + q] (rows in original order), division-free.
// TPR threads per row: thread (u, part) owns the columns q == part (mod TPR) of row u, so a
// column's elimination is spread over all TPR * rows threads.  The owner of column q scans it (it
// wrote it last); one barrier per column.  Returns k; s_u[t] / s_q[t] = local row / column of
// pivot t; chosen[u] = 1 for the rows picked.
constexpr int RT = 1024;      // threads of the pivot-search kernel

template <int TPR, class Ar>
__device__ uint32_t lds_rule(uint32_t* pan, uint32_t rows, uint32_t wc, uint32_t ws, uint8_t* chosen, const Ar& ar,
                             uint32_t* s_u, uint32_t* s_q) {
    static_assert((TPR & (TPR - 1)) == 0, "threads per row: a power of two");
    __shared__ uint32_t s_cand[2][RT / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, p = ar.p;
    const uint32_t u = tid / TPR, part = tid % TPR;
    const bool mine = u < rows;
    uint32_t* row = pan + u * ws;
    bool done = false;
    uint32_t k = 0;
    __syncthreads();                                   // the panel is in LDS
    if (rows == 0) return 0;                           // (uniform) no candidates: no column has a pivot
    for (uint32_t q = 0; q < wc; ++q) {
        // rows grow with the thread index, so the wave's smallest candidate is its lowest hit lane
        const uint64_t bal = __ballot(mine && part == (q % TPR) && !done && row[q] != 0);
        const uint32_t best = bal ? (wave * 64 + (uint32_t)__builtin_ctzll(bal)) / TPR : 0xffffffffu;
        if (lane == 0) s_cand[q & 1][wave] = best;
        __syncthreads();
        uint32_t bu = s_cand[q & 1][0];
#pragma unroll
        for (int w = 1; w < RT / 64; ++w) bu = min(bu, s_cand[q & 1][w]);
        if (bu == 0xffffffffu) continue;
        if (mine && !done) {
            if (u == bu) {
                done = true;
            } else {
                const uint32_t f = row[q];
                if (f != 0) {
                    const uint32_t* prow = pan + bu * ws;
                    const uint32_t v = prow[q], fneg = p - f;
                    // this thread's columns past q, four at a time with the reads first (row and
                    // prow share the LDS array: a store before the next load costs a round trip)
                    uint32_t q2 = q + 1 + ((part - (q + 1)) & (TPR - 1));
                    for (; q2 + 3 * TPR < wc; q2 += 4 * TPR) {
                        uint32_t x[4], y[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            x[e] = row[q2 + e * TPR];
                            y[e] = prow[q2 + e * TPR];
                        }
#pragma unroll
                        for (int e = 0; e < 4; ++e) row[q2 + e * TPR] = ar.dfree(v, x[e], fneg, y[e]);
                    }
                    for (; q2 < wc; q2 += TPR) row[q2] = ar.dfree(v, row[q2], fneg, prow[q2]);
                }
            }
        }
        if (tid == 0) { s_u[k] = bu; s_q[k] = q; }
        ++k;
    }
    if (mine && part == 0) chosen[u] = done ? 1 : 0;
    __syncthreads();
    return k;
}

// The rows u < rows with flag[u] set, ascending: emit(position, u) for each, positions 0, 1, ...
// A ballot per wave and the waves' counts in LDS (one barrier), instead of a one-thread loop.
template <class F>
__device__ __forceinline__ void block_compact(const uint8_t* flag, uint32_t rows, uint32_t* s_wcnt, F emit) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool f = tid < rows && flag[tid] != 0;
    const uint64_t bal = __ballot(f);
    if (lane == 0) s_wcnt[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w = 0; w < wave; ++w) off += s_wcnt[w];
    if (f) emit(off + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull)), tid);
}

// R = the rows of S not among the panel's k pivot rows sP[0..k) (any order), in order, and their
// original indices (the next Schur complement's rows_orig), by NT threads (thread ids 0 .. NT-1):
// row i goes to position i - #{t : sP[t] < i}.  Called by whichever block settles the panel.
template <int NT>
__device__ void rest_rows_of(uint32_t ma, uint32_t k, const uint32_t* sP, uint32_t* R, const uint32_t* rows_orig,
                             uint32_t* rows_orig_next) {
    uint32_t pv[LW];
#pragma unroll
    for (int t = 0; t < LW; ++t) pv[t] = (uint32_t)t < k ? sP[t] : 0xffffffffu;
    for (uint32_t i = threadIdx.x; i < ma; i += NT) {
        uint32_t below = 0;
        bool pivot = false;
#pragma unroll
        for (int t = 0; t < LW; ++t) {
            below += pv[t] < i ? 1u : 0u;
            pivot |= pv[t] == i;
        }
        if (!pivot) {
            R[i - below] = i;
            rows_orig_next[i - below] = rows_orig[i];
        }
    }
}

// The same for the rows r0 .. r0 + CH - 1 of S only, one row per thread (each block of the pivot
// search takes its own chunk, so the list is built by all of them at once).
__device__ __forceinline__ void rest_rows_chunk(uint32_t ma, uint32_t k, const uint32_t* sP, uint32_t* R,
                                                const uint32_t* rows_orig, uint32_t* rows_orig_next, uint32_t r0) {
    const uint32_t i = r0 + threadIdx.x;
    if (threadIdx.x >= CH || i >= ma) return;
    uint32_t below = 0;
    bool pivot = false;
#pragma unroll
    for (int t = 0; t < LW; ++t) {
        const uint32_t pv = (uint32_t)t < k ? sP[t] : 0xffffffffu;
        below += pv < i ? 1u : 0u;
        pivot |= pv == i;
    }
    if (!pivot) {
        R[i - below] = i;
        rows_orig_next[i - below] = rows_orig[i];
    }
}

// ---- the panel's pivot search: one launch --------------------------------------------------------
// One block per 256-row chunk of S.  Every block runs the rule on the first SPR = 64 rows of S alone
// ("spec"; every block gets the same answer).  If it finds a pivot in every panel column,
// those are the panel's pivots over all rows: by induction over the columns the earlier pivots agree,
// so every row is eliminated alike, and column q's pivot, the smallest row with a non-zero entry, lies
// among the first 64 rows, below every later row.  Then block 0 writes P, J and W = S[P, J]^{-1},
// every block writes its chunk of the row list R, and the last block to arrive advances the state
// (st->spec = 1).  Otherwise each block runs tournament level 0 on its chunk and the last block to
// arrive merges the lists (16 per group, level by level) and writes P, J, k and R.  So a panel costs
// one launch whichever way it is settled.
//
// The spec search, by SPW = 8 waves of the block (the others only keep the barriers): lane u of every
// such wave holds row u; wave w owns the panel columns q = 4w .. 4w + 3 (x[0..3]) and the coefficient
// slots t = 4w .. 4w + 3 (c[0..3]); sigma, the pivot flag and the row's own pivot entry d are kept by
// every such wave alike.  row_u = sigma_u orig_u + sum_t c_u[t] orig_{P_t}.  Per column q its owner
// finds the pivot (the lowest lane with a non-zero, one ballot) and publishes the column, the pivot
// lane and its entry v in LDS (two slots, one barrier per column); every wave reads its multiplier
// f = x_u[q], takes the pivot row's entries of its own columns and slots by readlane, and every other
// row with f != 0 takes the division-free Gauss-Jordan step row_u <- (v row_u - f row_b) R^{-1}:
// x[q'] for q' > q, sigma, c[t] for t < k and the new c[k] = -f sigma_b R^{-1}; x[q] becomes exactly
// 0 and every other column left of q is zero in the pivot row, so a row's own pivot entry d is the
// only one left of q that changes (scaled by v R^{-1}).  At the end row P_t is d_t e_t on the pivot
// columns, so M K = D with M[t][t'] = c_{P_t}[t'] (+ sigma on the diagonal) and K^{-1} = D^{-1} M.
// (Measured per column: ~0.9 us, bound by the 64-bit multiplies of the division-free steps; four
// waves with branch-free selects were slower, 1.25 us, sixteen waves no faster.)
// Called by every thread of the block (the barriers); s_res, s_u and s_q are complete on return.
#ifdef HX_PANEL_TIMERS
// Probe builds only: block 0's phase timestamps (wall_clock64, 100 MHz) for the first TIMER_PANELS
// panels of the process: 0 start, 1 loads done, 2 after step 0, 3 after step 15, 4 after step 31,
// 5 spec done, 6 row list done, 7 kernel end.
constexpr int TIMER_PANELS = 8, TIMER_SLOTS = 8;
__device__ long long g_panel_t[TIMER_PANELS * TIMER_SLOTS];
__device__ unsigned g_panel_i;
#define HX_TMARK(slot)                                                                                   \
    do {                                                                                                 \
        if (blockIdx.x == 0 && threadIdx.x == 0 && t_panel < TIMER_PANELS)                               \
            g_panel_t[t_panel * TIMER_SLOTS + (slot)] = wall_clock64();                                  \
    } while (0)
#else
#define HX_TMARK(slot) \
    do {               \
    } while (0)
#endif

constexpr int SPR = 64;
constexpr int SPW = 8;                                   // waves of the spec search
constexpr int SCW = LW / SPW;                            // columns (and coefficient slots) per wave
static_assert(SCW * SPW == LW && SPW * 64 <= RT, "the waves split the panel evenly");
constexpr int CH_LDS = CH * (LW + 1);                    // chunk panel words
constexpr int MG_LDS = GRP * LW * (LW + 1);              // merge panel words

template <class Ar>
__device__ void spec_block(const uint32_t* S, int64_t ld, LState* st, uint32_t wc, const Ar& ar, uint32_t* P,
                           uint32_t* J, uint32_t* Wg, uint32_t* s_u, uint32_t* s_q, uint32_t* s_res, bool writer,
                           unsigned t_panel) {
    (void)t_panel;
    __shared__ uint32_t s_col[2][SPR];
    __shared__ uint64_t s_bv[2];                             // the pivot lane (high word) and entry v (low word)
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, p = ar.p;
    const uint32_t ma = st->m_new;                           // the rows active before this panel's advance
    const uint32_t rows = min((uint32_t)SPR, ma);
    const bool act = w < SPW;                                // wave-uniform
    const bool mine = act && lane < rows;
    uint32_t x[SCW], c[SCW];
#pragma unroll
    for (int j = 0; j < SCW; ++j) {
        const uint32_t q = w * SCW + j;
        x[j] = (mine && q < wc) ? S[(int64_t)lane * ld + q] : 0u;
        c[j] = 0u;
    }
    uint32_t sig = 1u, d = 0u, myt = 0xffffffffu, k = 0;
    bool done = false, ok = true;
#ifdef HX_PANEL_TIMERS
    __builtin_amdgcn_s_waitcnt(0);
    HX_TMARK(1);
#endif
    // A rolled loop: no register array is indexed by the column (the owner's entry is picked by selects),
    // so nothing goes to scratch whatever the compiler does with the loop.
#pragma unroll 1
    for (uint32_t q = 0; q < wc; ++q) {
#ifdef HX_PANEL_TIMERS
        if (q == 1) HX_TMARK(2);
        if (q == 16) HX_TMARK(3);
#endif
        const uint32_t slot = q & 1u, jq = q % SCW;
        if (w == q / SCW) {
            uint32_t xq = x[0];
#pragma unroll
            for (int j = 1; j < SCW; ++j) xq = jq == (uint32_t)j ? x[j] : xq;
            const uint64_t bal = __ballot(mine && !done && xq != 0u);
            s_col[slot][lane] = xq;
            const uint32_t bu = bal ? (uint32_t)__builtin_ctzll(bal) : 0xffffffffu;
            const uint32_t v = bal ? __builtin_amdgcn_readlane(xq, bu) : 0u;
            if (lane == 0) s_bv[slot] = ((uint64_t)bu << 32) | v;
        }
        __syncthreads();
        const uint64_t bv = s_bv[slot];
        const uint32_t f = s_col[slot][lane];
        const uint32_t bu = __builtin_amdgcn_readfirstlane((uint32_t)(bv >> 32));
        if (bu == 0xffffffffu) {
            ok = false;                                      // a column without a pivot here: the tournament decides
            break;                                           // (uniform over the block)
        }
        if (act) {
            const uint32_t v = __builtin_amdgcn_readfirstlane((uint32_t)bv);
            uint32_t y[SCW], yc[SCW];
#pragma unroll
            for (int j = 0; j < SCW; ++j) {
                y[j] = __builtin_amdgcn_readlane(x[j], bu);
                yc[j] = __builtin_amdgcn_readlane(c[j], bu);
            }
            const uint32_t ys = __builtin_amdgcn_readlane(sig, bu);
            if (lane == bu) {
                done = true;
                d = v;
                myt = k;
            } else if (mine && f != 0u) {
                const uint32_t fneg = p - f;
#pragma unroll
                for (int j = 0; j < SCW; ++j) {
                    const uint32_t q2 = w * SCW + j;         // also the coefficient slot of c[j]
                    if (q2 > q) x[j] = ar.dfree(v, x[j], fneg, y[j]);
                    else if (q2 == q) x[j] = 0u;
                    if (q2 < k) c[j] = ar.dfree(v, c[j], fneg, yc[j]);
                    else if (q2 == k) c[j] = ar.dfree(v, 0u, fneg, ys);
                }
                sig = ar.dfree(v, sig, fneg, 0u);
                if (done) d = ar.dfree(v, d, fneg, 0u);
            }
        }
        if (threadIdx.x == 0) {
            s_u[k] = bu;
            s_q[k] = q;
        }
        ++k;
    }
    ok = ok && k == wc;
    HX_TMARK(4);
    if (threadIdx.x == 0) {
        s_res[0] = ok ? 1u : 0u;
        s_res[1] = k;
        s_res[2] = ma;
    }
    __syncthreads();
    if (!ok || !writer) return;
    // settled: the pivots and W = D^{-1} M (Montgomery form, prep), each wave its own slots
    if (threadIdx.x < k) {
        P[threadIdx.x] = s_u[threadIdx.x];
        J[threadIdx.x] = s_q[threadIdx.x];
    }
    if (done) {
        const uint32_t iv = ar.inv(d);
        if (iv == 0 && w == 0) atomicExch(&st->status, 1);
        const uint32_t dinv = ar.prep(iv);
#pragma unroll
        for (int j = 0; j < SCW; ++j) {
            const uint32_t t = w * SCW + j;
            if (t < k) {
                const uint32_t mval = t == myt ? addm(c[j], sig, p) : c[j];
                Wg[myt * LW + t] = ar.prep(ar.mul(mval, dinv));
            }
        }
    }
}

// Level 0: chunk ch of S's rows (ma rows active) -> its candidate rows (ascending S indices).
template <class Ar>
__device__ void chunk_block(const uint32_t* S, int64_t ld, uint32_t ma, uint32_t wc, const Ar& ar, uint32_t ch,
                            uint32_t* pan, uint32_t* cand, uint32_t* cnt) {
    __shared__ uint8_t chosen[CH];
    __shared__ uint32_t s_u[LW], s_q[LW];
    __shared__ uint32_t s_wcnt[RT / 64];
    const uint32_t r0 = ch * CH;
    if (r0 >= ma) {
        if (threadIdx.x == 0) cnt[ch] = 0;
        return;
    }
    const uint32_t rows = min((uint32_t)CH, ma - r0), ws = wc | 1u;
    uint32_t any = 0;
    for (uint32_t idx = threadIdx.x; idx < rows * wc; idx += RT) {
        const uint32_t u = idx / wc, q = idx % wc;
        const uint32_t v = S[(int64_t)(r0 + u) * ld + q];
        pan[u * ws + q] = v;
        any |= v;
    }
    // a chunk that is zero on the panel has no candidates (the Schur complement past the rank is zero)
    if (!__syncthreads_or(any != 0)) {
        if (threadIdx.x == 0) cnt[ch] = 0;
        return;
    }
    static_assert(CH * 4 <= RT, "four threads per chunk row");
    const uint32_t k = lds_rule<4>(pan, rows, wc, ws, chosen, ar, s_u, s_q);
    block_compact(chosen, rows, s_wcnt, [&](uint32_t pos, uint32_t u) { cand[ch * LW + pos] = r0 + u; });
    if (threadIdx.x == 0) cnt[ch] = k;
}

// Merge: group g of up to GRP candidate lists -> one list; the final level writes P, J, k, R and the
// state (ma rows active before the panel).
template <class Ar>
__device__ void merge_group(const uint32_t* S, int64_t ld, uint32_t wc, const Ar& ar, const uint32_t* cand_in,
                            const uint32_t* cnt_in, uint32_t nin, uint32_t g, uint32_t* cand_out, uint32_t* cnt_out,
                            bool final_level, uint32_t ma, uint32_t* P, uint32_t* J, LState* st, uint32_t* R,
                            const uint32_t* rows_orig, uint32_t* rows_orig_next, uint32_t* pan) {
    __shared__ uint32_t rowid[GRP * LW];
    __shared__ uint8_t chosen[GRP * LW];
    __shared__ uint32_t s_u[LW], s_q[LW], s_off[GRP + 1], s_P[LW];
    __shared__ uint32_t s_wcnt[RT / 64];
    const uint32_t g0 = g * GRP, ng = min((uint32_t)GRP, nin - g0);
    if (threadIdx.x < ng) s_off[threadIdx.x + 1] = cnt_in[g0 + threadIdx.x];   // the list lengths, in parallel
    __syncthreads();
    if (threadIdx.x == 0) {
        s_off[0] = 0;
        for (uint32_t i = 0; i < ng; ++i) s_off[i + 1] += s_off[i];
    }
    __syncthreads();
    const uint32_t rows = s_off[ng], ws = wc | 1u;
    for (uint32_t i = 0; i < ng; ++i)
        for (uint32_t t = threadIdx.x; t < s_off[i + 1] - s_off[i]; t += RT)
            rowid[s_off[i] + t] = cand_in[(g0 + i) * LW + t];
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < rows * wc; idx += RT) {
        const uint32_t u = idx / wc, q = idx % wc;
        pan[u * ws + q] = S[(int64_t)rowid[u] * ld + q];
    }
    static_assert(GRP * LW * 2 <= RT, "two threads per candidate row");
    const uint32_t k = lds_rule<2>(pan, rows, wc, ws, chosen, ar, s_u, s_q);
    block_compact(chosen, rows, s_wcnt, [&](uint32_t pos, uint32_t u) { cand_out[g * LW + pos] = rowid[u]; });
    if (threadIdx.x == 0) cnt_out[g] = k;
    if (final_level) {
        if (threadIdx.x < k) {
            P[threadIdx.x] = rowid[s_u[threadIdx.x]];
            J[threadIdx.x] = s_q[threadIdx.x];
            s_P[threadIdx.x] = rowid[s_u[threadIdx.x]];
        }
        __syncthreads();
        rest_rows_of<RT>(ma, k, s_P, R, rows_orig, rows_orig_next);
        if (threadIdx.x == 0) {
            st->k = k;
            st->m_new = ma - k;
        }
    }
    __syncthreads();
}

template <class Ar>
__global__ __launch_bounds__(RT) void rrp_pivot_kernel(const uint32_t* S, int64_t ld, LState* st, uint32_t wc, Ar ar,
                                                       uint32_t* P, uint32_t* J, uint32_t* R, const uint32_t* rows_orig,
                                                       uint32_t* rows_orig_next, uint32_t* Wg, uint32_t* cand0,
                                                       uint32_t* cnt0, uint32_t* cand1, uint32_t* cnt1, uint32_t nchunk) {
    __shared__ uint32_t pool[MG_LDS > CH_LDS ? MG_LDS : CH_LDS];
    __shared__ uint32_t s_u[LW], s_q[LW], s_res[3];
    __shared__ uint32_t s_last;
    // the panel chain (pivot search, X, the next panel's columns) is the critical path: its waves
    // take issue priority over the wide trailing update running beside them on the other stream
    __builtin_amdgcn_s_setprio(2);
    unsigned t_panel = 0;
#ifdef HX_PANEL_TIMERS
    __shared__ unsigned s_tp;
    if (blockIdx.x == 0 && threadIdx.x == 0) s_tp = atomicAdd(&g_panel_i, 1u);
    __syncthreads();
    t_panel = s_tp;
#endif
    HX_TMARK(0);
    // every block runs the spec search (identical results, block 0 writes them), so every block knows
    // at once whether the tournament is needed.  Nothing in the state changes before every block has
    // arrived (the last one advances it), so every block reads the same rows and row count.
    spec_block(S, ld, st, wc, ar, P, J, Wg, s_u, s_q, s_res, blockIdx.x == 0, t_panel);
    HX_TMARK(5);
    const bool settled = s_res[0] != 0;
    const uint32_t ks = s_res[1], ma = s_res[2];
    if (settled) {
        rest_rows_chunk(ma, ks, s_u, R, rows_orig, rows_orig_next, blockIdx.x * CH);   // this block's rows
        HX_TMARK(6);
    } else {
        chunk_block(S, ld, ma, wc, ar, blockIdx.x, pool, cand0, cnt0);               // tournament level 0
    }
    // arrival: every thread's writes reach the device before its block counts itself
    __threadfence();
    __syncthreads();
    HX_TMARK(7);
    if (threadIdx.x == 0) {
        const uint32_t n = __hip_atomic_fetch_add(&st->arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = n + 1 == gridDim.x ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();                                        // the other blocks' lists and rows are visible here
    if (settled) {
        // the state moves past the previous panel to this one
        if (threadIdx.x == 0) {
            st->arrive = 0u;
            st->r += st->k;
            st->m_a = ma;
            st->k = ks;
            st->m_new = ma - ks;
            st->spec = 1u;
            st->spec_hits += 1u;
        }
        return;
    }
    // the tournament decides: the advance past the previous panel, then the merge levels
    if (threadIdx.x == 0) {
        st->arrive = 0u;
        st->spec = 0u;
        st->r += st->k;
        st->m_a = ma;
        st->k = 0;
    }
    uint32_t nin = nchunk;
    int lv = 0;
    do {
        const uint32_t nout = (nin + GRP - 1) / GRP;
        const bool odd = (lv & 1) != 0;
        for (uint32_t g = 0; g < nout; ++g)
            merge_group(S, ld, wc, ar, odd ? cand1 : cand0, odd ? cnt1 : cnt0, nin, g, odd ? cand0 : cand1,
                        odd ? cnt0 : cnt1, nout == 1, ma, P, J, st, R, rows_orig, rows_orig_next, pool);
        __threadfence_block();
        __syncthreads();
        nin = nout;
        ++lv;
    } while (nin > 1);
}

// The k <= 32 products of a panel on the INT8 matrix cores: Cout[i, j] = Cin[c_rows[i], j] -
// sum_{q<k} A[a_rows[i], a_cols[q]] X[q, j] mod p.  A's entries are split into L balanced 8-bit
// limbs on the fly; X's limb planes come from panel_x (plane s, column j: 32 k-bytes).  One wave per
// 32-row block and ucb consecutive 32-column blocks: the row block's A fragment (gathered, split)
// and its Cin row indices are loaded once and serve every column block.  Per column block the X
// fragments are loaded before the Cin entries (so the MFMAs wait for the first only), and the L^2
// limb products of the one 32-deep k-step are taken one anti-diagonal at a time: each diagonal's
// int32 sum is folded into an int64 accumulator at once (sum_d D_d w_d, exact: below 2^63, the
// recombination of hx_limbs.h in another order), so 16 accumulator words serve all 2L - 1
// diagonals and more waves fit a SIMD.  Every output is written, also when k = 0 (Cout = Cin: the
// trailing update doubles as the compaction into the next Schur complement).  No barriers: a wave
// past the rows returns at once.
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
constexpr uint32_t UPD_UCB = 4;          // column blocks per wave of the wide products (the next panel's: 1)
struct UpdArgs {
    uint32_t m, n;                           // row bound (m_dev caps it) and columns
    uint32_t ucb;                            // 32-column blocks per wave
    uint32_t chain;                          // 1: on the panel chain (issue priority)
    const uint32_t* m_dev;
    const uint32_t* k_dev;                   // k <= 32
    const uint32_t* A;
    int64_t ars;
    const uint32_t* a_rows;                  // optional row gather
    const uint32_t* a_cols;                  // optional column gather
    const uint8_t* planes;                   // X's limb planes
    uint32_t pstride, pcol0;                 // bytes per plane; first plane column
    const uint32_t* Cin;
    int64_t cirs;
    const uint32_t* c_rows;                  // optional row gather on Cin
    uint32_t* Cout;
    int64_t cors;
    DevMod M;
    LimbConsts lc;
    uint32_t half;
};

template <int L>
__global__ __launch_bounds__(256) void upd32_kernel(UpdArgs u) {
    constexpr int ND = 2 * L - 1;
    if (u.chain) __builtin_amdgcn_s_setprio(2);     // the next panel's columns: on the panel chain
    const uint32_t m = min(*u.m_dev, u.m), k = min(*u.k_dev, 32u);
    if (k == 0 && u.Cin == u.Cout && u.c_rows == nullptr && u.cirs == u.cors) return;   // in place, nothing to do
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 31, lg = lane >> 5;
    const uint32_t ncg = (u.n + 32 * u.ucb - 1) / (32 * u.ucb);
    const uint32_t b0 = ((blockIdx.x / ncg) * 4 + wave) * 32, cg = blockIdx.x % ncg;
    if (b0 >= m) return;                     // (k = 0 still copies Cin to Cout: the compaction)
    const uint32_t p = u.M.p;
    // A fragment: row b0 + lr, k-bytes 16 lg .. 16 lg + 15, all loads first
    const uint32_t arow = b0 + lr;
    const bool alive = arow < m;
    const int64_t aoff = alive ? (u.a_rows ? (int64_t)u.a_rows[arow] : (int64_t)arow) * u.ars : 0;
    uint32_t av[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t t = 16 * lg + i;
        av[i] = (alive && t < k) ? u.A[aoff + (u.a_cols ? (int64_t)u.a_cols[t] : (int64_t)t)] : 0u;
    }
    // the Cin rows of this lane's 16 outputs (accumulator element e: row b0 + (e / 4) 8 + 4 lg + e % 4)
    uint32_t crow[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const uint32_t r = b0 + (e / 4) * 8 + lg * 4 + (e % 4);
        crow[e] = r < m ? (u.c_rows ? u.c_rows[r] : r) : 0u;
    }
    v4i fa[L];
#pragma unroll
    for (int s2 = 0; s2 < L; ++s2) fa[s2] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int32_t d[4] = {0, 0, 0, 0};
        limbs_of(av[i], p, u.half, L, d);
#pragma unroll
        for (int s2 = 0; s2 < L; ++s2) fa[s2][i >> 2] |= (int)(((uint32_t)d[s2] & 0xffu) << (8 * (i & 3)));
    }
    for (uint32_t cb = 0; cb < u.ucb; ++cb) {
        const uint32_t jb = (cg * u.ucb + cb) * 32;
        if (jb >= u.n) break;
        const uint32_t j = jb + lr;
        const bool jok = j < u.n;
        // the row addresses are rebuilt per column block from 32-bit row indices (kept opaque), not
        // hoisted as 32 pointers for the whole loop: fewer registers, more waves per SIMD
#pragma unroll
        for (int e = 0; e < 16; ++e) asm volatile("" : "+v"(crow[e]));
        uint32_t obase = b0 + lg * 4;
        asm volatile("" : "+v"(obase));
        v4i fb[L];
#pragma unroll
        for (int s2 = 0; s2 < L; ++s2)
            fb[s2] = jok ? *(const v4i*)(u.planes + (size_t)s2 * u.pstride + (size_t)(u.pcol0 + j) * 32 + 16 * lg)
                         : v4i{0, 0, 0, 0};
        uint32_t cv[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t r = obase + (e / 4) * 8 + (e % 4);
            cv[e] = (jok && r < m) ? u.Cin[(int64_t)crow[e] * u.cirs + j] : 0u;
        }
        int64_t xs[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) xs[e] = 0;
#pragma unroll
        for (int dg = 0; dg < ND; ++dg) {
            v16i acc;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] = 0;
#pragma unroll
            for (int s2 = 0; s2 < L; ++s2) {
                const int t = dg - s2;
                if (t >= 0 && t < L) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s2], fb[t], acc, 0, 0, 0);
            }
            const int64_t wd = limb_weight(dg, u.lc);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                xs[e] += (int64_t)acc[e] * wd;
                asm volatile("" : "+v"(xs[e]));  // folded here, not sunk to the epilogue: one accumulator live
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (!jok) continue;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t r = obase + (e / 4) * 8 + (e % 4);
            if (r >= m) continue;
            u.Cout[(int64_t)r * u.cors + j] = subm(cv[e], red64((uint64_t)xs[e] + u.lc.off, u.M), p);
        }
    }
}

// X[:, slab] = S[P, J]^{-1} S[P, slab]; also the pivots / pivot rows of this panel (block 0).
// Every workgroup reduces the same k x k pivot block beside its own SL-column slab: narrow slabs
// spread the (multiply-latency-bound) elimination over many CUs.
constexpr int SL = 16;
constexpr int PXT = 512;     // threads of panel_x_kernel (two waves per SIMD)
template <class Ar>
__global__ __launch_bounds__(PXT) void panel_x_kernel(const uint32_t* S, int64_t ld, uint32_t na, const uint32_t* P,
                                                     const uint32_t* J, LState* st, Ar ar, uint32_t* X, int64_t ldx,
                                                     uint32_t* U, int64_t ldu, uint32_t c, const uint32_t* rows_orig,
                                                     uint32_t* pivots, uint32_t* pivot_rows, uint8_t* Xp,
                                                     uint32_t pstride, uint32_t L, uint32_t half, int slot,
                                                     const uint32_t* Wg) {
    __shared__ uint32_t B[LW * (LW + SL)];
    __shared__ uint32_t s_inv[LW], s_q[LW];
    __builtin_amdgcn_s_setprio(2);                    // on the panel chain (see rrp_pivot_kernel)
    if (blockIdx.x == 0 && threadIdx.x == 0) {        // the sizes of this panel's products, slot i & 1
        st->snap_k[slot] = st->k;
        st->snap_m[slot] = st->m_new;
        st->snap_r[slot] = st->r;
    }
    const uint32_t k = st->k;
    if (k == 0) return;
    const uint32_t j0 = blockIdx.x * SL;
    if (j0 >= na) return;
    const uint32_t sw = min((uint32_t)SL, na - j0), bw = k + sw, p = ar.p;
    const uint32_t r = st->r;
    // the spec search settled the panel: W = S[P, J]^{-1} (Montgomery form) is ready, X = W S[P, slab]
    const bool wpath = st->spec != 0;
    for (uint32_t idx = threadIdx.x; idx < k * bw; idx += PXT) {
        const uint32_t t = idx / bw, j = idx % bw;
        const int64_t row = (int64_t)P[t] * ld;
        B[t * bw + j] = j < k ? (wpath ? Wg[t * LW + j] : S[row + J[j]]) : S[row + j0 + (j - k)];
    }
    if (threadIdx.x < k) s_q[threadIdx.x] = J[threadIdx.x];
    __syncthreads();
    if (wpath) {
        static_assert(LW * SL <= PXT, "one output per thread");
        const uint32_t t = threadIdx.x / SL, j = threadIdx.x % SL;
        uint32_t x = 0;
        if (t < k && j < sw)
            for (uint32_t t2 = 0; t2 < k; ++t2) x = addm(x, ar.mul(B[t2 * bw + k + j], B[t * bw + t2]), p);
        __syncthreads();
        if (t < k && j < sw) B[t * bw + k + j] = x;
        if (threadIdx.x < k) s_inv[threadIdx.x] = ar.prep(1u);
        __syncthreads();
    } else {
        // Division-free Gauss-Jordan in pivot order (column t of the block is pivot column J[t]):
        // row_i <- v row_i - f row_t for every row i != t with f = B[i, t] != 0.  Every leading block of
        // the pivot block is non-singular in this order, so the left block ends diagonal and the slab
        // part of row i is d_i (S[P, J]^-1 S[P, slab])_i.  Step t writes, in the rows i != t, only the
        // left columns past t, the slab and the diagonal entry (i, i) when i < t (row t is zero at
        // column i there); the left columns below t are zero and are neither written nor read again, so
        // no entry a step writes is read by another thread in that step: one buffer, one barrier per
        // step.  Sixteen threads share a row.
        {
            const uint32_t gi = threadIdx.x >> 4, gsub = threadIdx.x & 15;
            for (uint32_t t = 0; t < k; ++t) {
                if (gi < k && gi != t) {
                    uint32_t* row = B + gi * bw;
                    const uint32_t* trow = B + t * bw;
                    const uint32_t f = row[t];
                    if (f != 0) {
                        const uint32_t v = trow[t], fneg = p - f, nl = k - t - 1, ns = nl + sw;
                        const uint32_t nslot = ns + (gi < t ? 1u : 0u);
                        for (uint32_t sl = gsub; sl < nslot; sl += 16) {
                            if (sl < ns) {
                                const uint32_t col = sl < nl ? t + 1 + sl : k + (sl - nl);
                                row[col] = ar.dfree(v, row[col], fneg, trow[col]);
                            } else {
                                row[gi] = ar.dfree(v, row[gi], fneg, 0u);
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (threadIdx.x < k) {
            const uint32_t iv = invm(B[threadIdx.x * bw + threadIdx.x], p);
            if (iv == 0) atomicExch(&st->status, 1);
            s_inv[threadIdx.x] = ar.prep(iv);
        }
        __syncthreads();
    }
    // X[:, slab] = d^-1 (slab part); the slab's own pivot columns get the identity exactly (they
    // equal it already)
    for (uint32_t idx = threadIdx.x; idx < k * sw; idx += PXT) {
        const uint32_t t = idx / sw, j = idx % sw;
        uint32_t x = ar.mul(B[t * bw + k + j], s_inv[t]);
        for (uint32_t t2 = 0; t2 < k; ++t2)
            if (s_q[t2] == j0 + j) x = (t2 == t) ? 1u : 0u;
        X[(int64_t)t * ldx + j0 + j] = x;
        U[(int64_t)(r + t) * ldu + c + j0 + j] = x;
        B[t * bw + k + j] = x;
    }
    __syncthreads();
    // X's limb planes for the matrix-core products: plane s, column j0 + j holds the 32 k-bytes
    // (limb s of X[t, j0 + j], t = 0..31; zero past k), four k-bytes per store
    for (uint32_t idx = threadIdx.x; idx < 8 * sw; idx += PXT) {
        const uint32_t j = idx % sw, t4 = idx / sw;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t t = 4 * t4 + e;
            int32_t d[4] = {0, 0, 0, 0};
            limbs_of(t < k ? B[t * bw + k + j] : 0u, p, half, (int)L, d);
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2) w[s2] |= ((uint32_t)d[s2] & 0xffu) << (8 * e);
        }
        for (uint32_t s2 = 0; s2 < L; ++s2)
            *(uint32_t*)(Xp + (size_t)s2 * pstride + (size_t)(j0 + j) * 32 + 4 * t4) = w[s2];
    }
    if (blockIdx.x == 0 && threadIdx.x < k) {
        pivots[r + threadIdx.x] = c + s_q[threadIdx.x];
        pivot_rows[r + threadIdx.x] = rows_orig[P[threadIdx.x]];
    }
}

// The trailing update of a panel may still run (second stream) while the next panel's
// tournament rewrites k / m_new: it reads its own copy.
__global__ void advance_kernel(LState* st) {
    if (threadIdx.x == 0) {
        st->r += st->k;
        st->m_a = st->m_new;
        st->k = 0;
    }
}

__global__ void init_state_kernel(LState* st, uint32_t m, uint32_t* rows_orig, uint32_t* U, size_t ucount) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        st->m_a = m;
        st->r = 0;
        st->k = 0;
        st->m_new = m;
        st->status = 0;
        st->spec = 0;
        st->arrive = 0;
        st->spec_hits = 0;
    }
    if (i < m) rows_orig[i] = (uint32_t)i;
    for (size_t t = i; t < ucount; t += (size_t)gridDim.x * blockDim.x) U[t] = 0;
}

// T[t, q] = U[t, c + J[q]] for the r pivot rows stored before the panel (the Gauss-Jordan product's
// left operand; zero past the panel's k pivots).
__global__ void gj_gather_kernel(const uint32_t* U, int64_t ldu, uint32_t c, const uint32_t* J, const LState* st,
                                 int slot, uint32_t* T) {
    const uint32_t r = st->snap_r[slot], k = st->snap_k[slot];
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < r * LW; idx += gridDim.x * blockDim.x) {
        const uint32_t t = idx / LW, q = idx % LW;
        T[idx] = q < k ? U[(int64_t)t * ldu + c + J[q]] : 0u;
    }
}

inline uint32_t gsz(uint64_t n) { return (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536); }

// A second stream and two events for the lookahead, per calling stream (created once, reused by
// every elimination issued on that stream, so a call allocates nothing and can be captured).
// The streams beside the caller's (non-blocking, created once per caller stream): s2 takes the wide
// trailing updates, s3 the Gauss-Jordan gathers and products, sc the panel chain when the caller's
// stream is the legacy null stream.  (A CU-masked queue for s2, which leaves CUs free for the chain,
// was measured: 3 % faster alone, but a masked stream is a blocking one, so every launch on it waits
// for the null stream, and when the null stream shares a hardware queue with another side stream the
// updates serialise behind it -- 16.3 ms instead of 10 for C2 once other streams existed.)
struct SideStream {
    hipStream_t s2 = nullptr, s3 = nullptr, sc = nullptr;
    hipEvent_t ev_x = nullptr, ev_rest = nullptr, ev_gj[2] = {nullptr, nullptr}, ev_fork = nullptr,
               ev_join = nullptr;
};
SideStream side_for(hipStream_t s) {
    static std::mutex mu;
    static std::unordered_map<hipStream_t, SideStream> map;
    std::lock_guard<std::mutex> lk(mu);
    SideStream& ss = map[s];
    if (ss.sc == nullptr) {
        HX_HIP(hipStreamCreateWithFlags(&ss.s2, hipStreamNonBlocking));
        HX_HIP(hipStreamCreateWithFlags(&ss.s3, hipStreamNonBlocking));
        HX_HIP(hipStreamCreateWithFlags(&ss.sc, hipStreamNonBlocking));
        for (hipEvent_t* e : {&ss.ev_x, &ss.ev_rest, &ss.ev_gj[0], &ss.ev_gj[1], &ss.ev_fork, &ss.ev_join})
            HX_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    return ss;
}

// The results on the device: rank = r (0xffffffff when a pivot had no inverse), pivots, pivot rows.
__global__ void finish_kernel(const LState* st, const uint32_t* piv, const uint32_t* prow, uint32_t mn, uint32_t* d_rank,
                              uint32_t* d_piv, uint32_t* d_prow) {
    const uint32_t r = st->r, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *d_rank = st->status ? 0xffffffffu : r;
    if (i < r && i < mn) {
        if (d_piv) d_piv[i] = piv[i];
        if (d_prow) d_prow[i] = prow[i];
    }
}

// R = the rank rows of U (the rank read on the device), then zero rows.
__global__ void rref_out_dev_kernel(const uint32_t* U, uint32_t n, const LState* st, uint32_t m, uint32_t* R, int64_t ldr) {
    const uint32_t r = st->r;
    const uint64_t total = (uint64_t)m * n;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        R[(int64_t)i * ldr + j] = i < r ? U[(size_t)i * n + j] : 0u;
    }
}

// The buffers of one m x n elimination, in words, each rounded to 64 words (the order is the
// workspace layout of enqueue_large).
constexpr int WS_STATE = 15;                 // index of the device-side state (LState)
std::vector<size_t> large_sizes(uint32_t m, uint32_t n) {
    const uint32_t mn = std::min(m, n), nchunk = (m + CH - 1) / CH;
    std::vector<size_t> sizes;
    auto add = [&](size_t words) { sizes.push_back((words + 63) / 64 * 64); };
    add((size_t)m * n);           // 0 S0
    add((size_t)m * n);           // 1 S1
    add((size_t)mn * n);          // 2 U store
    add((size_t)2 * LW * n);      // 3 X copy (two panels: slot i & 1)
    add(m);                       // 4 rows_orig a
    add(m);                       // 5 rows_orig b
    add((size_t)2 * m);           // 6 R list (two panels: slot i & 1)
    add((size_t)nchunk * LW);     // 7 cand a
    add((size_t)nchunk * LW);     // 8 cand b
    add(nchunk);                  // 9 cnt a
    add(nchunk);                  // 10 cnt b
    add(2 * LW);                  // 11 P (two panels: slot i & 1)
    add(2 * LW);                  // 12 J
    add(mn);                      // 13 pivots
    add(mn);                      // 14 pivot rows
    add(16);                      // 15 state
    add((size_t)mn * LW);         // 16 Gauss-Jordan left operand T
    add((size_t)2 * 32 * n);      // 17 X's limb planes (two panels: slot i & 1; 4 planes x 32 x n bytes each)
    add((size_t)LW * LW);         // 18 W = S[P, J]^-1 from the spec search
    return sizes;
}

size_t large_workspace_words(uint32_t m, uint32_t n) {
    size_t total = 0;
    for (size_t x : large_sizes(m, n)) total += x;
    return total;
}

size_t large_state_offset_words(uint32_t m, uint32_t n) {
    const std::vector<size_t> sizes = large_sizes(m, n);
    size_t off = 0;
    for (int i = 0; i < WS_STATE; ++i) off += sizes[i];
    return off;
}

template <class Ar>
void enqueue_large(const hx_modctx& ctx, const Ar& ar, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda,
                   uint32_t* Rout, int64_t ldr, uint32_t* d_rank, uint32_t* d_piv, uint32_t* d_prow, uint32_t* base,
                   hipStream_t s, const SideStream& side) {
    const uint32_t mn = std::min(m, n);
    const uint32_t nchunk = (m + CH - 1) / CH;
    const std::vector<size_t> sizes = large_sizes(m, n);
    size_t total = 0;
    std::vector<size_t> off;
    for (size_t w : sizes) { off.push_back(total); total += w; }
    auto at = [&](int i) { return base + off[i]; };
    uint32_t* Sb[2] = {at(0), at(1)};
    uint32_t* U = at(2);
    uint32_t* Xb = at(3);
    uint32_t* T = at(16);
    uint8_t* Xpb = (uint8_t*)at(17);
    uint32_t* Wg = at(18);
    const uint32_t pstride = 32 * n;          // bytes per plane
    // the limb split of the k <= 32 products (balanced digits of residues mod p)
    const uint32_t L = ctx.p <= 255u ? 1 : (ctx.p <= 65279u ? 2 : (ctx.p <= 16711423u ? 3 : 4));
    const uint32_t half = (ctx.p - 1) / 2;
    const LimbConsts lc = limb_consts(ctx.p);
    const DevMod dm = devmod(ctx);
    uint64_t n_upd = 0;
    const uint64_t launches0 = stat_value(ST_KERNELS);
    auto upd = [&](const UpdArgs& u, hipStream_t st_) {
        const uint32_t grid = (u.m + 127) / 128 * ((u.n + 32 * u.ucb - 1) / (32 * u.ucb));
        if (grid == 0) return;
        ++n_upd;
        switch (L) {
            case 1: upd32_kernel<1><<<grid, 256, 0, st_>>>(u); break;
            case 2: upd32_kernel<2><<<grid, 256, 0, st_>>>(u); break;
            case 3: upd32_kernel<3><<<grid, 256, 0, st_>>>(u); break;
            default: upd32_kernel<4><<<grid, 256, 0, st_>>>(u); break;
        }
        after_launch("upd32_kernel");
    };
    uint32_t* ro[2] = {at(4), at(5)};
    uint32_t* Rlb = at(6);
    uint32_t* cand[2] = {at(7), at(8)};
    uint32_t* cnt[2] = {at(9), at(10)};
    uint32_t *Pb = at(11), *Jb = at(12), *piv = at(13), *prow = at(14);
    LState* st = (LState*)at(WS_STATE);
    // lookahead: the trailing update of panel i is split into the next panel's LW columns (on the
    // chain's stream) and the rest (on s2, overlapping panel i+1's pivot search); the Gauss-Jordan
    // gather and product run on s3.  Panel i+1's X waits for s2 (its rows are read whole); panel
    // i+2, which reuses panel i's slot (P, J, X, its limb planes, the sizes), waits for s3.
    hipStream_t sc = s == nullptr ? side.sc : s;            // the panel chain
    hipStream_t s2 = side.s2, s3 = side.s3;
    if (sc != s) {
        HX_HIP(hipEventRecord(side.ev_fork, s));
        HX_HIP(hipStreamWaitEvent(sc, side.ev_fork, 0));
    }
    bool rest_pending = false, gj_pending[2] = {false, false};
    {
        HX_HIP(hipMemcpy2DAsync(Sb[0], (size_t)n * 4, A, (size_t)lda * 4, (size_t)n * 4, m, hipMemcpyDeviceToDevice, sc));
        init_state_kernel<<<gsz(std::max<size_t>(m, (size_t)mn * n)), 256, 0, sc>>>(st, m, ro[0], U, (size_t)mn * n);
        after_launch("init_state_kernel");
        int cur = 0;
        for (uint32_t c = 0, panel = 0; c < n; c += LW, ++panel) {
            const uint32_t na = n - c, wc = std::min<uint32_t>(LW, na);
            const uint32_t* S = Sb[cur];
            const int slot = (int)(panel & 1);
            uint32_t* P = Pb + slot * LW;
            uint32_t* J = Jb + slot * LW;
            uint32_t* X = Xb + (size_t)slot * LW * n;
            uint8_t* Xp = Xpb + (size_t)slot * 4 * pstride;
            uint32_t* Rl = Rlb + (size_t)slot * m;
            // panel i - 2's Gauss-Jordan product still reads this slot's J, X, planes and sizes
            if (gj_pending[slot]) {
                HX_HIP(hipStreamWaitEvent(sc, side.ev_gj[slot], 0));
                gj_pending[slot] = false;
            }
            // 1. the pivot search (reads only this panel's LW columns, already up to date): the search
            //    over the first SPR rows, which settles most panels alone, else the tournament
            rrp_pivot_kernel<Ar><<<nchunk, RT, 0, sc>>>(S, n, st, wc, ar, P, J, Rl, ro[cur], ro[cur ^ 1], Wg, cand[0],
                                                        cnt[0], cand[1], cnt[1], nchunk);
            after_launch("rrp_pivot_kernel");
            // the previous panel's update must be complete before its rows are read whole
            if (rest_pending) {
                HX_HIP(hipStreamWaitEvent(sc, side.ev_rest, 0));
                rest_pending = false;
            }
            // 2. X = S[P,J]^{-1} S[P,:] into the U store and the contiguous copy
            panel_x_kernel<Ar><<<(na + SL - 1) / SL, PXT, 0, sc>>>(S, n, na, P, J, st, ar, X, n, U, n, c, ro[cur], piv,
                                                                   prow, Xp, pstride, L, half, slot, Wg);
            after_launch("panel_x_kernel");
            // 3. the next Schur complement, compacted: first the next panel's columns, then the rest.  The
            //    other streams start after the next panel's columns (no marker between the chain's kernels)
            bool x_marked = false;
            auto mark_x = [&] {
                if (!x_marked) HX_HIP(hipEventRecord(side.ev_x, sc));   // X, the row list and the sizes are ready
                x_marked = true;
            };
            if (na > wc) {
                const uint32_t nt = na - wc, n1 = std::min<uint32_t>(LW, nt);
                // S'[i, j] = S[R_i, wc + j] - sum_t S[R_i, J_t] X[t, wc + j] on the matrix cores
                UpdArgs u;
                u.m = m;
                u.m_dev = &st->snap_m[slot];
                u.k_dev = &st->snap_k[slot];
                u.A = S;
                u.ars = n;
                u.a_rows = Rl;
                u.a_cols = J;
                u.planes = Xp;
                u.pstride = pstride;
                u.Cin = S + wc;
                u.cirs = n;
                u.c_rows = Rl;
                u.cors = n;
                u.M = dm;
                u.lc = lc;
                u.half = half;
                u.n = n1;
                u.ucb = 1;
                u.chain = 1;
                u.pcol0 = wc;
                u.Cout = Sb[cur ^ 1];
                upd(u, sc);
                if (nt > n1) {
                    mark_x();
                    HX_HIP(hipStreamWaitEvent(s2, side.ev_x, 0));
                    u.n = nt - n1;
                    u.ucb = UPD_UCB;
                    u.chain = 0;
                    u.pcol0 = wc + n1;
                    u.Cin = S + wc + n1;
                    u.Cout = Sb[cur ^ 1] + n1;
                    upd(u, s2);
                    HX_HIP(hipEventRecord(side.ev_rest, s2));
                    rest_pending = true;
                }
                cur ^= 1;
            }
            // 4. Gauss-Jordan on the stored pivot rows (at most min(c, mn) of them), on s3; not needed
            //    when only the rank, pivots and pivot rows are asked for
            const uint32_t rmax = Rout ? std::min<uint32_t>(c, mn) : 0u;
            if (rmax > 0) {
                mark_x();
                HX_HIP(hipStreamWaitEvent(s3, side.ev_x, 0));
                gj_gather_kernel<<<gsz((uint64_t)rmax * LW), 256, 0, s3>>>(U, n, c, J, st, slot, T);
                after_launch("gj_gather_kernel");
                // U[0:r, c + j] -= sum_t T[., t] X[t, j] on the matrix cores (in place)
                UpdArgs u;
                u.m = rmax;
                u.n = na;
                u.ucb = UPD_UCB;
                u.chain = 0;
                u.m_dev = &st->snap_r[slot];
                u.k_dev = &st->snap_k[slot];
                u.A = T;
                u.ars = LW;
                u.a_rows = nullptr;
                u.a_cols = nullptr;
                u.planes = Xp;
                u.pstride = pstride;
                u.pcol0 = 0;
                u.Cin = U + c;
                u.cirs = n;
                u.c_rows = nullptr;
                u.Cout = U + c;
                u.cors = n;
                u.M = dm;
                u.lc = lc;
                u.half = half;
                upd(u, s3);
                HX_HIP(hipEventRecord(side.ev_gj[slot], s3));
                gj_pending[slot] = true;
            }
        }
        advance_kernel<<<1, 64, 0, sc>>>(st);           // past the last panel (the pivot search does it for the others)
        after_launch("advance_kernel");
        // every product on s2 and s3 (the last Gauss-Jordan product included) before U is read
        if (rest_pending) HX_HIP(hipStreamWaitEvent(sc, side.ev_rest, 0));
        for (int sl = 0; sl < 2; ++sl)
            if (gj_pending[sl]) HX_HIP(hipStreamWaitEvent(sc, side.ev_gj[sl], 0));
        // the results, on the device: rank (0xffffffff when a pivot had no inverse), pivots, pivot rows, R
        finish_kernel<<<gsz(std::max<uint32_t>(mn, 1)), 256, 0, sc>>>(st, piv, prow, mn, d_rank, d_piv, d_prow);
        after_launch("finish_kernel");
        if (Rout) {
            rref_out_dev_kernel<<<gsz((uint64_t)m * n), 256, 0, sc>>>(U, n, st, m, Rout, ldr);
            after_launch("rref_out_dev_kernel");
        }
        if (sc != s) {
            HX_HIP(hipEventRecord(side.ev_join, sc));
            HX_HIP(hipStreamWaitEvent(s, side.ev_join, 0));
        }
    }
    elim_stat_add(EL_CALLS, 1);
    elim_stat_add(EL_PANELS, (n + LW - 1) / LW);
    elim_stat_add(EL_PANEL_COLUMNS, n);
    elim_stat_add(EL_GEMMS, n_upd);
    elim_stat_add(EL_LAUNCHES, stat_value(ST_KERNELS) - launches0);
}

}  // namespace

// The whole-GPU path takes matrices with at least 1024 rows and m * n >= 2^21, and (force) every
// shape the one-workgroup kernel cannot hold (more than 4096 rows, or rows too wide for its LDS
// panel); the rest run in one workgroup (hx_rref_batch.hip).
size_t rref_large_workspace_bytes(uint32_t m, uint32_t n) { return large_workspace_words(m, n) * 4; }

// Enqueues the elimination on s: no allocation, no host synchronisation (capturable).
void rref_large_enqueue(const hx_modctx& ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                        int64_t ldr, uint32_t* d_rank, uint32_t* d_piv, uint32_t* d_prow, void* ws, size_t ws_bytes,
                        hipStream_t s) {
    HX_CHECK(ws != nullptr && ws_bytes >= rref_large_workspace_bytes(m, n), "rref: workspace too small");
    HX_CHECK(d_rank != nullptr, "rref: rank is NULL");
    const SideStream side = side_for(s);
    if (ctx.p & 1u)
        enqueue_large(ctx, make_mont(ctx.p), m, n, A, lda, R, ldr, d_rank, d_piv, d_prow, (uint32_t*)ws, s, side);
    else
        enqueue_large(ctx, make_barrett(ctx.p), m, n, A, lda, R, ldr, d_rank, d_piv, d_prow, (uint32_t*)ws, s, side);
}

bool rref_large(const hx_modctx& ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                int64_t ldr, uint32_t* rank, uint32_t* pivots, uint32_t* pivot_rows, hipStream_t s, bool force) {
    if (!force && (m < 1024 || (uint64_t)m * n < (1ull << 21))) return false;
    // the workspace (and three result arrays) from the arena: after the first call of a shape the block
    // is reused, no device allocation
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t mn = std::min(m, n);
    const size_t wsb = rref_large_workspace_bytes(m, n), res_words = 2 * (size_t)std::max<uint32_t>(mn, 1) + 64;
    uint32_t* block = (uint32_t*)arena_alloc(wsb + res_words * 4);
    elim_peak_bytes(wsb + res_words * 4);
    try {
        uint32_t* d_res = (uint32_t*)((char*)block + wsb);
        uint32_t* d_piv = d_res + 64;
        uint32_t* d_prow = d_piv + std::max<uint32_t>(mn, 1);
        rref_large_enqueue(ctx, m, n, A, lda, R, ldr, d_res, d_piv, d_prow, block, wsb, s);
        uint32_t h_rank = 0, h_spec = 0;
        HX_HIP(hipMemcpyAsync(&h_rank, d_res, 4, hipMemcpyDeviceToHost, s));
        const LState* d_st = (const LState*)(block + large_state_offset_words(m, n));
        HX_HIP(hipMemcpyAsync(&h_spec, &d_st->spec_hits, 4, hipMemcpyDeviceToHost, s));
        HX_HIP(hipStreamSynchronize(s));
        elim_stat_add(EL_SPEC_PANELS, h_spec);
#ifdef HX_PANEL_TIMERS
        {
            long long t[TIMER_PANELS * TIMER_SLOTS];
            HX_HIP(hipMemcpyFromSymbol(t, HIP_SYMBOL(g_panel_t), sizeof(t)));
            for (int pi = 0; pi < TIMER_PANELS; ++pi) {
                std::fprintf(stderr, "panel-timers %d:", pi);
                for (int j = 1; j < TIMER_SLOTS; ++j)
                    std::fprintf(stderr, " %.2f", (t[pi * TIMER_SLOTS + j] - t[pi * TIMER_SLOTS]) / 100.0);
                std::fprintf(stderr, " us\n");
            }
        }
#endif
        if (h_rank == 0xffffffffu) throw Error{HX_E_NOT_INVERTIBLE, "a pivot had no inverse mod p (is p prime?)"};
        *rank = h_rank;
        if (h_rank) {
            if (pivots) HX_HIP(hipMemcpyAsync(pivots, d_piv, (size_t)h_rank * 4, hipMemcpyDeviceToHost, s));
            if (pivot_rows) HX_HIP(hipMemcpyAsync(pivot_rows, d_prow, (size_t)h_rank * 4, hipMemcpyDeviceToHost, s));
            HX_HIP(hipStreamSynchronize(s));
        }
    } catch (...) {
        (void)hipStreamSynchronize(s);
        arena_free(block);
        throw;
    }
    arena_free(block);
    elim_seconds_add(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return true;
}

}  // namespace hx
