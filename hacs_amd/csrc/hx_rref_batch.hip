// hx_rref_batch.hip -- RREF over F_p for many independent, uneven matrices in ONE persistent launch.
//
// One 256-thread workgroup eliminates one job at a time and pulls the next job index from an atomic
// counter (jobs pre-sorted by decreasing work), so thousands of jobs of shapes 16..1024 share the
// chip without a launch per job.  Per job (in place on its buffer, like elim::rref, DenseElim.h:149-151):
//
//   forward, in column panels of w <= 32 columns:
//     (a) the panel of every still-unused row goes to LDS; per column the pivot is the unused row of
//         SMALLEST ORIGINAL INDEX with a non-zero Schur entry (the reference's pivot_rows rule,
//         Cuda/CudaKernels.h:255-265; rows never move, so list order == original order), found by a
//         wave-min + one barrier; the other rows are eliminated division-free inside the panel
//         (x <- (v x - f y) R^-1, Montgomery): only zero tests are read from this copy.
//     (c) the k pivot rows of the panel (their pre-panel values, columns c..n) go to LDS and are reduced
//         to X = S[P,J]^{-1} S[P, c:n]: division-free forward over the pivot order, one inverse per
//         pivot (all in parallel), back substitution -- X has the identity at the panel's pivot columns.
//     (d) trailing update of every other unused row: S[i, j] -= sum_t S[i, J_t] X[t, j] for j past the
//         panel on the INT8 matrix cores (both operands split into balanced 8-bit limbs on the fly,
//         exact int32 limb products, exact recombination; trailing_mfma) -- the Schur complement
//         w.r.t. the panel pivots; the panel's own columns of those rows become zero and are dropped.
//         In the same pass the pivot rows stored so far (U rows 0 .. r-1) are reduced against this
//         panel's pivots, U[t, c:n] -= U[t, c + J] X (Gauss-Jordan), so they are already RREF rows
//         when the forward pass ends and no back substitution is left.
//     (e) X is stored as the next k rows of the U workspace (pivot order).
//   The rank rows of U are written over the job's matrix, then zero rows (RREF with zero rows last).
//
// Every value is exact; the RREF, its pivot columns and the smallest-row pivot_rows are unique
// functions of the input, so the result is bit-identical to the CPU engine's.
#include <algorithm>
#include <numeric>
#include <vector>

#include "hx_elim.h"
#include "hx_limbs.h"

namespace hx {
namespace {

constexpr int RT = 256;   // threads per job workgroup
constexpr int RW = 32;    // widest panel

struct RJob {
    uint32_t m, n;
    uint32_t* A;
    int64_t lda;
    uint32_t* pivots;
    uint32_t* pivot_rows;
    uint32_t* rank;
    uint32_t* U;          // min(m, n) x n, row stride n
    uint32_t panel;       // widest panel for this job (a tuning choice)
};

struct RArgs {
    const RJob* jobs;
    const uint32_t* order;
    uint32_t njobs;
    uint32_t* next;
    int* status;
    uint32_t region;      // LDS words for the panel / pivot-row block
    uint32_t mmax, nmax;
    DevMod M;
    LimbConsts lc;        // the limb split of the MFMA trailing update (hx_limbs.h)
    uint32_t half;        // (p - 1) / 2
    int L;                // limbs per residue (1..4)
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// (d) on the INT8 matrix cores, k <= 32 (one 32-deep MFMA k-step), for two row sets:
//   the unused rows u:        S[list[u], c + j] -= sum_{t<k} S[list[u], c + J_t] X[t, j],  j in [wc, nb)
//   the stored pivot rows t:  U[t, c + j]       -= sum_{t'<k} U[t, c + J_t'] X[t', j],  j in [0, nb)
// (the second set is the Gauss-Jordan half: the earlier pivot rows are reduced against this panel's
// pivots as the panel is made, so no back substitution is left at the end; X is the identity on
// the pivot columns, so those entries become exactly zero).  Both operands are split into L
// balanced 8-bit limbs on the fly (the X fragments from LDS, the row fragments from HBM), the L^2
// limb products accumulate per anti-diagonal in int32, and the exact recombination (hx_limbs.h)
// gives the residue subtracted from the entry.  A wave takes a 32-row block of either set, loads
// the block's k coefficients into registers before any write, then walks the 32-column blocks.
template <int L>
__device__ void trailing_mfma(const RJob& job, const RArgs& a, const uint32_t* B, bool planes, uint32_t nb,
                              uint32_t wc, uint32_t k, uint32_t c, uint32_t mu, const uint32_t* list,
                              const uint8_t* chosen, const uint32_t* s_q, uint32_t r) {
    constexpr int ND = 2 * L - 1;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 31, lg = lane >> 5;
    const uint32_t p = a.M.p, half = a.half;
    const uint32_t nrb_u = nb > wc ? (mu + 31) / 32 : 0, nrb = nrb_u + (r + 31) / 32;
    // the limb planes of 16 residues (one per k-byte) into L fragment registers
    auto split_into = [&](uint32_t x, int i, v4i (&f)[L]) {
        int32_t d[4] = {0, 0, 0, 0};
        limbs_of(x, p, half, L, d);
#pragma unroll
        for (int s = 0; s < L; ++s) f[s][i >> 2] |= (int)(((uint32_t)d[s] & 0xffu) << (8 * (i & 3)));
    };
    for (uint32_t rb = wave; rb < nrb; rb += RT / 64) {
        const bool urows = rb >= nrb_u;                       // a block of stored pivot rows
        const uint32_t b0 = (urows ? rb - nrb_u : rb) * 32, nrows = urows ? r : mu, jbeg = urows ? 0 : wc;
        uint32_t* base = urows ? job.U : job.A;
        const int64_t ld = urows ? (int64_t)job.n : job.lda;
        auto live_row = [&](uint32_t u) { return u < nrows && (urows || !chosen[u]); };
        // this lane's row of the block and its 16 k-values (k-bytes 16 lg .. 16 lg + 15)
        const uint32_t u = b0 + lr;
        const bool live = live_row(u);
        const int64_t rowoff = live ? (urows ? (int64_t)u : (int64_t)list[u]) * ld + c : 0;
        v4i fa[L];
#pragma unroll
        for (int s = 0; s < L; ++s) fa[s] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t t = 16 * lg + i;
            split_into((live && t < k) ? base[rowoff + s_q[t]] : 0u, i, fa);
        }
        // the 16 output rows this lane touches (accumulator element e), as matrix row indices;
        // NOROW for rows that take no update
        constexpr uint32_t NOROW = 0xffffffffu;
        uint32_t orow[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t ur = b0 + (e / 4) * 8 + lg * 4 + (e % 4);
            orow[e] = live_row(ur) ? (urows ? ur : list[ur]) : NOROW;
        }
        uint32_t* const basec = base + c;
        const uint32_t ncb = (nb - jbeg + 31) / 32;
        for (uint32_t cb = 0; cb < ncb; ++cb) {
            const uint32_t j = jbeg + cb * 32 + lr;
            // the block's entries are loaded before the MFMAs, so their latency overlaps them
            uint32_t cv[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) cv[e] = (j < nb && orow[e] != NOROW) ? basec[(int64_t)orow[e] * ld + j] : 0u;
            v4i fb[L];
#pragma unroll
            for (int s = 0; s < L; ++s) fb[s] = v4i{0, 0, 0, 0};
            if (planes) {
                if (j < nb) {
                    const uint8_t* P8 = (const uint8_t*)B + 32 * j + 16 * lg;
#pragma unroll
                    for (int s = 0; s < L; ++s) fb[s] = *(const v4i*)(P8 + (size_t)s * 32 * nb);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t t = 16 * lg + i;
                    split_into((j < nb && t < k) ? B[t * nb + j] : 0u, i, fb);
                }
            }
            v16i acc[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[d][e] = 0;
#pragma unroll
            for (int s = 0; s < L; ++s)
#pragma unroll
                for (int t = 0; t < L; ++t)
                    acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s], fb[t], acc[s + t], 0, 0, 0);
            if (j < nb) {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    if (orow[e] == NOROW) continue;
                    int32_t D[ND];
                    bool nzd = false;
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        D[d] = acc[d][e];
                        nzd |= D[d] != 0;
                    }
                    if (!nzd) continue;                           // a zero update leaves the entry
                    basec[(int64_t)orow[e] * ld + j] = subm(cv[e], limb_recombine<ND>(D, a.lc, a.M), p);
                }
            }
        }
    }
}

// row[q2] <- (v row[q2] - f prow[q2]) R^-1 for q2 in (q, wc), four columns at a time with their
// eight reads issued before the four writes: row and prow lie in the same LDS array, so a loop that
// stores each entry before the next loads would wait one LDS round trip per column.
template <class Ar>
__device__ __forceinline__ void elim_row(uint32_t* row, const uint32_t* prow, uint32_t q, uint32_t wc, uint32_t v,
                                         uint32_t fneg, const Ar& ar) {
    uint32_t q2 = q + 1;
    for (; q2 + 4 <= wc; q2 += 4) {
        uint32_t x[4], y[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            x[e] = row[q2 + e];
            y[e] = prow[q2 + e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) row[q2 + e] = ar.dfree(v, x[e], fneg, y[e]);
    }
    for (; q2 < wc; ++q2) row[q2] = ar.dfree(v, row[q2], fneg, prow[q2]);
}

// (c) X = S[P, J]^{-1} S[P, c:n] for the panel's k pivot rows P (pivot columns J = c + s_q).
//   c1: [K | I] with K = S[P, J] into LDS (two 32 x 64 buffers at the start of the panel region,
//       whose panel copy is dead by now);
//   c2: division-free Gauss-Jordan in the pivot order, row_i <- K_tt row_i - K_it row_t (i != t):
//       every leading principal block of K is non-singular in this order (each pivot was chosen
//       non-zero after eliminating the earlier ones), the left block ends diagonal, D, and the
//       right block is D K^{-1}; a row left alone (K_it = 0) keeps its own scale, which is fine,
//       since only the ratio right / left of a row is used.  One barrier per step (the buffers
//       alternate), then the k inverses of D at once;
//   c3: X = W S[P, c:n] on the INT8 matrix cores, W = D^{-1} (D K^{-1}) split into limbs once per
//       wave (all waves hold the same A fragment), S[P, c + j] gathered from HBM per 32-column
//       block; the exact recombination is X's entry.  X overwrites the region (after a barrier,
//       so every wave holds W first), k rows of nb words.
template <int L, class Ar>
__device__ void pivot_solve(const RJob& job, const RArgs& a, const Ar& ar, uint32_t* pan, uint32_t k, uint32_t c,
                            uint32_t nb, uint32_t r, bool planes, const uint32_t* s_P, const uint32_t* s_q,
                            uint32_t* s_inv) {
    constexpr int ND = 2 * L - 1, GS = 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 31, lg = lane >> 5;
    const uint32_t p = ar.p;
    uint32_t* G0 = pan;
    uint32_t* G1 = pan + 32 * GS;
    {   // K's k^2 <= 4 x 256 entries gathered with all loads in flight, then [K | I] stored
        const uint32_t kk = k * k;
        uint32_t v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t idx = e * RT + tid, i = idx / k, j = idx - i * k;
            v[e] = idx < kk ? job.A[(int64_t)s_P[i] * job.lda + c + s_q[j]] : 0u;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t idx = e * RT + tid, i = idx / k, j = idx - i * k;
            if (idx < kk) {
                G0[i * GS + j] = v[e];
                G0[i * GS + k + j] = i == j ? 1u : 0u;
            }
        }
    }
    __syncthreads();
    // Step t touches only the entries it can change.  Before step t a row i holds explicitly: the
    // left columns [t, k), the right columns [0, t), its own diagonal entry (i, i) when i < t and
    // its own identity entry (i, k + i) when i >= t; every other entry is zero (left columns < t
    // were eliminated, right columns > t are still identity columns).  A zero stays zero under
    // row_i <- v row_i - f row_t (row t is zero there too), so step t writes k + 1 entries per row:
    // left (t, k), right [0, t] (right column t read as zero in rows i != t) and the own entry
    // (row t is zero at both own positions of a row i != t).  Eight threads share a row.
    const uint32_t gi = tid >> 3, gsub = tid & 7;
    for (uint32_t t = 0; t < k; ++t) {
        const uint32_t* src = (t & 1) ? G1 : G0;
        uint32_t* dst = (t & 1) ? G0 : G1;
        if (gi < k) {
            const uint32_t* srow = src + gi * GS;
            const uint32_t* trow = src + t * GS;
            uint32_t* drow = dst + gi * GS;
            const uint32_t v = trow[t], f = srow[t], fneg = p - f, nl = k - t - 1;
            const bool upd = gi != t && f != 0;
            for (uint32_t sl = gsub; sl <= k; sl += 8) {
                const bool own = sl == k;
                const uint32_t col = sl < nl ? t + 1 + sl : (!own ? k + (sl - nl) : (gi <= t ? gi : k + gi));
                const uint32_t x = (col == k + t && gi != t) ? 0u : srow[col];
                const uint32_t y = own ? 0u : trow[col];
                drow[col] = upd ? ar.dfree(v, x, fneg, y) : x;
            }
        }
        __syncthreads();
    }
    const uint32_t* Gk = (k & 1) ? G1 : G0;
    if (tid < k) {
        const uint32_t iv = invm(Gk[tid * GS + tid], p);
        if (iv == 0) atomicExch(a.status, 1);
        s_inv[tid] = ar.prep(iv);
    }
    __syncthreads();
    auto split_into = [&](uint32_t x, int i, v4i (&f)[L]) {
        int32_t d[4] = {0, 0, 0, 0};
        limbs_of(x, p, a.half, L, d);
#pragma unroll
        for (int s = 0; s < L; ++s) f[s][i >> 2] |= (int)(((uint32_t)d[s] & 0xffu) << (8 * (i & 3)));
    };
    // W's fragment: row lr, k-bytes 16 lg .. 16 lg + 15
    v4i fa[L];
#pragma unroll
    for (int s = 0; s < L; ++s) fa[s] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t t2 = 16 * lg + i;
        split_into((lr < k && t2 < k) ? ar.mul(Gk[lr * GS + k + t2], s_inv[lr]) : 0u, i, fa);
    }
    __syncthreads();                                  // every wave holds W: the region may be overwritten
    uint32_t* B = pan;
    const uint32_t ncb = (nb + 31) / 32;
    for (uint32_t cb = wave; cb < ncb; cb += RT / 64) {
        const uint32_t j = cb * 32 + lr;
        v4i fb[L];
#pragma unroll
        for (int s = 0; s < L; ++s) fb[s] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t t2 = 16 * lg + i;
            split_into((j < nb && t2 < k) ? job.A[(int64_t)s_P[t2] * job.lda + c + j] : 0u, i, fb);
        }
        v16i acc[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[d][e] = 0;
#pragma unroll
        for (int s = 0; s < L; ++s)
#pragma unroll
            for (int t = 0; t < L; ++t)
                acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s], fb[t], acc[s + t], 0, 0, 0);
        if (j >= nb) continue;
        uint32_t word[4][L];                          // limb bytes of rows g*8 + lg*4 .. +3, per plane
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t t = (e / 4) * 8 + lg * 4 + (e % 4);
            int32_t D[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) D[d] = acc[d][e];
            const uint32_t x = t < k ? limb_recombine<ND>(D, a.lc, a.M) : 0u;
            if (t < k) job.U[(size_t)(r + t) * job.n + c + j] = x;
            if (planes) {
                int32_t dg[4] = {0, 0, 0, 0};
                limbs_of(x, p, a.half, L, dg);
#pragma unroll
                for (int sp = 0; sp < L; ++sp) {
                    const uint32_t byte = ((uint32_t)dg[sp] & 0xffu) << (8 * (e % 4));
                    word[e / 4][sp] = (e % 4) == 0 ? byte : (word[e / 4][sp] | byte);
                }
            } else if (t < k) {
                B[t * nb + j] = x;
            }
        }
        if (planes) {
            // plane sp, column j: 32 k-bytes (rows t = 0..31) at byte sp * 32 nb + 32 j + t
            uint8_t* P8 = (uint8_t*)B;
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int sp = 0; sp < L; ++sp)
                    *(uint32_t*)(P8 + (size_t)sp * 32 * nb + 32 * j + g * 8 + lg * 4) = word[g][sp];
        }
    }
    __syncthreads();
}

template <class Ar, int L>
__device__ void rref_one(const RJob& job, const Ar& ar, const RArgs& a, uint32_t* sm) {
    __shared__ uint32_t s_cand[2][RT / 64];
    __shared__ uint32_t s_P[RW], s_q[RW], s_inv[RW];
    __shared__ uint32_t s_keep[RT / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t m = job.m, n = job.n, p = ar.p;
    const int64_t lda = job.lda;
    uint32_t* pan = sm;
    uint32_t* list = sm + a.region;
    uint8_t* chosen = (uint8_t*)(list + a.mmax);
    if (m == 0 || n == 0) {
        if (tid == 0) *job.rank = 0;
        return;
    }
    uint32_t w = job.panel ? min((uint32_t)RW, job.panel) : (uint32_t)RW;
    while (w > 1 && ((uint64_t)m * (w | 1) > a.region || (uint64_t)w * n > a.region)) w >>= 1;
    for (uint32_t i = tid; i < m; i += RT) list[i] = i;
    __syncthreads();

    uint32_t mu = m, r = 0;
    for (uint32_t c = 0; c < n && mu > 0; c += w) {
        const uint32_t wc = min(w, n - c), ws = wc | 1u, nb = n - c;
        // (a) panel of the unused rows, pivot search by the smallest-row rule
        for (uint32_t i0 = 0; i0 < mu * wc; i0 += 8 * RT) {      // eight loads per thread in flight
            uint32_t v[8], at[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t idx = i0 + e * RT + tid;
                if (idx < mu * wc) {
                    const uint32_t u = wc == RW ? idx / RW : idx / wc, q = idx - u * wc;
                    v[e] = job.A[(int64_t)list[u] * lda + c + q];
                    at[e] = u * ws + q;
                }
            }
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (i0 + e * RT + tid < mu * wc) pan[at[e]] = v[e];
        }
        for (uint32_t u = tid; u < mu; u += RT) chosen[u] = 0;
        __syncthreads();
        uint32_t k = 0;
        for (uint32_t q = 0; q < wc; ++q) {
            // the wave's smallest candidate row: rounds of 256 rows in order, the lowest hit lane of
            // the first round with a hit (a wave-uniform loop)
            uint32_t best = 0xffffffffu;
            for (uint32_t u0 = 0; u0 < mu; u0 += RT) {
                const uint32_t u = u0 + tid;
                const uint64_t bal = __ballot(u < mu && !chosen[u] && pan[u * ws + q] != 0);
                if (bal) {
                    best = u0 + wave * 64 + (uint32_t)__builtin_ctzll(bal);
                    break;
                }
            }
            if (lane == 0) s_cand[q & 1][wave] = best;
            __syncthreads();
            const uint32_t bu = min(min(s_cand[q & 1][0], s_cand[q & 1][1]), min(s_cand[q & 1][2], s_cand[q & 1][3]));
            if (bu == 0xffffffffu) continue;
            const uint32_t* prow = pan + bu * ws;
            const uint32_t v = prow[q];
            for (uint32_t u = tid; u < mu; u += RT) {
                if (u == bu) { chosen[u] = 1; continue; }
                if (chosen[u]) continue;
                uint32_t* row = pan + u * ws;
                const uint32_t f = row[q];
                if (f == 0) continue;
                const uint32_t fneg = p - f;
                elim_row(row, prow, q, wc, v, fneg, ar);
            }
            if (tid == 0) { s_P[k] = list[bu]; s_q[k] = q; }
            ++k;
        }
        __syncthreads();
        if (k == 0) continue;

        // (c) X = W S[P, c:n] with W = S[P, J]^{-1}: the k x k pivot block is inverted by
        //     division-free Gauss-Jordan on [K | I] in LDS (one barrier per step), then the product
        //     runs on the matrix cores into LDS; X is exactly the identity at the pivot columns.
        uint32_t* B = pan;
        // X's limb planes (L x nb x 32 bytes) replace X in LDS when they fit (then the trailing
        // pass reads its B fragments directly); X itself goes straight to U rows r .. r+k-1
        const bool planes = (uint64_t)8 * L * nb <= a.region;
        pivot_solve<L>(job, a, ar, pan, k, c, nb, r, planes, s_P, s_q, s_inv);

        // (d) trailing update of the other unused rows, S[i, j] -= sum_t S[i, J_t] X[t, j], and the
        // reduction of the stored pivot rows 0 .. r-1 against this panel's pivots (MFMA)
        trailing_mfma<L>(job, a, B, planes, nb, wc, k, c, mu, list, chosen, s_q, r);
        // (e) this panel's pivot rows are X, already in U rows r .. r+k-1 from column c on (the
        // final copy writes their zeros left of c); pivots, pivot rows
        if (tid < k) {
            job.pivots[r + tid] = c + s_q[tid];
            job.pivot_rows[r + tid] = s_P[tid];
        }
        __syncthreads();
        // stable compaction of the unused list (chosen rows leave), 256 positions per round: a
        // ballot per wave gives each kept row its place (writes only ever move left, past rows
        // this round has already read)
        {
            uint32_t base = 0;
            for (uint32_t u0 = 0; u0 < mu; u0 += RT) {
                const uint32_t u = u0 + tid;
                const bool keep = u < mu && !chosen[u];
                const uint32_t val = keep ? list[u] : 0u;
                const uint64_t bal = __ballot(keep);
                if (lane == 0) s_keep[wave] = (uint32_t)__popcll(bal);
                __syncthreads();
                uint32_t off = base, tot = 0;
#pragma unroll
                for (uint32_t w = 0; w < RT / 64; ++w) {
                    off += w < wave ? s_keep[w] : 0u;
                    tot += s_keep[w];
                }
                if (keep) list[off + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = val;
                base += tot;
                __syncthreads();
            }
        }
        // U rows r .. r+k-1 start at column c: kept in the list's free tail (positions mu-k .. mu-1,
        // past the compacted list) as list[m - 1 - row]
        if (tid < k) list[m - 1 - (r + tid)] = c;
        mu -= k;
        r += k;
    }
    __syncthreads();

    // the RREF over the job's matrix: rank rows of U, then zero rows
    for (uint32_t i = wave; i < m; i += RT / 64) {          // a wave per row, four words per lane in flight
        uint32_t* Ai = job.A + (int64_t)i * lda;
        const uint32_t* Ui = job.U + (size_t)i * n;
        const uint32_t c0 = i < r ? list[m - 1 - i] : n;     // zeros left of the row's panel
        for (uint32_t j0 = lane; j0 < n; j0 += 256) {
            uint32_t v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = (j0 + 64 * e >= c0 && j0 + 64 * e < n) ? Ui[j0 + 64 * e] : 0u;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (j0 + 64 * e < n) Ai[j0 + 64 * e] = v[e];
        }
    }
    if (tid == 0) *job.rank = r;
    __syncthreads();
}

template <class Ar, int L>
__global__ __launch_bounds__(RT) __attribute__((amdgpu_waves_per_eu(2, 8))) void rref_batch_kernel(RArgs a, Ar ar) {
    extern __shared__ __attribute__((aligned(16))) uint32_t sm[];
    __shared__ uint32_t s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t jj = s_job;
        __syncthreads();
        if (jj >= a.njobs) break;
        rref_one<Ar, L>(a.jobs[a.order[jj]], ar, a, sm);
    }
}

__global__ void zero_u32_kernel(uint32_t* x, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

struct RBatch {
    hx_modctx ctx;
    uint32_t njobs = 0;
    RJob* d_jobs = nullptr;
    uint32_t* d_order = nullptr;
    uint32_t* d_counters = nullptr;     // [0] next job, [1] status
    uint32_t* d_U = nullptr;
    uint32_t region = 0, mmax = 0, nmax = 0;
    size_t lds = 0;
    uint32_t grid = 0;
    uint32_t per_cu = 1;                // workgroups per CU the LDS size was chosen for (1 or 2)
    uint32_t planes_out = 0;            // jobs whose first panel keeps X's limb planes outside LDS
};

constexpr size_t kLdsCap = 156 * 1024;

}  // namespace

// Shared with the single-matrix entry point.
int rref_batch_create_impl(const hx_modctx& ctx, const hx_rref_job* jobs, uint32_t njobs, int mode, void** out);
int rref_batch_run_impl(void* b, hipStream_t s);
void rref_batch_destroy_impl(void* b);

// Whether one m x n job runs well in one workgroup: at most 4096 rows, and an LDS region (after the
// row lists) that holds a panel of at least 16 columns -- the kernel halves its panel width until
// m (w | 1) and w n fit the region, so a wider row would run in panels of a few columns (a
// 600 x 37000 job: one column per panel, 37000 panels on one workgroup).  Such shapes take the
// whole-GPU path instead.
bool rref_batch_fits(uint32_t m, uint32_t n) {
    if (m > 4096) return false;
    const size_t m4 = (std::max<uint32_t>(m, 1) + 3) / 4 * 4, n4 = (std::max<uint32_t>(n, 1) + 3) / 4 * 4;
    const size_t fixed = m4 * 4 + m4 + 64;
    if (fixed + 4 * 64 >= kLdsCap) return false;
    const size_t cap_words = (kLdsCap - fixed) / 4;
    return m4 * 17 <= cap_words && 16 * n4 <= cap_words && (size_t)2 * 32 * 64 <= cap_words;
}

int rref_batch_create_impl(const hx_modctx& ctx, const hx_rref_job* jobs, uint32_t njobs, int mode, void** out) {
    HX_CHECK(ctx.p >= 2 && ctx.p < 0x80000000u, "rref: needs a prime 2 <= p < 2^31");
    HX_CHECK(mode == HX_RREF_BATCH_AUTO || mode == HX_RREF_BATCH_ONE_PER_CU || mode == HX_RREF_BATCH_TWO_PER_CU,
             "rref batch: mode must be HX_RREF_BATCH_AUTO, _ONE_PER_CU or _TWO_PER_CU");
    HX_CHECK(jobs != nullptr || njobs == 0, "jobs is NULL");
    RBatch* b = new RBatch();
    try {
        b->ctx = ctx;
        b->njobs = njobs;
        std::vector<RJob> hj(njobs);
        size_t ucount = 0;
        uint32_t mmax = 1, nmax = 1;
        for (uint32_t i = 0; i < njobs; ++i) {
            const hx_rref_job& j = jobs[i];
            HX_CHECK(j.m == 0 || j.n == 0 || (j.A && j.pivots && j.pivot_rows && j.rank), "rref job: NULL pointer");
            HX_CHECK(j.lda >= (int64_t)j.n, "rref job: lda < n");
            HX_CHECK(j.m <= 4096, "rref job: m > 4096 (use hx_rref_u32)");
            HX_CHECK(j.rank != nullptr, "rref job: rank is NULL");
            HX_CHECK(j.panel == 0 || j.panel == 8 || j.panel == 16 || j.panel == 32, "rref job: panel must be 0, 8, 16 or 32");
            hj[i] = RJob{j.m, j.n, j.A, j.lda, j.pivots, j.pivot_rows, j.rank, nullptr, j.panel};
            ucount += (size_t)std::min(j.m, j.n) * j.n;
            mmax = std::max(mmax, j.m);
            nmax = std::max(nmax, j.n);
        }
        b->mmax = (mmax + 3) / 4 * 4;
        b->nmax = (nmax + 3) / 4 * 4;
        const size_t fixed = (size_t)b->mmax * 4 + b->mmax + 64;      // the unused-row list and flags
        HX_CHECK(fixed + 4 * 64 < kLdsCap, "rref batch: job too tall for the LDS row lists");
        // the panel copy (m x (w|1)), the pivot rows' X (w x n), the two [K | I] buffers (2 x 32 x 64)
        size_t want = std::max<size_t>(std::max<size_t>((size_t)mmax * (RW | 1), (size_t)RW * nmax), 2 * 32 * 64);
        // Two workgroups per CU (half the LDS each) when every job still gets a panel of at least 16
        // columns there AND the batch holds enough work to keep both busy: each workgroup then runs
        // at ~0.57x the speed of one alone on its CU, so the largest job's own time grows while the
        // total throughput rises; with total work >= 512x the largest job's, the throughput term
        // dominates (measured on the 6667 echelon jobs of the mixed stream: 110 -> 103 ms).
        double wsum = 0.0, wmax = 0.0;
        for (const RJob& jj : hj) {
            const double w = (double)jj.m * jj.n * std::min(jj.m, jj.n);
            wsum += w;
            wmax = std::max(wmax, w);
        }
        const size_t half_words = kLdsCap / 2 > fixed ? (kLdsCap / 2 - fixed) / 4 : 0;
        const bool two_fits = (size_t)mmax * 17 <= half_words && (size_t)16 * nmax <= half_words &&
                              half_words >= 2 * 32 * 64;
        HX_CHECK(mode != HX_RREF_BATCH_TWO_PER_CU || two_fits, "rref batch: jobs too large for two workgroups per CU");
        const bool two_per_cu = mode == HX_RREF_BATCH_TWO_PER_CU ||
                                (mode == HX_RREF_BATCH_AUTO && two_fits && wsum >= 512.0 * wmax);
        b->per_cu = two_per_cu ? 2 : 1;
        const size_t cap_words = ((two_per_cu ? kLdsCap / 2 : kLdsCap) - fixed) / 4;
        b->region = (uint32_t)std::min(want, cap_words);
        HX_CHECK(b->region >= std::max<size_t>(std::max<size_t>(mmax + 1, nmax), 2 * 32 * 64),
                 "rref batch: shape exceeds the LDS panel");
        b->lds = (size_t)b->region * 4 + fixed;
        {
            // the kernel's own test (X's limb planes in LDS iff 8 L nb <= region, nb = n - c) at c = 0
            const uint32_t pp = ctx.p;
            const uint32_t L = pp <= 255u ? 1 : (pp <= 65279u ? 2 : (pp <= 16711423u ? 3 : 4));
            for (const RJob& jj : hj)
                if (jj.m && jj.n && (uint64_t)8 * L * jj.n > b->region) ++b->planes_out;
        }
        HX_HIP(hipMalloc(&b->d_U, std::max<size_t>(ucount, 1) * 4));
        size_t off = 0;
        for (uint32_t i = 0; i < njobs; ++i) {
            hj[i].U = b->d_U + off;
            off += (size_t)std::min(hj[i].m, hj[i].n) * hj[i].n;
        }
        std::vector<uint32_t> order(njobs);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            const double wx = (double)hj[x].m * hj[x].n * std::min(hj[x].m, hj[x].n);
            const double wy = (double)hj[y].m * hj[y].n * std::min(hj[y].m, hj[y].n);
            return wx > wy;
        });
        HX_HIP(hipMalloc(&b->d_jobs, std::max<size_t>(njobs, 1) * sizeof(RJob)));
        HX_HIP(hipMalloc(&b->d_order, std::max<size_t>(njobs, 1) * sizeof(uint32_t)));
        HX_HIP(hipMalloc(&b->d_counters, 2 * sizeof(uint32_t)));
        if (njobs) {
            HX_HIP(hipMemcpy(b->d_jobs, hj.data(), njobs * sizeof(RJob), hipMemcpyHostToDevice));
            HX_HIP(hipMemcpy(b->d_order, order.data(), njobs * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
        HX_HIP(hipMemset(b->d_counters, 0, 2 * sizeof(uint32_t)));
        int cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) cus = prop.multiProcessorCount;
        const uint32_t per_cu = std::max<uint32_t>(1, (uint32_t)(160 * 1024 / std::max<size_t>(b->lds, 1)));
        b->grid = std::max<uint32_t>(1, std::min<uint32_t>(njobs, (uint32_t)cus * std::min<uint32_t>(per_cu, 4)));
        static bool attr_done = false;
        if (!attr_done) {
            for (const void* f : {(const void*)rref_batch_kernel<ArMont, 1>, (const void*)rref_batch_kernel<ArMont, 2>,
                                  (const void*)rref_batch_kernel<ArMont, 3>, (const void*)rref_batch_kernel<ArMont, 4>,
                                  (const void*)rref_batch_kernel<ArBarrett, 1>, (const void*)rref_batch_kernel<ArBarrett, 2>,
                                  (const void*)rref_batch_kernel<ArBarrett, 3>, (const void*)rref_batch_kernel<ArBarrett, 4>})
                (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsCap);
            attr_done = true;
        }
    } catch (...) {
        rref_batch_destroy_impl(b);
        throw;
    }
    *out = b;
    return HX_OK;
}

int rref_batch_run_impl(void* vb, hipStream_t s) {
    RBatch* b = (RBatch*)vb;
    if (b->njobs == 0) return HX_OK;
    zero_u32_kernel<<<1, 64, 0, s>>>(b->d_counters, 2);   // job counter and status (a kernel, not a memset: capturable)
    after_launch("zero_u32_kernel");
    RArgs a;
    a.jobs = b->d_jobs;
    a.order = b->d_order;
    a.njobs = b->njobs;
    a.next = b->d_counters;
    a.status = (int*)(b->d_counters + 1);
    a.region = b->region;
    a.mmax = b->mmax;
    a.nmax = b->nmax;
    a.M = devmod(b->ctx);
    a.lc = limb_consts(b->ctx.p);
    a.half = (b->ctx.p - 1) / 2;
    const uint32_t pp = b->ctx.p;                       // limbs from p itself (balanced digits)
    a.L = pp <= 255u ? 1 : (pp <= 65279u ? 2 : (pp <= 16711423u ? 3 : 4));
    // one kernel per limb count (the MFMA fragments' size is a compile-time constant)
    auto launch = [&](auto ar) {
        using Ar = decltype(ar);
        switch (a.L) {
            case 1: rref_batch_kernel<Ar, 1><<<b->grid, RT, b->lds, s>>>(a, ar); break;
            case 2: rref_batch_kernel<Ar, 2><<<b->grid, RT, b->lds, s>>>(a, ar); break;
            case 3: rref_batch_kernel<Ar, 3><<<b->grid, RT, b->lds, s>>>(a, ar); break;
            default: rref_batch_kernel<Ar, 4><<<b->grid, RT, b->lds, s>>>(a, ar); break;
        }
    };
    if (b->ctx.p & 1u)
        launch(make_mont(b->ctx.p));
    else
        launch(make_barrett(b->ctx.p));
    after_launch("rref_batch_kernel");
    return HX_OK;
}

int rref_batch_status_impl(void* vb, hipStream_t s) {
    RBatch* b = (RBatch*)vb;
    uint32_t st = 0;
    HX_HIP(hipMemcpyAsync(&st, b->d_counters + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HX_HIP(hipStreamSynchronize(s));
    if (st != 0) throw Error{HX_E_NOT_INVERTIBLE, "a pivot had no inverse mod p (is p prime?)"};
    return HX_OK;
}

void rref_batch_destroy_impl(void* vb) {
    RBatch* b = (RBatch*)vb;
    if (!b) return;
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_order) (void)hipFree(b->d_order);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_U) (void)hipFree(b->d_U);
    delete b;
}

}  // namespace hx

using namespace hx;

extern "C" {

HX_API int hx_rref_batch_create(const hx_modctx* ctx, const hx_rref_job* jobs, uint32_t njobs, int mode,
                                hx_rref_batch* out) {
    return guard("hx_rref_batch_create", [&] {
        HX_CHECK(ctx != nullptr && out != nullptr, "NULL argument");
        void* b = nullptr;
        rref_batch_create_impl(*ctx, jobs, njobs, mode, &b);
        *out = b;
    });
}
HX_API int hx_rref_batch_info_get(hx_rref_batch vb, hx_rref_batch_info* out) {
    return guard("hx_rref_batch_info_get", [&] {
        HX_CHECK(vb != nullptr && out != nullptr, "NULL argument");
        const RBatch* b = (const RBatch*)vb;
        out->workgroups_per_cu = b->per_cu;
        out->grid = b->grid;
        out->lds_bytes = b->lds;
        out->region_words = b->region;
        out->jobs_planes_outside_lds = b->planes_out;
        out->njobs = b->njobs;
    });
}
HX_API int hx_rref_batch_run(hx_rref_batch b, hx_stream s) {
    return guard("hx_rref_batch_run", [&] {
        HX_CHECK(b != nullptr, "NULL batch");
        rref_batch_run_impl(b, (hipStream_t)s);
    });
}
HX_API int hx_rref_batch_status(hx_rref_batch b, hx_stream s) {
    return guard("hx_rref_batch_status", [&] {
        HX_CHECK(b != nullptr, "NULL batch");
        rref_batch_status_impl(b, (hipStream_t)s);
    });
}
HX_API int hx_rref_batch_destroy(hx_rref_batch b) {
    return guard("hx_rref_batch_destroy", [&] { rref_batch_destroy_impl(b); });
}

}  // extern "C"
