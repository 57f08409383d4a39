// hx_howell.hip -- the Howell form over Z/2^k (1 <= k <= 32) of many uneven matrices in ONE
// persistent launch (ring2k::howell, Binary/Ring2k.h:959-969; ENGINE_API.md Sec. 13.5).
//
// One 256-thread workgroup per job, jobs pulled from an atomic counter (largest first).  Per job the
// Sec. 13.5 pass runs on the rows of A (copied into the job's work rows), in native u32 wraparound
// (Z/2^k is the quotient of Z/2^32, Ring2k.h:17-22; only valuations, zero tests and the output look
// at the value mod 2^k):
//   per column c: the active row of MINIMAL valuation, ties to the topmost active row (active order =
//   input rows, then closure rows in creation order; a chosen row leaves the list); the pivot row is
//   scaled by the inverse of its unit part (Newton, to the full 32-bit width) so it leads with 2^v;
//   every other active row r gets r -= ((r[c] mod 2^k) >> v) * P (exact: v is minimal); when v > 0 the
//   closure row 2^(k-v) * P joins the active list at the bottom.
// Then reduce above, pivots left to right: H[i'] -= ((H[i'][c_i] mod 2^k) >> v_i) * H[i] for i' < i.
// The Howell form is unique (ENGINE_API.md Sec. 13.5), so this equals the reference's echelon-first
// route (echelon_blocked + howell_pass) entry for entry.
#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "hx_common.h"

namespace hx {
namespace {

constexpr int HT = 256;

struct HJob {
    uint32_t m, n;
    const uint32_t* A;
    int64_t lda;
    uint32_t* H;
    uint32_t* pivots;
    uint32_t* pivot_v;
    uint32_t* h;
    uint32_t* work;       // (m + n) x n rows, then (m + n) active ids, then n pivot row ids
};

struct HArgs {
    const HJob* jobs;
    const uint32_t* order;
    uint32_t njobs;
    uint32_t* next;
    uint32_t k;
};

__device__ __forceinline__ uint32_t inv_odd(uint32_t u) {
    uint32_t x = u;   // u * u == 1 mod 8: 3 correct bits, each Newton step doubles them
#pragma unroll
    for (int i = 0; i < 4; ++i) x = x * (2u - u * x);
    return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off));
    return x;
}

__device__ void howell_one(const HJob& job, uint32_t k) {
    __shared__ uint32_t s_red[2][HT / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t m = job.m, n = job.n;
    const uint32_t mask = k >= 32 ? 0xffffffffu : ((1u << k) - 1u);
    if (n == 0) {
        if (tid == 0) *job.h = 0;
        return;
    }
    uint32_t* rows = job.work;                               // row id i at rows + i * n
    uint32_t* act = job.work + (size_t)(m + n) * n;          // active order: row ids, or ~0u once chosen
    uint32_t* prow = act + (m + n);                          // pivot row ids in pivot order
    for (uint64_t idx = tid; idx < (uint64_t)m * n; idx += HT) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        rows[(size_t)i * n + j] = job.A[(int64_t)i * job.lda + j] & mask;
    }
    for (uint32_t i = tid; i < m; i += HT) act[i] = i;
    __syncthreads();
    uint32_t na = m, h = 0, ncl = 0;
    for (uint32_t c = 0; c < n; ++c) {
        // 1. minimal valuation, ties to the topmost active row: key = v << 24 | position
        uint32_t best = 0xffffffffu;
        for (uint32_t pos = tid; pos < na; pos += HT) {
            const uint32_t id = act[pos];
            if (id == 0xffffffffu) continue;
            const uint32_t x = rows[(size_t)id * n + c] & mask;
            if (x == 0) continue;
            const uint32_t v = (uint32_t)__builtin_ctz(x);
            best = min(best, (v << 24) | pos);
        }
        best = wave_min_u32(best);
        if (lane == 0) s_red[c & 1][wave] = best;
        __syncthreads();
        best = min(min(s_red[c & 1][0], s_red[c & 1][1]), min(s_red[c & 1][2], s_red[c & 1][3]));
        if (best == 0xffffffffu) continue;
        const uint32_t bpos = best & 0xffffffu, v = best >> 24;
        const uint32_t pid = act[bpos];
        uint32_t* P = rows + (size_t)pid * n;
        // 2. scale by the inverse of the unit part (P[c] == 2^v afterwards)
        const uint32_t ui = inv_odd(P[c] >> v);
        __syncthreads();                                     // everyone has read act[bpos] and P[c]
        for (uint32_t j = c + tid; j < n; j += HT) P[j] = (P[j] * ui) & mask;
        if (tid == 0) act[bpos] = 0xffffffffu;
        __syncthreads();
        // 3. eliminate every other active row (one wave per row, lanes along the row)
        for (uint32_t pos = wave; pos < na; pos += HT / 64) {
            const uint32_t id = act[pos];
            if (id == 0xffffffffu) continue;
            uint32_t* R = rows + (size_t)id * n;
            const uint32_t x = R[c] & mask;
            if (x == 0) continue;
            const uint32_t q = x >> v;
            for (uint32_t j = c + lane; j < n; j += 64) R[j] -= q * P[j];
        }
        // 4. the closure row 2^(k-v) P at the bottom of the active list
        if (v > 0) {
            const uint32_t cid = m + ncl;
            uint32_t* Z = rows + (size_t)cid * n;
            const uint32_t sh = k - v;
            for (uint32_t j = tid; j < n; j += HT) Z[j] = j <= c ? 0u : (sh >= 32 ? 0u : (P[j] << sh));
            if (tid == 0) act[na] = cid;
            ++na;
            ++ncl;
        }
        if (tid == 0) {
            prow[h] = pid;
            job.pivots[h] = c;
            job.pivot_v[h] = v;
        }
        ++h;
        __syncthreads();
    }
    // H rows in pivot order (masked), then reduce above, pivots left to right
    for (uint64_t idx = tid; idx < (uint64_t)h * n; idx += HT) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        job.H[(size_t)i * n + j] = rows[(size_t)prow[i] * n + j] & mask;
    }
    __syncthreads();
    for (uint32_t i = 1; i < h; ++i) {
        const uint32_t ci = job.pivots[i], vi = job.pivot_v[i];
        const uint32_t* Hi = job.H + (size_t)i * n;
        for (uint32_t ip = wave; ip < i; ip += HT / 64) {
            uint32_t* Hp = job.H + (size_t)ip * n;
            const uint32_t q = (Hp[ci] & mask) >> vi;
            if (q == 0) continue;
            for (uint32_t j = ci + lane; j < n; j += 64) Hp[j] = (Hp[j] - q * Hi[j]) & mask;
        }
        __syncthreads();
    }
    if (tid == 0) *job.h = h;
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Panel-blocked pass (jobs with m + n <= kBlockedRows).  The same pass, reordered:
//  * per panel of HW columns, the steps of the pass run on the panel columns only, in LDS
//    (pan: the active rows' panel entries), recording for every active row the multiplier q of
//    every step (qm) -- LU's L, with the closure rows inside it: a closure row made at step s
//    has trailing part 2^(k-v_s) T_s, i.e. the row "0 - (-(2^(k-v_s))) T_s", so it starts from
//    zero with qm = -2^(k-v_s) at step s and then collects its own later multipliers;
//  * then the trailing columns are brought up to date in one sweep, one thread per column:
//    the panel's pivot rows in step order, T_t = ui_t (row_t - sum_{s<t} qm[row_t][s] T_s)
//    (kept in registers), and every other active row gets row -= sum_t qm[row][t] T_t.
//  Every trailing entry is read and written once per panel instead of once per column.
//  Reduce-above is blocked the same way: per block of HW pivots (left to right) the
//  multipliers of every row above come from a small triangular recurrence on the block's
//  pivot columns, then one sweep over the columns applies the block (block rows are read
//  before they are updated, so the update uses their values from before the block, as the
//  column-at-a-time order does).
// ------------------------------------------------------------------------------------------
constexpr uint32_t HW = 16;            // panel width / reduce-above block
constexpr uint32_t kBlockedRows = 1024;
constexpr uint32_t NONE = 0xffffffffu;

// Panel entries and multipliers are stored in elements of `eb` bytes: 2 when k <= 16 (every stored
// value is reduced mod 2^k), else 4; the halved store lets the largest job class (m + n <= 1024)
// run two workgroups per CU.  The panel region holds at least HW x HT words: during the trailing
// sweep it holds the panel's pivot rows over HT columns instead (the panel entries are dead then).
__host__ __device__ constexpr uint32_t pan_rows(uint32_t rc, uint32_t eb) {
    return rc * eb >= HT * 4u ? rc : HT * 4u / eb;
}
__host__ __device__ constexpr size_t blocked_lds_bytes(uint32_t rc, uint32_t eb) {
    return (size_t)pan_rows(rc, eb) * HW * eb + (size_t)rc * HW * eb + (size_t)rc * 4 * 2 + rc;
}

// Multiplication of the pass mod 2^32 (only the value mod 2^k matters).  For k <= 24 every
// multiplier operand is kept reduced mod 2^k (< 2^24), and the full-rate 24-bit multiply gives
// the exact low 32 bits of the product; otherwise the quarter-rate 32-bit multiply.
template <bool SMALLK>
__device__ __forceinline__ uint32_t kmul(uint32_t a, uint32_t b) {
    if constexpr (SMALLK) return __umul24(a, b);
    else return a * b;
}

template <bool SMALLK, class PT>
__device__ void howell_blocked(const HJob& job, uint32_t k, uint32_t rc, uint32_t* lds) {
    __shared__ uint32_t s_red[2][HT / 64];
    __shared__ uint32_t s_pid[HW], s_ui[HW], s_v[HW], s_c[HW];
    __shared__ uint32_t s_hb[HW * HW];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t m = job.m, n = job.n;
    const uint32_t mask = k >= 32 ? 0xffffffffu : ((1u << k) - 1u);
    if (n == 0) {
        if (tid == 0) *job.h = 0;
        return;
    }
    PT* pan = (PT*)lds;                        // [rc][HW] panel entries of active rows
    uint32_t* s_T = lds;                       // [HW][HT] the panel's pivot rows over HT trailing columns
                                               // (phase B only, when the panel entries are dead)
    PT* qm = pan + (size_t)pan_rows(rc, sizeof(PT)) * HW;   // [rc][HW] multipliers (later: reduce-above q's)
    uint32_t* act = (uint32_t*)(qm + (size_t)rc * HW);     // active order: row ids, NONE once chosen
    uint32_t* prow = act + rc;                 // pivot row ids in pivot order
    uint8_t* nz = (uint8_t*)(prow + rc);       // row has a non-zero multiplier in this panel
    uint32_t* rows = job.work;                 // row id i at rows + i * n
    for (uint32_t i = wave; i < m; i += HT / 64) {     // a wave per row, four words per lane in flight
        const uint32_t* Ai = job.A + (int64_t)i * job.lda;
        uint32_t* Ri = rows + (size_t)i * n;
        for (uint32_t j0 = lane; j0 < n; j0 += 256) {
            uint32_t v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = j0 + 64 * e < n ? Ai[j0 + 64 * e] : 0u;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (j0 + 64 * e < n) Ri[j0 + 64 * e] = v[e] & mask;
        }
    }
    for (uint32_t i = tid; i < m; i += HT) act[i] = i;
    __syncthreads();
    uint32_t na = m, h = 0, ncl = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += HW) {
        const uint32_t wc = min(HW, n - c0), c1 = c0 + wc;
        for (uint32_t idx = tid; idx < na * HW; idx += HT) {
            const uint32_t pos = idx / HW, t = idx % HW, id = act[pos];
            if (id == NONE) continue;
            pan[id * HW + t] = t < wc ? rows[(size_t)id * n + c0 + t] & mask : 0u;
            qm[id * HW + t] = 0;
            if (t == 0) nz[id] = 0;
        }
        if (tid < HW) s_pid[tid] = NONE;
        const uint32_t ncl0 = ncl;
        __syncthreads();
        // (A) the pass on the panel columns.  Row position pos belongs to thread pos % HT for the
        // whole panel (scan, elimination, closure rows appended there, the chosen mark), and the
        // pivot row is left unscaled (the eliminating multiplier is q u^-1 instead), so the only
        // exchange per step is the pivot search: one barrier per step.
        uint32_t chosen = 0;                                   // bit j: position tid + j HT picked
        for (uint32_t t = 0; t < wc; ++t) {
            uint32_t best = NONE;
            for (uint32_t pos = tid, j = 0; pos < na; pos += HT, ++j) {
                const uint32_t id = act[pos];                  // NONE: chosen in an earlier panel
                if (id == NONE || (chosen >> j & 1u)) continue;
                const uint32_t x = pan[id * HW + t] & mask;
                if (x != 0) best = min(best, ((uint32_t)__builtin_ctz(x) << 24) | pos);
            }
            best = wave_min_u32(best);
            if (lane == 0) s_red[t & 1][wave] = best;
            __syncthreads();
            best = min(min(s_red[t & 1][0], s_red[t & 1][1]), min(s_red[t & 1][2], s_red[t & 1][3]));
            if (best == NONE) continue;
            const uint32_t bpos = best & 0xffffffu, v = best >> 24, pid = act[bpos];
            const PT* P = pan + pid * HW;                      // unscaled: P[t] = u 2^v
            const uint32_t ui = inv_odd((P[t] & mask) >> v) & mask;
            if (bpos % HT == tid) chosen |= 1u << (bpos / HT);
            if (tid == 0) {
                s_pid[t] = pid;
                s_ui[t] = ui;
                prow[h] = pid;
                job.pivots[h] = c0 + t;
                job.pivot_v[h] = v;
            }
            for (uint32_t pos = tid, j = 0; pos < na; pos += HT, ++j) {
                const uint32_t id = act[pos];
                if (id == NONE || (chosen >> j & 1u)) continue;
                PT* R = pan + id * HW;
                const uint32_t q = (R[t] & mask) >> v;
                if (q == 0) continue;
                qm[id * HW + t] = q;                           // w.r.t. the scaled pivot row
                nz[id] = 1;
                const uint32_t qe = kmul<SMALLK>(q, ui) & mask;
                uint32_t t2 = t;                               // four columns per round, reads first
                for (; t2 + 4 <= wc; t2 += 4) {
                    uint32_t rv[4], pv[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        rv[e] = R[t2 + e];
                        pv[e] = P[t2 + e];
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) R[t2 + e] = (rv[e] - kmul<SMALLK>(qe, pv[e])) & mask;
                }
                for (; t2 < wc; ++t2) R[t2] = (R[t2] - kmul<SMALLK>(qe, P[t2])) & mask;
            }
            if (v > 0) {                                       // closure row 2^(k-v) u^-1 P at the bottom
                const uint32_t cid = m + ncl, sh = k - v;      // 1 <= sh <= 31
                if (na % HT == tid) {
                    act[na] = cid;
                    nz[cid] = 1;
                    for (uint32_t t2 = 0; t2 < HW; ++t2) {
                        pan[cid * HW + t2] = (t2 <= t || t2 >= wc) ? 0u : (kmul<SMALLK>(P[t2], ui) << sh) & mask;
                        qm[cid * HW + t2] = t2 == t ? (0u - (1u << sh)) & mask : 0u;
                    }
                }
                ++na;
                ++ncl;
            }
            ++h;
        }
        __syncthreads();
        for (uint32_t pos = tid, j = 0; pos < na; pos += HT, ++j)
            if (chosen >> j & 1u) act[pos] = NONE;
        // the panel's pivot rows keep their panel entries; new closure rows start from zero
        for (uint32_t idx = tid; idx < HW * HW; idx += HT) {
            const uint32_t t = idx / HW, t2 = idx % HW, pid = s_pid[t];
            if (pid != NONE && t2 < wc) rows[(size_t)pid * n + c0 + t2] = kmul<SMALLK>(pan[pid * HW + t2], s_ui[t]) & mask;
        }
        const uint32_t tw = n - c1;
        for (uint64_t idx = tid; idx < (uint64_t)(ncl - ncl0) * tw; idx += HT)
            rows[(size_t)(m + ncl0 + idx / tw) * n + c1 + idx % tw] = 0;
        __syncthreads();
        // (B) the trailing columns, 256 at a time: (B1) one thread per column brings the panel's
        // pivot rows up to date (T_t, in step order) into s_T and back to their rows; (B2) every
        // wave sweeps a quarter of the active rows over the 256 columns, 4 columns per lane (64
        // apart: coalesced) and 8 rows at a time, so 32 independent loads are in flight per lane.
        for (uint32_t j0 = c1; j0 < n; j0 += HT) {
            const uint32_t j = j0 + tid;
            if (j < n) {
                // all HW pivot-row entries are loaded before any is written back (the rows are
                // distinct), so the loads are in flight together instead of one after another
                uint32_t T[HW];
#pragma unroll
                for (uint32_t t = 0; t < HW; ++t) {
                    const uint32_t pid = s_pid[t];
                    T[t] = pid == NONE ? 0u : rows[(size_t)pid * n + j];
                }
#pragma unroll
                for (uint32_t t = 0; t < HW; ++t) {
                    const uint32_t pid = s_pid[t];
                    if (pid == NONE) { T[t] = 0; continue; }
                    uint32_t x = T[t];
                    const PT* qr = qm + pid * HW;
#pragma unroll
                    for (uint32_t s2 = 0; s2 < t; ++s2) x -= kmul<SMALLK>(qr[s2], T[s2]);
                    T[t] = kmul<SMALLK>(x & mask, s_ui[t] & mask) & mask;
                }
#pragma unroll
                for (uint32_t t = 0; t < HW; ++t) {
                    const uint32_t pid = s_pid[t];
                    if (pid != NONE) rows[(size_t)pid * n + j] = T[t];
                }
#pragma unroll
                for (uint32_t t = 0; t < HW; ++t) s_T[t * HT + tid] = T[t];
            }
            __syncthreads();
            uint32_t Tr[HW][4];
#pragma unroll
            for (uint32_t t = 0; t < HW; ++t)
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) Tr[t][u] = s_T[t * HT + lane + 64 * u];
            bool in[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) in[u] = j0 + lane + 64 * u < n;
            // groups of 8 active rows per wave; the next group's ids and entries are loaded before
            // this group's entries are written back (different rows)
            uint32_t ids[8], xs[8][4];
            auto load_group = [&](uint32_t pos, uint32_t (&gid)[8], uint32_t (&gx)[8][4]) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const uint32_t id = pos + r < na ? act[pos + r] : NONE;
                    gid[r] = (id != NONE && nz[id]) ? id : NONE;
                }
#pragma unroll
                for (int r = 0; r < 8; ++r)
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        gx[r][u] = (gid[r] != NONE && in[u]) ? rows[(size_t)gid[r] * n + j0 + lane + 64 * u] : 0u;
            };
            if (wave * 8 < na) load_group(wave * 8, ids, xs);
            for (uint32_t pos = wave * 8; pos < na; pos += 32) {
                uint32_t nids[8], nxs[8][4];
                if (pos + 32 < na) load_group(pos + 32, nids, nxs);
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    if (ids[r] == NONE) continue;
                    const PT* qr = qm + ids[r] * HW;
                    uint32_t q[HW];
#pragma unroll
                    for (uint32_t t = 0; t < HW; ++t) q[t] = qr[t];
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        if (!in[u]) continue;
                        uint32_t x = xs[r][u];
#pragma unroll
                        for (uint32_t t = 0; t < HW; ++t) x -= kmul<SMALLK>(q[t], Tr[t][u]);
                        rows[(size_t)ids[r] * n + j0 + lane + 64 * u] = x;
                    }
                }
                if (pos + 32 < na) {
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        ids[r] = nids[r];
#pragma unroll
                        for (uint32_t u = 0; u < 4; ++u) xs[r][u] = nxs[r][u];
                    }
                }
            }
            __syncthreads();
        }
    }
    // H rows in pivot order (masked).  Left of its pivot column a pivot row is zero mod 2^k; its
    // stored entries there may be stale (panel entries are written back for pivot rows only).
    for (uint32_t i = wave; i < h; i += HT / 64) {     // a wave per row, four words per lane in flight
        const uint32_t* Ri = rows + (size_t)prow[i] * n;
        uint32_t* Hi = job.H + (size_t)i * n;
        const uint32_t pc = job.pivots[i];
        for (uint32_t j0 = lane; j0 < n; j0 += 256) {
            uint32_t v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t j = j0 + 64 * e;
                v[e] = (j < n && j >= pc) ? Ri[j] & mask : 0u;
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (j0 + 64 * e < n) Hi[j0 + 64 * e] = v[e];
        }
    }
    __syncthreads();
    // blocked reduce above
    for (uint32_t i0 = 0; i0 < h; i0 += HW) {
        const uint32_t bw = min(HW, h - i0), i1 = i0 + bw;
        if (tid < HW) {
            s_c[tid] = tid < bw ? job.pivots[i0 + tid] : 0u;
            s_v[tid] = tid < bw ? job.pivot_v[i0 + tid] : 0u;
        }
        __syncthreads();
        for (uint32_t idx = tid; idx < HW * HW; idx += HT) {
            const uint32_t t = idx / HW, t2 = idx % HW;
            s_hb[idx] = (t < bw && t2 < bw) ? job.H[(size_t)(i0 + t) * n + s_c[t2]] : 0u;
        }
        __syncthreads();
        for (uint32_t ip = tid; ip < i1; ip += HT) {
            uint32_t cur[HW];
#pragma unroll
            for (uint32_t t2 = 0; t2 < HW; ++t2) cur[t2] = t2 < bw ? job.H[(size_t)ip * n + s_c[t2]] : 0u;
#pragma unroll 1
            for (uint32_t t = 0; t < HW; ++t) {
                uint32_t q = 0;
                if (t < bw && i0 + t > ip) q = (cur[t] & mask) >> s_v[t];
#pragma unroll
                for (uint32_t t2 = t + 1; t2 < HW; ++t2) cur[t2] -= kmul<SMALLK>(q, s_hb[t * HW + t2]);
                qm[ip * HW + t] = q;
            }
        }
        __syncthreads();
        for (uint32_t j = s_c[0] + tid; j < n; j += HT) {
            uint32_t Hr[HW];
#pragma unroll
            for (uint32_t t = 0; t < HW; ++t) Hr[t] = t < bw ? job.H[(size_t)(i0 + t) * n + j] : 0u;
            // rows in groups of 8; the next group's entries are loaded before this group's are
            // written back (different rows), so a load's latency overlaps the previous group
            uint32_t xs[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) xs[u] = u < i1 ? job.H[(size_t)u * n + j] : 0u;
            for (uint32_t ip0 = 0; ip0 < i1; ip0 += 8) {
                uint32_t xn[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u)
                    xn[u] = ip0 + 8 + u < i1 ? job.H[(size_t)(ip0 + 8 + u) * n + j] : 0u;
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    if (ip0 + u >= i1) continue;
                    const PT* qr = qm + (ip0 + u) * HW;
                    uint32_t x = xs[u];
#pragma unroll
                    for (uint32_t t = 0; t < HW; ++t) x -= kmul<SMALLK>(qr[t], Hr[t]);
                    job.H[(size_t)(ip0 + u) * n + j] = x & mask;
                }
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) xs[u] = xn[u];
            }
        }
        __syncthreads();
    }
    if (tid == 0) *job.h = h;
    __syncthreads();
}

// Persistent: one workgroup per job of this row-capacity class, jobs from the class's counter.
template <bool SMALLK, class PT>
__global__ __launch_bounds__(HT) void howell_blocked_kernel(HArgs a, uint32_t rc) {
    extern __shared__ __attribute__((aligned(16))) uint32_t hlds[];
    __shared__ uint32_t s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t jj = s_job;
        __syncthreads();
        if (jj >= a.njobs) break;
        howell_blocked<SMALLK, PT>(a.jobs[a.order[jj]], a.k, rc, hlds);
    }
}

// The blocked kernel for this k: 16-bit stores for k <= 16, the 24-bit multiply for k <= 24.
inline const void* blocked_kernel_for(uint32_t k) {
    if (k <= 16) return (const void*)howell_blocked_kernel<true, uint16_t>;
    if (k <= 24) return (const void*)howell_blocked_kernel<true, uint32_t>;
    return (const void*)howell_blocked_kernel<false, uint32_t>;
}
inline uint32_t blocked_elt_bytes(uint32_t k) { return k <= 16 ? 2u : 4u; }

__global__ __launch_bounds__(HT) void howell_batch_kernel(HArgs a) {
    __shared__ uint32_t s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t jj = s_job;
        __syncthreads();
        if (jj >= a.njobs) break;
        howell_one(a.jobs[a.order[jj]], a.k);
    }
}

// ------------------------------------------------------------------------------------------
// Z/ell^k for an odd prime ell, q = ell^k < 2^32 (elim::howell, DenseElim.cpp:1450 -> howell_rows
// :1130-1300).  The same pass on canonical residues in [0, q): the valuation by the divisibility test
// of an odd divisor (x is a multiple of ell iff x ell^-1 mod 2^32 <= (2^32 - 1) / ell, and then x / ell
// is that product), the exact quotients x / ell^v as x (ell^-1)^v mod 2^32, the unit part's inverse
// mod q by Euclid, the products mod q by Barrett (DevMod), the closure row ell^(k-v) P, and reduce
// above with floor(H[i'][c_i] / ell^v_i).
// ------------------------------------------------------------------------------------------
struct EllCtx {
    uint32_t ell, k, q;
    uint32_t ellinv;                     // ell^-1 mod 2^32
    uint32_t lim;                        // (2^32 - 1) / ell
    uint32_t pw[33];                     // ell^v, v <= k
    uint32_t ipw[33];                    // (ell^-1)^v mod 2^32
    DevMod M;                            // Barrett constants of q
};

__device__ __forceinline__ uint32_t ell_val(uint32_t x, const EllCtx& e) {   // x != 0, canonical
    uint32_t v = 0;
    while (v < e.k) {
        const uint32_t t = x * e.ellinv;
        if (t > e.lim) break;
        x = t;
        ++v;
    }
    return v;
}

// ET threads per job: the largest jobs (m + n up to 1024 rows) set the batch's makespan, and their
// row eliminations spread over ET / 64 waves.
constexpr int ET = 1024;

__device__ void howell_ell_one(const HJob& job, const EllCtx& e) {
    __shared__ uint32_t s_red[2][ET / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t m = job.m, n = job.n, q = e.q;
    if (n == 0) {
        if (tid == 0) *job.h = 0;
        return;
    }
    uint32_t* rows = job.work;                               // row id i at rows + i * n
    uint32_t* act = job.work + (size_t)(m + n) * n;          // active order: row ids, or ~0u once chosen
    uint32_t* prow = act + (m + n);                          // pivot row ids in pivot order
    for (uint64_t idx = tid; idx < (uint64_t)m * n; idx += ET) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        rows[(size_t)i * n + j] = red64(job.A[(int64_t)i * job.lda + j], e.M);
    }
    for (uint32_t i = tid; i < m; i += ET) act[i] = i;
    __syncthreads();
    uint32_t na = m, h = 0, ncl = 0;
    for (uint32_t c = 0; c < n; ++c) {
        // 1. minimal valuation, ties to the topmost active row: key = v << 24 | position
        uint32_t best = 0xffffffffu;
        for (uint32_t pos = tid; pos < na; pos += ET) {
            const uint32_t id = act[pos];
            if (id == 0xffffffffu) continue;
            const uint32_t x = rows[(size_t)id * n + c];
            if (x == 0) continue;
            best = min(best, (ell_val(x, e) << 24) | pos);
        }
        best = wave_min_u32(best);
        if (lane == 0) s_red[c & 1][wave] = best;
        __syncthreads();
        best = s_red[c & 1][0];
#pragma unroll
        for (int w2 = 1; w2 < ET / 64; ++w2) best = min(best, s_red[c & 1][w2]);
        if (best == 0xffffffffu) continue;
        const uint32_t bpos = best & 0xffffffu, v = best >> 24;
        const uint32_t pid = act[bpos];
        uint32_t* P = rows + (size_t)pid * n;
        // 2. scale by the inverse of the unit part (P[c] == ell^v afterwards)
        const uint32_t ui = invm(P[c] * e.ipw[v], q);
        __syncthreads();                                     // everyone has read act[bpos] and P[c]
        for (uint32_t j = c + tid; j < n; j += ET) P[j] = mulm(P[j], ui, e.M);
        if (tid == 0) act[bpos] = 0xffffffffu;
        __syncthreads();
        // 3. eliminate every other active row exactly: one wave per group of four rows, lanes along the
        //    rows, the four rows' loads in flight together (the rows are independent given P)
        for (uint32_t pos0 = 4 * wave; pos0 < na; pos0 += 4 * (ET / 64)) {
            uint32_t* R[4];
            uint32_t qf[4];
            bool live[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t id = pos0 + u < na ? act[pos0 + u] : 0xffffffffu;
                live[u] = id != 0xffffffffu;
                R[u] = rows + (size_t)(live[u] ? id : pid) * n;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t x = live[u] ? R[u][c] : 0u;
                live[u] = x != 0;
                qf[u] = x * e.ipw[v];                        // x / ell^v, exact
            }
            if (!(live[0] || live[1] || live[2] || live[3])) continue;
            for (uint32_t j = c + lane; j < n; j += 64) {
                const uint32_t pj = P[j];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (live[u]) R[u][j] = subm(R[u][j], mulm(qf[u], pj, e.M), q);
            }
        }
        // 4. the closure row ell^(k-v) P at the bottom of the active list
        if (v > 0) {
            const uint32_t cid = m + ncl, f = e.pw[e.k - v];
            uint32_t* Z = rows + (size_t)cid * n;
            for (uint32_t j = tid; j < n; j += ET) Z[j] = j <= c ? 0u : mulm(P[j], f, e.M);
            if (tid == 0) act[na] = cid;
            ++na;
            ++ncl;
        }
        if (tid == 0) {
            prow[h] = pid;
            job.pivots[h] = c;
            job.pivot_v[h] = v;
        }
        ++h;
        __syncthreads();
    }
    for (uint64_t idx = tid; idx < (uint64_t)h * n; idx += ET) {
        const uint32_t i = (uint32_t)(idx / n), j = (uint32_t)(idx % n);
        job.H[(size_t)i * n + j] = rows[(size_t)prow[i] * n + j];
    }
    __syncthreads();
    for (uint32_t i = 1; i < h; ++i) {
        const uint32_t ci = job.pivots[i], d = e.pw[job.pivot_v[i]];
        const uint32_t* Hi = job.H + (size_t)i * n;
        for (uint32_t ip = wave; ip < i; ip += ET / 64) {
            uint32_t* Hp = job.H + (size_t)ip * n;
            const uint32_t qq = Hp[ci] / d;
            if (qq == 0) continue;
            for (uint32_t j = ci + lane; j < n; j += 64) Hp[j] = subm(Hp[j], mulm(qq, Hi[j], e.M), q);
        }
        __syncthreads();
    }
    if (tid == 0) *job.h = h;
    __syncthreads();
}

__global__ __launch_bounds__(ET, 8) void howell_ell_kernel(HArgs a, EllCtx e) {      // two jobs per CU (64 VGPRs)
    __shared__ uint32_t s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next, 1u);
        __syncthreads();
        const uint32_t jj = s_job;
        __syncthreads();
        if (jj >= a.njobs) break;
        howell_ell_one(a.jobs[a.order[jj]], e);
    }
}

__global__ void zero_counters_kernel(uint32_t* x, uint32_t n) {
    if (threadIdx.x < n) x[threadIdx.x] = 0;
}

// Job classes by row capacity m + n: three blocked classes (LDS sized to the class) and the
// unblocked kernel for the rest.  Each class is one persistent launch with its own counter.
constexpr int NCLASS = 4;
constexpr uint32_t kClassRows[NCLASS - 1] = {256, 512, kBlockedRows};

struct HClass {
    uint32_t njobs = 0, grid = 0, rc = 0;
    size_t lds = 0;
    uint32_t* d_order = nullptr;
};

struct HBatch {
    uint32_t k = 0, njobs = 0;
    uint32_t ell = 2;                    // odd: the Z/ell^k kernel for every job
    EllCtx ectx{};
    HJob* d_jobs = nullptr;
    uint32_t* d_next = nullptr;          // one counter per class
    HClass cls[NCLASS];
    // the two smaller classes run on a second stream beside the larger ones (forked from and joined
    // back into the caller's stream by events, so a capture on the caller's stream records both)
    hipStream_t s2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

void destroy(HBatch* b) {
    if (!b) return;
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_next) (void)hipFree(b->d_next);
    for (HClass& c : b->cls)
        if (c.d_order) (void)hipFree(c.d_order);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->s2) (void)hipStreamDestroy(b->s2);
    delete b;
}

EllCtx ell_ctx(uint32_t ell, uint32_t k) {
    EllCtx e{};
    e.ell = ell;
    e.k = k;
    uint32_t inv = ell;                                          // ell^-1 mod 2^32 by Newton (ell odd)
    for (int i = 0; i < 5; ++i) inv *= 2u - ell * inv;
    e.ellinv = inv;
    e.lim = 0xffffffffu / ell;
    e.pw[0] = 1;
    e.ipw[0] = 1;
    for (uint32_t v = 1; v <= k; ++v) {
        e.pw[v] = e.pw[v - 1] * ell;
        e.ipw[v] = e.ipw[v - 1] * inv;
    }
    e.q = e.pw[k];
    const unsigned __int128 two64 = (unsigned __int128)1 << 64;
    e.M = DevMod{e.q, (uint64_t)(two64 / e.q), (uint64_t)(two64 % e.q)};
    return e;
}

// A batch over Z/2^k (ell = 2: the blocked classes and the unblocked kernel) or Z/ell^k (ell odd).
HBatch* batch_create(uint32_t ell, uint32_t k, const hx_howell_job* jobs, uint32_t njobs) {
    HBatch* b = new HBatch();
    try {
        b->k = k;
        b->njobs = njobs;
        b->ell = ell;
        if (ell != 2) b->ectx = ell_ctx(ell, k);
        std::vector<HJob> hj(njobs);
        for (uint32_t i = 0; i < njobs; ++i) {
            const hx_howell_job& j = jobs[i];
            HX_CHECK(j.n == 0 || (j.H && j.pivots && j.pivot_v && j.work), "howell job: NULL pointer");
            HX_CHECK(j.h != nullptr, "howell job: h is NULL");
            HX_CHECK(j.m == 0 || j.A != nullptr, "howell job: A is NULL");
            HX_CHECK(j.lda >= (int64_t)j.n, "howell job: lda < n");
            HX_CHECK((uint64_t)j.m + j.n < (1u << 24), "howell job: too many rows");
            hj[i] = HJob{j.m, j.n, j.A, j.lda, j.H, j.pivots, j.pivot_v, j.h, j.work};
        }
        std::vector<uint32_t> order(njobs);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return (double)hj[x].m * hj[x].n * hj[x].n > (double)hj[y].m * hj[y].n * hj[y].n;
        });
        std::vector<uint32_t> per[NCLASS];
        for (uint32_t i : order) {
            const uint32_t rows = hj[i].m + hj[i].n;
            int c = NCLASS - 1;                              // Z/ell^k, ell odd: every job unblocked
            for (int t = 0; ell == 2 && t < NCLASS - 1; ++t)
                if (rows <= kClassRows[t]) { c = t; break; }
            per[c].push_back(i);
        }
        HX_HIP(hipMalloc(&b->d_jobs, std::max<size_t>(njobs, 1) * sizeof(HJob)));
        HX_HIP(hipMalloc(&b->d_next, NCLASS * sizeof(uint32_t)));
        if (njobs) HX_HIP(hipMemcpy(b->d_jobs, hj.data(), njobs * sizeof(HJob), hipMemcpyHostToDevice));
        int cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) cus = prop.multiProcessorCount;
        static std::once_flag once;
        std::call_once(once, [] {
            for (uint32_t kk : {16u, 24u, 32u})
                (void)hipFuncSetAttribute(blocked_kernel_for(kk), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)blocked_lds_bytes(kBlockedRows, blocked_elt_bytes(kk)));
        });
        for (int c = 0; c < NCLASS; ++c) {
            HClass& cl = b->cls[c];
            cl.njobs = (uint32_t)per[c].size();
            if (!cl.njobs) continue;
            HX_HIP(hipMalloc(&cl.d_order, cl.njobs * sizeof(uint32_t)));
            HX_HIP(hipMemcpy(cl.d_order, per[c].data(), cl.njobs * sizeof(uint32_t), hipMemcpyHostToDevice));
            int per_cu = 4;
            if (c < NCLASS - 1) {
                uint32_t rc = 0;
                for (uint32_t i : per[c]) rc = std::max(rc, hj[i].m + hj[i].n);
                cl.rc = rc;
                cl.lds = blocked_lds_bytes(rc, blocked_elt_bytes(k));
                int nb = 0;
                const void* kern = blocked_kernel_for(k);
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, HT, cl.lds) == hipSuccess &&
                    nb > 0)
                    per_cu = nb;
                else
                    per_cu = 1;
            }
            cl.grid = std::max<uint32_t>(1, std::min<uint32_t>(cl.njobs, (uint32_t)cus * per_cu));
        }
        HX_HIP(hipStreamCreateWithFlags(&b->s2, hipStreamNonBlocking));
        HX_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
        HX_HIP(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
    } catch (...) {
        destroy(b);
        throw;
    }
    return b;
}

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API size_t hx_howell_work_words(uint32_t m, uint32_t n) {
    return (size_t)(m + n) * n + (m + n) + n;
}

HX_API int hx_howell_batch_create(uint32_t k, const hx_howell_job* jobs, uint32_t njobs, hx_howell_batch* out) {
    return guard("hx_howell_batch_create", [&] {
        HX_CHECK(k >= 1 && k <= 32, "howell: need 1 <= k <= 32");
        HX_CHECK(out != nullptr && (jobs != nullptr || njobs == 0), "NULL argument");
        *out = batch_create(2, k, jobs, njobs);
    });
}

HX_API int hx_howell_batch_create_ell(uint32_t ell, uint32_t k, const hx_howell_job* jobs, uint32_t njobs,
                                      hx_howell_batch* out) {
    return guard("hx_howell_batch_create_ell", [&] {
        HX_CHECK(out != nullptr && (jobs != nullptr || njobs == 0), "NULL argument");
        HX_CHECK(ell >= 2 && k >= 1, "howell: need ell >= 2 and k >= 1");
        uint64_t q = 1;
        for (uint32_t v = 0; v < k; ++v) {
            q *= ell;
            HX_CHECK(q <= 0xffffffffull, "howell: ell^k must be below 2^32");
        }
        for (uint32_t d = 2; (uint64_t)d * d <= ell; ++d) HX_CHECK(ell % d != 0, "howell: ell must be prime");
        *out = batch_create(ell, k, jobs, njobs);
    });
}

HX_API int hx_howell_batch_run(hx_howell_batch vb, hx_stream s) {
    return guard("hx_howell_batch_run", [&] {
        HX_CHECK(vb != nullptr, "NULL batch");
        HBatch* b = (HBatch*)vb;
        if (b->njobs == 0) return;
        const hipStream_t s1 = (hipStream_t)s;
        zero_counters_kernel<<<1, 64, 0, s1>>>(b->d_next, NCLASS);
        after_launch("zero_counters_kernel");
        // class 1 (m + n <= 512) on the second stream beside classes 3 and 2 (the largest jobs
        // first) on the caller's stream, then class 0 there: the two heavy blocked classes share
        // the CUs and the small jobs follow the one that ends first
        HX_HIP(hipEventRecord(b->ev_fork, s1));
        HX_HIP(hipStreamWaitEvent(b->s2, b->ev_fork, 0));
        for (int c = NCLASS - 1; c >= 0; --c) {
            const HClass& cl = b->cls[c];
            if (!cl.njobs) continue;
            const hipStream_t sc = c == 1 ? b->s2 : s1;
            HArgs a{b->d_jobs, cl.d_order, cl.njobs, b->d_next + c, b->k};
            if (b->ell != 2) {
                howell_ell_kernel<<<cl.grid, ET, 0, sc>>>(a, b->ectx);
                after_launch("howell_ell_kernel");
            } else if (c == NCLASS - 1) {
                howell_batch_kernel<<<cl.grid, HT, 0, sc>>>(a);
                after_launch("howell_batch_kernel");
            } else {
                if (b->k <= 16)
                    howell_blocked_kernel<true, uint16_t><<<cl.grid, HT, cl.lds, sc>>>(a, cl.rc);
                else if (b->k <= 24)
                    howell_blocked_kernel<true, uint32_t><<<cl.grid, HT, cl.lds, sc>>>(a, cl.rc);
                else
                    howell_blocked_kernel<false, uint32_t><<<cl.grid, HT, cl.lds, sc>>>(a, cl.rc);
                after_launch("howell_blocked_kernel");
            }
        }
        HX_HIP(hipEventRecord(b->ev_join, b->s2));
        HX_HIP(hipStreamWaitEvent(s1, b->ev_join, 0));
    });
}

HX_API int hx_howell_batch_destroy(hx_howell_batch vb) {
    return guard("hx_howell_batch_destroy", [&] { destroy((HBatch*)vb); });
}

}  // extern "C"
