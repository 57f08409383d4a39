// hx_ring2k_lu.hip -- the host driver of the blocked LU over Z/2^32 / Z/2^64 on a device session
// (the device route of ring2k::echelon_blocked, Binary/Ring2kDeviceLU.h, with the MinValuation rule),
// the Ops table's plan / store_budget / verify, and the route's counters (Binary/Ring2kDevice.h:86-101).
//
// The driver uses only the session's C-ABI (hx_ring2k_open / put / get / step / fetch / close), the
// way the reference's driver uses its Ops table.  Per panel the host factors the panel's columns
// (rows r .. m-1): per column the row of least valuation mod 2^k from the next pivot row down (the
// first one; a unit ends the search), swapped up together with its L slots; every row below gets the
// slot f = (x >> v) u^-1 (0 when x is 0 mod 2^k) and its panel columns lose f times the pivot row.
// The device then gets the factored panel, the swaps and L (step), and updates the rest.  The
// matrix, rank, pivots and valuations are exactly those of the host code (echelon_blocked's serial
// route), which the oracle restates and the tests compare entrywise.
//
// A window narrower than n (plan() under a memory budget) runs in phases: a phase uploads the window
// [c, c + cw) at the next panel column, factors panels inside it, downloads it; the columns right of
// it then pass through the device in slabs receiving the phase's panels again (swaps, TRSM, product,
// replayed from a host record of each panel's swaps and L), and the columns left of it get the phase's
// swaps on the host as one row permutation.  Every column receives the same operations in the same
// order as in the host code, so the tiling cannot change a value.
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hx_common.h"

namespace hx {
namespace {

std::atomic<uint64_t> g_r2[16];
enum {
    R2_CALLS, R2_DEVICE, R2_RESIDENT, R2_TILED, R2_SLABS, R2_REPLAYS, R2_FALLBACKS, R2_PANELS, R2_LOOKAHEADS,
    R2_VERIFIED, R2_NS_OPEN, R2_NS_FETCH, R2_NS_PANEL, R2_NS_STEP, R2_NS_HOST, R2_NS_CLOSE
};
void note(int id, uint64_t v = 1) { g_r2[id].fetch_add(v, std::memory_order_relaxed); }

void ok(int rc) {
    if (rc != HX_OK) throw Error{rc, hx_last_error()};
}

template <class E>
struct Wrap {
    static constexpr uint32_t bits = 8 * sizeof(E);
    static uint32_t val(E x, uint32_t k) {          // valuation of x mod 2^k (k when x = 0 mod 2^k)
        const E mk = k >= bits ? (E)~(E)0 : (E)(((E)1 << k) - 1);
        const E y = x & mk;
        if (y == 0) return k;
        const uint32_t c = (uint32_t)__builtin_ctzll((unsigned long long)y);
        return c < k ? c : k;
    }
    static E inv_odd(E u) {                        // u^-1 mod 2^bits, Newton from 3 correct bits
        E x = u;
        for (uint32_t b = 3; b < bits; b *= 2) x = (E)(x * (E)((E)2 - (E)(u * x)));
        return x;
    }
};

// The record of one panel of a phase (tiled route): its pivot rows, swaps and L block (rows r .. m-1, ld np).
template <class E>
struct PanelRecord {
    uint32_t r = 0, np = 0;
    std::vector<uint32_t> swaps;
    std::vector<E> L;
};

// The host factor of one panel: P holds the panel's pw columns (first column c) on the rows r .. r+rows-1,
// eliminated column by column with the MinValuation rule mod 2^k; the swaps (global row pairs) go to
// `swaps`, the L slots to the rows' slots of L (m x panel, swapped with their rows), each pivot column and
// valuation to pivots / pivot_v at npiv (advanced).  Returns the panel's pivot count.
template <class E>
uint32_t factor_panel(uint32_t k, std::vector<E>& P, uint32_t rows, uint32_t pw, std::vector<E>& L, uint32_t r,
                      uint32_t panel, std::vector<uint32_t>& swaps, uint32_t c, uint32_t* pivots, uint32_t* pivot_v,
                      uint32_t& npiv) {
    using W = Wrap<E>;
    uint32_t np = 0;
    swaps.clear();
    for (uint32_t j = 0; j < pw && np < rows; ++j) {
        uint32_t best = rows, bv = k;
        for (uint32_t i = np; i < rows; ++i) {
            const uint32_t v = W::val(P[(size_t)i * pw + j], k);
            if (v < bv) {
                bv = v;
                best = i;
                if (v == 0) break;
            }
        }
        if (best == rows) continue;
        if (best != np) {
            std::swap_ranges(P.begin() + (size_t)best * pw, P.begin() + (size_t)(best + 1) * pw,
                             P.begin() + (size_t)np * pw);
            std::swap_ranges(L.begin() + (size_t)(r + best) * panel, L.begin() + (size_t)(r + best + 1) * panel,
                             L.begin() + (size_t)(r + np) * panel);
            swaps.push_back(r + best);
            swaps.push_back(r + np);
        }
        const E* prow = P.data() + (size_t)np * pw;
        const E ui = W::inv_odd((E)(prow[j] >> bv));
        for (uint32_t i = np + 1; i < rows; ++i) {
            E* row = P.data() + (size_t)i * pw;
            E& slot = L[(size_t)(r + i) * panel + np];
            if (W::val(row[j], k) == k) {
                slot = (E)0;
                continue;
            }
            const E f = (E)((E)(row[j] >> bv) * ui);
            slot = f;
            for (uint32_t q = j; q < pw; ++q) row[q] = (E)(row[q] - (E)(f * prow[q]));
        }
        pivots[npiv] = c + j;
        pivot_v[npiv] = bv;
        ++npiv;
        ++np;
    }
    return np;
}

template <class E>
void echelon_run(uint32_t k, uint32_t m, uint32_t n, E* A, size_t lda, uint32_t panel, uint32_t climit, uint32_t cw,
                 uint64_t store_cap, uint32_t* rank_out, uint32_t* pivots, uint32_t* pivot_v) {
    const bool tiled = cw < n;
    using clk = std::chrono::steady_clock;
    auto tick = clk::now();
    auto lap = [&](int id) {
        const auto t = clk::now();
        note(id, (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t - tick).count());
        tick = t;
    };
    hx_ring2k sess = nullptr;
    ok(hx_ring2k_open(sizeof(E), m, cw, panel, &sess));
    lap(R2_NS_OPEN);
    uint32_t nslabs = 0;
    try {
        std::vector<E> L((size_t)m * panel, (E)0);      // L slots of every row (swapped with their rows)
        std::vector<E> P;                               // the panel's columns, rows r .. m-1
        std::vector<uint32_t> swaps;
        std::vector<PanelRecord<E>> recs;
        uint32_t r = 0, c = 0, npiv = 0;
        while (c < climit && r < m) {
            const uint32_t s0 = c, s1 = std::min<uint32_t>(s0 + cw, n), w = s1 - s0;
            ++nslabs;
            recs.clear();
            uint64_t store = 0;
            ok(hx_ring2k_put(sess, A + s0, lda, 0, m, 0, w));
            lap(R2_NS_OPEN);
            const uint32_t cend = std::min(climit, s1);
            bool have_next = false;
            while (c < cend && r < m) {
                const uint32_t c1 = std::min(c + panel, climit), pw = c1 - c, rows = m - r;
                P.assign((size_t)rows * pw, (E)0);
                if (have_next) {
                    ok(hx_ring2k_fetch(sess, P.data(), pw));
                    note(R2_LOOKAHEADS);
                } else {
                    ok(hx_ring2k_get(sess, P.data(), pw, r, rows, c - s0, pw));
                }
                lap(R2_NS_FETCH);
                const uint32_t np = factor_panel<E>(k, P, rows, pw, L, r, panel, swaps, c, pivots, pivot_v, npiv);
                lap(R2_NS_PANEL);
                const uint32_t rn = r + np;
                const uint64_t lbytes = (uint64_t)rows * np * sizeof(E);
                const uint64_t next_bytes = rn < m ? (uint64_t)(m - rn) * panel * sizeof(E) : 0;
                const bool budget_ok = !tiled || s1 >= n || store_cap == 0 || store + lbytes + next_bytes <= store_cap;
                const bool next = c1 < cend && rn < m && budget_ok;
                const uint32_t nc1 = std::min(c1 + panel, climit);
                hx_ring2k_step_args a{};
                a.m = m;
                a.r = r;
                a.np = np;
                a.pc0 = c - s0;
                a.pc1 = c1 - s0;
                a.panel = P.data();
                a.ldp = pw;
                a.swaps = swaps.data();
                a.nswaps = (uint32_t)(swaps.size() / 2);
                a.L = L.data() + (size_t)r * panel;
                a.ldl = panel;
                a.u0 = (np > 0 && c1 < s1) ? c1 - s0 : w;
                a.la0 = next ? c1 - s0 : 0;
                a.la1 = next ? nc1 - s0 : 0;
                a.w = w;
                ok(hx_ring2k_step(sess, &a));
                note(R2_PANELS);
                have_next = next;
                lap(R2_NS_STEP);
                if (tiled && (np > 0 || !swaps.empty())) {
                    PanelRecord<E> pr;
                    pr.r = r;
                    pr.np = np;
                    pr.swaps = swaps;
                    if (s1 < n) {
                        pr.L.resize((size_t)rows * np);
                        for (uint32_t i = 0; i < rows; ++i)
                            std::memcpy(pr.L.data() + (size_t)i * np, L.data() + (size_t)(r + i) * panel, sizeof(E) * np);
                        store += lbytes;
                    }
                    recs.push_back(std::move(pr));
                }
                lap(R2_NS_HOST);
                r = rn;
                c = c1;
                if (!budget_ok) break;          // the record is full: end the phase at this panel boundary
            }
            ok(hx_ring2k_get(sess, A + s0, lda, 0, m, 0, w));
            lap(R2_NS_CLOSE);
            if (!tiled || recs.empty()) continue;
            // the columns right of the window: slabs through the device, the phase's panels replayed
            for (uint32_t t0 = s1; t0 < n; t0 += cw) {
                const uint32_t tw = std::min(cw, n - t0);
                ++nslabs;
                ok(hx_ring2k_put(sess, A + t0, lda, 0, m, 0, tw));
                lap(R2_NS_OPEN);
                for (const PanelRecord<E>& pr : recs) {
                    hx_ring2k_step_args a{};
                    a.m = m;
                    a.r = pr.r;
                    a.np = pr.np;
                    a.swaps = pr.swaps.data();
                    a.nswaps = (uint32_t)(pr.swaps.size() / 2);
                    a.L = pr.L.data();
                    a.ldl = pr.np;
                    a.u0 = 0;
                    a.w = tw;
                    ok(hx_ring2k_step(sess, &a));
                    note(R2_REPLAYS);
                }
                lap(R2_NS_HOST);
                ok(hx_ring2k_get(sess, A + t0, lda, 0, m, 0, tw));
                lap(R2_NS_CLOSE);
            }
            // the columns left of the window: the phase's swaps as one row permutation, new[i] = old[src[i]]
            if (s0 > 0) {
                std::vector<uint32_t> src(m);
                for (uint32_t i = 0; i < m; ++i) src[i] = i;
                bool any = false;
                for (const PanelRecord<E>& pr : recs)
                    for (size_t t = 0; t + 1 < pr.swaps.size(); t += 2) {
                        std::swap(src[pr.swaps[t]], src[pr.swaps[t + 1]]);
                        any = true;
                    }
                if (any) {
                    std::vector<E> old((size_t)m * s0);
                    for (uint32_t i = 0; i < m; ++i) std::memcpy(old.data() + (size_t)i * s0, A + (size_t)i * lda, sizeof(E) * s0);
                    for (uint32_t i = 0; i < m; ++i)
                        if (src[i] != i) std::memcpy(A + (size_t)i * lda, old.data() + (size_t)src[i] * s0, sizeof(E) * s0);
                }
                lap(R2_NS_HOST);
            }
        }
        *rank_out = r;
    } catch (...) {
        hx_ring2k_close(sess);
        throw;
    }
    hx_ring2k_close(sess);
    note(R2_DEVICE);
    note(tiled ? R2_TILED : R2_RESIDENT);
    if (tiled) note(R2_SLABS, nslabs);
}

uint32_t plan_window(uint32_t elem, uint32_t m, uint32_t n, uint32_t panel, uint64_t budget) {
    if ((elem != 4 && elem != 8) || m == 0 || n == 0 || panel == 0 || panel > 64) return 0;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        budget = (uint64_t)(fr * 0.9);
    }
    const uint64_t fixed = hx_ring2k_session_bytes(elem, m, 0, panel);
    if (budget <= fixed) return 0;
    const uint64_t cols = (budget - fixed) / ((uint64_t)m * elem);
    if (cols >= n) return n;
    const uint64_t cw = cols / panel * panel;
    return cw >= panel ? (uint32_t)cw : 0;
}

// ---- the Smith valuations over Z/2^k by levels (ring2k::smith, Binary/Ring2k.h:701-841) ----------------
// The matrix stays resident in one session window for every level; level s works on the block of rows and
// columns [R, m) x [R, n) (R = the pivots of the levels before), read mod 2^(k - s):
//   1. one wave per 64 columns of a row packs the block's low bits into a bit matrix (a ballot per wave)
//      and flags whether any entry is non-zero mod 2^(k - s) (none: the valuations are complete);
//   2. the F_2 RREF of the bit matrix (hx_f2_rref) gives the level's rank r0 and its pivot columns;
//   3. those columns move to the front of the block (a gather through a row-chunk buffer);
//   4. the session's LU over the first r0 columns of the block: the host-factored panels find an odd pivot
//      in every one of them (MinValuation and FirstOdd choose the same row there: the first odd one), and
//      the device updates the rest of the block;
//   5. the Schur complement, rows and columns [R + r0, ..), is even: it is halved in place (a kernel that
//      also flags an odd entry, which would be an internal error) and becomes the next level's block.
// The pivots of level s contribute r0 valuations s.  The Smith form of A^T is that of A, so a matrix with
// more rows than columns is transposed on the device first (fewer rows for the F_2 search and the panels).

template <class E>
__global__ __launch_bounds__(256) void smith_pack_kernel(const E* __restrict__ win, uint32_t ld, uint32_t R,
                                                         uint32_t rows, uint32_t cols, uint32_t ldw, E kmask,
                                                         uint64_t* __restrict__ bits, uint32_t* __restrict__ flags) {
    const uint64_t wv = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (wv >= (uint64_t)rows * ldw) return;               // the wave index is uniform: whole waves leave
    const uint32_t i = (uint32_t)(wv / ldw), g = (uint32_t)(wv % ldw), j = g * 64 + lane;
    const E x = j < cols ? win[(size_t)(R + i) * ld + R + j] : (E)0;
    const uint64_t odd = __ballot((int)(x & (E)1));
    const uint64_t nz = __ballot((int)((x & kmask) != 0));
    if (lane == 0) {
        bits[(size_t)i * ldw + g] = odd;
        if (nz) flags[0] = 1u;
    }
}

// tmp row i = window row R + i0 + i, columns R + order[j] (j < cols)
template <class E>
__global__ __launch_bounds__(256) void smith_gather_kernel(const E* __restrict__ win, uint32_t ld, uint32_t R, uint32_t i0,
                                                           uint32_t cols, const uint32_t* __restrict__ order,
                                                           E* __restrict__ tmp) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x, i = blockIdx.y;
    if (j >= cols) return;
    tmp[(size_t)i * cols + j] = win[(size_t)(R + i0 + i) * ld + R + order[j]];
}

// window rows R + blockIdx.y, columns [R, R + cols): x -> x >> 1, flags[1] = 1 on an odd entry
template <class E>
__global__ __launch_bounds__(256) void smith_halve_kernel(E* __restrict__ win, uint32_t ld, uint32_t R, uint32_t cols,
                                                          uint32_t* __restrict__ flags) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x, i = blockIdx.y;
    if (j >= cols) return;
    E* x = win + (size_t)(R + i) * ld + R + j;
    const E v = *x;
    if (v & (E)1) flags[1] = 1u;
    *x = v >> 1;
}

// dst (cols x rows, row stride ldd) = src^T, src rows x cols (row stride cols); 64 x 64 tiles through LDS
template <class E>
__global__ __launch_bounds__(256) void smith_transpose_kernel(const E* __restrict__ src, uint32_t rows, uint32_t cols,
                                                              E* __restrict__ dst, uint32_t ldd) {
    __shared__ E t[64][65];
    const uint32_t r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;     // x: row tiles (the long side)
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (uint32_t y = ty; y < 64; y += 4) {
        const uint32_t r = r0 + y, c = c0 + tx;
        if (r < rows && c < cols) t[y][tx] = src[(size_t)r * cols + c];
    }
    __syncthreads();
    for (uint32_t y = ty; y < 64; y += 4) {
        const uint32_t c = c0 + y, r = r0 + tx;
        if (r < rows && c < cols) dst[(size_t)c * ldd + r] = t[tx][y];
    }
}

// A device buffer from the arena, returned once its stream has drained (also when an error unwinds).
struct ArenaBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    ArenaBuf(size_t bytes, hipStream_t st) : p(arena_alloc(bytes)), s(st) {}
    ~ArenaBuf() {
        if (!p) return;
        (void)hipStreamSynchronize(s);
        arena_free(p);
    }
    ArenaBuf(const ArenaBuf&) = delete;
    ArenaBuf& operator=(const ArenaBuf&) = delete;
};

// A: am x an host words (row stride lda).  The device block is m x n with m <= n (A or A^T).
template <class E>
void smith_run(uint32_t k, uint32_t am, uint32_t an, const E* A, size_t lda, uint32_t panel, uint32_t* sigma,
               uint32_t* nsigma, uint32_t* level_ranks, uint32_t* nlevels) {
    const bool tr = am > an;
    const uint32_t m = tr ? an : am, n = tr ? am : an;
    HX_CHECK(m <= 16384, "ring2k smith: min(m, n) > 16384 is not supported (the F_2 rank profile's row limit)");
    HX_CHECK(plan_window(sizeof(E), m, n, panel, 0) == n, "ring2k smith: the matrix does not fit in device memory");
    hx_ring2k sess = nullptr;
    ok(hx_ring2k_open(sizeof(E), m, n, panel, &sess));
    try {
        const Ring2kWindow W = ring2k_window(sess);
        const hipStream_t st = W.s;
        E* win = (E*)W.win;
        if (!tr) {
            ok(hx_ring2k_put(sess, A, lda, 0, m, 0, n));
        } else {
            ArenaBuf src(sizeof(E) * (size_t)am * an, st);
            if (lda == an)
                ok(hx_upload(src.p, A, sizeof(E) * (size_t)am * an, (hx_stream)st));
            else
                staged_copy_2d(src.p, sizeof(E) * an, A, sizeof(E) * lda, sizeof(E) * an, am, true, st);
            smith_transpose_kernel<E><<<dim3((am + 63) / 64, (an + 63) / 64), 256, 0, st>>>((const E*)src.p, am, an, win, n);
            after_launch("smith_transpose_kernel");
            HX_HIP(hipStreamSynchronize(st));
        }
        const uint32_t ldw0 = (n + 63) / 64;
        ArenaBuf bits(sizeof(uint64_t) * (size_t)m * ldw0, st), flags(256, st), d_order(sizeof(uint32_t) * n, st);
        const size_t chunk = std::max<size_t>(1, std::min<size_t>(m, ((size_t)64 << 20) / (sizeof(E) * n)));
        ArenaBuf tmp(sizeof(E) * chunk * n, st);
        uint32_t* fl = (uint32_t*)flags.p;
        std::vector<E> L((size_t)m * panel, (E)0), P;
        std::vector<uint32_t> swaps, piv(m), order(n), lpiv(panel), lpv(panel);
        uint32_t R = 0, ns = 0, nl = 0;
        for (uint32_t s = 0; s < k && R < m && R < n; ++s) {
            const uint32_t kk = k - s, rows = m - R, cols = n - R, ldw = (cols + 63) / 64;
            const E kmask = kk >= 8 * sizeof(E) ? (E)~(E)0 : (E)(((E)1 << kk) - 1);
            // 1. the low bits and the zero test
            HX_HIP(hipMemsetAsync(fl, 0, 2 * sizeof(uint32_t), st));
            const uint64_t waves = (uint64_t)rows * ldw;
            smith_pack_kernel<E><<<(uint32_t)((waves + 3) / 4), 256, 0, st>>>(win, n, R, rows, cols, ldw, kmask,
                                                                              (uint64_t*)bits.p, fl);
            after_launch("smith_pack_kernel");
            uint32_t hf[2] = {0, 0};
            ok(hx_download(hf, fl, sizeof(hf), (hx_stream)st));
            if (!hf[0]) break;
            // 2. the F_2 rank profile
            uint32_t r0 = 0;
            ok(hx_f2_rref(rows, cols, (uint64_t*)bits.p, ldw, &r0, piv.data(), (hx_stream)st));
            level_ranks[s] = r0;
            nl = s + 1;
            for (uint32_t t = 0; t < r0; ++t) sigma[ns++] = s;
            if (r0 > 0) {
                // 3. the pivot columns to the front (pivots ascend: none moves when the last is r0 - 1)
                if (piv[r0 - 1] != r0 - 1) {
                    for (uint32_t j = 0, t = 0, q = r0; j < cols; ++j) {
                        if (t < r0 && piv[t] == j)
                            order[t++] = j;
                        else
                            order[q++] = j;
                    }
                    ok(hx_upload(d_order.p, order.data(), sizeof(uint32_t) * cols, (hx_stream)st));
                    for (uint32_t i0 = 0; i0 < rows; i0 += (uint32_t)chunk) {
                        const uint32_t nr = std::min<uint32_t>((uint32_t)chunk, rows - i0);
                        smith_gather_kernel<E><<<dim3((cols + 255) / 256, nr), 256, 0, st>>>(
                            win, n, R, i0, cols, (const uint32_t*)d_order.p, (E*)tmp.p);
                        after_launch("smith_gather_kernel");
                        HX_HIP(hipMemcpy2DAsync(win + (size_t)(R + i0) * n + R, sizeof(E) * n, tmp.p, sizeof(E) * cols,
                                                sizeof(E) * cols, nr, hipMemcpyDeviceToDevice, st));
                    }
                }
                // 4. the odd-pivot LU over the block's first r0 columns
                const uint32_t climit = R + r0;
                uint32_t r = R, c = R;
                bool have_next = false;
                while (c < climit) {
                    const uint32_t c1 = std::min(c + panel, climit), pw = c1 - c, prow = m - r;
                    P.assign((size_t)prow * pw, (E)0);
                    if (have_next) {
                        ok(hx_ring2k_fetch(sess, P.data(), pw));
                        note(R2_LOOKAHEADS);
                    } else {
                        ok(hx_ring2k_get(sess, P.data(), pw, r, prow, c, pw));
                    }
                    uint32_t cnt = 0;
                    const uint32_t np = factor_panel<E>(kk, P, prow, pw, L, r, panel, swaps, c, lpiv.data(), lpv.data(), cnt);
                    bool agree = np == pw;
                    for (uint32_t t = 0; agree && t < np; ++t) agree = lpiv[t] == c + t && lpv[t] == 0;
                    if (!agree) {
                        size_t nz = 0;
                        for (const E& x : P) nz += x != 0;
                        throw Error{HX_E_RUNTIME, "ring2k smith: the odd-pivot LU disagrees with the F_2 rank profile (level " +
                                                      std::to_string(s) + ", columns " + std::to_string(c) + ".." +
                                                      std::to_string(c1) + ", " + std::to_string(np) + " pivots, " +
                                                      std::to_string(nz) + " non-zero panel entries of " +
                                                      std::to_string(P.size()) + ", fetched by " +
                                                      (have_next ? "lookahead)" : "get)")};
                    }
                    const uint32_t nc1 = std::min(c1 + panel, climit);
                    const bool next = c1 < climit;
                    hx_ring2k_step_args a{};
                    a.m = m;
                    a.r = r;
                    a.np = np;
                    a.pc0 = c;
                    a.pc1 = c1;
                    a.panel = P.data();
                    a.ldp = pw;
                    a.swaps = swaps.data();
                    a.nswaps = (uint32_t)(swaps.size() / 2);
                    a.L = L.data() + (size_t)r * panel;
                    a.ldl = panel;
                    a.u0 = c1;
                    a.la0 = next ? c1 : 0;
                    a.la1 = next ? nc1 : 0;
                    a.w = n;
                    ok(hx_ring2k_step(sess, &a));
                    note(R2_PANELS);
                    have_next = next;
                    r += np;
                    c = c1;
                }
            }
            // 5. the even Schur complement, halved in place
            const uint32_t Rn = R + r0;
            if (Rn < m && Rn < n) {
                (void)ring2k_window(sess);              // after the last step's lookahead copy, if any
                smith_halve_kernel<E><<<dim3((n - Rn + 255) / 256, m - Rn), 256, 0, st>>>(win, n, Rn, n - Rn, fl);
                after_launch("smith_halve_kernel");
                ok(hx_download(hf, fl, sizeof(hf), (hx_stream)st));
                if (hf[1]) throw Error{HX_E_RUNTIME, "ring2k smith: the Schur complement is not even"};
            }
            R = Rn;
        }
        HX_HIP(hipStreamSynchronize(st));
        *nsigma = ns;
        *nlevels = nl;
    } catch (...) {
        hx_ring2k_close(sess);
        throw;
    }
    hx_ring2k_close(sess);
}

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API uint32_t hx_ring2k_plan(uint32_t elem, uint32_t m, uint32_t n, uint32_t panel, uint64_t budget_bytes) {
    return plan_window(elem, m, n, panel, budget_bytes);
}

HX_API uint64_t hx_ring2k_store_budget(void) {
    const char* v = std::getenv("HX_RING2K_HOST_MB");
    if (v == nullptr || *v == 0) return 0;
    return (uint64_t)std::strtoull(v, nullptr, 10) << 20;
}

HX_API int hx_ring2k_verify(void) {
    const char* v = std::getenv("HX_VERIFY_DEVICE");
    return (v != nullptr && *v != 0 && *v != '0') ? 1 : 0;
}

HX_API int hx_ring2k_echelon(uint32_t elem, uint32_t k, uint32_t m, uint32_t n, void* A, size_t lda, uint32_t panel,
                             uint32_t col_limit, uint64_t budget_bytes, uint64_t store_bytes, uint32_t* rank,
                             uint32_t* pivots, uint32_t* pivot_v) {
    return guard("hx_ring2k_echelon", [&] {
        HX_CHECK(elem == 4 || elem == 8, "elem must be 4 or 8");
        HX_CHECK(k >= 1 && k <= 8 * elem, "need 1 <= k <= the word's bits");
        HX_CHECK(rank != nullptr && (m == 0 || n == 0 || (A != nullptr && lda >= n && pivots && pivot_v)),
                 "NULL argument or lda < n");
        HX_CHECK(panel >= 1 && panel <= 64, "panel must be 1 .. 64");
        *rank = 0;
        if (m == 0 || n == 0) return;
        note(R2_CALLS);
        const uint32_t climit = std::min(col_limit, n);
        if (climit == 0) return;
        const uint32_t cw = plan_window(elem, m, n, panel, budget_bytes);
        if (cw == 0) {
            note(R2_FALLBACKS);
            throw Error{HX_E_NOMEM, "the device declines: not one panel-wide window fits the memory budget"};
        }
        if (elem == 4)
            echelon_run<uint32_t>(k, m, n, (uint32_t*)A, lda, panel, climit, cw, store_bytes, rank, pivots, pivot_v);
        else
            echelon_run<uint64_t>(k, m, n, (uint64_t*)A, lda, panel, climit, cw, store_bytes, rank, pivots, pivot_v);
    });
}

HX_API int hx_ring2k_smith(uint32_t elem, uint32_t k, uint32_t m, uint32_t n, const void* A, size_t lda,
                           uint32_t panel, uint32_t* sigma, uint32_t* nsigma, uint32_t* level_ranks, uint32_t* nlevels) {
    return guard("hx_ring2k_smith", [&] {
        HX_CHECK(elem == 4 || elem == 8, "elem must be 4 or 8");
        HX_CHECK(k >= 1 && k <= 8 * elem, "need 1 <= k <= the word's bits");
        HX_CHECK(nsigma != nullptr && nlevels != nullptr, "NULL argument");
        *nsigma = 0;
        *nlevels = 0;
        if (m == 0 || n == 0) return;
        HX_CHECK(A != nullptr && lda >= n && sigma != nullptr && level_ranks != nullptr, "NULL argument or lda < n");
        if (panel == 0) panel = 64;
        HX_CHECK(panel <= 64, "panel must be 0 (64) or 1 .. 64");
        if (elem == 4)
            smith_run<uint32_t>(k, m, n, (const uint32_t*)A, lda, panel, sigma, nsigma, level_ranks, nlevels);
        else
            smith_run<uint64_t>(k, m, n, (const uint64_t*)A, lda, panel, sigma, nsigma, level_ranks, nlevels);
    });
}

HX_API int hx_ring2k_stats_get(hx_ring2k_stats* out) {
    return guard("hx_ring2k_stats_get", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        uint64_t* f = &out->calls;
        for (int i = 0; i < 16; ++i) f[i] = g_r2[i].load(std::memory_order_relaxed);
    });
}

HX_API int hx_ring2k_stats_reset(void) {
    for (auto& c : g_r2) c.store(0, std::memory_order_relaxed);
    return HX_OK;
}

}  // extern "C"
