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

template <class E>
void echelon_run(uint32_t k, uint32_t m, uint32_t n, E* A, size_t lda, uint32_t panel, uint32_t climit, uint32_t cw,
                 uint64_t store_cap, uint32_t* rank_out, uint32_t* pivots, uint32_t* pivot_v) {
    using W = Wrap<E>;
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
                // the panel, eliminated column by column (rows of P are global rows r + i)
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
