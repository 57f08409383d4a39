// ref_driver.cpp -- a C-ABI shim over the reference's own CPU matrix engine.
//
// TEST INFRASTRUCTURE ONLY (the checker and the "reference" CPU baseline).
// oracle/Makefile compiles this file together with the reference's
// src/ModArithmetic sources, where they lie, into oracle/_ref/libhacsref.so.
// Nothing here re-implements an algorithm: every entry point copies plain
// arrays into the reference's DenseBuf / ring2k::Mat types and calls the
// reference function named in its comment.

#include <cstdint>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

#include "DenseBuf.h"
#include "DenseElim.h"
#include "DenseView.h"
#include "Fnv.h"
#include "Kernels.h"
#include "ModCtx.h"
#include "RefLinAlg.h"
#include "Binary/Ring2k.h"

using engine::DenseBuf;
using engine::DenseView;
using engine::ModCtx;

namespace {

DenseBuf<uint32_t> to_buf(const uint32_t* A, uint32_t m, uint32_t n, int64_t lda) {
    DenseBuf<uint32_t> B(m, n);
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t j = 0; j < n; ++j) B.at(i, j) = A[static_cast<int64_t>(i) * lda + j];
    return B;
}

}  // namespace

extern "C" {

// engine::FnvHasher (Fnv.h:76-83): feed() each 64-bit word in order from the offset basis.
uint64_t hxr_fnv_words(const uint64_t* w, size_t n) {
    engine::FnvHasher h;
    for (size_t i = 0; i < n; ++i) h.feed(w[i]);
    return h.value();
}

// engine::FnvBytes (Fnv.h:58-64) from the offset basis over n bytes.
uint64_t hxr_fnv_bytes(const char* b, size_t n) { return engine::FnvBytes(engine::kFnvOffset64, std::string(b, n)); }

// engine::kernels::gemm<uint32_t> (Kernels.cpp:1417-1483), split over row tiles of C on
// `threads` host threads as the tape executor splits output tiles (Engine/Executor.cpp:69-100).
// A, B, C row-major with leading dimensions; op 0 Set, 1 Add, 2 Sub. Returns 0 or -1 on an exception.
// engine::kernels::gemm<uint64_t> (Kernels.h:184) on row-major arrays, one thread.
int hxr_gemm_u64(uint64_t p, uint32_t m, uint32_t n, uint32_t k, const uint64_t* A, const uint64_t* B, uint64_t* C,
                 int op) {
    try {
        const ModCtx<uint64_t> ctx(p);
        DenseBuf<uint64_t> a(m, k), b(k, n), c(m, n);
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t t = 0; t < k; ++t) a.at(i, t) = A[static_cast<size_t>(i) * k + t];
        for (uint32_t t = 0; t < k; ++t)
            for (uint32_t j = 0; j < n; ++j) b.at(t, j) = B[static_cast<size_t>(t) * n + j];
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t j = 0; j < n; ++j) c.at(i, j) = C[static_cast<size_t>(i) * n + j];
        engine::kernels::gemm<uint64_t>(ctx, DenseView<const uint64_t>(a), DenseView<const uint64_t>(b),
                                        DenseView<uint64_t>(c), static_cast<engine::kernels::GemmOp>(op));
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t j = 0; j < n; ++j) C[static_cast<size_t>(i) * n + j] = c.at(i, j);
        return 0;
    } catch (...) {
        return -1;
    }
}

int hxr_gemm_u32(uint32_t p, uint32_t m, uint32_t n, uint32_t k, const uint32_t* A, int64_t lda, const uint32_t* B,
                 int64_t ldb, uint32_t* C, int64_t ldc, int op, int threads) {
    try {
        const ModCtx<uint32_t> ctx(p);
        DenseBuf<uint32_t> a = to_buf(A, m, k, lda), b = to_buf(B, k, n, ldb), c = to_buf(C, m, n, ldc);
        if (threads < 1) threads = 1;
        const uint32_t tile = (m + threads - 1) / threads;
        std::vector<std::thread> pool;
        std::exception_ptr err;
        for (int t = 0; t < threads; ++t) {
            const uint32_t r0 = t * tile;
            if (r0 >= m) break;
            const uint32_t nr = std::min(tile, m - r0);
            pool.emplace_back([&, r0, nr]() {
                try {
                    engine::kernels::gemm<uint32_t>(ctx, DenseView<const uint32_t>(a, r0, nr, 0, k),
                                                    DenseView<const uint32_t>(b), DenseView<uint32_t>(c, r0, nr, 0, n),
                                                    static_cast<engine::kernels::GemmOp>(op));
                } catch (...) {
                    err = std::current_exception();
                }
            });
        }
        for (auto& th : pool) th.join();
        if (err) return -1;
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t j = 0; j < n; ++j) C[static_cast<int64_t>(i) * ldc + j] = c.at(i, j);
        return 0;
    } catch (...) {
        return -1;
    }
}

// engine::elim::rref<uint32_t> (DenseElim.cpp:920-930): M (m x n, row stride ld) is replaced by its
// RREF (zero rows at the bottom); pivots receives the rank pivot columns. Returns the rank, -1 on error.
int64_t hxr_rref_u32(uint32_t p, uint32_t m, uint32_t n, uint32_t* M, int64_t ld, uint32_t* pivots) {
    try {
        const ModCtx<uint32_t> ctx(p);
        DenseBuf<uint32_t> a = to_buf(M, m, n, ld);
        const engine::elim::RrefResult r = engine::elim::rref<uint32_t>(ctx, DenseView<uint32_t>(a));
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t j = 0; j < n; ++j) M[static_cast<int64_t>(i) * ld + j] = a.at(i, j);
        for (uint32_t t = 0; t < r.rank; ++t) pivots[t] = r.pivots[t];
        return r.rank;
    } catch (...) {
        return -1;
    }
}

// engine::elim::column_rank_profile<uint32_t> (DenseElim.cpp:932-951) of A (not modified).
int64_t hxr_column_rank_profile_u32(uint32_t p, uint32_t m, uint32_t n, const uint32_t* A, int64_t ld,
                                    uint32_t* pivots) {
    try {
        const ModCtx<uint32_t> ctx(p);
        DenseBuf<uint32_t> a = to_buf(A, m, n, ld);
        const engine::elim::RrefResult r = engine::elim::column_rank_profile<uint32_t>(ctx, DenseView<uint32_t>(a));
        for (uint32_t t = 0; t < r.rank; ++t) pivots[t] = r.pivots[t];
        return r.rank;
    } catch (...) {
        return -1;
    }
}

// engine::ref::rref<uint32_t> (RefLinAlg.cpp:140-175), in place; returns the rank.
int64_t hxr_ref_rref_u32(uint32_t p, uint32_t m, uint32_t n, uint32_t* M, int64_t ld, uint32_t* pivots) {
    try {
        DenseBuf<uint32_t> a = to_buf(M, m, n, ld);
        const std::vector<uint32_t> piv = engine::ref::rref<uint32_t>(p, DenseView<uint32_t>(a));
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t j = 0; j < n; ++j) M[static_cast<int64_t>(i) * ld + j] = a.at(i, j);
        for (size_t t = 0; t < piv.size(); ++t) pivots[t] = piv[t];
        return static_cast<int64_t>(piv.size());
    } catch (...) {
        return -1;
    }
}

// engine::ring2k::howell on the rung `rung` (16: uint16_t, the rung select_rung picks for k <= 16,
// Ring2k.cpp:368-380; 32: uint32_t) -- Binary/Ring2k.h:959-969. H receives h x n (row stride n).
// Returns h, -1 on error.
int64_t hxr_ring2k_howell(uint32_t k, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, int rung, uint32_t* H,
                          uint32_t* pivots, uint32_t* pivot_v) {
    try {
        auto run = [&](auto tag) -> int64_t {
            using E = decltype(tag);
            engine::ring2k::Mat<E> a(m, n);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < n; ++j) a.at(i, j) = static_cast<E>(A[static_cast<int64_t>(i) * lda + j]);
            const engine::ring2k::HowellForm<E> h =
                engine::ring2k::howell<engine::ring2k::Traits<E>>(std::move(a), k);
            const uint32_t hr = static_cast<uint32_t>(h.pivots.size());
            for (uint32_t i = 0; i < hr; ++i) {
                pivots[i] = h.pivots[i];
                pivot_v[i] = h.pivot_v[i];
                for (uint32_t j = 0; j < n; ++j) H[static_cast<size_t>(i) * n + j] = h.H.at(i, j);
            }
            return hr;
        };
        if (rung == 16) return run(uint16_t{});
        return run(uint32_t{});
    } catch (...) {
        return -1;
    }
}

// engine::ring2k::echelon_blocked<Traits<E>> (Binary/Ring2k.h:538-694), MinValuation, on the u32 (elem 4) or u64
// (elem 8) rung: A (m x n words of elem bytes, row stride lda) in place; returns the rank, pivots / pivot_v filled.
int64_t hxr_r2_echelon(uint32_t elem, uint32_t k, uint32_t m, uint32_t n, void* A, int64_t lda, uint32_t panel,
                       uint32_t climit, uint32_t* pivots, uint32_t* pivot_v) {
    try {
        auto run = [&](auto tag) -> int64_t {
            using E = decltype(tag);
            E* src = static_cast<E*>(A);
            engine::ring2k::Mat<E> a(m, n);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < n; ++j) a.at(i, j) = src[static_cast<int64_t>(i) * lda + j];
            const engine::ring2k::EchelonInfo e = engine::ring2k::echelon_blocked<engine::ring2k::Traits<E>>(
                a, k, engine::ring2k::PivotRule::MinValuation, panel, climit);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < n; ++j) src[static_cast<int64_t>(i) * lda + j] = a.at(i, j);
            for (size_t t = 0; t < e.pivots.size(); ++t) {
                pivots[t] = e.pivots[t];
                pivot_v[t] = e.pivot_v[t];
            }
            return e.rank;
        };
        if (elem == 8) return run(uint64_t{});
        return run(uint32_t{});
    } catch (...) {
        return -1;
    }
}

// engine::ring2k::smith (valuation levels, Binary/Ring2k.h:701) or smith_naive (global minimal-valuation
// pivoting) on a copy of A over Z/2^k (elem 4: k <= 32, elem 8: k <= 64): sigma ascending (room for
// min(m, n)) and, for smith, level_ranks (room for k; *nlevels set).  Returns the number of invariants.
int64_t hxr_smith(uint32_t elem, uint32_t k, uint32_t m, uint32_t n, const void* A, int64_t lda, int naive,
                  uint32_t* sigma, uint32_t* level_ranks, uint32_t* nlevels) {
    try {
        auto run = [&](auto tag) -> int64_t {
            using E = decltype(tag);
            const E* src = static_cast<const E*>(A);
            engine::ring2k::Mat<E> a(m, n);
            for (uint32_t i = 0; i < m; ++i)
                for (uint32_t j = 0; j < n; ++j) a.at(i, j) = src[static_cast<int64_t>(i) * lda + j];
            const engine::ring2k::SmithResult r =
                naive ? engine::ring2k::smith_naive<engine::ring2k::Traits<E>>(std::move(a), k)
                      : engine::ring2k::smith<engine::ring2k::Traits<E>>(std::move(a), k);
            for (size_t t = 0; t < r.sigma.size(); ++t) sigma[t] = r.sigma[t];
            if (nlevels) *nlevels = static_cast<uint32_t>(r.level_ranks.size());
            if (level_ranks)
                for (size_t t = 0; t < r.level_ranks.size(); ++t) level_ranks[t] = r.level_ranks[t];
            return static_cast<int64_t>(r.sigma.size());
        };
        if (elem == 8) return run(uint64_t{});
        return run(uint32_t{});
    } catch (...) {
        return -1;
    }
}

// engine::ref::howell<uint32_t>(2, k, M) (RefLinAlg.cpp:231-290), the naive Sec. 13.5 pass; k <= 31.
int64_t hxr_ref_howell(uint32_t k, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* H,
                       uint32_t* pivots, uint32_t* pivot_v) {
    try {
        DenseBuf<uint32_t> a = to_buf(A, m, n, lda);
        const engine::ref::RefHowell<uint32_t> h = engine::ref::howell<uint32_t>(2u, k, DenseView<const uint32_t>(a));
        for (size_t i = 0; i < h.rows.size(); ++i) {
            pivots[i] = h.pivots[i];
            pivot_v[i] = h.pivot_v[i];
            for (uint32_t j = 0; j < n; ++j) H[i * n + j] = h.rows[i][j];
        }
        return static_cast<int64_t>(h.rows.size());
    } catch (...) {
        return -1;
    }
}

// engine::elim::howell (DenseElim.cpp:1450) with ModCtx::prime_power(2, k), k >= 2.
int64_t hxr_elim_howell(uint32_t k, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* H,
                        uint32_t* pivots, uint32_t* pivot_v) {
    try {
        const ModCtx<uint32_t> ctx = ModCtx<uint32_t>::prime_power(2u, k);
        DenseBuf<uint32_t> a = to_buf(A, m, n, lda);
        const engine::elim::HowellResult<uint32_t> h = engine::elim::howell<uint32_t>(ctx, DenseView<const uint32_t>(a));
        const uint32_t hr = h.rows();
        for (uint32_t i = 0; i < hr; ++i) {
            pivots[i] = h.pivots[i];
            pivot_v[i] = h.pivot_v[i];
            for (uint32_t j = 0; j < n; ++j) H[static_cast<size_t>(i) * n + j] = h.H.at(i, j);
        }
        return hr;
    } catch (...) {
        return -1;
    }
}

// engine::elim::howell (DenseElim.cpp:1450) over Z/ell^k: ModCtx::prime_power(ell, k), k >= 2.
int64_t hxr_elim_howell_ell(uint32_t ell, uint32_t k, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda,
                            uint32_t* H, uint32_t* pivots, uint32_t* pivot_v) {
    try {
        const ModCtx<uint32_t> ctx = ModCtx<uint32_t>::prime_power(ell, k);
        DenseBuf<uint32_t> a = to_buf(A, m, n, lda);
        const engine::elim::HowellResult<uint32_t> h = engine::elim::howell<uint32_t>(ctx, DenseView<const uint32_t>(a));
        const uint32_t hr = h.rows();
        for (uint32_t i = 0; i < hr; ++i) {
            pivots[i] = h.pivots[i];
            pivot_v[i] = h.pivot_v[i];
            for (uint32_t j = 0; j < n; ++j) H[static_cast<size_t>(i) * n + j] = h.H.at(i, j);
        }
        return hr;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
