// hx_common.h -- internal helpers shared by the HIP sources of libhacs_mi355x.so:
// error plumbing for the C-ABI and the device-side exact modular arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/hacs_mi355x.h"

namespace hx {

// ---- host: thread-local last error (the C form of CudaError, Cuda/CudaCtx.h:40-54) ----
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define HX_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::hx::Error{HX_E_RUNTIME, std::string(#call) + ": " + hipGetErrorString(e_) + " (" + \
                                                __FILE__ + ":" + std::to_string(__LINE__) + ")"};      \
    } while (0)

#define HX_CHECK(cond, msg)                                                                   \
    do {                                                                                      \
        if (!(cond)) throw ::hx::Error{HX_E_ARGUMENT, std::string(msg)};                      \
    } while (0)

// Runs body, converting hx::Error / std::exception into a status code + last error.
template <class F>
int guard(const char* where, F&& body) {
    try {
        body();
        return HX_OK;
    } catch (const Error& e) {
        return fail(e.code, std::string(where) + ": " + e.msg);
    } catch (const std::exception& e) {
        return fail(HX_E_RUNTIME, std::string(where) + ": " + e.what());
    }
}

// Process-wide diagnostic counters (launch_stats, Cuda/CudaKernels.h:299-306; transfer counters,
// Cuda/CudaCtx.h:94-101).  Relaxed atomics: counts only, never synchronisation.
enum StatId {
    ST_KERNELS, ST_GEMMS, ST_BATCHED, ST_LIMB, ST_H2D_BYTES, ST_D2H_BYTES, ST_D2D_BYTES, ST_H2D_CALLS,
    ST_D2H_CALLS, ST_MALLOCS, ST_FREES, ST_SPLITK, ST_COUNT
};
void stat_add(StatId id, uint64_t n);
uint64_t stat_value(StatId id);

// Elimination counters (ImmaElimStats, Cuda/CudaKernels.h:289-298), diagnostics only.
enum ElimStatId { EL_CALLS, EL_PANELS, EL_PANEL_COLUMNS, EL_LAUNCHES, EL_GEMMS, EL_H2D, EL_D2H, EL_SPEC_PANELS, EL_COUNT };
void elim_stat_add(ElimStatId id, uint64_t n);
void elim_peak_bytes(size_t bytes);          // keeps the largest single call's device footprint
void elim_seconds_add(double seconds);

// Checks the launch just issued (and counts it).
inline void after_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{HX_E_RUNTIME, std::string(what) + ": " + hipGetErrorString(e)};
    stat_add(ST_KERNELS, 1);
}

// Whether the byte ranges spanned by two views intersect (windows_overlap, cuda_kernels.cu:384-394,
// taken conservatively over the bounding byte ranges); eb = bytes per element of each view.
inline bool views_overlap(const hx_view& a, const hx_view& b, int eba = 4, int ebb = 4) {
    if (a.rows == 0 || a.cols == 0 || b.rows == 0 || b.cols == 0) return false;
    auto span = [](const hx_view& v, int eb, uintptr_t& lo, uintptr_t& hi) {
        const int64_t i1 = (int64_t)(v.rows - 1) * v.rs, j1 = (int64_t)(v.cols - 1) * v.cs;
        int64_t mn = 0, mx = 0;
        for (int64_t e : {i1, j1, i1 + j1}) {
            mn = e < mn ? e : mn;
            mx = e > mx ? e : mx;
        }
        lo = (uintptr_t)v.base + mn * eb;
        hi = (uintptr_t)v.base + mx * eb + eb;
    };
    uintptr_t alo, ahi, blo, bhi;
    span(a, eba, alo, ahi);
    span(b, ebb, blo, bhi);
    return !(ahi <= blo || bhi <= alo);
}

// Device arena (host side): hx_malloc / hx_free and every library-owned buffer go through it.
void* arena_alloc(size_t bytes);
// In-place RREF of m rows (row_words 64-bit words each, contiguous) with a one-pass routine that takes at most
// `cap` rows: the RREF of the rows so far (r <= cap nonzero rows) is stacked with the next cap - r rows and
// reduced again until every row has been taken, then written back (zero rows below).  Exact for any field:
// the RREF is a function of the row space alone, and stacking keeps the row space.  `one(rows, W, rank,
// pivots)` reduces W's first rows in place (pivots only asked for on the last pass); fails with `what` when
// the rank reaches cap while rows are left.
void rref_stacked(uint32_t m, uint64_t* A, size_t row_words, uint32_t cap, uint32_t* rank, uint32_t* pivots,
                  hipStream_t st, const std::function<void(uint32_t, uint64_t*, uint32_t*, uint32_t*)>& one,
                  const char* what);

void arena_free(void* p);
// Whether s is being captured into a graph (then nothing may be allocated or freed for it).
bool stream_capturing(hipStream_t s);
// Library-owned scratch of at least `bytes`, private to stream s and one purpose (below); grows
// (after waiting for s) but never shrinks; nullptr when it would have to grow inside a capture.
enum ScratchPurpose { SCRATCH_LIMB = 0, SCRATCH_MLP = 1, SCRATCH_RREF = 2, SCRATCH_SPLIT = 3 };
void* stream_scratch(hipStream_t s, int purpose, size_t bytes);
void stream_scratch_release(hipStream_t s);
// A rows x width-byte block between pageable host memory and the device (either direction, pitches in
// bytes) through the runtime's pinned staging chunks, ordered on st; returns once it has landed.
void staged_copy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                    bool to_dev, hipStream_t st);
// hx_upload / hx_download's status turned into an Error.
inline void host_status(int rc) {
    if (rc != HX_OK) throw Error{rc, hx_last_error()};
}

// A ring2k session's resident window (m x cw words, row stride cw) and its compute stream, for the
// drivers that run their own kernels between steps (the Smith levels).  The stream has been made to
// wait for any lookahead copy still reading the window.
struct Ring2kWindow {
    void* win;
    uint32_t m, cw;
    hipStream_t s;
};
Ring2kWindow ring2k_window(hx_ring2k sess);

// ---- device: exact arithmetic mod p, p < 2^32 (ModCtx.h:202-333) ----
struct DevMod {
    uint32_t p;
    uint64_t bm;        // floor(2^64 / p)
    uint64_t t64;       // 2^64 mod p
};

// x mod p for every x < 2^64 (ModCtx.h:323-329: q = mulhi(x, m), one correction).
__device__ __forceinline__ uint32_t red64(uint64_t x, const DevMod& M) {
    const uint64_t q = __umul64hi(x, M.bm);
    uint64_t r = x - q * (uint64_t)M.p;
    if (r >= M.p) r -= M.p;
    return (uint32_t)r;
}
__device__ __forceinline__ uint32_t addm(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t r = a + b;
    if (r < a || r >= p) r -= p;
    return r;
}
__device__ __forceinline__ uint32_t subm(uint32_t a, uint32_t b, uint32_t p) { return a >= b ? a - b : a + (p - b); }
__device__ __forceinline__ uint32_t mulm(uint32_t a, uint32_t b, const DevMod& M) {
    return red64((uint64_t)a * b, M);
}
// (hi * 2^64 + lo) mod p for hi < 2^32: the exact 96-bit accumulator of the plain kernels.
__device__ __forceinline__ uint32_t red96(uint32_t hi, uint64_t lo, const DevMod& M) {
    const uint32_t r = red64(lo, M);
    if (hi == 0) return r;
    return addm(r, red64((uint64_t)hi * M.t64, M), M.p);
}
// a^{-1} mod p by the extended Euclidean algorithm (ModArithmetic.cpp:54-78); 0 if gcd(a, p) != 1.
__device__ __forceinline__ uint32_t invm(uint32_t a, uint32_t p) {
    int64_t r0 = p, r1 = a % p, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const uint32_t q = (uint32_t)r0 / (uint32_t)r1;
        const int64_t r2 = r0 - (int64_t)q * r1, s2 = s0 - (int64_t)q * s1;
        r0 = r1; r1 = r2; s0 = s1; s1 = s2;
    }
    if (r0 != 1) return 0;
    int64_t s = s0 % (int64_t)p;
    if (s < 0) s += p;
    return (uint32_t)s;
}
__device__ __forceinline__ uint32_t apply_op(uint32_t c, uint32_t r, int op, uint32_t p) {
    return op == HX_OP_SET ? r : op == HX_OP_ADD ? addm(c, r, p) : subm(c, r, p);
}

inline DevMod devmod(const hx_modctx& c) { return DevMod{c.p, c.barrett_m, c.two64_mod}; }

}  // namespace hx
