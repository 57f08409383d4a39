// hx_ring2k.hip -- the device side of the blocked LU over Z/2^32 and Z/2^64 (the ring2k::device
// Ops of Binary/Ring2kDevice.h:49-83; the reference's device session is Cuda/CudaRing2k.h:28-46).
//
// A session keeps a column window of the matrix resident in HBM: m rows x cw columns of native
// words (4 or 8 bytes), row-major with row stride cw.  Wraparound arithmetic is the words' own, so
// every entry gets the one value the host code gives it, whatever the tiling or summation order.
//
// One step (a panel the host has factored) is, on the session's compute stream:
//   1. the panel's factored block lands in the window (one contiguous copy into device staging, then a
//      kernel places rows r..m-1 of [pc0, pc1)), its L slots and the composed row permutation land in
//      device staging;
//   2. the panel's row swaps reach every other window column: the swap list is composed on the host
//      into one permutation of the rows it touches, applied per 64-column strip through LDS;
//   3. the unit-lower solve on the pivot rows r..r+np-1 as the product L11^-1 X (L11^-1 from the host):
//      one 256-thread workgroup per 64-column strip with the strip and L11^-1 in LDS;
//   4. the trailing update C[r+np.., u0..w) -= L21 U12 as a tiled wraparound product: 128 x 128
//      output tiles, the whole K (= np <= 64) of both operands staged in LDS once, 8 x 8 outputs per
//      thread.  The next panel's columns [la0, la1) are updated first and packed; a second stream
//      copies them to pinned host memory while the rest of the product runs (the lookahead fetch).
// Host staging is pinned and double-buffered by step parity (an event guards each slot's reuse), so
// step() returns as soon as its inputs are copied into staging.
//
// The host driver of the whole LU (hx_ring2k_echelon) is in hx_ring2k_lu.hip.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hx_common.h"

namespace hx {
namespace {

constexpr uint32_t kMaxPanel = 64;      // panel widths the kernels stage in LDS
constexpr int kStrip = 64;              // columns per strip (swaps, TRSM)
constexpr int kTile = 128;              // GEMM output tile edge
constexpr int kTileThreads = 256;       // 16 x 16 threads, 8 x 8 outputs each

// ---- kernels ------------------------------------------------------------------------------------

// Rows rows[0..nt) of the window get new(rows[i]) = old(src[i]) on every column of [0, w) outside
// [skip0, skip1).  One workgroup per 64-column strip; the touched rows' strip is staged in LDS.
template <class E>
__global__ __launch_bounds__(256) void r2_permute_kernel(E* win, uint32_t ld, uint32_t w, uint32_t skip0,
                                                         uint32_t skip1, const uint32_t* rows, const uint32_t* src_slot,
                                                         uint32_t nt) {
    extern __shared__ __align__(16) unsigned char r2_lds[];
    E* buf = (E*)r2_lds;                                  // nt x kStrip
    const uint32_t c0 = blockIdx.x * kStrip;
    const uint32_t lane = threadIdx.x % kStrip, grp = threadIdx.x / kStrip;   // 4 row groups
    const uint32_t c = c0 + lane;
    const bool live = c < w && (c < skip0 || c >= skip1);
    for (uint32_t t = grp; t < nt; t += 4) buf[t * kStrip + lane] = live ? win[(size_t)rows[t] * ld + c] : E(0);
    __syncthreads();
    if (!live) return;
    for (uint32_t t = grp; t < nt; t += 4) win[(size_t)rows[t] * ld + c] = buf[src_slot[t] * kStrip + lane];
}

// A rows x cols block of the window at (r0, c0) to / from a contiguous buffer (to_win: buffer -> window).
// Narrow blocks (a panel: 64 words a row) move between host and device as one contiguous copy plus this
// kernel, not as a strided copy of one short row after another.
template <class E>
__global__ __launch_bounds__(256) void r2_block_kernel(E* win, uint32_t ld, uint32_t r0, uint32_t rows, uint32_t c0,
                                                       uint32_t cols, E* buf, int to_win) {
    const size_t total = (size_t)rows * cols;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        const uint32_t i = (uint32_t)(idx / cols), j = (uint32_t)(idx % cols);
        E* wp = win + (size_t)(r0 + i) * ld + c0 + j;
        if (to_win)
            *wp = buf[idx];
        else
            buf[idx] = *wp;
    }
}

// The unit-lower solve on the pivot rows r .. r+np-1 of the columns [u0, w) as one product:
// X <- L11^-1 X, with L11^-1 (np x np, unit lower, row-major) computed by the host (the forward
// substitution's result in the ring, bit for bit).  One 256-thread workgroup per 64-column strip: the
// strip's X and L11^-1 in LDS, lane = column, the four waves take rows t = wave, wave + 4, ...; no
// chain of dependent steps.
template <class E>
__global__ __launch_bounds__(256) void r2_trsm_kernel(E* win, uint32_t ld, uint32_t r, uint32_t np, uint32_t u0,
                                                      uint32_t w, const E* Linv) {
    extern __shared__ __align__(16) unsigned char r2_lds[];
    E* X = (E*)r2_lds;                                   // np x kStrip
    E* Li = X + (size_t)np * kStrip;                     // np x np
    const uint32_t lane = threadIdx.x % kStrip, grp = threadIdx.x / kStrip;
    const uint32_t c = u0 + blockIdx.x * kStrip + lane;
    const bool live = c < w;
    for (uint32_t t = grp; t < np; t += 4) X[t * kStrip + lane] = live ? win[(size_t)(r + t) * ld + c] : E(0);
    for (uint32_t i = threadIdx.x; i < np * np; i += 256) Li[i] = Linv[i];
    __syncthreads();
    if (!live) return;
    for (uint32_t t = grp + (grp == 0 ? 4 : 0); t < np; t += 4) {    // row 0 is unchanged
        const E* lt = Li + (size_t)t * np;
        E acc = X[t * kStrip + lane];                                  // L11^-1[t][t] = 1
        for (uint32_t s = 0; s < t; ++s) acc += lt[s] * X[s * kStrip + lane];
        win[(size_t)(r + t) * ld + c] = acc;
    }
}

// C (rows [r0, r1) of the window, columns [c0, c1)) -= L21 U12 with L21 = Ls rows (i - lrow0), K = np,
// U12 = window rows r .. r+np-1 of the same columns.  One 128 x 128 tile per workgroup; A (128 x np)
// and B (np x 128) staged whole in LDS (K <= 64), 8 x 8 outputs per thread.
template <class E>
__global__ __launch_bounds__(kTileThreads) void r2_gemm_kernel(E* win, uint32_t ld, uint32_t r, uint32_t np, uint32_t r0,
                                                               uint32_t r1, uint32_t c0, uint32_t c1, const E* Ls,
                                                               uint32_t ldl, uint32_t lrow0, uint32_t ntc) {
    extern __shared__ __align__(16) unsigned char r2_lds[];
    E* As = (E*)r2_lds;                                 // np x (kTile + 1): As[k][i], transposed (padded row)
    E* Bs = As + (size_t)np * (kTile + 1);              // np x kTile
    const uint32_t tm = blockIdx.x / ntc, tn = blockIdx.x % ntc;
    const uint32_t i0 = r0 + tm * kTile, j0 = c0 + tn * kTile;
    const uint32_t tid = threadIdx.x;
    for (uint32_t idx = tid; idx < kTile * np; idx += kTileThreads) {
        const uint32_t i = idx / np, k = idx % np;
        As[k * (kTile + 1) + i] = (i0 + i < r1) ? Ls[(size_t)(i0 + i - lrow0) * ldl + k] : E(0);
    }
    for (uint32_t idx = tid; idx < kTile * np; idx += kTileThreads) {
        const uint32_t k = idx / kTile, j = idx % kTile;
        Bs[k * kTile + j] = (j0 + j < c1) ? win[(size_t)(r + k) * ld + j0 + j] : E(0);
    }
    __syncthreads();
    const uint32_t tx = tid % 16, ty = tid / 16;        // columns tx + 16 q, rows ty + 16 p
    E acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = E(0);
    for (uint32_t k = 0; k < np; ++k) {
        E av[8], bv[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) av[a] = As[k * (kTile + 1) + ty + 16 * a];
#pragma unroll
        for (int b = 0; b < 8; ++b) bv[b] = Bs[k * kTile + tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) acc[a][b] += av[a] * bv[b];
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const uint32_t i = i0 + ty + 16 * a;
        if (i >= r1) continue;
        E* row = win + (size_t)i * ld;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint32_t j = j0 + tx + 16 * b;
            if (j < c1) row[j] -= acc[a][b];
        }
    }
}

// ---- session ------------------------------------------------------------------------------------

struct HostSlot {
    void* buf = nullptr;          // pinned: panel block | L block (m x panel words each) | permutation (4 x 64 u32)
    hipEvent_t done = nullptr;    // the copies out of this slot have completed
    bool used = false;
};

struct Session {
    uint32_t elem = 4, m = 0, cw = 0, panel = 0;
    void* win = nullptr;          // m x cw words
    void* d_L = nullptr;          // m x panel words (L slots of the step, rows r .. m-1)
    void* d_Linv = nullptr;       // panel x panel words: L11^-1 of the step
    void* d_pin = nullptr;        // m x panel words: the factored panel on its way into the window
    void* d_la = nullptr;         // m x panel words: the lookahead block on its way to the host
    uint32_t* d_perm = nullptr;   // 4 x panel: touched rows | their source slots
    void* h_la = nullptr;         // pinned lookahead block, m x panel words
    HostSlot slot[2];
    int parity = 0;
    hipStream_t s = nullptr, s_copy = nullptr;
    hipEvent_t ev_la_src = nullptr, ev_la_done = nullptr;
    bool la_pending = false;
    uint32_t la_rows = 0, la_cols = 0;
    size_t slot_bytes = 0;
};

size_t words_bytes(uint32_t elem, size_t words) { return words * elem; }

void session_free(Session* S) {
    if (!S) return;
    if (S->s) (void)hipStreamSynchronize(S->s);
    if (S->s_copy) (void)hipStreamSynchronize(S->s_copy);
    for (HostSlot& h : S->slot) {
        if (h.done) (void)hipEventDestroy(h.done);
        if (h.buf) (void)hipHostFree(h.buf);
    }
    if (S->h_la) (void)hipHostFree(S->h_la);
    if (S->ev_la_src) (void)hipEventDestroy(S->ev_la_src);
    if (S->ev_la_done) (void)hipEventDestroy(S->ev_la_done);
    if (S->win) arena_free(S->win);
    if (S->d_L) arena_free(S->d_L);
    if (S->d_Linv) arena_free(S->d_Linv);
    if (S->d_pin) arena_free(S->d_pin);
    if (S->d_la) arena_free(S->d_la);
    if (S->d_perm) arena_free(S->d_perm);
    if (S->s) (void)hipStreamDestroy(S->s);
    if (S->s_copy) (void)hipStreamDestroy(S->s_copy);
    delete S;
}

Session* session_open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) {
    HX_CHECK(elem == 4 || elem == 8, "ring2k: elem must be 4 or 8");
    HX_CHECK(m > 0 && cw > 0, "ring2k: empty window");
    HX_CHECK(panel >= 1 && panel <= kMaxPanel, "ring2k: panel must be 1 .. 64");
    Session* S = new Session();
    try {
        S->elem = elem;
        S->m = m;
        S->cw = cw;
        S->panel = panel;
        S->win = arena_alloc(words_bytes(elem, (size_t)m * cw));
        S->d_L = arena_alloc(words_bytes(elem, (size_t)m * panel));
        S->d_Linv = arena_alloc(words_bytes(elem, (size_t)kMaxPanel * kMaxPanel));
        S->d_pin = arena_alloc(words_bytes(elem, (size_t)m * panel));
        S->d_la = arena_alloc(words_bytes(elem, (size_t)m * panel));
        S->d_perm = (uint32_t*)arena_alloc(sizeof(uint32_t) * 4 * kMaxPanel);
        S->slot_bytes = 2 * words_bytes(elem, (size_t)m * panel) + sizeof(uint32_t) * 4 * kMaxPanel +
                        words_bytes(elem, (size_t)kMaxPanel * kMaxPanel);
        for (HostSlot& h : S->slot) {
            HX_HIP(hipHostMalloc(&h.buf, S->slot_bytes, hipHostMallocDefault));
            HX_HIP(hipEventCreateWithFlags(&h.done, hipEventDisableTiming));
        }
        HX_HIP(hipHostMalloc(&S->h_la, words_bytes(elem, (size_t)m * panel), hipHostMallocDefault));
        HX_HIP(hipStreamCreateWithFlags(&S->s, hipStreamNonBlocking));
        HX_HIP(hipStreamCreateWithFlags(&S->s_copy, hipStreamNonBlocking));
        HX_HIP(hipEventCreateWithFlags(&S->ev_la_src, hipEventDisableTiming));
        HX_HIP(hipEventCreateWithFlags(&S->ev_la_done, hipEventDisableTiming));
    } catch (...) {
        session_free(S);
        throw;
    }
    return S;
}

void copy_block(Session* S, void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols, bool to_dev) {
    HX_CHECK(r0 + (uint64_t)rows <= S->m && c0 + (uint64_t)cols <= S->cw, "ring2k: block outside the window");
    HX_CHECK(rows == 0 || cols == 0 || (host != nullptr && ldh >= cols), "ring2k: bad host block");
    if (rows == 0 || cols == 0) return;
    const size_t e = S->elem;
    char* d = (char*)S->win + ((size_t)r0 * S->cw + c0) * e;
    // the lookahead copy (second stream) may still read the window
    HX_HIP(hipStreamWaitEvent(S->s, S->ev_la_done, 0));
    if (cols == S->cw && ldh == cols) {
        // whole rows on both sides: one contiguous block through the runtime's pinned staging
        const size_t bytes = (size_t)rows * cols * e;
        host_status(to_dev ? hx_upload(d, host, bytes, (hx_stream)S->s) : hx_download(host, d, bytes, (hx_stream)S->s));
        return;
    }
    // a strided block through the runtime's pinned staging (a 2D copy straight into pageable memory has
    // returned from its stream synchronisation before its bytes had landed)
    if (to_dev)
        staged_copy_2d(d, (size_t)S->cw * e, host, ldh * e, (size_t)cols * e, rows, true, S->s);
    else
        staged_copy_2d(host, ldh * e, d, (size_t)S->cw * e, (size_t)cols * e, rows, false, S->s);
}

template <class E>
void launch_gemm(Session* S, uint32_t r, uint32_t np, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1, const E* Ls,
                 uint32_t ldl, uint32_t lrow0, hipStream_t st) {
    if (r1 <= r0 || c1 <= c0 || np == 0) return;
    const uint32_t ntm = (r1 - r0 + kTile - 1) / kTile, ntn = (c1 - c0 + kTile - 1) / kTile;
    const size_t lds = sizeof(E) * ((size_t)np * (kTile + 1) + (size_t)np * kTile);
    r2_gemm_kernel<E><<<ntm * ntn, kTileThreads, lds, st>>>((E*)S->win, S->cw, r, np, r0, r1, c0, c1, Ls, ldl, lrow0,
                                                             ntn);
    after_launch("r2_gemm_kernel");
}

template <class E>
void gemm_attrs() {
    static std::once_flag once;
    std::call_once(once, [] {
        const int most = (int)(sizeof(E) * ((size_t)kMaxPanel * (kTile + 1) + (size_t)kMaxPanel * kTile));
        (void)hipFuncSetAttribute((const void*)r2_gemm_kernel<E>, hipFuncAttributeMaxDynamicSharedMemorySize, most);
        (void)hipFuncSetAttribute((const void*)r2_permute_kernel<E>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(E) * 2 * kMaxPanel * kStrip));
        (void)hipFuncSetAttribute((const void*)r2_trsm_kernel<E>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(E) * ((size_t)kMaxPanel * kStrip + (size_t)kMaxPanel * kMaxPanel)));
    });
}

uint32_t block_grid(size_t words) { return (uint32_t)std::min<size_t>(2048, (words + 255) / 256); }

// The lookahead block (rows rb .. m-1 of the columns [la0, la1), as the stream has them now) to the
// pinned host block: packed on the compute stream, then one contiguous copy on the copy stream while the
// compute stream goes on.  The next step (and any window copy) waits for ev_la_done first, so the packed
// buffer is not overwritten before its copy has left.
template <class E>
void lookahead_copy(Session* S, uint32_t rb, uint32_t la0, uint32_t la1) {
    const size_t e = sizeof(E);
    S->la_rows = S->m - rb;
    S->la_cols = la1 - la0;
    const size_t words = (size_t)S->la_rows * S->la_cols;
    r2_block_kernel<E><<<block_grid(words), 256, 0, S->s>>>((E*)S->win, S->cw, rb, S->la_rows, la0, S->la_cols,
                                                             (E*)S->d_la, 0);
    after_launch("r2_block_kernel");
    HX_HIP(hipEventRecord(S->ev_la_src, S->s));
    HX_HIP(hipStreamWaitEvent(S->s_copy, S->ev_la_src, 0));
    HX_HIP(hipMemcpyAsync(S->h_la, S->d_la, words * e, hipMemcpyDeviceToHost, S->s_copy));
    HX_HIP(hipEventRecord(S->ev_la_done, S->s_copy));
    stat_add(ST_D2H_BYTES, (uint64_t)words * e);
    S->la_pending = true;
}

template <class E>
void step_impl(Session* S, const hx_ring2k_step_args& a) {
    const uint32_t m = S->m, w = a.w ? a.w : S->cw;
    HX_CHECK(a.m == m, "ring2k step: m differs from the session's");
    HX_CHECK(w <= S->cw, "ring2k step: w beyond the window");
    HX_CHECK(a.r <= m && a.np <= S->panel && a.r + (uint64_t)a.np <= m, "ring2k step: bad pivot rows");
    HX_CHECK(a.nswaps <= S->panel, "ring2k step: more swaps than panel columns");
    HX_CHECK(a.pc0 <= a.pc1 && a.pc1 <= w, "ring2k step: bad panel columns");
    HX_CHECK(a.la0 <= a.la1 && a.la1 <= w && a.la1 - a.la0 <= S->panel, "ring2k step: bad lookahead columns");
    gemm_attrs<E>();
    const uint32_t rows = m - a.r;
    const bool update = a.np > 0 && a.u0 < w;
    // ---- host staging: the L block and the composed permutation ----
    HostSlot& hs = S->slot[S->parity];
    S->parity ^= 1;
    if (hs.used) HX_HIP(hipEventSynchronize(hs.done));
    E* hP = (E*)hs.buf;
    E* hL = hP + (size_t)m * S->panel;
    E* hLi = hL + (size_t)m * S->panel;                         // np x np: L11^-1
    uint32_t* hperm = (uint32_t*)(hLi + (size_t)kMaxPanel * kMaxPanel);
    if (update) {
        HX_CHECK(a.L != nullptr && a.ldl >= a.np, "ring2k step: L missing");
        const E* L = (const E*)a.L;
        for (uint32_t i = 0; i < rows; ++i) std::memcpy(hL + (size_t)i * a.np, L + (size_t)i * a.ldl, sizeof(E) * a.np);
        // L11^-1 by columns: Li[t][j] = -sum_{j <= q < t} L11[t][q] Li[q][j], Li[j][j] = 1 (wraparound)
        const uint32_t np = a.np;
        for (uint32_t t = 0; t < np; ++t) {
            const E* lt = L + (size_t)t * a.ldl;
            E* it = hLi + (size_t)t * np;
            for (uint32_t j = 0; j < t; ++j) {
                E acc = lt[j];
                for (uint32_t q = j + 1; q < t; ++q) acc += lt[q] * hLi[(size_t)q * np + j];
                it[j] = (E)(E(0) - acc);
            }
            it[t] = E(1);
            for (uint32_t j = t + 1; j < np; ++j) it[j] = E(0);
        }
    }
    // the swaps, in order, composed into one permutation of the rows they touch: new(row) = old(src(row))
    std::vector<uint32_t> touched, srcrow;
    if (a.nswaps) {
        HX_CHECK(a.swaps != nullptr, "ring2k step: swaps missing");
        std::vector<std::pair<uint32_t, uint32_t>> cur;   // (row, the original row whose entries it now holds)
        auto find = [&](uint32_t row) -> size_t {
            for (size_t i = 0; i < cur.size(); ++i)
                if (cur[i].first == row) return i;
            cur.push_back({row, row});
            return cur.size() - 1;
        };
        for (uint32_t t = 0; t < a.nswaps; ++t) {
            const uint32_t x = a.swaps[2 * t], y = a.swaps[2 * t + 1];
            HX_CHECK(x < m && y < m, "ring2k step: swap row out of range");
            if (x == y) continue;
            const size_t ix = find(x), iy = find(y);
            std::swap(cur[ix].second, cur[iy].second);
        }
        std::sort(cur.begin(), cur.end());
        for (auto& pr : cur) touched.push_back(pr.first);
        for (auto& pr : cur)
            srcrow.push_back((uint32_t)(std::lower_bound(touched.begin(), touched.end(), pr.second) - touched.begin()));
    }
    const uint32_t nt = (uint32_t)touched.size();
    HX_CHECK(nt <= 2 * kMaxPanel, "ring2k step: the swaps touch more than 128 rows");
    std::memcpy(hperm, touched.data(), sizeof(uint32_t) * nt);
    std::memcpy(hperm + 2 * kMaxPanel, srcrow.data(), sizeof(uint32_t) * nt);

    hipStream_t s = S->s;
    // the previous lookahead copy reads the window: the swaps and updates below come after it
    HX_HIP(hipStreamWaitEvent(s, S->ev_la_done, 0));
    // 1. the factored panel into the window (rows r .. m-1 of its columns), through the slot
    if (a.panel && a.pc1 > a.pc0 && rows > 0) {
        const uint32_t pw = a.pc1 - a.pc0;
        HX_CHECK(pw <= S->panel && a.ldp >= pw, "ring2k step: panel wider than the session's, or ldp < width");
        const size_t e = sizeof(E);
        for (uint32_t i = 0; i < rows; ++i)
            std::memcpy(hP + (size_t)i * pw, (const E*)a.panel + (size_t)i * a.ldp, e * pw);
        HX_HIP(hipMemcpyAsync(S->d_pin, hP, e * pw * rows, hipMemcpyHostToDevice, s));
        r2_block_kernel<E><<<block_grid((size_t)rows * pw), 256, 0, s>>>((E*)S->win, S->cw, a.r, rows, a.pc0, pw,
                                                                         (E*)S->d_pin, 1);
        after_launch("r2_block_kernel");
        stat_add(ST_H2D_BYTES, (uint64_t)rows * pw * e);
    }
    if (nt) HX_HIP(hipMemcpyAsync(S->d_perm, hperm, sizeof(uint32_t) * 4 * kMaxPanel, hipMemcpyHostToDevice, s));
    if (update) {
        HX_HIP(hipMemcpyAsync(S->d_L, hL, sizeof(E) * (size_t)rows * a.np, hipMemcpyHostToDevice, s));
        HX_HIP(hipMemcpyAsync(S->d_Linv, hLi, sizeof(E) * (size_t)a.np * a.np, hipMemcpyHostToDevice, s));
    }
    HX_HIP(hipEventRecord(hs.done, s));
    hs.used = true;
    // 2. the swaps on every other window column
    if (nt) {
        r2_permute_kernel<E><<<(w + kStrip - 1) / kStrip, 256, sizeof(E) * nt * kStrip, s>>>(
            (E*)S->win, S->cw, w, a.pc0, a.pc1, S->d_perm, S->d_perm + 2 * kMaxPanel, nt);
        after_launch("r2_permute_kernel");
    }
    // 3. TRSM on the pivot rows, 4. the trailing product (lookahead columns first)
    S->la_pending = false;
    if (update) {
        const E* Ls = (const E*)S->d_L;          // row i of the block = global row r + i
        if (a.np > 1) {
            const size_t lds = sizeof(E) * ((size_t)a.np * kStrip + (size_t)a.np * a.np);
            r2_trsm_kernel<E><<<(w - a.u0 + kStrip - 1) / kStrip, 256, lds, s>>>((E*)S->win, S->cw, a.r, a.np, a.u0, w,
                                                                                (const E*)S->d_Linv);
            after_launch("r2_trsm_kernel");
        }
        const uint32_t rb = a.r + a.np;
        const bool la = a.la1 > a.la0 && rb < m;
        if (la) {
            HX_CHECK(a.la0 >= a.u0, "ring2k step: lookahead columns left of the update");
            launch_gemm<E>(S, a.r, a.np, rb, m, a.la0, a.la1, Ls + (size_t)a.np * a.np, a.np, rb, s);
            lookahead_copy<E>(S, rb, a.la0, a.la1);
            launch_gemm<E>(S, a.r, a.np, rb, m, a.u0, a.la0, Ls + (size_t)a.np * a.np, a.np, rb, s);
            launch_gemm<E>(S, a.r, a.np, rb, m, a.la1, w, Ls + (size_t)a.np * a.np, a.np, rb, s);
        } else {
            launch_gemm<E>(S, a.r, a.np, rb, m, a.u0, w, Ls + (size_t)a.np * a.np, a.np, rb, s);
        }
    } else if (a.la1 > a.la0 && a.r + a.np < m) {
        // no update, but the caller still wants the next panel's columns back
        lookahead_copy<E>(S, a.r + a.np, a.la0, a.la1);
    }
}

void fetch_impl(Session* S, void* host, size_t ldh) {
    HX_CHECK(S->la_pending, "ring2k fetch: the last step had no lookahead");
    HX_CHECK(host != nullptr && ldh >= S->la_cols, "ring2k fetch: bad host block");
    HX_HIP(hipEventSynchronize(S->ev_la_done));
    const size_t e = S->elem;
    for (uint32_t i = 0; i < S->la_rows; ++i)
        std::memcpy((char*)host + (size_t)i * ldh * e, (const char*)S->h_la + (size_t)i * S->la_cols * e, S->la_cols * e);
    S->la_pending = false;
}

}  // namespace

Ring2kWindow ring2k_window(hx_ring2k sess) {
    HX_CHECK(sess != nullptr, "NULL session");
    Session* S = (Session*)sess;
    HX_HIP(hipStreamWaitEvent(S->s, S->ev_la_done, 0));
    return Ring2kWindow{S->win, S->m, S->cw, S->s};
}

}  // namespace hx

using namespace hx;

extern "C" {

HX_API size_t hx_ring2k_session_bytes(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) {
    // the window, the L slots, the panel and lookahead staging, L11^-1, the permutation
    return (size_t)elem * ((size_t)m * cw + 3 * (size_t)m * panel + (size_t)kMaxPanel * kMaxPanel) +
           sizeof(uint32_t) * 4 * kMaxPanel;
}

HX_API size_t hx_ring2k_staging_bytes(uint32_t elem, uint32_t m, uint32_t panel) {
    return 3 * (size_t)elem * m * panel + 2 * (sizeof(uint32_t) * 4 * kMaxPanel + (size_t)elem * kMaxPanel * kMaxPanel);
}

HX_API int hx_ring2k_open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel, hx_ring2k* out) {
    return guard("hx_ring2k_open", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        *out = session_open(elem, m, cw, panel);
    });
}

HX_API int hx_ring2k_put(hx_ring2k s, const void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0,
                         uint32_t cols) {
    return guard("hx_ring2k_put", [&] {
        HX_CHECK(s != nullptr, "NULL session");
        copy_block((Session*)s, (void*)host, ldh, r0, rows, c0, cols, true);
    });
}

HX_API int hx_ring2k_get(hx_ring2k s, void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols) {
    return guard("hx_ring2k_get", [&] {
        HX_CHECK(s != nullptr, "NULL session");
        copy_block((Session*)s, host, ldh, r0, rows, c0, cols, false);
    });
}

HX_API int hx_ring2k_step(hx_ring2k s, const hx_ring2k_step_args* a) {
    return guard("hx_ring2k_step", [&] {
        HX_CHECK(s != nullptr && a != nullptr, "NULL argument");
        Session* S = (Session*)s;
        if (S->elem == 4)
            step_impl<uint32_t>(S, *a);
        else
            step_impl<uint64_t>(S, *a);
    });
}

HX_API int hx_ring2k_fetch(hx_ring2k s, void* host, size_t ldh) {
    return guard("hx_ring2k_fetch", [&] {
        HX_CHECK(s != nullptr, "NULL session");
        fetch_impl((Session*)s, host, ldh);
    });
}

HX_API int hx_ring2k_close(hx_ring2k s) {
    try {
        session_free((Session*)s);
    } catch (...) {
    }
    return HX_OK;
}

HX_API int hx_ring2k_gemm_sub_host(uint32_t elem, void* C, size_t ldc, const void* L, size_t ldl, const void* U,
                                   size_t ldu, uint32_t m, uint32_t n, uint32_t k) {
    return guard("hx_ring2k_gemm_sub_host", [&] {
        HX_CHECK(elem == 4 || elem == 8, "elem must be 4 or 8");
        if (m == 0 || n == 0 || k == 0) return;
        HX_CHECK(C && L && U && ldc >= n && ldl >= k && ldu >= n, "bad operands");
        // a session whose window holds U (rows 0 .. k-1) over C (rows k .. k+m-1); K in slices of <= 64
        const uint32_t kc_max = std::min<uint32_t>(k, kMaxPanel);
        Session* S = session_open(elem, k + m, n, kc_max);
        try {
            const size_t e = elem;
            copy_block(S, (void*)U, ldu, 0, k, 0, n, true);
            copy_block(S, C, ldc, k, m, 0, n, true);
            std::vector<char> Lb((size_t)m * kc_max * e);
            for (uint32_t k0 = 0; k0 < k; k0 += kMaxPanel) {
                const uint32_t kc = std::min<uint32_t>(kMaxPanel, k - k0);
                for (uint32_t i = 0; i < m; ++i)
                    std::memcpy(Lb.data() + (size_t)i * kc * e, (const char*)L + ((size_t)i * ldl + k0) * e, kc * e);
                host_status(hx_upload(S->d_L, Lb.data(), (size_t)m * kc * e, (hx_stream)S->s));
                if (elem == 4) {
                    gemm_attrs<uint32_t>();
                    launch_gemm<uint32_t>(S, k0, kc, k, k + m, 0, n, (const uint32_t*)S->d_L, kc, k, S->s);
                } else {
                    gemm_attrs<uint64_t>();
                    launch_gemm<uint64_t>(S, k0, kc, k, k + m, 0, n, (const uint64_t*)S->d_L, kc, k, S->s);
                }
                HX_HIP(hipStreamSynchronize(S->s));     // Lb is refilled by the next slice
            }
            copy_block(S, C, ldc, k, m, 0, n, false);
        } catch (...) {
            session_free(S);
            throw;
        }
        session_free(S);
    });
}

}  // extern "C"
