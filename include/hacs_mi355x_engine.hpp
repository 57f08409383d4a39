// hacs_mi355x_engine.hpp -- header-only C++17 facade over include/hacs_mi355x.h with the shapes of the
// reference's device layer (namespace engine::cuda: Cuda/CudaCtx.h, Cuda/DeviceBuf.h, Cuda/CudaKernels.h,
// Cuda/CudaRing2k.h), so a caller written against that surface reads the same here:
//
//   engine::cuda::CudaError            -> hxe::CudaError (std::runtime_error, carries the C status)
//   engine::cuda::DeviceView<Elt>      -> hxe::DeviceView<Elt>     {base, rs, cs, rows, cols}
//   engine::kernels::GemmOp            -> hxe::GemmOp              {Set, Add, Sub}
//   engine::cuda::gemm / gemm_nt / schur_update / GemmBatch / trsm_unit / imma_rref
//                                      -> hxe::gemm / gemm_nt / schur_update / GemmBatch / trsm_unit / rref
//   engine::cuda::ring2k::{open, put, get, step, fetch, close}   -> hxe::ring2k::{...}
//
// Every function forwards to the C-ABI and rethrows a failure as CudaError.  No CUDA, no HIP and no
// torch type appears here: the facade compiles with any C++17 compiler and links libhacs_mi355x.so.
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "hacs_mi355x.h"

namespace hxe {

class CudaError : public std::runtime_error {
public:
    CudaError(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != HX_OK) throw CudaError(rc, hx_last_error());
}

inline bool available() { return hx_available() != 0; }
inline const char* unavailable_reason() { return hx_unavailable_reason(); }

struct Stream {
    hx_stream h = nullptr;
    bool valid() const { return h != nullptr; }
};
inline Stream create_stream() {
    Stream s;
    check(hx_stream_create(&s.h));
    return s;
}
inline void destroy_stream(Stream s) { check(hx_stream_destroy(s.h)); }
inline void sync(Stream s) { check(hx_stream_sync(s.h)); }

enum class GemmOp : int { Set = HX_OP_SET, Add = HX_OP_ADD, Sub = HX_OP_SUB };

// The device-relevant part of ModCtx<uint32_t> (ModCtx.h:202-250).
struct ModCtx32 {
    hx_modctx c{};
    explicit ModCtx32(uint32_t p) { check(hx_modctx_init(p, &c)); }
    uint32_t p() const { return c.p; }
    uint64_t k_max() const { return c.k_max; }
};

template <class Elt>
struct DeviceView {
    Elt* base = nullptr;
    int64_t rs = 0, cs = 0;
    uint32_t rows = 0, cols = 0;
    DeviceView() = default;
    DeviceView(Elt* b, int64_t rs_, int64_t cs_, uint32_t r, uint32_t c) : base(b), rs(rs_), cs(cs_), rows(r), cols(c) {}
    // mutable -> const, as DeviceView's converting constructor (Cuda/DeviceBuf.h:68-75)
    template <class E2, class = typename std::enable_if<std::is_const<Elt>::value &&
                                                        std::is_same<typename std::remove_const<Elt>::type, E2>::value>::type>
    DeviceView(const DeviceView<E2>& o) : base(o.base), rs(o.rs), cs(o.cs), rows(o.rows), cols(o.cols) {}
    DeviceView sub(uint32_t r0, uint32_t nr, uint32_t c0, uint32_t nc) const {
        if ((uint64_t)r0 + nr > rows || (uint64_t)c0 + nc > cols) throw CudaError(HX_E_ARGUMENT, "DeviceView::sub: window exceeds view");
        return DeviceView(base + r0 * rs + c0 * cs, rs, cs, nr, nc);
    }
    DeviceView transposed_view() const { return DeviceView(base, cs, rs, cols, rows); }
    hx_view c_view() const { return hx_view{(void*)base, rs, cs, rows, cols}; }
};

// A device buffer (row-major, ld = cols) owning its memory (DeviceBuf's device side).
class DeviceBuf {
public:
    DeviceBuf(uint32_t rows, uint32_t cols) : rows_(rows), cols_(cols) {
        check(hx_malloc((void**)&p_, (size_t)rows * cols * sizeof(uint32_t)));
    }
    ~DeviceBuf() { hx_free(p_); }
    DeviceBuf(const DeviceBuf&) = delete;
    DeviceBuf& operator=(const DeviceBuf&) = delete;
    uint32_t* dptr() { return p_; }
    DeviceView<uint32_t> view() { return {p_, (int64_t)cols_, 1, rows_, cols_}; }
    void upload(const uint32_t* host, Stream s) { check(hx_upload(p_, host, bytes(), s.h)); }
    void download(uint32_t* host, Stream s) const { check(hx_download(host, p_, bytes(), s.h)); }
    size_t bytes() const { return (size_t)rows_ * cols_ * sizeof(uint32_t); }

private:
    uint32_t rows_, cols_;
    uint32_t* p_ = nullptr;
};

struct GemmOpts {                 // GemmOpts (CudaKernels.h:64-79)
    int variant = HX_GEMM_AUTO;   // HX_GEMM_PLAIN / HX_GEMM_LIMB / AUTO
    void* workspace = nullptr;    // the limb planes; allocate before a capture
    size_t workspace_bytes = 0;
    uint32_t splits = 0;          // split-K of the plain kernel: 0 = automatic, 1 = never
    void* split_workspace = nullptr;
    size_t split_workspace_bytes = 0;
};

inline void gemm(const ModCtx32& ctx, DeviceView<const uint32_t> A, DeviceView<const uint32_t> B,
                 DeviceView<uint32_t> C, GemmOp op, Stream s, const GemmOpts& o = GemmOpts{}) {
    hx_gemm_opts co{o.variant, 0, o.workspace, o.workspace_bytes, nullptr, nullptr,
                    o.splits, o.split_workspace, o.split_workspace_bytes};
    check(hx_gemm_u32(&ctx.c, hx_view{(void*)A.base, A.rs, A.cs, A.rows, A.cols},
                      hx_view{(void*)B.base, B.rs, B.cs, B.rows, B.cols}, C.c_view(), (int)op, s.h, &co));
}
inline void gemm_nt(const ModCtx32& ctx, DeviceView<const uint32_t> A, DeviceView<const uint32_t> B,
                    DeviceView<uint32_t> C, GemmOp op, Stream s, const GemmOpts& o = GemmOpts{}) {
    gemm(ctx, A, B.transposed_view(), C, op, s, o);
}
inline void schur_update(const ModCtx32& ctx, DeviceView<const uint32_t> L, DeviceView<const uint32_t> U,
                         DeviceView<uint32_t> C, Stream s, const GemmOpts& o = GemmOpts{}) {
    gemm(ctx, L, U, C, GemmOp::Sub, s, o);
}
inline size_t gemm_workspace_bytes(const ModCtx32& ctx, uint32_t m, uint32_t n, uint32_t k, int variant) {
    return hx_gemm_workspace_bytes(&ctx.c, m, n, k, variant);
}

// GemmBatch: push uneven (row-major) triples, upload once, run (capturable).
class GemmBatch {
public:
    ~GemmBatch() { clear(); }
    void push(DeviceView<const uint32_t> A, DeviceView<const uint32_t> B, DeviceView<uint32_t> C, GemmOp op,
              int variant = HX_GEMM_AUTO) {
        if (A.cs != 1 || B.cs != 1 || C.cs != 1) throw CudaError(HX_E_ARGUMENT, "GemmBatch: row-major views only");
        hx_gemm_job j{};
        j.m = A.rows; j.k = A.cols; j.n = B.cols; j.op = (int)op; j.variant = variant;
        j.A = A.base; j.B = B.base; j.C = C.base; j.lda = A.rs; j.ldb = B.rs; j.ldc = C.rs;
        jobs_.push_back(j);
    }
    void upload(const ModCtx32& ctx, int variant = HX_GEMM_AUTO) {
        clear_handle();
        check(hx_gemm_batch_create(&ctx.c, jobs_.data(), (uint32_t)jobs_.size(), variant, &h_));
    }
    void run(Stream s) const { check(hx_gemm_batch_run(h_, s.h)); }
    size_t size() const { return jobs_.size(); }
    void clear() {
        clear_handle();
        jobs_.clear();
    }

private:
    void clear_handle() {
        if (h_) hx_gemm_batch_destroy(h_);
        h_ = nullptr;
    }
    std::vector<hx_gemm_job> jobs_;
    hx_gemm_batch h_ = nullptr;
};

// ---- level-1/2 and tape support (CudaKernels.h:186-238), LaunchStats (CudaKernels.h:299-306) ----
template <class E>
inline hx_view cview(DeviceView<E> v) { return hx_view{(void*)v.base, v.rs, v.cs, v.rows, v.cols}; }

inline void gemv(const ModCtx32& ctx, DeviceView<const uint32_t> A, DeviceView<const uint32_t> x,
                 DeviceView<uint32_t> y, GemmOp op, Stream s) {
    check(hx_gemv_u32(&ctx.c, cview(A), cview(x), cview(y), (int)op, s.h));
}
inline void axpy(const ModCtx32& ctx, uint32_t alpha, DeviceView<const uint32_t> x, DeviceView<uint32_t> y,
                 bool subtract, Stream s) {
    check(hx_axpy_u32(&ctx.c, alpha, cview(x), cview(y), subtract ? 1 : 0, s.h));
}
inline void scal(const ModCtx32& ctx, uint32_t alpha, DeviceView<uint32_t> x, Stream s) {
    check(hx_scal_u32(&ctx.c, alpha, cview(x), s.h));
}
inline void copy(DeviceView<const uint32_t> src, DeviceView<uint32_t> dst, Stream s) {
    check(hx_copy_u32(cview(src), cview(dst), s.h));
}
inline void transpose(DeviceView<const uint32_t> src, DeviceView<uint32_t> dst, Stream s) {
    check(hx_transpose_u32(cview(src), cview(dst), s.h));
}
inline void fill(DeviceView<uint32_t> x, uint32_t value, Stream s) { check(hx_fill_u32(cview(x), value, s.h)); }
inline void axpy_identity(const ModCtx32& ctx, uint32_t alpha, DeviceView<uint32_t> y, bool subtract, Stream s) {
    check(hx_axpy_identity_u32(&ctx.c, alpha, cview(y), subtract ? 1 : 0, s.h));
}
inline void gather(const ModCtx32& ctx, uint32_t alpha, uint32_t beta, DeviceView<const uint32_t> a, uint32_t row_off,
                   const uint32_t* cols, DeviceView<uint32_t> c, bool accumulate, Stream s,
                   const uint32_t* rows = nullptr) {
    check(hx_gather_u32(&ctx.c, alpha, beta, cview(a), row_off, cols, cview(c), accumulate ? 1 : 0, rows, s.h));
}
inline void canonicalise(const ModCtx32& ctx, DeviceView<uint32_t> v, Stream s) {
    check(hx_canonicalise_u32(&ctx.c, cview(v), s.h));
}
inline void canonicalise(const ModCtx32& ctx, DeviceView<const uint64_t> acc, DeviceView<uint32_t> dst, Stream s) {
    check(hx_canonicalise_acc64_u32(&ctx.c, cview(acc), cview(dst), s.h));
}
struct LaunchStats {
    uint64_t kernel_launches = 0, gemm_launches = 0, batched_instances = 0;
    uint64_t imma_launches = 0;   // the limb-MFMA route's launches
};
inline LaunchStats launch_stats() {
    hx_stats st;
    check(hx_stats_get(&st));
    return LaunchStats{st.kernel_launches, st.gemm_launches, st.batched_instances, st.limb_launches};
}
inline void reset_launch_stats() { check(hx_stats_reset()); }

inline void trsm_unit(const ModCtx32& ctx, DeviceView<const uint32_t> T, DeviceView<uint32_t> B, bool upper,
                      Stream s) {
    check(hx_trsm_unit_u32(&ctx.c, hx_view{(void*)T.base, T.rs, T.cs, T.rows, T.cols}, B.c_view(), upper ? 1 : 0,
                           s.h));
}

// imma_rref's result (CudaKernels.h:269-273) for every prime p < 2^31.
struct RrefResult {
    uint32_t rank = 0;
    std::vector<uint32_t> pivots;       // pivot columns, ascending
    std::vector<uint32_t> pivot_rows;   // the original row of pivot t (smallest-original-row rule)
};

// Device-resident form: A (m x n, row stride lda) on the device; R (>= m rows, row stride ldr) receives the
// RREF with zero rows last, or nullptr for rank / pivots / pivot_rows only.
inline RrefResult rref(const ModCtx32& ctx, uint32_t m, uint32_t n, const uint32_t* A, int64_t lda, uint32_t* R,
                       int64_t ldr, Stream s) {
    RrefResult r;
    const uint32_t mn = m < n ? m : n;
    r.pivots.resize(mn ? mn : 1);
    r.pivot_rows.resize(mn ? mn : 1);
    check(hx_rref_u32(&ctx.c, m, n, A, lda, R, ldr, &r.rank, r.pivots.data(), r.pivot_rows.data(), s.h));
    r.pivots.resize(r.rank);
    r.pivot_rows.resize(r.rank);
    return r;
}

// The row-slab form of imma_rref (CudaKernels.h:276-288): src(r0, nr, &ld) returns host rows, sink(rank, &ld)
// returns where the rank x n RREF rows go (an empty sink: rank / pivots / pivot_rows only).
using RowSource = std::function<const uint32_t*(uint32_t r0, uint32_t nr, int64_t* ld)>;
using RowSink = std::function<uint32_t*(uint32_t rank, int64_t* ld)>;
inline RrefResult rref(const ModCtx32& ctx, uint32_t m, uint32_t n, const RowSource& src, const RowSink& sink,
                       uint32_t slab_rows = 0) {
    struct Ctx {
        const RowSource* src;
        const RowSink* sink;
    } uc{&src, &sink};
    auto csrc = [](void* u, uint32_t r0, uint32_t nr, int64_t* ld) -> const uint32_t* {
        return (*((Ctx*)u)->src)(r0, nr, ld);
    };
    auto csink = [](void* u, uint32_t rank, int64_t* ld) -> uint32_t* { return (*((Ctx*)u)->sink)(rank, ld); };
    RrefResult r;
    const uint32_t mn = m < n ? m : n;
    r.pivots.resize(mn ? mn : 1);
    r.pivot_rows.resize(mn ? mn : 1);
    check(hx_rref_u32_host(&ctx.c, m, n, csrc, sink ? (hx_row_sink)csink : nullptr, &uc, slab_rows, &r.rank,
                           r.pivots.data(), r.pivot_rows.data()));
    r.pivots.resize(r.rank);
    r.pivot_rows.resize(r.rank);
    return r;
}

namespace ring2k {   // Cuda/CudaRing2k.h:28-46
inline size_t session_bytes(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) {
    return hx_ring2k_session_bytes(elem, m, cw, panel);
}
inline size_t staging_bytes(uint32_t elem, uint32_t m, uint32_t panel) { return hx_ring2k_staging_bytes(elem, m, panel); }
inline void* open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) {
    hx_ring2k s = nullptr;
    check(hx_ring2k_open(elem, m, cw, panel, &s));
    return s;
}
inline void put(void* s, const void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols) {
    check(hx_ring2k_put(s, host, ldh, r0, rows, c0, cols));
}
inline void get(void* s, void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols) {
    check(hx_ring2k_get(s, host, ldh, r0, rows, c0, cols));
}
inline void step(void* s, const hx_ring2k_step_args& a) { check(hx_ring2k_step(s, &a)); }
inline void fetch(void* s, void* host, size_t ldh) { check(hx_ring2k_fetch(s, host, ldh)); }
inline void close(void* s) noexcept { (void)hx_ring2k_close(s); }
}  // namespace ring2k

}  // namespace hxe
