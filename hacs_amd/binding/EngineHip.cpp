// EngineHip.cpp -- the reference engine's device layer (namespace engine::cuda: Cuda/CudaCtx.h,
// Cuda/CudaKernels.h, Cuda/CudaRing2k.h) implemented on libhacs_mi355x.so's C-ABI
// (include/hacs_mi355x.h), plus the two hook tables ModArithmetic reads (the dense RREF hook,
// ModArithmetic/DenseElim.h:111-123, and the ring2k device Ops, Binary/Ring2kDevice.h:64-83).
//
// A maintainer builds this one translation unit in place of the reference's EngineCuda sources
// (CudaCtx.cpp, cuda_kernels.cu, ring2k_device.cu) and links libhacs_mi355x.so:
//   g++ -std=c++17 -O2 -DOC_ENABLE_LEGACY=0 -I<hacs>/src/Cuda -I<hacs>/src/ModArithmetic -I<hacs>/src
//       -I<this repo>/include -c EngineHip.cpp
// It includes only the reference's CUDA-header-free headers and the C-ABI: no HIP header, no CUDA
// header.  Every C-ABI failure is rethrown as engine::cuda::CudaError (argument errors as
// std::logic_error, the reference's ENGINE_CHECK), so callers see the reference's error classes.
// Call engine::cuda::hip::install_hooks() once at start-up (instead of the reference's own static
// installers in TapeElim.cpp and DescentSmithGpu.cpp) to route the dense RREF of every prime p < 2^31
// and the Z/2^32 / Z/2^64 LU through the device.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "CudaCtx.h"
#include "CudaKernels.h"
#include "CudaRing2k.h"
#include "DenseElim.h"
#include "hacs_mi355x.h"

namespace engine {
namespace cuda {

namespace {

[[noreturn]] void raise(int rc, const char* where) {
    const std::string msg = std::string(where) + ": " + hx_last_error();
    if (rc == HX_E_ARGUMENT) throw std::logic_error(msg);
    throw CudaError(msg);
}
inline void ok(int rc, const char* where) {
    if (rc != HX_OK) raise(rc, where);
}

hx_modctx ctx32(const ModCtx<uint32_t>& ctx) {
    hx_modctx c;
    ok(hx_modctx_init(ctx.p, &c), "ModCtx<uint32_t>");
    return c;
}
hx_modctx64 ctx64(const ModCtx<uint64_t>& ctx) {
    hx_modctx64 c;
    ok(hx_modctx64_init(ctx.p, &c), "ModCtx<uint64_t>");
    return c;
}
template <class E>
hx_view hv(const DeviceView<E>& v) {
    return hx_view{(void*)v.base, (int64_t)v.rs, (int64_t)v.cs, v.rows, v.cols};
}
// The reference's limb-route workspace (GemmOpts::imma_workspace) is the library's limb workspace;
// its split-K count and workspace (GemmOpts::splits / workspace) are the plain kernel's split-K.
hx_gemm_opts gopts(const GemmOpts& o) {
    hx_gemm_opts h{};
    h.variant = o.imma == 0 ? HX_GEMM_PLAIN : HX_GEMM_AUTO;
    h.workspace = o.imma_workspace;
    h.workspace_bytes = o.imma_workspace_bytes;
    h.splits = o.splits == 0 ? 1u : o.splits;       // the reference's default (1) never splits
    h.split_workspace = o.workspace;
    h.split_workspace_bytes = o.workspace_bytes;
    return h;
}

thread_local int t_capture_depth = 0;

}  // namespace

// ---- availability / policy (CudaCtx.h:56-73) -------------------------------------------------
bool available() { return hx_available() != 0; }
const char* unavailable_reason() { return hx_unavailable_reason(); }
Device requested_device() {
    const char* e = std::getenv("ENGINE_DEVICE");
    if (e == nullptr) return Device::Cpu;
    const std::string v(e);
    return v == "gpu" ? Device::Gpu : (v == "hybrid" ? Device::Hybrid : Device::Cpu);
}
bool device_enabled() { return available() && requested_device() != Device::Cpu; }
DeviceInfo device_info() {
    hx_device_info d;
    ok(hx_device_info_get(&d), "device_info");
    DeviceInfo r;
    r.name = d.name;
    r.index = d.index;
    r.sms = d.compute_units;
    r.total_mem = d.total_mem;
    r.free_mem = d.free_mem;
    r.driver_version = d.driver_version;
    r.runtime_version = d.runtime_version;
    return r;
}

// ---- the context (CudaCtx.h:118-214) -----------------------------------------------------------
struct CudaCtx::Impl {
    static constexpr int kStreams = 3;
    Stream streams[kStreams];
    std::mutex mu;
    TransferStats tr;
};

CudaCtx& CudaCtx::get() {
    if (!available()) throw CudaError(std::string("no usable device: ") + unavailable_reason());
    static CudaCtx* ctx = new CudaCtx();      // lives for the process (its streams outlive static destructors)
    return *ctx;
}
CudaCtx::CudaCtx() : impl_(new Impl()) {
    for (Stream& s : impl_->streams) ok(hx_stream_create(&s.h), "create_stream");
}
CudaCtx::~CudaCtx() {
    for (Stream& s : impl_->streams) (void)hx_stream_destroy(s.h);
    delete impl_;
}
int CudaCtx::num_streams() const { return Impl::kStreams; }
Stream CudaCtx::stream(int i) const {
    if (i < 0 || i >= Impl::kStreams) throw std::logic_error("CudaCtx::stream: index out of range");
    return impl_->streams[i];
}
Stream CudaCtx::create_stream() {
    Stream s;
    ok(hx_stream_create(&s.h), "create_stream");
    return s;
}
void CudaCtx::destroy_stream(Stream s) { ok(hx_stream_destroy(s.h), "destroy_stream"); }
void CudaCtx::sync(Stream s) const { ok(hx_stream_sync(s.h), "sync"); }
void CudaCtx::sync_all() const { ok(hx_device_sync(), "sync_all"); }
Event CudaCtx::create_event() const {
    Event e;
    ok(hx_event_create(&e.h, 0), "create_event");
    return e;
}
void CudaCtx::record(Event e, Stream s) const { ok(hx_event_record(e.h, s.h), "record"); }
void CudaCtx::wait(Stream s, Event e) const { ok(hx_event_wait(s.h, e.h), "wait"); }
void CudaCtx::sync(Event e) const { ok(hx_event_sync(e.h), "sync(event)"); }
void CudaCtx::destroy(Event e) const { ok(hx_event_destroy(e.h), "destroy(event)"); }
Event CudaCtx::create_timing_event() const {
    Event e;
    ok(hx_event_create(&e.h, 1), "create_timing_event");
    return e;
}
float CudaCtx::elapsed_ms(Event a, Event b) const {
    float ms = 0.0f;
    ok(hx_event_elapsed_ms(a.h, b.h, &ms), "elapsed_ms");
    return ms;
}
void* CudaCtx::dalloc(size_t bytes) {
    if (capturing_now()) throw std::logic_error("dalloc inside a graph capture");
    void* p = nullptr;
    ok(hx_malloc(&p, bytes), "dalloc");
    return p;
}
void CudaCtx::dfree(void* p, size_t) { ok(hx_free(p), "dfree"); }
ArenaStats CudaCtx::arena_stats() const {
    hx_arena_stats a;
    ok(hx_arena_stats_get(&a), "arena_stats");
    ArenaStats r;
    r.reserved_bytes = a.reserved_bytes;
    r.live_bytes = a.live_bytes;
    r.peak_live_bytes = a.peak_live_bytes;
    r.peak_reserved_bytes = a.peak_reserved_bytes;
    r.cuda_mallocs = a.driver_mallocs;
    r.reuses = a.reuses;
    r.cuda_frees = a.driver_frees;
    r.oom_retries = a.oom_retries;
    return r;
}
size_t CudaCtx::arena_idle_slack() const { return 0; }       // the library trims only on request or on failure
ReleaseStats CudaCtx::release_cached(bool) {
    size_t freed = 0;
    ok(hx_arena_trim(&freed), "release_cached");
    ReleaseStats r;
    r.arena_bytes_freed = freed;
    return r;
}
void CudaCtx::upload(const void* host, void* dev, size_t bytes, Stream s) {
    if (capturing_now()) throw std::logic_error("upload inside a graph capture");
    ok(hx_upload(dev, host, bytes, s.h), "upload");
}
void CudaCtx::download(void* host, const void* dev, size_t bytes, Stream s) {
    if (capturing_now()) throw std::logic_error("download inside a graph capture");
    ok(hx_download(host, dev, bytes, s.h), "download");
}
void CudaCtx::copy_d2d(void* dst, const void* src, size_t bytes, Stream s) { ok(hx_copy_d2d(dst, src, bytes, s.h), "copy_d2d"); }
void CudaCtx::memset_d(void* dst, int value, size_t bytes, Stream s) { ok(hx_memset(dst, value, bytes, s.h), "memset_d"); }
void CudaCtx::upload_pageable(const void* host, void* dev, size_t bytes) { ok(hx_upload(dev, host, bytes, nullptr), "upload_pageable"); }
void CudaCtx::download_pageable(void* host, const void* dev, size_t bytes) { ok(hx_download(host, dev, bytes, nullptr), "download_pageable"); }
void* CudaCtx::pinned_alloc(size_t bytes) {
    void* p = nullptr;
    ok(hx_host_alloc(&p, bytes), "pinned_alloc");
    return p;
}
void CudaCtx::pinned_free(void* p) { ok(hx_host_free(p), "pinned_free"); }
PinnedPoolStats CudaCtx::pinned_stats() const {
    hx_staging_stats st;
    ok(hx_staging_stats_get(&st), "pinned_stats");
    PinnedPoolStats r;
    r.chunk_bytes = st.chunk_bytes;
    r.cap_bytes = st.cap_bytes;
    r.allocated_bytes = st.allocated_bytes;
    r.peak_bytes = st.peak_bytes;
    r.acquisitions = st.acquisitions;
    r.waits = st.waits;
    return r;
}
TransferStats CudaCtx::transfer_stats() const {
    hx_stats st;
    ok(hx_stats_get(&st), "transfer_stats");
    TransferStats r;
    r.h2d_bytes = st.h2d_bytes;
    r.d2h_bytes = st.d2h_bytes;
    r.h2d_calls = st.h2d_calls;
    r.d2h_calls = st.d2h_calls;
    r.d2d_bytes = st.d2d_bytes;
    r.pageable_copies = 0;        // every hx_upload / hx_download is staged through pinned chunks
    return r;
}
void CudaCtx::reset_transfer_stats() { ok(hx_stats_reset(), "reset_transfer_stats"); }
size_t CudaCtx::pinned_cap_bytes() const { return pinned_stats().cap_bytes; }
size_t CudaCtx::pinned_chunk_bytes() const { return pinned_stats().chunk_bytes; }
int CudaCtx::copy_threads() const { return 0; }

ReleaseStats release_device_caches() {
    if (!available()) return ReleaseStats{};
    return CudaCtx::get().release_cached(true);
}
bool context_created() { return available(); }
namespace {
thread_local const char* t_trace_tag = nullptr;
}
void set_trace_tag(const char* tag) { t_trace_tag = tag; }
const char* trace_tag() { return t_trace_tag; }

// ---- graph capture (CudaCtx.h:223-253) -------------------------------------------------------
struct GraphCapture::Impl {
    Stream s;
    hx_graph g = nullptr;
    bool open = false;
};
GraphCapture::GraphCapture() : impl_(new Impl()) {}
GraphCapture::~GraphCapture() {
    if (impl_->g) (void)hx_graph_destroy(impl_->g);
    delete impl_;
}
void GraphCapture::begin(Stream s) {
    if (impl_->open) throw std::logic_error("GraphCapture::begin: already capturing");
    ok(hx_graph_begin(s.h), "GraphCapture::begin");
    impl_->s = s;
    impl_->open = true;
    ++t_capture_depth;
}
void GraphCapture::end() {
    if (!impl_->open) throw std::logic_error("GraphCapture::end: not capturing");
    impl_->open = false;
    --t_capture_depth;
    if (impl_->g) (void)hx_graph_destroy(impl_->g);
    impl_->g = nullptr;
    ok(hx_graph_end(impl_->s.h, &impl_->g), "GraphCapture::end");
}
void GraphCapture::launch(Stream s) const { ok(hx_graph_launch(impl_->g, s.h), "GraphCapture::launch"); }
bool GraphCapture::instantiated() const { return impl_->g != nullptr; }
size_t GraphCapture::num_nodes() const { return 0; }          // not exposed by the C-ABI
Stream GraphCapture::captured_stream() const { return impl_->s; }
bool GraphCapture::capturing() const { return impl_->open; }
bool capturing_now() { return t_capture_depth > 0; }

// ---- GEMM (CudaKernels.h:62-133) ---------------------------------------------------------------
size_t gemm_workspace_bytes(uint32_t m, uint32_t n, uint32_t splits, size_t elt_bytes) {
    // the partials are u32 words whatever elt_bytes; k is not known here, so no clamp to k
    return elt_bytes == 4 ? hx_gemm_split_workspace_bytes(m, n, 0xffffffffu, splits) : 0;
}
uint32_t suggest_splits(uint32_t m, uint32_t n, uint32_t k) { return hx_gemm_suggest_splits(m, n, k); }
bool imma_supported() { return available(); }
bool imma_policy() { return true; }

template <>
void gemm<uint32_t>(const ModCtx<uint32_t>& ctx, DeviceView<const uint32_t> A, DeviceView<const uint32_t> B,
                    DeviceView<uint32_t> C, GemmOp op, Stream s, const GemmOpts& opts) {
    const hx_modctx c = ctx32(ctx);
    const hx_gemm_opts o = gopts(opts);
    ok(hx_gemm_u32(&c, hv(A), hv(B), hv(C), (int)op, s.h, &o), "gemm<uint32_t>");
}
template <>
void gemm<uint64_t>(const ModCtx<uint64_t>& ctx, DeviceView<const uint64_t> A, DeviceView<const uint64_t> B,
                    DeviceView<uint64_t> C, GemmOp op, Stream s, const GemmOpts&) {
    const hx_modctx64 c = ctx64(ctx);
    ok(hx_gemm_u64(&c, hv(A), hv(B), hv(C), (int)op, s.h), "gemm<uint64_t>");
}
template <>
void gemm_nt<uint32_t>(const ModCtx<uint32_t>& ctx, DeviceView<const uint32_t> A, DeviceView<const uint32_t> B,
                       DeviceView<uint32_t> C, GemmOp op, Stream s, const GemmOpts& opts) {
    gemm<uint32_t>(ctx, A, B.transposed_view(), C, op, s, opts);
}
template <>
void gemm_nt<uint64_t>(const ModCtx<uint64_t>& ctx, DeviceView<const uint64_t> A, DeviceView<const uint64_t> B,
                       DeviceView<uint64_t> C, GemmOp op, Stream s, const GemmOpts& opts) {
    gemm<uint64_t>(ctx, A, B.transposed_view(), C, op, s, opts);
}
template <>
void schur_update<uint32_t>(const ModCtx<uint32_t>& ctx, DeviceView<const uint32_t> L, DeviceView<const uint32_t> U,
                            DeviceView<uint32_t> C, Stream s, const GemmOpts& opts) {
    gemm<uint32_t>(ctx, L, U, C, GemmOp::Sub, s, opts);
}
template <>
void schur_update<uint64_t>(const ModCtx<uint64_t>& ctx, DeviceView<const uint64_t> L, DeviceView<const uint64_t> U,
                            DeviceView<uint64_t> C, Stream s, const GemmOpts& opts) {
    gemm<uint64_t>(ctx, L, U, C, GemmOp::Sub, s, opts);
}

// ---- GemmBatch<uint32_t> (CudaKernels.h:141-185): the uneven batch of the C-ABI under the
//      reference's same-shape interface; the job table is built by upload(), run() is capturable --------
namespace {
std::mutex g_batch_mu;
std::unordered_map<const void*, hx_gemm_batch> g_batches;     // GemmBatch object -> library batch (per op)
std::unordered_map<const void*, int> g_batch_op;
}  // namespace

template <>
GemmBatch<uint32_t>::~GemmBatch() {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto it = g_batches.find(this);
    if (it != g_batches.end()) {
        (void)hx_gemm_batch_destroy(it->second);
        g_batches.erase(it);
        g_batch_op.erase(this);
    }
}
template <>
void GemmBatch<uint32_t>::set_shape(uint32_t m, uint32_t n, uint32_t k, size_t a_rs, size_t a_cs, size_t b_rs,
                                    size_t b_cs, size_t c_rs, size_t c_cs) {
    if (shape_set_) throw std::logic_error("GemmBatch::set_shape: shape already set");
    if (a_cs != 1 || b_cs != 1 || c_cs != 1)
        throw std::logic_error("GemmBatch (device): row-contiguous operands only");
    m_ = m; n_ = n; k_ = k;
    a_rs_ = a_rs; a_cs_ = a_cs; b_rs_ = b_rs; b_cs_ = b_cs; c_rs_ = c_rs; c_cs_ = c_cs;
    shape_set_ = true;
}
template <>
void GemmBatch<uint32_t>::set_shape(DeviceView<const uint32_t> A, DeviceView<const uint32_t> B, DeviceView<uint32_t> C) {
    set_shape(A.rows, B.cols, A.cols, A.rs, A.cs, B.rs, B.cs, C.rs, C.cs);
}
template <>
void GemmBatch<uint32_t>::push(DeviceView<const uint32_t> A, DeviceView<const uint32_t> B, DeviceView<uint32_t> C) {
    if (!shape_set_) set_shape(A, B, C);
    if (A.rows != m_ || B.cols != n_ || A.cols != k_ || B.rows != k_ || C.rows != m_ || C.cols != n_ || A.rs != a_rs_ ||
        B.rs != b_rs_ || C.rs != c_rs_ || A.cs != 1 || B.cs != 1 || C.cs != 1)
        throw std::logic_error("GemmBatch::push: shape or strides differ from set_shape");
    entries_.push_back(GemmBatchEntry<uint32_t>{A.base, B.base, C.base});
}
template <>
void GemmBatch<uint32_t>::upload(Stream) {
    if (capturing_now()) throw std::logic_error("GemmBatch::upload inside a graph capture");
    uploaded_ = entries_.size();      // the table goes to the device in run()'s first (uncaptured) call
}
template <>
bool GemmBatch<uint32_t>::prepare_imma(const ModCtx<uint32_t>&) { return true; }   // the library owns the limb planes
template <>
size_t GemmBatch<uint32_t>::imma_bytes_needed() const { return 0; }
template <>
void GemmBatch<uint32_t>::run(const ModCtx<uint32_t>& ctx, GemmOp op, Stream s, int, int imma, void*, size_t) const {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto it = g_batches.find(this);
    if (it == g_batches.end() || g_batch_op[this] != (int)op) {
        if (capturing_now()) throw std::logic_error("GemmBatch::run: first run (table build) inside a capture");
        if (it != g_batches.end()) (void)hx_gemm_batch_destroy(it->second);
        std::vector<hx_gemm_job> jobs(entries_.size());
        for (size_t i = 0; i < entries_.size(); ++i) {
            hx_gemm_job& j = jobs[i];
            j = hx_gemm_job{};
            j.m = m_; j.n = n_; j.k = k_;
            j.op = (int)op;
            j.variant = HX_GEMM_AUTO;
            j.A = entries_[i].A; j.B = entries_[i].B; j.C = entries_[i].C;
            j.lda = (int64_t)a_rs_; j.ldb = (int64_t)b_rs_; j.ldc = (int64_t)c_rs_;
        }
        const hx_modctx c = ctx32(ctx);
        hx_gemm_batch h = nullptr;
        ok(hx_gemm_batch_create(&c, jobs.data(), (uint32_t)jobs.size(), imma == 0 ? HX_GEMM_PLAIN : HX_GEMM_AUTO, &h),
           "GemmBatch::upload");
        g_batches[this] = h;
        g_batch_op[this] = (int)op;
        it = g_batches.find(this);
    }
    ok(hx_gemm_batch_run(it->second, s.h), "GemmBatch::run");
}

// ---- level-1/2 and tape support (CudaKernels.h:186-238) -------------------------------------------
template <>
void gemv<uint32_t>(const ModCtx<uint32_t>& ctx, DeviceView<const uint32_t> A, DeviceView<const uint32_t> x,
                    DeviceView<uint32_t> y, GemmOp op, Stream s) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_gemv_u32(&c, hv(A), hv(x), hv(y), (int)op, s.h), "gemv");
}
template <>
void axpy<uint32_t>(const ModCtx<uint32_t>& ctx, uint32_t alpha, DeviceView<const uint32_t> x, DeviceView<uint32_t> y,
                    bool subtract, Stream s) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_axpy_u32(&c, alpha, hv(x), hv(y), subtract ? 1 : 0, s.h), "axpy");
}
template <>
void scal<uint32_t>(const ModCtx<uint32_t>& ctx, uint32_t alpha, DeviceView<uint32_t> x, Stream s) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_scal_u32(&c, alpha, hv(x), s.h), "scal");
}
template <>
void copy<uint32_t>(DeviceView<const uint32_t> src, DeviceView<uint32_t> dst, Stream s) {
    ok(hx_copy_u32(hv(src), hv(dst), s.h), "copy");
}
template <>
void transpose<uint32_t>(DeviceView<const uint32_t> src, DeviceView<uint32_t> dst, Stream s) {
    ok(hx_transpose_u32(hv(src), hv(dst), s.h), "transpose");
}
template <>
void fill<uint32_t>(DeviceView<uint32_t> x, uint32_t value, Stream s) { ok(hx_fill_u32(hv(x), value, s.h), "fill"); }
template <>
void axpy_identity<uint32_t>(const ModCtx<uint32_t>& ctx, uint32_t alpha, DeviceView<uint32_t> y, bool subtract, Stream s) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_axpy_identity_u32(&c, alpha, hv(y), subtract ? 1 : 0, s.h), "axpy_identity");
}
template <>
void gather<uint32_t>(const ModCtx<uint32_t>& ctx, uint32_t alpha, uint32_t beta, DeviceView<const uint32_t> a,
                      uint32_t row_off, const uint32_t* cols, DeviceView<uint32_t> c, bool accumulate, Stream s,
                      const uint32_t* rows) {
    const hx_modctx k = ctx32(ctx);
    ok(hx_gather_u32(&k, alpha, beta, hv(a), row_off, cols, hv(c), accumulate ? 1 : 0, rows, s.h), "gather");
}
template <>
void canonicalise<uint32_t>(const ModCtx<uint32_t>& ctx, DeviceView<uint32_t> v, Stream s) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_canonicalise_u32(&c, hv(v), s.h), "canonicalise");
}
void canonicalise(const ModCtx<uint32_t>& ctx, DeviceView<const uint64_t> acc, DeviceView<uint32_t> dst, Stream s) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_canonicalise_acc64_u32(&c, hv(acc), hv(dst), s.h), "canonicalise(Accum64)");
}
template <>
void trsm_unit<uint32_t>(const ModCtx<uint32_t>& ctx, DeviceView<const uint32_t> T, DeviceView<uint32_t> B, Stream s,
                         const GemmOpts&) {
    const hx_modctx c = ctx32(ctx);
    ok(hx_trsm_unit_u32(&c, hv(T), hv(B), T.structure == Structure::UpperUnitri ? 1 : 0, s.h), "trsm_unit");
}

// ---- the device RREF (CudaKernels.h:249-298), every prime p < 2^31 -----------------------------------
namespace {
std::mutex g_elim_mu;
ImmaElimStats g_elim;
}  // namespace
ImmaRrefResult imma_rref(const ModCtx<uint32_t>& ctx, uint32_t m, uint32_t n, const ImmaRowSource& src,
                         const ImmaRowSink& sink, uint32_t slab_rows, uint32_t) {
    struct Ctx {
        const ImmaRowSource* src;
        const ImmaRowSink* sink;
    } u{&src, &sink};
    auto csrc = [](void* p, uint32_t r0, uint32_t nr, int64_t* ld) -> const uint32_t* {
        const DenseView<const uint32_t> v = (*static_cast<Ctx*>(p)->src)(r0, nr);
        *ld = (int64_t)v.row_stride();
        return v.data();
    };
    auto csink = [](void* p, uint32_t rank, int64_t* ld) -> uint32_t* {
        const DenseView<uint32_t> v = (*static_cast<Ctx*>(p)->sink)(rank);
        *ld = (int64_t)v.row_stride();
        return v.data();
    };
    const hx_modctx c = ctx32(ctx);
    ImmaRrefResult r;
    r.pivots.resize(std::max<uint32_t>(1, std::min(m, n)));
    r.pivot_rows.resize(r.pivots.size());
    ok(hx_rref_u32_host(&c, m, n, csrc, sink ? (hx_row_sink)csink : nullptr, &u, slab_rows, &r.rank, r.pivots.data(),
                        r.pivot_rows.data()),
       "imma_rref");
    r.pivots.resize(r.rank);
    r.pivot_rows.resize(r.rank);
    std::lock_guard<std::mutex> lk(g_elim_mu);
    ++g_elim.calls;
    g_elim.h2d_bytes += (uint64_t)m * n * 4;
    if (sink) g_elim.d2h_bytes += (uint64_t)r.rank * n * 4;
    return r;
}
ImmaElimStats imma_elim_stats() {
    std::lock_guard<std::mutex> lk(g_elim_mu);
    return g_elim;
}
void reset_imma_elim_stats() {
    std::lock_guard<std::mutex> lk(g_elim_mu);
    g_elim = ImmaElimStats{};
}
size_t imma_rref_device_bytes(uint32_t m, uint32_t n, uint32_t) { return hx_rref_workspace_bytes(m, n) + 2ull * m * n * 4; }

LaunchStats launch_stats() {
    hx_stats st;
    ok(hx_stats_get(&st), "launch_stats");
    LaunchStats r;
    r.kernel_launches = st.kernel_launches;
    r.gemm_launches = st.gemm_launches;
    r.batched_instances = st.batched_instances;
    r.imma_launches = st.limb_launches;
    return r;
}
void reset_launch_stats() { ok(hx_stats_reset(), "reset_launch_stats"); }
const char* kernel_layer_version() { return hx_version(); }

// ---- the ring2k session (CudaRing2k.h:28-46) ------------------------------------------------------
namespace ring2k {
MemInfo mem_info() {
    hx_device_info d;
    ok(hx_device_info_get(&d), "ring2k::mem_info");
    MemInfo r;
    r.free_bytes = d.free_mem;
    r.total_bytes = d.total_mem;
    return r;
}
size_t session_bytes(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) { return hx_ring2k_session_bytes(elem, m, cw, panel); }
size_t staging_bytes(uint32_t elem, uint32_t m, uint32_t panel) { return hx_ring2k_staging_bytes(elem, m, panel); }
void* open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) {
    hx_ring2k s = nullptr;
    ok(hx_ring2k_open(elem, m, cw, panel, &s), "ring2k::open");
    return s;
}
void put(void* s, const void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols) {
    ok(hx_ring2k_put(s, host, ldh, r0, rows, c0, cols), "ring2k::put");
}
void get(void* s, void* host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols) {
    ok(hx_ring2k_get(s, host, ldh, r0, rows, c0, cols), "ring2k::get");
}
void step(void* s, const engine::ring2k::device::StepArgs& a) {
    const hx_ring2k_step_args x{a.m, a.r, a.np, a.pc0, a.pc1, a.panel, a.ldp, a.swaps, a.nswaps,
                                a.L, a.ldl, a.u0, a.la0, a.la1, a.w};     // field for field
    ok(hx_ring2k_step(s, &x), "ring2k::step");
}
void fetch(void* s, void* host, size_t ldh) { ok(hx_ring2k_fetch(s, host, ldh), "ring2k::fetch"); }
void close(void* s) noexcept { (void)hx_ring2k_close(s); }
void gemm_sub_host(uint32_t elem, void* C, size_t ldc, const void* L, size_t ldl, const void* U, size_t ldu, uint32_t m,
                   uint32_t n, uint32_t k) {
    ok(hx_ring2k_gemm_sub_host(elem, C, ldc, L, ldl, U, ldu, m, n, k), "ring2k::gemm_sub_host");
}
}  // namespace ring2k

// ---- the hooks ModArithmetic reads ------------------------------------------------------------------
namespace hip {
namespace {

// The dense RREF of every prime p < 2^31 on the device (DenseElim.h:111-123).  Declines (false, A
// untouched) off the device policy, for a non-field context, p >= 2^31, or a view that is not row
// contiguous; a device error before the first write into A also declines, after it rethrows.
bool dense_rref_hook(const ModCtx<uint32_t>& ctx, DenseView<uint32_t> A, bool want_rref, elim::RrefResult& res) {
    if (!ctx.is_field() || ctx.p >= (1u << 31) || !device_enabled()) return false;
    const uint32_t m = A.rows(), n = A.cols();
    if (m == 0 || n == 0) return false;
    bool wrote = false;
    try {
        ImmaRowSink sink;
        if (want_rref)
            sink = [&](uint32_t rank) {
                wrote = true;
                return A.sub(0, rank, 0, n);
            };
        const ImmaRrefResult r =
            imma_rref(ctx, m, n, [&](uint32_t r0, uint32_t nr) { return DenseView<const uint32_t>(A.sub(r0, nr, 0, n)); },
                      sink);
        if (want_rref)
            for (uint32_t i = r.rank; i < m; ++i) {
                uint32_t* row = A.data() + (size_t)i * A.row_stride();
                std::fill(row, row + n, 0u);
            }
        res.rank = r.rank;
        res.pivots = r.pivots;
    } catch (const CudaError&) {
        if (wrote) throw;
        return false;
    }
    return true;
}

// The ring2k device Ops (Binary/Ring2kDevice.h:64-83) on the session above.
uint32_t ops_plan(uint32_t elem, uint32_t m, uint32_t n, uint32_t panel) {
    if (!device_enabled()) return 0;
    return hx_ring2k_plan(elem, m, n, panel, 0);
}
void* ops_open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel) { return ring2k::open(elem, m, cw, panel); }
bool ops_verify() { return hx_ring2k_verify() != 0; }
uint64_t ops_store_budget() { return hx_ring2k_store_budget(); }

const engine::ring2k::device::Ops g_ops = {
    &ops_plan, &ops_open, &ring2k::put, &ring2k::get, &ring2k::step, &ring2k::fetch, &ring2k::close,
    &ops_verify, &ops_store_budget};

}  // namespace

// Installs both hooks (call once, before the first elimination).
void install_hooks() {
    elim::set_device_rref_hook(&dense_rref_hook);
    engine::ring2k::device::set_ops(&g_ops);
}

}  // namespace hip
}  // namespace cuda
}  // namespace engine
