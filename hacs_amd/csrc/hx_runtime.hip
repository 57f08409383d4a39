// hx_runtime.hip -- the runtime part of the C-ABI: availability, errors, the modulus,
// streams, events, memory, transfers, graph capture and device input generation.
// Replaces engine::cuda::CudaCtx / GraphCapture (Cuda/CudaCtx.h:51-253) for one
// process on one GPU: a caching device arena, per-stream library scratch, and pageable
// transfers staged through two pinned chunks (pinned host buffers go straight to the DMA).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "hx_common.h"

namespace hx {

static thread_local std::string g_last_error;

static std::atomic<uint64_t> g_stat[ST_COUNT];

void stat_add(StatId id, uint64_t n) { g_stat[id].fetch_add(n, std::memory_order_relaxed); }
uint64_t stat_value(StatId id) { return g_stat[id].load(std::memory_order_relaxed); }

static std::atomic<uint64_t> g_elim[EL_COUNT];
static std::atomic<uint64_t> g_elim_peak{0};
static std::mutex g_elim_mu;
static double g_elim_seconds = 0.0;
void elim_stat_add(ElimStatId id, uint64_t n) { g_elim[id].fetch_add(n, std::memory_order_relaxed); }
void elim_peak_bytes(size_t bytes) {
    uint64_t cur = g_elim_peak.load(std::memory_order_relaxed);
    while (bytes > cur && !g_elim_peak.compare_exchange_weak(cur, bytes, std::memory_order_relaxed)) {
    }
}
void elim_seconds_add(double seconds) {
    std::lock_guard<std::mutex> lk(g_elim_mu);
    g_elim_seconds += seconds;
}

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

namespace {

struct Avail {
    bool ok = false;
    std::string reason;
};

const Avail& avail() {
    static Avail a = [] {
        Avail r;
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            r.reason = std::string("hipGetDeviceCount: ") + hipGetErrorString(e);
        } else if (n <= 0) {
            r.reason = "no HIP device";
        } else {
            hipDeviceProp_t prop;
            e = hipGetDeviceProperties(&prop, 0);
            if (e != hipSuccess) {
                r.reason = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
            } else if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
                r.reason = std::string("device 0 is ") + prop.gcnArchName + ", this build targets gfx950 only";
            } else {
                r.ok = true;
            }
        }
        return r;
    }();
    return a;
}

// ---- device arena (the caching allocator of Cuda/CudaCtx.h:104-168, designed afresh) ----------
// Blocks are rounded up to a granule (512 B below 1 MiB, 2 MiB above) and cached on free in
// exact-size free lists; a request first takes a cached block of its rounded size.  A failed
// hipMalloc is answered by a trim (every cached block returned to the driver) and one retry.
// Nothing here is legal inside a stream capture (hipMalloc / hipFree are not): callers allocate
// before a capture begins.
struct Arena {
    std::mutex mu;
    std::unordered_map<size_t, std::vector<void*>> free_lists;   // rounded size -> cached blocks
    std::unordered_map<void*, size_t> live;                      // block -> rounded size
    hx_arena_stats st{};
};
Arena& arena() {
    static Arena* a = new Arena();   // never destroyed: blocks may be freed from static destructors
    return *a;
}

size_t arena_round(size_t bytes) {
    const size_t g = bytes < (1u << 20) ? 512 : (2u << 20);
    return (bytes + g - 1) / g * g;
}

size_t arena_trim_locked(Arena& a) {
    size_t freed = 0;
    for (auto& kv : a.free_lists) {
        for (void* p : kv.second) {
            (void)hipFree(p);
            freed += kv.first;
            a.st.reserved_bytes -= kv.first;
            ++a.st.driver_frees;
        }
        kv.second.clear();
    }
    a.free_lists.clear();
    return freed;
}

// ---- per-stream library scratch --------------------------------------------------------------
// One buffer per (stream, purpose): work queued on a stream only ever uses that stream's buffer,
// so calls on different streams never share bytes.  A buffer grows by first waiting for the work
// already queued on its stream (the only stream that reads it), then returning it to the arena.
struct ScratchKey {
    hipStream_t s;
    int purpose;
    bool operator==(const ScratchKey& o) const { return s == o.s && purpose == o.purpose; }
};
struct ScratchKeyHash {
    size_t operator()(const ScratchKey& k) const { return std::hash<void*>()((void*)k.s) * 31 + (size_t)k.purpose; }
};
struct ScratchSlot {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_scratch_mu;
std::unordered_map<ScratchKey, ScratchSlot, ScratchKeyHash> g_scratch;

// ---- pinned staging for pageable transfers (the CopyPool role, Cuda/CudaCtx.cpp:158-230) -------------
// Two pinned chunks of 16 MiB, allocated on first use: a transfer moves through them alternately, so
// the host memcpy of one chunk overlaps the DMA of the other (an event per chunk says when its DMA is
// done).  Host buffers that are already pinned skip the staging.  One transfer at a time (a mutex).
struct Staging {
    std::mutex mu;
    void* chunk[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};
    hx_staging_stats st{};
};
constexpr size_t kStageChunk = (size_t)16 << 20;
Staging& staging() {
    static Staging* s = new Staging();
    return *s;
}
void staging_init_locked(Staging& S) {
    if (S.chunk[0]) return;
    for (int i = 0; i < 2; ++i) {
        HX_HIP(hipHostMalloc(&S.chunk[i], kStageChunk, hipHostMallocDefault));
        HX_HIP(hipEventCreateWithFlags(&S.done[i], hipEventDisableTiming));
    }
    S.st.chunk_bytes = kStageChunk;
    S.st.cap_bytes = 2 * kStageChunk;
    S.st.allocated_bytes = 2 * kStageChunk;
    S.st.peak_bytes = 2 * kStageChunk;
}
bool host_is_pinned(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}
// Waits for chunk i's previous DMA (counted as a wait when it had not finished).
void chunk_acquire(Staging& S, int i) {
    ++S.st.acquisitions;
    if (!S.busy[i]) return;
    if (hipEventQuery(S.done[i]) == hipErrorNotReady) ++S.st.waits;
    (void)hipGetLastError();
    HX_HIP(hipEventSynchronize(S.done[i]));
    S.busy[i] = false;
}

}  // namespace

// A 2D block between pageable host memory and the device through the pinned chunks: row chunks
// packed into a chunk (pitch = width) and moved by one 2D DMA each, alternating chunks so that the
// host copy of one overlaps the DMA of the other.  Returns once every byte has landed.
void staged_copy_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                    bool to_dev, hipStream_t st) {
    if (width == 0 || rows == 0) return;
    HX_CHECK(width <= kStageChunk, "staged copy: one row exceeds a staging chunk");
    Staging& S = staging();
    std::lock_guard<std::mutex> lk(S.mu);
    staging_init_locked(S);
    const size_t rc = kStageChunk / width;
    const size_t nchunks = (rows + rc - 1) / rc;
    if (to_dev) {
        for (size_t c = 0; c < nchunks; ++c) {
            const int i = (int)(c & 1);
            const size_t r0 = c * rc, nr = std::min(rc, rows - r0);
            chunk_acquire(S, i);
            char* ch = (char*)S.chunk[i];
            for (size_t r = 0; r < nr; ++r) std::memcpy(ch + r * width, (const char*)src + (r0 + r) * spitch, width);
            HX_HIP(hipMemcpy2DAsync((char*)dst + r0 * dpitch, dpitch, ch, width, width, nr, hipMemcpyHostToDevice, st));
            HX_HIP(hipEventRecord(S.done[i], st));
            S.busy[i] = true;
        }
        HX_HIP(hipStreamSynchronize(st));
    } else {
        // chunk c's DMA is queued before chunk c-1 is copied out of its pinned buffer
        for (size_t c = 0; c <= nchunks; ++c) {
            if (c < nchunks) {
                const int i = (int)(c & 1);
                const size_t r0 = c * rc, nr = std::min(rc, rows - r0);
                chunk_acquire(S, i);
                HX_HIP(hipMemcpy2DAsync(S.chunk[i], width, (const char*)src + r0 * spitch, spitch, width, nr,
                                        hipMemcpyDeviceToHost, st));
                HX_HIP(hipEventRecord(S.done[i], st));
                S.busy[i] = true;
            }
            if (c > 0) {
                const int j = (int)((c - 1) & 1);
                const size_t r0 = (c - 1) * rc, nr = std::min(rc, rows - r0);
                HX_HIP(hipEventSynchronize(S.done[j]));
                S.busy[j] = false;
                const char* ch = (const char*)S.chunk[j];
                for (size_t r = 0; r < nr; ++r) std::memcpy((char*)dst + (r0 + r) * dpitch, ch + r * width, width);
            }
        }
    }
    stat_add(to_dev ? ST_H2D_BYTES : ST_D2H_BYTES, (uint64_t)rows * width);
    stat_add(to_dev ? ST_H2D_CALLS : ST_D2H_CALLS, 1);
}

void* arena_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    Arena& a = arena();
    const size_t r = arena_round(bytes);
    std::lock_guard<std::mutex> lk(a.mu);
    void* p = nullptr;
    auto it = a.free_lists.find(r);
    if (it != a.free_lists.end() && !it->second.empty()) {
        p = it->second.back();
        it->second.pop_back();
        ++a.st.reuses;
    } else {
        if (hipMalloc(&p, r) != hipSuccess) {
            (void)hipGetLastError();
            ++a.st.oom_retries;
            arena_trim_locked(a);
            if (hipMalloc(&p, r) != hipSuccess) {
                (void)hipGetLastError();
                throw Error{HX_E_NOMEM, "device allocation of " + std::to_string(r) + " bytes failed (after a trim)"};
            }
        }
        ++a.st.driver_mallocs;
        a.st.reserved_bytes += r;
        a.st.peak_reserved_bytes = std::max(a.st.peak_reserved_bytes, a.st.reserved_bytes);
    }
    a.live[p] = r;
    a.st.live_bytes += r;
    a.st.peak_live_bytes = std::max(a.st.peak_live_bytes, a.st.live_bytes);
    return p;
}

void arena_free(void* p) {
    if (p == nullptr) return;
    Arena& a = arena();
    std::lock_guard<std::mutex> lk(a.mu);
    auto it = a.live.find(p);
    if (it == a.live.end()) throw Error{HX_E_ARGUMENT, "pointer was not allocated by this library (or freed twice)"};
    const size_t r = it->second;
    a.live.erase(it);
    a.st.live_bytes -= r;
    a.free_lists[r].push_back(p);
}

bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

void* stream_scratch(hipStream_t s, int purpose, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchSlot& sl = g_scratch[ScratchKey{s, purpose}];
    if (sl.bytes >= bytes) return sl.ptr;
    if (stream_capturing(s)) return nullptr;          // no allocation inside a capture
    if (sl.ptr) {
        HX_HIP(hipStreamSynchronize(s));               // the queued work that reads the old buffer is done
        arena_free(sl.ptr);
        sl.ptr = nullptr;
        sl.bytes = 0;
    }
    const size_t want = bytes + bytes / 4 + 4096;
    sl.ptr = arena_alloc(want);
    sl.bytes = want;
    return sl.ptr;
}

void stream_scratch_release(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    for (auto it = g_scratch.begin(); it != g_scratch.end();) {
        if (it->first.s == s) {
            arena_free(it->second.ptr);
            it = g_scratch.erase(it);
        } else {
            ++it;
        }
    }
}

void rref_stacked(uint32_t m, uint64_t* A, size_t row_words, uint32_t cap, uint32_t* rank, uint32_t* pivots,
                  hipStream_t st, const std::function<void(uint32_t, uint64_t*, uint32_t*, uint32_t*)>& one,
                  const char* what) {
    const size_t rb = row_words * 8;
    struct Buf {                     // released on every exit, the stream drained first
        hipStream_t st;
        uint64_t* p = nullptr;
        ~Buf() {
            if (p) {
                (void)hipStreamSynchronize(st);
                arena_free(p);
            }
        }
    } W{st};
    W.p = (uint64_t*)arena_alloc((size_t)cap * rb);
    uint32_t r = 0, next = 0;
    while (next < m) {
        const uint32_t t = std::min<uint32_t>(cap - r, m - next);
        HX_CHECK(t > 0, what);
        HX_HIP(hipMemcpyAsync(W.p + (size_t)r * row_words, A + (size_t)next * row_words, (size_t)t * rb,
                              hipMemcpyDeviceToDevice, st));
        next += t;
        uint32_t rr = 0;
        one(r + t, W.p, &rr, next >= m ? pivots : nullptr);
        r = rr;
    }
    if (r > 0) HX_HIP(hipMemcpyAsync(A, W.p, (size_t)r * rb, hipMemcpyDeviceToDevice, st));
    if (r < m) HX_HIP(hipMemsetAsync(A + (size_t)r * row_words, 0, (size_t)(m - r) * rb, st));
    *rank = r;
}

}  // namespace hx

using namespace hx;

namespace {

__global__ void random_entries_kernel(uint32_t* out, uint32_t rows, uint32_t cols, int64_t ld, uint32_t ell,
                                      uint64_t seed) {
    const uint64_t total = (uint64_t)rows * cols;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        // SplitMix64's state after idx + 1 draws is seed + (idx + 1) * gamma: every draw is independent.
        uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t i = (uint32_t)(idx / cols), j = (uint32_t)(idx % cols);
        out[(int64_t)i * ld + j] = (uint32_t)(z % ell);
    }
}

__global__ void random_sparse_kernel(uint32_t* out, uint32_t rows, uint32_t cols, int64_t ld, uint32_t p,
                                     uint64_t seed) {
    const uint64_t total = (uint64_t)rows * cols;
    const uint32_t vals[4] = {(p - 2 % p) % p, p - 1, 1 % p, 2 % p};
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t i = (uint32_t)(idx / cols), j = (uint32_t)(idx % cols);
        out[(int64_t)i * ld + j] = ((z >> 32) % 20 == 0) ? vals[z & 3] : 0u;
    }
}

}  // namespace

extern "C" {

HX_API int hx_available(void) { return avail().ok ? 1 : 0; }
HX_API const char* hx_unavailable_reason(void) { return avail().ok ? "" : avail().reason.c_str(); }
HX_API const char* hx_last_error(void) { return g_last_error.c_str(); }
HX_API const char* hx_version(void) { return "hacs_mi355x 0.1 (gfx950)"; }

HX_API int hx_device_info_get(hx_device_info* out) {
    return guard("hx_device_info_get", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        if (!avail().ok) throw Error{HX_E_UNAVAILABLE, avail().reason};
        hipDeviceProp_t prop;
        HX_HIP(hipGetDeviceProperties(&prop, 0));
        std::memset(out, 0, sizeof(*out));
        std::strncpy(out->name, prop.name, sizeof(out->name) - 1);
        std::strncpy(out->arch, prop.gcnArchName, sizeof(out->arch) - 1);
        out->index = 0;
        out->compute_units = prop.multiProcessorCount;
        out->clock_khz = prop.clockRate;
        size_t fr = 0, tot = 0;
        HX_HIP(hipMemGetInfo(&fr, &tot));
        out->total_mem = tot;
        out->free_mem = fr;
        HX_HIP(hipRuntimeGetVersion(&out->runtime_version));
        HX_HIP(hipDriverGetVersion(&out->driver_version));
    });
}

HX_API int hx_modctx_init(uint32_t p, hx_modctx* out) {
    return guard("hx_modctx_init", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        HX_CHECK(p >= 2, "modulus must be >= 2");
        out->p = p;
        out->barrett_m = (uint64_t)(((unsigned __int128)1 << 64) / p);
        const uint64_t d = (uint64_t)(p - 1) * (p - 1);
        out->k_max = d == 0 ? UINT64_MAX : UINT64_MAX / d;
        out->two64_mod = (uint64_t)(((unsigned __int128)1 << 64) % p);
        // A balanced residue |x| <= (p - 1) / 2 fits L signed 8-bit digits when it is at most
        // 127 * (256^L - 1) / 255 (the positive end of the digit range).
        const uint64_t half = (uint64_t)(p - 1) / 2;
        uint32_t L = 0;
        for (uint32_t l = 1; l <= 4; ++l) {
            uint64_t cap = 0;
            for (uint32_t t = 0; t < l; ++t) cap = cap * 256 + 127;
            if (half <= cap) { L = l; break; }
        }
        out->limbs = L;
    });
}

HX_API int hx_stream_create(hx_stream* s) {
    return guard("hx_stream_create", [&] {
        hipStream_t h;
        HX_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
        *s = (hx_stream)h;
    });
}
HX_API int hx_stream_destroy(hx_stream s) {
    return guard("hx_stream_destroy", [&] {
        HX_HIP(hipStreamSynchronize((hipStream_t)s));   // its queued work may still read its scratch
        stream_scratch_release((hipStream_t)s);
        HX_HIP(hipStreamDestroy((hipStream_t)s));
    });
}
HX_API int hx_stream_sync(hx_stream s) {
    return guard("hx_stream_sync", [&] { HX_HIP(hipStreamSynchronize((hipStream_t)s)); });
}
HX_API int hx_device_sync(void) {
    return guard("hx_device_sync", [&] { HX_HIP(hipDeviceSynchronize()); });
}
HX_API int hx_event_create(hx_event* e, int timing) {
    return guard("hx_event_create", [&] {
        hipEvent_t h;
        HX_HIP(hipEventCreateWithFlags(&h, timing ? hipEventDefault : hipEventDisableTiming));
        *e = (hx_event)h;
    });
}
HX_API int hx_event_record(hx_event e, hx_stream s) {
    return guard("hx_event_record", [&] { HX_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s)); });
}
HX_API int hx_event_wait(hx_stream s, hx_event e) {
    return guard("hx_event_wait", [&] { HX_HIP(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0)); });
}
HX_API int hx_event_sync(hx_event e) {
    return guard("hx_event_sync", [&] { HX_HIP(hipEventSynchronize((hipEvent_t)e)); });
}
HX_API int hx_event_elapsed_ms(hx_event a, hx_event b, float* ms) {
    return guard("hx_event_elapsed_ms", [&] { HX_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b)); });
}
HX_API int hx_event_destroy(hx_event e) {
    return guard("hx_event_destroy", [&] { HX_HIP(hipEventDestroy((hipEvent_t)e)); });
}

HX_API int hx_stats_get(hx_stats* out) {
    return guard("hx_stats_get", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        auto g = [](StatId id) { return g_stat[id].load(std::memory_order_relaxed); };
        out->kernel_launches = g(ST_KERNELS);
        out->gemm_launches = g(ST_GEMMS);
        out->batched_instances = g(ST_BATCHED);
        out->limb_launches = g(ST_LIMB);
        out->h2d_bytes = g(ST_H2D_BYTES);
        out->d2h_bytes = g(ST_D2H_BYTES);
        out->d2d_bytes = g(ST_D2D_BYTES);
        out->h2d_calls = g(ST_H2D_CALLS);
        out->d2h_calls = g(ST_D2H_CALLS);
        out->mallocs = g(ST_MALLOCS);
        out->frees = g(ST_FREES);
        out->splitk_launches = g(ST_SPLITK);
    });
}
HX_API int hx_stats_reset(void) {
    for (auto& c : g_stat) c.store(0, std::memory_order_relaxed);
    return HX_OK;
}

HX_API int hx_elim_stats_get(hx_elim_stats* out) {
    return guard("hx_elim_stats_get", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        auto g = [](ElimStatId id) { return g_elim[id].load(std::memory_order_relaxed); };
        out->calls = g(EL_CALLS);
        out->panels = g(EL_PANELS);
        out->panel_columns = g(EL_PANEL_COLUMNS);
        out->kernel_launches = g(EL_LAUNCHES);
        out->gemms = g(EL_GEMMS);
        out->h2d_bytes = g(EL_H2D);
        out->d2h_bytes = g(EL_D2H);
        out->spec_panels = g(EL_SPEC_PANELS);
        out->peak_device_bytes = (size_t)g_elim_peak.load(std::memory_order_relaxed);
        std::lock_guard<std::mutex> lk(g_elim_mu);
        out->seconds = g_elim_seconds;
    });
}
HX_API int hx_elim_stats_reset(void) {
    for (auto& c : g_elim) c.store(0, std::memory_order_relaxed);
    g_elim_peak.store(0, std::memory_order_relaxed);
    std::lock_guard<std::mutex> lk(g_elim_mu);
    g_elim_seconds = 0.0;
    return HX_OK;
}

HX_API int hx_malloc(void** ptr, size_t bytes) {
    return guard("hx_malloc", [&] {
        HX_CHECK(ptr != nullptr, "ptr is NULL");
        *ptr = nullptr;
        if (bytes == 0) return;
        *ptr = arena_alloc(bytes);
        stat_add(ST_MALLOCS, 1);
    });
}
HX_API int hx_free(void* ptr) {
    return guard("hx_free", [&] {
        if (ptr) {
            arena_free(ptr);
            stat_add(ST_FREES, 1);
        }
    });
}
HX_API int hx_arena_stats_get(hx_arena_stats* out) {
    return guard("hx_arena_stats_get", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        Arena& a = arena();
        std::lock_guard<std::mutex> lk(a.mu);
        *out = a.st;
    });
}
HX_API int hx_arena_trim(size_t* freed) {
    return guard("hx_arena_trim", [&] {
        HX_CHECK(!stream_capturing(nullptr), "hx_arena_trim: not legal during a capture of the null stream");
        Arena& a = arena();
        std::lock_guard<std::mutex> lk(a.mu);
        HX_HIP(hipDeviceSynchronize());
        const size_t f = arena_trim_locked(a);
        if (freed) *freed = f;
    });
}
HX_API int hx_upload(void* dev, const void* host, size_t bytes, hx_stream s) {
    return guard("hx_upload", [&] {
        if (bytes == 0) return;
        HX_CHECK(dev != nullptr && host != nullptr, "NULL pointer");
        HX_CHECK(!stream_capturing((hipStream_t)s), "hx_upload is not legal inside a stream capture");
        const hipStream_t st = (hipStream_t)s;
        if (host_is_pinned(host)) {
            HX_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
        } else {
            Staging& S = staging();
            std::lock_guard<std::mutex> lk(S.mu);
            staging_init_locked(S);
            int i = 0;
            for (size_t off = 0; off < bytes; off += kStageChunk, i ^= 1) {
                const size_t n = std::min(kStageChunk, bytes - off);
                chunk_acquire(S, i);
                std::memcpy(S.chunk[i], (const char*)host + off, n);
                HX_HIP(hipMemcpyAsync((char*)dev + off, S.chunk[i], n, hipMemcpyHostToDevice, st));
                HX_HIP(hipEventRecord(S.done[i], st));
                S.busy[i] = true;
            }
        }
        // the host buffer is free again when this returns; the device side is stream-ordered (as the
        // reference's upload), so a later kernel on s reads the bytes
        HX_HIP(hipStreamSynchronize(st));
        stat_add(ST_H2D_BYTES, bytes);
        stat_add(ST_H2D_CALLS, 1);
    });
}
HX_API int hx_download(void* host, const void* dev, size_t bytes, hx_stream s) {
    return guard("hx_download", [&] {
        if (bytes == 0) return;
        HX_CHECK(dev != nullptr && host != nullptr, "NULL pointer");
        HX_CHECK(!stream_capturing((hipStream_t)s), "hx_download is not legal inside a stream capture");
        const hipStream_t st = (hipStream_t)s;
        if (host_is_pinned(host)) {
            HX_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
            HX_HIP(hipStreamSynchronize(st));
        } else {
            Staging& S = staging();
            std::lock_guard<std::mutex> lk(S.mu);
            staging_init_locked(S);
            // chunk c's DMA is queued before chunk c-1 is copied out of its pinned buffer
            const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
            for (size_t c = 0; c <= nchunks; ++c) {
                if (c < nchunks) {
                    const int i = (int)(c & 1);
                    chunk_acquire(S, i);
                    const size_t off = c * kStageChunk, n = std::min(kStageChunk, bytes - off);
                    HX_HIP(hipMemcpyAsync(S.chunk[i], (const char*)dev + off, n, hipMemcpyDeviceToHost, st));
                    HX_HIP(hipEventRecord(S.done[i], st));
                    S.busy[i] = true;
                }
                if (c > 0) {
                    const int j = (int)((c - 1) & 1);
                    const size_t off = (c - 1) * kStageChunk, n = std::min(kStageChunk, bytes - off);
                    HX_HIP(hipEventSynchronize(S.done[j]));
                    S.busy[j] = false;
                    std::memcpy((char*)host + off, S.chunk[j], n);
                }
            }
        }
        stat_add(ST_D2H_BYTES, bytes);
        stat_add(ST_D2H_CALLS, 1);
    });
}
HX_API int hx_host_alloc(void** ptr, size_t bytes) {
    return guard("hx_host_alloc", [&] {
        HX_CHECK(ptr != nullptr, "ptr is NULL");
        *ptr = nullptr;
        if (bytes == 0) return;
        HX_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    });
}
HX_API int hx_host_free(void* ptr) {
    return guard("hx_host_free", [&] {
        if (ptr) HX_HIP(hipHostFree(ptr));
    });
}
HX_API int hx_staging_stats_get(hx_staging_stats* out) {
    return guard("hx_staging_stats_get", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        Staging& S = staging();
        std::lock_guard<std::mutex> lk(S.mu);
        *out = S.st;
        if (S.st.chunk_bytes == 0) {      // not yet allocated: report the configuration
            out->chunk_bytes = kStageChunk;
            out->cap_bytes = 2 * kStageChunk;
        }
    });
}
HX_API int hx_copy_d2d(void* dst, const void* src, size_t bytes, hx_stream s) {
    return guard("hx_copy_d2d", [&] {
        if (bytes == 0) return;
        HX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
        stat_add(ST_D2D_BYTES, bytes);
    });
}
HX_API int hx_memset(void* dst, int value, size_t bytes, hx_stream s) {
    return guard("hx_memset", [&] {
        if (bytes == 0) return;
        HX_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)s));
    });
}

HX_API int hx_graph_begin(hx_stream s) {
    return guard("hx_graph_begin", [&] {
        HX_HIP(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeThreadLocal));
    });
}
HX_API int hx_graph_end(hx_stream s, hx_graph* g) {
    return guard("hx_graph_end", [&] {
        hipGraph_t graph;
        HX_HIP(hipStreamEndCapture((hipStream_t)s, &graph));
        hipGraphExec_t exec;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) throw Error{HX_E_RUNTIME, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
        *g = (hx_graph)exec;
    });
}
HX_API int hx_graph_launch(hx_graph g, hx_stream s) {
    return guard("hx_graph_launch", [&] { HX_HIP(hipGraphLaunch((hipGraphExec_t)g, (hipStream_t)s)); });
}
HX_API int hx_graph_destroy(hx_graph g) {
    return guard("hx_graph_destroy", [&] { HX_HIP(hipGraphExecDestroy((hipGraphExec_t)g)); });
}

HX_API int hx_random_entries(uint32_t* out, uint32_t rows, uint32_t cols, int64_t ld, uint32_t ell, uint64_t seed,
                             hx_stream s) {
    return guard("hx_random_entries", [&] {
        HX_CHECK(ell >= 1, "ell must be >= 1");
        HX_CHECK(ld >= (int64_t)cols, "ld < cols");
        if ((uint64_t)rows * cols == 0) return;
        const uint64_t total = (uint64_t)rows * cols;
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
        random_entries_kernel<<<blocks, 256, 0, (hipStream_t)s>>>(out, rows, cols, ld, ell, seed);
        after_launch("random_entries_kernel");
    });
}

HX_API int hx_random_sparse(uint32_t* out, uint32_t rows, uint32_t cols, int64_t ld, uint32_t p, uint64_t seed,
                            hx_stream s) {
    return guard("hx_random_sparse", [&] {
        HX_CHECK(p >= 3, "p must be >= 3");
        HX_CHECK(ld >= (int64_t)cols, "ld < cols");
        if ((uint64_t)rows * cols == 0) return;
        const uint64_t total = (uint64_t)rows * cols;
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
        random_sparse_kernel<<<blocks, 256, 0, (hipStream_t)s>>>(out, rows, cols, ld, p, seed);
        after_launch("random_sparse_kernel");
    });
}

}  // extern "C"
