// hx_gemm_noalloc.hip -- hx_gemm_u32_noalloc: the GEMM entry point that allocates nothing.
// For callers that cannot allocate device memory (code under graph capture, tight serving loops,
// harnesses that forbid it): the limb route's planes live in a fixed file-scope device array, K is
// chunked to fit it, and the plain route needs no scratch at all.  No counterpart in the reference.
#include <mutex>
#include <string>

#include "hx_gemm.h"

using namespace hx;

// ---- the no-allocation GEMM's limb planes -------------------------------------------------
// A file-scope device array: each GPU the library runs on holds its own copy until the process
// ends, and every hx_gemm_u32_noalloc call on that GPU uses it -- calls that overlap on one GPU
// (two streams, two host threads) race on it.  The prep kernels write every byte a product reads
// (zero padding included), so nothing carries over from the previous call.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) int8_t scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_noalloc_mu;
constexpr int kMaxDevices = 64;
void* g_noalloc_ptr[kMaxDevices] = {};    // the array's device address, per device

// The array's address on the current device, looked up once per device.  Host code never passes
// the array's name to a kernel or a copy (on the host it is a stand-in, not a device address).
void* noalloc_scratch() {
    int dev = 0;
    HX_HIP(hipGetDevice(&dev));
    HX_CHECK(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        HX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        HX_CHECK(ptr != nullptr, "hipGetSymbolAddress returned no address for the no-allocation scratch");
        g_noalloc_ptr[dev] = ptr;
    }
    return g_noalloc_ptr[dev];
}

GemmArgs view_args(hx_view A, hx_view B, hx_view C, int op) {
    GemmArgs g;
    g.m = A.rows;
    g.k = A.cols;
    g.n = B.cols;
    g.A = (const uint32_t*)A.base;
    g.ars = A.rs;
    g.acs = A.cs;
    g.B = (const uint32_t*)B.base;
    g.brs = B.rs;
    g.bcs = B.cs;
    g.Cin = (const uint32_t*)C.base;
    g.cirs = C.rs;
    g.cics = C.cs;
    g.Cout = (uint32_t*)C.base;
    g.cors = C.rs;
    g.cocs = C.cs;
    g.op = op;
    return g;
}

}  // namespace

extern "C" {

HX_API size_t hx_gemm_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

HX_API int hx_gemm_u32_noalloc(const hx_modctx* ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s,
                               const hx_gemm_opts* opts) {
    return guard("hx_gemm_u32_noalloc", [&] {
        HX_CHECK(ctx != nullptr && ctx->p >= 2, "bad modulus context");
        HX_CHECK(A.cols == B.rows && C.rows == A.rows && C.cols == B.cols, "gemm: shape mismatch");
        HX_CHECK(op == HX_OP_SET || op == HX_OP_ADD || op == HX_OP_SUB, "gemm: bad op");
        HX_CHECK(!views_overlap(C, A) && !views_overlap(C, B), "gemm: C aliases an operand");
        HX_CHECK(C.rows == 0 || C.cols == 0 || C.base != nullptr, "gemm: C is NULL");
        int variant = opts ? opts->variant : HX_GEMM_AUTO;
        HX_CHECK(variant == HX_GEMM_AUTO || variant == HX_GEMM_PLAIN || variant == HX_GEMM_LIMB, "gemm: bad variant");
        HX_CHECK(variant != HX_GEMM_LIMB || ctx->limbs > 0, "gemm: the limb route needs p <= 4278124287");
        const GemmArgs g = view_args(A, B, C, op);
        const bool automatic = variant == HX_GEMM_AUTO;
        if (automatic) variant = gemm_auto_variant(*ctx, g.m, g.n, g.k);
        const hipEvent_t ev0 = opts ? (hipEvent_t)opts->main_begin : nullptr;
        const hipEvent_t ev1 = opts ? (hipEvent_t)opts->main_end : nullptr;
        if (variant == HX_GEMM_LIMB && g.m != 0 && g.n != 0) {
            const bool fits = limb_k_chunk_for(*ctx, g.m, g.n, g.k, kNoAllocScratchBytes) > 0;
            if (!fits && automatic) {
                variant = HX_GEMM_PLAIN;
            } else {
                HX_CHECK(fits, "hx_gemm_u32_noalloc: the limb planes of one 64-deep K chunk exceed the " +
                                   std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch (needs limbs * (round_up(m, 64) + "
                                   "round_up(n, 128)) * 64 bytes at most)");
                gemm_limb(*ctx, g, (hipStream_t)s, noalloc_scratch(), kNoAllocScratchBytes, opts ? opts->tile : 0,
                          ev0, ev1);
            }
        }
        if (variant != HX_GEMM_LIMB) gemm_plain(*ctx, g, (hipStream_t)s, 0, ev0, ev1);
        if (g.m != 0 && g.n != 0) {
            stat_add(ST_GEMMS, 1);
            if (variant == HX_GEMM_LIMB) stat_add(ST_LIMB, 1);
        }
    });
}

}  // extern "C"
