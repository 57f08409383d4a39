// hx_gemm64.hip -- C op= A B mod p for 64-bit residues, every modulus 2 <= p < 2^64 (the device
// u64 GEMM of Cuda/cuda_kernels.cu:104-164; the CPU path of record is kernels::gemm<uint64_t>,
// ModArithmetic/Kernels.h:184).
//
// Exact sums, one reduction per output: each 64 x 64 -> 128-bit product is formed from four 32 x 32
// -> 64 partial products and added into a 160-bit accumulator (w0 + w1 2^64 + w2 2^128, exact for
// K < 2^32), and the sum is reduced once: T = w0 + w1 2^64 + w2 2^128 mod p.  For odd p by Montgomery
// (R = 2^64): REDC(w1 mod p, w0) = (w1 2^64 + w0) R^-1, multiplied by R^2 mod p and reduced again;
// the w2 term is w2 (2^128 mod p).  For even p (no Montgomery inverse) by binary long division of
// the 128-bit words.  No floating point; the bytes equal the oracle's exact integer arithmetic.
//
// Tiling: 64 x 64 outputs per 256-thread workgroup, 4 x 4 per thread (rows ty + 16 i, columns
// tx + 16 j), K staged by 16 through a double-buffered LDS tile (A transposed, padded rows), the
// next stage's global loads issued before the current stage's products.  VALU-bound: ~14 integer
// operations per mod-p multiply-add.
#include <cstring>

#include "hx_common.h"

namespace hx {
namespace {

constexpr int G64_TILE = 64, G64_BK = 16, G64_T = 256;

struct Mod64 {
    uint64_t p, ninv, r2, two64;    // ninv = -p^-1 mod 2^64 (odd p); r2 = 2^128 mod p; two64 = 2^64 mod p
    int odd;
};

struct Gemm64Args {
    uint32_t m, n, k;
    const uint64_t* A;
    int64_t ars, acs;
    const uint64_t* B;
    int64_t brs, bcs;
    uint64_t* C;
    int64_t crs, ccs;
    int op;
};

__device__ __forceinline__ uint64_t add64m(uint64_t a, uint64_t b, uint64_t p) {
    const uint64_t r = a + b;
    return (r < a || r >= p) ? r - p : r;
}
__device__ __forceinline__ uint64_t sub64m(uint64_t a, uint64_t b, uint64_t p) { return a >= b ? a - b : a + (p - b); }

// (hi 2^64 + lo) R^-1 mod p for hi < p, odd p (Montgomery, R = 2^64).
__device__ __forceinline__ uint64_t redc128(uint64_t hi, uint64_t lo, const Mod64& M) {
    const uint64_t m = lo * M.ninv;
    const uint64_t mlo = m * M.p, mhi = __umul64hi(m, M.p);
    const uint64_t s = lo + mlo;                       // == 0 mod 2^64 by construction
    const uint64_t carry = s < lo ? 1u : 0u;
    // hi + mhi + carry < 2p; it may pass 2^64 when p > 2^63: then subtract p in 128-bit terms
    const uint64_t t1 = hi + mhi;
    const uint64_t c1 = t1 < hi ? 1u : 0u;
    const uint64_t t = t1 + carry;
    const uint64_t c2 = t < t1 ? 1u : 0u;
    return (c1 | c2 || t >= M.p) ? t - M.p : t;
}
__device__ __forceinline__ uint64_t mont_mul64(uint64_t a, uint64_t b, const Mod64& M) {
    return redc128(__umul64hi(a, b), a * b, M);
}
// (hi 2^64 + lo) mod p for any p >= 2 by binary long division (the even-modulus path).
__device__ uint64_t mod128_div(uint64_t hi, uint64_t lo, uint64_t p) {
    uint64_t r = hi % p;
    for (int b = 63; b >= 0; --b) {
        const uint64_t top = r >> 63;
        r = (r << 1) | ((lo >> b) & 1u);
        if (top || r >= p) r -= p;
    }
    return r;
}
__device__ __forceinline__ uint64_t reduce160(uint64_t w0, uint64_t w1, uint32_t w2, const Mod64& M) {
    uint64_t r;
    if (M.odd) {
        const uint64_t h = w1 >= M.p ? w1 % M.p : w1;
        r = mont_mul64(redc128(h, w0, M), M.r2, M);          // (h 2^64 + w0) mod p
        if (w2) r = add64m(r, mont_mul64(mont_mul64((uint64_t)w2, M.r2, M), M.r2, M), M.p);   // w2 2^128
    } else {
        r = mod128_div(w1, w0, M.p);
        if (w2) {
            // w2 2^128 mod p = ((w2 2^64 mod p) 2^64) mod p
            const uint64_t t = mod128_div((uint64_t)w2, 0, M.p);
            r = add64m(r, mod128_div(t, 0, M.p), M.p);
        }
    }
    return r;
}

// acc += a b exactly (a, b < 2^64; 160-bit accumulator w0, w1, w2)
__device__ __forceinline__ void mac128(uint64_t& w0, uint64_t& w1, uint32_t& w2, uint64_t a, uint64_t b) {
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    w0 += lo;
    const uint64_t c0 = w0 < lo ? 1u : 0u;
    const uint64_t h = hi + c0;                           // hi <= 2^64 - 2, so no wrap
    w1 += h;
    w2 += w1 < h ? 1u : 0u;
}

__global__ __launch_bounds__(G64_T) void gemm64_kernel(Gemm64Args g, Mod64 M) {
    __shared__ uint64_t As[2][G64_BK][G64_TILE + 1];
    __shared__ uint64_t Bs[2][G64_BK][G64_TILE];
    const uint32_t ntn = (g.n + G64_TILE - 1) / G64_TILE;
    const uint32_t i0 = (blockIdx.x / ntn) * G64_TILE, j0 = (blockIdx.x % ntn) * G64_TILE;
    const uint32_t tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    uint64_t w0[4][4], w1[4][4];
    uint32_t w2[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            w0[a][b] = 0;
            w1[a][b] = 0;
            w2[a][b] = 0;
        }
    // loader roles: A tile 64 rows x 16 k (thread: k = tid % 16, rows tid / 16 + 16 e), B tile 16 k x 64
    // columns (thread: column tid % 64, k = tid / 64 + 4 e) -- unit-stride directions first
    uint64_t ra[4], rb[4];
    auto load = [&](uint32_t k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t row = i0 + tid / 16 + 16 * e, kk = k0 + tid % 16;
            ra[e] = (row < g.m && kk < g.k) ? g.A[(int64_t)row * g.ars + (int64_t)kk * g.acs] : 0u;
            const uint32_t col = j0 + tid % 64, kb = k0 + tid / 64 + 4 * e;
            rb[e] = (col < g.n && kb < g.k) ? g.B[(int64_t)kb * g.brs + (int64_t)col * g.bcs] : 0u;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            As[buf][tid % 16][tid / 16 + 16 * e] = ra[e];
            Bs[buf][tid / 64 + 4 * e][tid % 64] = rb[e];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (uint32_t k0 = 0; k0 < g.k; k0 += G64_BK) {
        const bool more = k0 + G64_BK < g.k;
        if (more) load(k0 + G64_BK);                       // in flight during the products below
#pragma unroll 4
        for (int kk = 0; kk < G64_BK; ++kk) {
            uint64_t av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[buf][kk][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[buf][kk][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) mac128(w0[a][b], w1[a][b], w2[a][b], av[a], bv[b]);
        }
        if (more) {
            store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const uint32_t i = i0 + ty + 16 * a;
        if (i >= g.m) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t j = j0 + tx + 16 * b;
            if (j >= g.n) continue;
            const uint64_t r = reduce160(w0[a][b], w1[a][b], w2[a][b], M);
            uint64_t* c = g.C + (int64_t)i * g.crs + (int64_t)j * g.ccs;
            *c = g.op == HX_OP_SET ? r : g.op == HX_OP_ADD ? add64m(*c, r, M.p) : sub64m(*c, r, M.p);
        }
    }
}

Mod64 mod64_of(const hx_modctx64& c) { return Mod64{c.p, c.ninv, c.r2, c.two64, (int)(c.p & 1u)}; }

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API int hx_modctx64_init(uint64_t p, hx_modctx64* out) {
    return guard("hx_modctx64_init", [&] {
        HX_CHECK(out != nullptr, "out is NULL");
        HX_CHECK(p >= 2, "modulus must be >= 2");
        std::memset(out, 0, sizeof(*out));
        out->p = p;
        out->two64 = (uint64_t)(((unsigned __int128)1 << 64) % p);
        const unsigned __int128 t = (unsigned __int128)out->two64 * out->two64;
        out->r2 = (uint64_t)(t % p);
        if (p & 1u) {
            uint64_t inv = p;                       // p^-1 mod 2^64 by Newton (p odd: right to 3 bits)
            for (int i = 0; i < 6; ++i) inv *= 2u - p * inv;
            out->ninv = 0u - inv;
        }
    });
}

HX_API int hx_gemm_u64(const hx_modctx64* ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s) {
    return guard("hx_gemm_u64", [&] {
        HX_CHECK(ctx != nullptr && ctx->p >= 2, "bad modulus context");
        HX_CHECK(A.cols == B.rows && C.rows == A.rows && C.cols == B.cols, "gemm: shape mismatch");
        HX_CHECK(op == HX_OP_SET || op == HX_OP_ADD || op == HX_OP_SUB, "gemm: bad op");
        HX_CHECK(!views_overlap(C, A, 8, 8) && !views_overlap(C, B, 8, 8), "gemm: C aliases an operand");
        HX_CHECK(C.rows == 0 || C.cols == 0 || C.base != nullptr, "gemm: C is NULL");
        HX_CHECK((uint64_t)A.cols < (1ull << 32), "gemm: k must be below 2^32");
        if (C.rows == 0 || C.cols == 0) return;
        Gemm64Args g{C.rows, C.cols, A.cols, (const uint64_t*)A.base, A.rs, A.cs, (const uint64_t*)B.base, B.rs, B.cs,
                     (uint64_t*)C.base, C.rs, C.cs, op};
        const uint32_t tiles = ((g.m + G64_TILE - 1) / G64_TILE) * ((g.n + G64_TILE - 1) / G64_TILE);
        gemm64_kernel<<<tiles, G64_T, 0, (hipStream_t)s>>>(g, mod64_of(*ctx));
        after_launch("gemm64_kernel");
        stat_add(ST_GEMMS, 1);
    });
}

}  // extern "C"
