// hx_limbs.h -- the balanced signed 8-bit limb split of residues mod p and the exact recombination
// of limb-product anti-diagonals, shared by the limb GEMM (hx_gemm.hip) and the MFMA trailing update
// inside the batched elimination (hx_rref_batch.hip).
//
//  * A canonical residue x < p (p <= 4278124287) is replaced by its balanced representative
//    v = x or x - p, |v| <= (p-1)/2, written in L signed base-256 digits d_s in [-128, 127]:
//    v = sum_s d_s 256^s (L = 4 covers p <= 4278124287, 3 covers p <= 16711423, 2 p <= 65279, 1 p <= 255).
//  * A product of two such numbers is sum_d 256^d D_d with D_d = sum_{s+t=d} a_s b_t; summed over K
//    terms, |D_d| <= pairs(d) K 2^14 (< 2^31 for K <= limb_k_chunk(L)).  Recombination:
//    x = sum_{d<=3} D_d 2^(8d) + sum_{d>=4} D_d w_d with w_d = 2^(8d) mod p balanced into (-p/2, p/2]:
//    |x| < 6 * 2^60 + 2^56 < 2^63, then one Barrett reduction of x + p ceil(2^63 / p) in [0, 2^64).
#pragma once

#include "hx_common.h"

namespace hx {

// Balanced signed base-256 digits of a canonical residue x (x < p): L digits in [-128, 127].
__device__ __forceinline__ void limbs_of(uint32_t x, uint32_t p, uint32_t half, int L, int32_t* d) {
    int32_t v = x > half ? (int32_t)((int64_t)x - (int64_t)p) : (int32_t)x;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        if (s >= L - 1) break;
        const int32_t t = ((v + 128) & 255) - 128;
        d[s] = t;
        v = (v - t) >> 8;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (s == L - 1) d[s] = v;
}

// The recombination constants for modulus p.
struct LimbConsts {
    uint64_t off;     // p * ceil(2^63 / p)
    int32_t w[3];     // 2^(8d) mod p, balanced, d = 4, 5, 6
};

inline LimbConsts limb_consts(uint32_t p) {
    LimbConsts c;
    c.off = ((1ull << 63) / p + 1) * (uint64_t)p;
    uint64_t w = 1;
    for (int d = 1; d <= 6; ++d) {
        w = (w << 8) % p;
        if (d >= 4) c.w[d - 4] = (int32_t)(w <= p / 2 ? (int64_t)w : (int64_t)w - (int64_t)p);
    }
    return c;
}

// The weight of anti-diagonal d in the recombination: 2^(8d) for d <= 3, w_d (balanced) above.
__device__ __forceinline__ int64_t limb_weight(int d, const LimbConsts& c) {
    return d <= 3 ? (int64_t)(1u << (8 * d)) : (int64_t)c.w[d - 4];
}

// sum_d 256^d D_d mod p from the ND = 2L - 1 anti-diagonal sums of one output.
template <int ND>
__device__ __forceinline__ uint32_t limb_recombine(const int32_t (&D)[ND], const LimbConsts& c, const DevMod& M) {
    int64_t x = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        if (d <= 3)
            x += (int64_t)D[d] * (int64_t)(1u << (8 * d));
        else
            x += (int64_t)D[d] * (int64_t)c.w[d - 4];
    }
    return red64((uint64_t)x + c.off, M);
}

}  // namespace hx
