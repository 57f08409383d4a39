// hx_elim.h -- device arithmetic for the eliminations (Montgomery / Barrett policies).
#pragma once

#include "hx_common.h"

namespace hx {

// Montgomery arithmetic with R = 2^32 for an odd modulus p < 2^31.
//   redc(T) = T R^{-1} mod p for T < p R  (T + m p < 2 p R < 2^64, so nothing wraps)
// The division-free elimination step (v x - f y) R^{-1} scales a whole row by the unit R^{-1} v^{-1}
// relative to the true Schur-complement row: zero tests, ranks and pivot choices are unaffected,
// and every pivot row is normalised by its own leading entry before it is used as a value.
struct ArMont {
    uint32_t p, pinv, r2;   // pinv = -p^{-1} mod 2^32, r2 = 2^64 mod p
    __device__ __forceinline__ uint32_t redc(uint64_t T) const {
        const uint32_t m = (uint32_t)T * pinv;
        uint32_t t = (uint32_t)((T + (uint64_t)m * p) >> 32);
        return t >= p ? t - p : t;
    }
    // (v x + fneg y) R^{-1}, for v, x, fneg, y < p
    __device__ __forceinline__ uint32_t dfree(uint32_t v, uint32_t x, uint32_t fneg, uint32_t y) const {
        return redc((uint64_t)v * x + (uint64_t)fneg * y);
    }
    __device__ __forceinline__ uint32_t prep(uint32_t f) const { return redc((uint64_t)f * r2); }   // f R
    __device__ __forceinline__ uint32_t mul(uint32_t y, uint32_t fprep) const { return redc((uint64_t)y * fprep); }
    // a^{-1} mod p, 0 when a has none: a^(p-2) by Montgomery square-and-multiply (the same steps in every
    // lane, no data-dependent loop), checked; a failed check (p not prime) falls back to Euclid.
    __device__ uint32_t inv(uint32_t a) const {
        if (a == 0) return 0;
        uint32_t e = p - 2, base = prep(a), acc = prep(1u);
        while (e) {
            if (e & 1u) acc = mul(acc, base);
            base = mul(base, base);
            e >>= 1;
        }
        const uint32_t r = mul(acc, 1u);
        return (uint64_t)r * a % p == 1u ? r : invm(a, p);
    }
};

// Barrett arithmetic (any p < 2^31, used for even p, i.e. p = 2).
struct ArBarrett {
    uint32_t p;
    uint64_t bm;
    __device__ __forceinline__ uint32_t red(uint64_t x) const {
        const uint64_t q = __umul64hi(x, bm);
        uint64_t r = x - q * (uint64_t)p;
        if (r >= p) r -= p;
        return (uint32_t)r;
    }
    __device__ __forceinline__ uint32_t dfree(uint32_t v, uint32_t x, uint32_t fneg, uint32_t y) const {
        return red((uint64_t)v * x + (uint64_t)fneg * y);
    }
    __device__ __forceinline__ uint32_t prep(uint32_t f) const { return f; }
    __device__ __forceinline__ uint32_t mul(uint32_t y, uint32_t fprep) const { return red((uint64_t)y * fprep); }
    __device__ uint32_t inv(uint32_t a) const { return invm(a, p); }
};

inline ArMont make_mont(uint32_t p) {
    uint32_t inv = p;
    for (int i = 0; i < 5; ++i) inv *= 2u - p * inv;
    ArMont a;
    a.p = p;
    a.pinv = 0u - inv;
    a.r2 = (uint32_t)(((unsigned __int128)1 << 64) % p);
    return a;
}
inline ArBarrett make_barrett(uint32_t p) {
    ArBarrett a;
    a.p = p;
    a.bm = (uint64_t)(((unsigned __int128)1 << 64) / p);
    return a;
}

}  // namespace hx
