// hx_gemm.hip -- exact GEMM mod p on gfx950: C op= A * B mod p for uint32 residues.
//
// Two device kernels compute the same bytes (the product is the canonical residue of an
// exact integer sum, so every correct kernel agrees bit for bit; ENGINE_API.md:93-102):
//
//  * plain_kernel: u32 x u32 products accumulated EXACTLY on the VALU in a 96-bit
//    accumulator (u64 + a u32 carry count), one Barrett reduction per output cell at the
//    end.  The device twin of the reference's u32 GEMM (Cuda/cuda_kernels.cu:195-300,
//    which instead flushes a u64 sum by Barrett every k_max products, k_max = 4 for 31-bit p).
//
//  * limb route (limb_prep_kernel + limb_kernel<L>): every residue x is replaced by its
//    balanced representative (|x| <= (p-1)/2, the reference's `bal` trick,
//    Cuda/ImmaKernels.cuh:15-18) written in L signed base-256 digits in [-128, 127]:
//        x == sum_s d_s 256^s (mod p).
//    Then A*B = sum_{s,t} 256^{s+t} A_s B_t, and each limb product A_s B_t is an int8 GEMM
//    on the MFMA cores (v_mfma_i32_32x32x32_i8, exact int32 accumulation).  Products on one
//    anti-diagonal d = s + t share one int32 accumulator D_d: |D_d| <= K * L * 128 * 128
//    < 2^31 for K <= limb_k_chunk(L), so nothing wraps.  The epilogue recombines
//    sum_d 256^d D_d exactly in int64 (Horner, reduced mod p every three steps) and applies
//    the op.  L = 4 covers every p <= 4278124287 (all 31-bit moduli); L = 3 covers
//    p <= 16711423, L = 2 p <= 65279, L = 1 p <= 255.
//    The operands are split into limb planes once per call (limb_prep_kernel, HBM-bound:
//    4 bytes in, L bytes out per element), so the MFMA loop streams int8 planes only.
//
// No floating point anywhere; no atomics on results.
#include "hx_gemm.h"
#include "hx_limbs.h"

#include <algorithm>
#include <numeric>
#include <mutex>

namespace hx {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t ld_a(const GemmArgs& g, uint32_t i, uint32_t t) {
    const int64_t r = g.a_rows ? (int64_t)g.a_rows[i] : (int64_t)i;
    const int64_t c = g.a_cols ? (int64_t)g.a_cols[t] : (int64_t)t;
    return g.A[r * g.ars + c * g.acs];
}
__device__ __forceinline__ uint32_t ld_cin(const GemmArgs& g, uint32_t i, uint32_t j) {
    const int64_t r = g.c_rows ? (int64_t)g.c_rows[i] : (int64_t)i;
    return g.Cin[r * g.cirs + (int64_t)j * g.cics];
}

// XCD-aware tile order: consecutive workgroups are dealt round-robin over the 8 XCDs, so
// workgroup b is remapped to let every XCD own one contiguous run of tiles; inside a run,
// tiles go down groups of 8 tile-rows first, so the A rows and B columns a run touches stay
// in that XCD's L2.
__device__ __forceinline__ void tile_of(uint32_t b, uint32_t ntm, uint32_t ntn, uint32_t& tm, uint32_t& tn) {
    const uint32_t T = ntm * ntn, per = T / 8;
    uint32_t t = b;
    if (b < per * 8) t = (b % 8) * per + b / 8;
    const uint32_t G = 8, width = G * ntn;
    const uint32_t group = t / width, first = group * G;
    const uint32_t gm = min(ntm - first, G);
    const uint32_t in = t % width;
    tm = first + in % gm;
    tn = in / gm;
}

// ------------------------------------------------------------------------------------------
// plain kernel: 64 x 64 tile, 256 threads (4 x 4 outputs each), BK = 32, register prefetch.
// ------------------------------------------------------------------------------------------
constexpr int PM = 64, PN = 64, PK = 32;

// Split-K (GemmOpts::splits, CudaKernels.h:64-66): workgroup (tile, z) sums the K range
// [z kc, min(k, (z + 1) kc)) exactly and writes its canonical partial to part[(tile splits + z)
// PM PN + cell]; the tile's last split to arrive (an agent-scope counter per tile, zero at launch)
// adds the partials in the order z = 0, 1, ... and applies op to C -- no second launch.  The
// bytes equal the unsplit product's (every partial is exact, the sum is reduced mod p once more).
struct SplitK {
    uint32_t* part;
    uint32_t* arrive;
    uint32_t splits, kc;
};

template <bool SPLIT>
__device__ __forceinline__ void plain_tile(const GemmArgs& g, const DevMod& M, uint32_t tm, uint32_t tn,
                                           const SplitK& sk = SplitK{}, uint32_t tile = 0, uint32_t z = 0) {
    __shared__ uint32_t As[2][PK][PM + 1];
    __shared__ uint32_t Bs[2][PK][PN + 1];
    const uint32_t m = g.m_dev ? min(*g.m_dev, g.m) : g.m;
    const uint32_t kall = g.k_dev ? min(*g.k_dev, g.k) : g.k;
    const uint32_t kb = SPLIT ? z * sk.kc : 0u;
    const uint32_t k = SPLIT ? min(kall, kb + sk.kc) : kall;      // this workgroup's K range: [kb, k)
    const uint32_t n = g.n;
    const uint32_t m0 = tm * PM, n0 = tn * PN;
    if (m0 >= m) return;
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;

    uint64_t lo[4][4];
    uint32_t hi[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) { lo[i][j] = 0; hi[i][j] = 0; }

    // Each thread loads the same 8 A rows (tid / PK + 8 t) and the same B column (tid % PN) at every
    // k-step: their offsets, row gathers included, are computed once; per k-step only the column
    // gather (one index) remains.
    int64_t a_row[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const uint32_t row = m0 + (tid + 256 * t) / PK;
        a_row[t] = row < m ? (g.a_rows ? (int64_t)g.a_rows[row] : (int64_t)row) * g.ars : -1;
    }
    const uint32_t a_c = tid % PK, b_c = tid % PN, b_r0 = tid / PN;
    const bool b_ok = n0 + b_c < n;
    const uint32_t* Bcol = g.B + (int64_t)(n0 + b_c) * g.bcs;
    uint32_t ra[8], rb[8];
    auto fetch = [&](uint32_t k0) {
        const uint32_t kc = k0 + a_c;
        const bool k_ok = kc < k;
        const int64_t acol = k_ok ? (g.a_cols ? (int64_t)g.a_cols[kc] : (int64_t)kc) * g.acs : 0;
#pragma unroll
        for (int t = 0; t < 8; ++t) ra[t] = (k_ok && a_row[t] >= 0) ? g.A[a_row[t] + acol] : 0u;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const uint32_t kr = k0 + b_r0 + 4 * t;          // (tid + 256 t) / PN
            rb[t] = (b_ok && kr < k) ? Bcol[(int64_t)kr * g.brs] : 0u;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int e = tid + 256 * t;
            As[buf][e % PK][e / PK] = ra[t];
            Bs[buf][e / PN][e % PN] = rb[t];
        }
    };
    const uint32_t KT = k > kb ? (k - kb + PK - 1) / PK : 0u;
    if (KT > 0) {
        fetch(kb);
        stash(0);
    }
    __syncthreads();
    for (uint32_t kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) fetch(kb + (kt + 1) * PK);
#pragma unroll 4
        for (int kk = 0; kk < PK; ++kk) {
            uint32_t a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[buf][kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[buf][kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint64_t prod = (uint64_t)a[i] * b[j];
                    const uint64_t s = lo[i][j] + prod;
                    hi[i][j] += (s < prod) ? 1u : 0u;
                    lo[i][j] = s;
                }
        }
        if (kt + 1 < KT) stash(buf ^ 1);
        __syncthreads();
    }
    if (SPLIT) {
        // this split's partial, then the arrival; the last split of the tile combines
        uint32_t* const mine = sk.part + ((size_t)tile * sk.splits + z) * (PM * PN);
        uint32_t r[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                r[i][j] = red96(hi[i][j], lo[i][j], M);
                mine[(ty + 16 * i) * PN + tx + 16 * j] = r[i][j];
            }
        __shared__ uint32_t s_last;
        __threadfence();
        __syncthreads();
        if (tid == 0)
            s_last = __hip_atomic_fetch_add(&sk.arrive[tile], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
                     sk.splits;
        __syncthreads();
        if (!s_last) return;
        __threadfence();                                   // every split's partial is visible here
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t row = m0 + ty + 16 * i;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t col = n0 + tx + 16 * j;
                if (row >= m || col >= n) continue;
                const uint32_t cell = (ty + 16 * i) * PN + tx + 16 * j;
                uint64_t sum = 0;                          // < 64 p < 2^38
                for (uint32_t zz = 0; zz < sk.splits; ++zz)
                    sum += zz == z ? r[i][j] : sk.part[((size_t)tile * sk.splits + zz) * (PM * PN) + cell];
                const uint32_t v = red64(sum, M);
                const uint32_t c = g.op == HX_OP_SET ? 0u : ld_cin(g, row, col);
                g.Cout[(int64_t)row * g.cors + (int64_t)col * g.cocs] = apply_op(c, v, g.op, M.p);
            }
        }
        if (tid == 0) sk.arrive[tile] = 0u;               // (the host clears the counters before each launch too)
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t row = m0 + ty + 16 * i;
        if (row >= m) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t col = n0 + tx + 16 * j;
            if (col >= n) continue;
            const uint32_t r = red96(hi[i][j], lo[i][j], M);
            const uint32_t c = g.op == HX_OP_SET ? 0u : ld_cin(g, row, col);
            g.Cout[(int64_t)row * g.cors + (int64_t)col * g.cocs] = apply_op(c, r, g.op, M.p);
        }
    }
}

__global__ __launch_bounds__(256) void plain_kernel(GemmArgs g, DevMod M) {
    const uint32_t ntm = (g.m + PM - 1) / PM, ntn = (g.n + PN - 1) / PN;
    uint32_t tm, tn;
    tile_of(blockIdx.x, ntm, ntn, tm, tn);
    plain_tile<false>(g, M, tm, tn);
}

// Split-K launch: blockIdx = z * tiles + b (every tile's split z before any split z + 1).
__global__ __launch_bounds__(256) void plain_split_kernel(GemmArgs g, DevMod M, SplitK sk) {
    const uint32_t ntm = (g.m + PM - 1) / PM, ntn = (g.n + PN - 1) / PN, tiles = ntm * ntn;
    const uint32_t z = blockIdx.x / tiles;
    uint32_t tm, tn;
    tile_of(blockIdx.x % tiles, ntm, ntn, tm, tn);
    plain_tile<true>(g, M, tm, tn, sk, tm * ntn + tn, z);
}

__global__ void clear_words_kernel(uint32_t* w, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 0u;
}

// The workgroup's job in a batch: prefix[j] = first flat tile of job j (prefix[njobs] = total).
__device__ __forceinline__ uint32_t job_of(const uint32_t* prefix, uint32_t njobs, uint32_t b) {
    uint32_t lo = 0, hi = njobs;          // largest j with prefix[j] <= b
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (prefix[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void plain_batch_kernel(const GemmArgs* jobs, const uint32_t* prefix, uint32_t njobs,
                                                          DevMod M) {
    const uint32_t j = job_of(prefix, njobs, blockIdx.x);
    const GemmArgs g = jobs[j];
    const uint32_t local = blockIdx.x - prefix[j], ntn = (g.n + PN - 1) / PN;
    plain_tile<false>(g, M, local / ntn, local % ntn);
}

// ------------------------------------------------------------------------------------------
// limb route
// ------------------------------------------------------------------------------------------
struct PrepArgs {
    const uint32_t* src;
    int64_t s_r, s_c;             // source strides along output rows / columns
    const uint32_t* rlist;        // optional gathers (output row r reads source row rlist[r])
    const uint32_t* clist;
    uint32_t R, K;                // true output rows / columns (upper bounds if r_dev / k_dev)
    const uint32_t* r_dev;
    const uint32_t* k_dev;
    uint32_t Kp;                  // padded row length of a plane (multiple of 64)
    int8_t* out;
    size_t plane;                 // bytes per plane
    uint32_t p, half;
    int L;
};

// The operand tile of prep_tile (64 x 64 of the output-row x k layout) into T; zero past R x K.  Without
// gathers or device-side sizes each thread walks one base pointer by a uniform stride (no per-entry
// 64-bit index products); either way a wave reads 64 consecutive entries of the source's contiguous
// direction.
__device__ __forceinline__ void prep_load(const PrepArgs& a, uint32_t r0, uint32_t c0, uint32_t (*T)[65]) {
    const int tid = threadIdx.x;
    const bool ccontig = a.s_c == 1 && a.clist == nullptr;
    if (a.rlist == nullptr && a.clist == nullptr && a.r_dev == nullptr && a.k_dev == nullptr) {
        const uint32_t i = tid & 63, o = tid >> 6;           // i along the contiguous direction
        const uint32_t r = ccontig ? o : i, c = ccontig ? i : o;
        const int64_t step = 4 * (ccontig ? a.s_r : a.s_c);  // the other coordinate advances by 4 per entry
        const uint32_t* src = a.src + (int64_t)(r0 + r) * a.s_r + (int64_t)(c0 + c) * a.s_c;
        const bool in_i = ccontig ? c0 + c < a.K : r0 + r < a.R;
        const uint32_t lim = ccontig ? (a.R > r0 + r ? a.R - r0 - r : 0) : (a.K > c0 + c ? a.K - c0 - c : 0);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const uint32_t v = in_i && 4u * t < lim ? src[t * step] : 0u;
            if (ccontig) T[r + 4 * t][c] = v;
            else T[r][c + 4 * t] = v;
        }
        return;
    }
    const uint32_t R = a.r_dev ? min(*a.r_dev, a.R) : a.R;
    const uint32_t K = a.k_dev ? min(*a.k_dev, a.K) : a.K;
#pragma unroll 4
    for (int t = 0; t < 16; ++t) {
        const int e = tid + 256 * t;
        const uint32_t r = ccontig ? e / 64 : e % 64, c = ccontig ? e % 64 : e / 64;
        uint32_t v = 0;
        if (r0 + r < R && c0 + c < K) {
            const int64_t sr = a.rlist ? (int64_t)a.rlist[r0 + r] : (int64_t)(r0 + r);
            const int64_t sc = a.clist ? (int64_t)a.clist[c0 + c] : (int64_t)(c0 + c);
            v = a.src[sr * a.s_r + sc * a.s_c];
        }
        T[r][c] = v;
    }
}

// One 64 x 64 tile of the output (rows x k) per workgroup: coalesced loads along the source's
// contiguous direction into LDS, then every thread writes 16 k-bytes of one row per plane.
__device__ __forceinline__ void prep_tile(const PrepArgs& a, uint32_t bx, uint32_t by) {
    __shared__ uint32_t T[64][65];
    const uint32_t r0 = by * 64, c0 = bx * 64;
    prep_load(a, r0, c0, T);
    __syncthreads();
    const int tid = threadIdx.x, r = tid / 4, q = tid % 4;
    uint32_t w[4][4];   // [plane][word]: 16 k-bytes of this row per plane
#pragma unroll
    for (int wd = 0; wd < 4; ++wd) {
#pragma unroll
        for (int s = 0; s < 4; ++s) w[s][wd] = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int32_t d[4] = {0, 0, 0, 0};
            limbs_of(T[r][16 * q + 4 * wd + b], a.p, a.half, a.L, d);
#pragma unroll
            for (int s = 0; s < 4; ++s) w[s][wd] |= ((uint32_t)d[s] & 0xffu) << (8 * b);
        }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s >= a.L) break;
        v4i vw;
#pragma unroll
        for (int wd = 0; wd < 4; ++wd) vw[wd] = (int)w[s][wd];
        *(v4i*)(a.out + (size_t)s * a.plane + (size_t)(r0 + r) * a.Kp + c0 + 16 * q) = vw;
    }
}

__global__ __launch_bounds__(256) void limb_prep_kernel(PrepArgs a) { prep_tile(a, blockIdx.x, blockIdx.y); }

// Batched prep: prefix over the (Kp / 64) x (Rp / 64) tile grids of every operand.
__global__ __launch_bounds__(256) void prep_batch_kernel(const PrepArgs* preps, const uint32_t* prefix, uint32_t n) {
    const uint32_t j = job_of(prefix, n, blockIdx.x);
    const PrepArgs a = preps[j];
    const uint32_t local = blockIdx.x - prefix[j], gx = a.Kp / 64;
    prep_tile(a, local % gx, local / gx);
}

constexpr int LM = 64, LN = 128, LK = 64;      // workgroup tile; k-depth of one LDS stage
constexpr int LROW = 80;                        // LDS bytes per (plane, row): 64 k-bytes + 16 pad

struct LimbArgs {
    const int8_t* A8;
    const int8_t* B8;
    size_t a_plane, b_plane;
    uint32_t Kp;                 // k-bytes per plane row (multiple of 64)
    uint32_t m, n;               // true output shape (m an upper bound if m_dev)
    const uint32_t* m_dev;
    uint32_t ntm, ntn;
    const uint32_t* Cin;
    int64_t cirs, cics;
    const uint32_t* c_rows;
    uint32_t* Cout;
    int64_t cors, cocs;
    int op;
    LimbConsts lc;               // the recombination constants (hx_limbs.h)
    DevMod M;
};

inline void limb_epilogue_consts(uint32_t p, LimbArgs& la) { la.lc = limb_consts(p); }

// Epilogue of both limb kernels: exact recombination sum_d 256^d D_d of one wave's 32 x 64
// block (two 32 x 32 MFMA blocks; lane = column, 16 accumulator rows per lane), op into C.
template <int L, int NCB = 2>
__device__ __forceinline__ void limb_store(const LimbArgs& a, const v16i (&acc)[NCB][2 * L - 1], uint32_t m,
                                           uint32_t r0, uint32_t c0, int lane) {
    constexpr int ND = 2 * L - 1;
    const int lr = lane & 31, lg = lane >> 5;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const uint32_t col = c0 + cb * 32 + lr;
        if (col >= a.n) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t row = r0 + (r / 4) * 8 + lg * 4 + (r % 4);
            if (row >= m) continue;
            int32_t D[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) D[d] = acc[cb][d][r];
            const uint32_t res = limb_recombine<ND>(D, a.lc, a.M);   // exact (bounds in hx_limbs.h)
            uint32_t c = 0;
            if (a.op != HX_OP_SET) {
                const int64_t rr = a.c_rows ? (int64_t)a.c_rows[row] : (int64_t)row;
                c = a.Cin[rr * a.cirs + (int64_t)col * a.cics];
            }
            a.Cout[(int64_t)row * a.cors + (int64_t)col * a.cocs] = apply_op(c, res, a.op, a.M.p);
        }
    }
}

template <int L>
__device__ __forceinline__ void limb_tile(const LimbArgs& a, uint32_t tm, uint32_t tn, uint8_t* smem) {
    constexpr int A_BYTES = L * LM * LROW, B_BYTES = L * LN * LROW, STAGE = A_BYTES + B_BYTES;
    constexpr int ND = 2 * L - 1;
    const uint32_t m = a.m_dev ? min(*a.m_dev, a.m) : a.m;
    const uint32_t m0 = tm * LM, n0 = tn * LN;
    if (m0 >= m) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    v16i acc[2][ND];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[cb][d][r] = 0;

    // global -> register staging: L A-chunks and 2L B-chunks of 16 bytes per thread
    v4i ga[L], gb[2 * L];
    const int8_t* Abase = a.A8 + (size_t)m0 * a.Kp;
    const int8_t* Bbase = a.B8 + (size_t)n0 * a.Kp;
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            const int c = tid + 256 * i;
            const int s = c / 256, row = (c % 256) / 4, q = c % 4;
            ga[i] = *(const v4i*)(Abase + (size_t)s * a.a_plane + (size_t)row * a.Kp + k0 + 16 * q);
        }
#pragma unroll
        for (int i = 0; i < 2 * L; ++i) {
            const int c = tid + 256 * i;
            const int s = c / 512, col = (c % 512) / 4, q = c % 4;
            gb[i] = *(const v4i*)(Bbase + (size_t)s * a.b_plane + (size_t)col * a.Kp + k0 + 16 * q);
        }
    };
    auto stash = [&](int buf) {
        uint8_t* sa = smem + buf * STAGE;
        uint8_t* sb = sa + A_BYTES;
#pragma unroll
        for (int i = 0; i < L; ++i) {
            const int c = tid + 256 * i;
            const int s = c / 256, row = (c % 256) / 4, q = c % 4;
            *(v4i*)(sa + (s * LM + row) * LROW + 16 * q) = ga[i];
        }
#pragma unroll
        for (int i = 0; i < 2 * L; ++i) {
            const int c = tid + 256 * i;
            const int s = c / 512, col = (c % 512) / 4, q = c % 4;
            *(v4i*)(sb + (s * LN + col) * LROW + 16 * q) = gb[i];
        }
    };

    const uint32_t KT = a.Kp / LK;
    if (KT > 0) {
        fetch(0);
        stash(0);
    }
    __syncthreads();
    const int lr = lane & 31, lg = lane >> 5;
    for (uint32_t kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) fetch((kt + 1) * LK);
        const uint8_t* sa = smem + buf * STAGE;
        const uint8_t* sb = sa + A_BYTES;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            v4i fa[L], fb[2][L];
#pragma unroll
            for (int s = 0; s < L; ++s)
                fa[s] = *(const v4i*)(sa + (s * LM + wm * 32 + lr) * LROW + ks * 32 + lg * 16);
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int s = 0; s < L; ++s)
                    fb[cb][s] = *(const v4i*)(sb + (s * LN + wn * 64 + cb * 32 + lr) * LROW + ks * 32 + lg * 16);
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int s = 0; s < L; ++s)
#pragma unroll
                    for (int t = 0; t < L; ++t)
                        acc[cb][s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s], fb[cb][t], acc[cb][s + t], 0, 0, 0);
        }
        if (kt + 1 < KT) stash(buf ^ 1);
        __syncthreads();
    }

    limb_store<L>(a, acc, m, m0 + wm * 32, n0 + wn * 64, lane);
}

// Direct-to-LDS variant of limb_tile: the global -> LDS copy is done by global_load_lds
// (16 B per lane, 1 KB per wave instruction) into a ring of three stages, two stages in flight
// while the MFMAs consume the third.  A stage holds, per plane, the tile's rows as 64-byte
// k-strips; inside a row the four 16-byte chunks are stored at chunk q ^ swz(row).  A fragment
// read takes one logical chunk from 32 consecutive rows; with the swizzle any 8 aligned rows
// (one 128-byte bank line = 2 rows) and any 16 aligned rows (256 bytes = 4 rows) hit pairwise
// different (row within the bank line, chunk) pairs, so a 16-byte read phase is conflict-free.
// The loop is unrolled over the three stage arrays so that every LDS access names its array
// statically; the waits are explicit (vmcnt counts this wave's outstanding DMA instructions:
// 3L per stage) and the workgroup barrier carries no memory fence of its own.
__device__ __forceinline__ int swz(int row) { return ((row >> 2) & 3) ^ (row & 2); }

// Immediate of s_waitcnt: wait until at most `vm` vector-memory and `lgkm` LDS operations of this
// wave are outstanding (expcnt not waited on).
constexpr int wait_imm(int vm, int lgkm) { return (vm & 15) | ((vm >> 4) << 14) | (7 << 4) | ((lgkm & 15) << 8); }

// One wave's MFMA operands for one 32-deep k-step: the A fragment of its 32 rows and the B
// fragments of its two 32-column blocks, every limb plane.
template <int L>
struct LFrag {
    v4i a[L];
    v4i b[2][L];
};

// The DMA of K-stage kt (64 k-bytes of every plane row of the tile) into stage buffer st: 12L
// blocks of 1 KB (16 rows x 64 bytes), 3L per wave, one global_load_lds (16 B per lane) each.
// (Written with a per-block A / B branch on purpose: in this form the compiler adds no waits of
// its own on the in-flight copies before the LDS reads; the kernels' explicit waits and
// barriers order every copy before the reads of its stage.)
template <int L>
__device__ __forceinline__ void stage_issue(const LimbArgs& a, const int8_t* Abase, const int8_t* Bbase, int wave,
                                            uint32_t kt, uint8_t* st) {
    constexpr int SA = L * LM * 64, PER_WAVE = 3 * L;
    const size_t k0 = (size_t)kt * LK;
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
        const int blk = wave * PER_WAVE + i;
        const int8_t* src;
        uint8_t* dst;
        if (blk < 4 * L) {
            const int s = blk >> 2, g = blk & 3;
            src = Abase + s * a.a_plane + (size_t)(16 * g) * a.Kp + k0;
            dst = st + s * (LM * 64) + g * 1024;
        } else {
            const int s = (blk - 4 * L) >> 3, g = (blk - 4 * L) & 7;
            src = Bbase + s * a.b_plane + (size_t)(16 * g) * a.Kp + k0;
            dst = st + SA + s * (LN * 64) + g * 1024;
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
}

template <int L>
__device__ __forceinline__ void stage_frag(const uint8_t* st, int ks, int wm, int wn, int lane, LFrag<L>& f) {
    constexpr int SA = L * LM * 64;
    const int lr = lane & 31, q = ks * 2 + (lane >> 5);
    const int ra = wm * 32 + lr;
#pragma unroll
    for (int s = 0; s < L; ++s) f.a[s] = *(const v4i*)(st + s * (LM * 64) + ra * 64 + 16 * (q ^ swz(ra)));
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int rb = wn * 64 + cb * 32 + lr;
#pragma unroll
        for (int s = 0; s < L; ++s) f.b[cb][s] = *(const v4i*)(st + SA + s * (LN * 64) + rb * 64 + 16 * (q ^ swz(rb)));
    }
}

template <int L>
__device__ __forceinline__ void frag_mfma(const LFrag<L>& f, v16i (&acc)[2][2 * L - 1]) {
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int s = 0; s < L; ++s)
#pragma unroll
            for (int t = 0; t < L; ++t)
                acc[cb][s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.a[s], f.b[cb][t], acc[cb][s + t], 0, 0, 0);
}

__device__ __forceinline__ void wg_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int L>
__device__ __forceinline__ void limb_tile2(const LimbArgs& a, uint32_t tm, uint32_t tn) {
    constexpr int SA = L * LM * 64, SB = L * LN * 64, STAGE = SA + SB;
    constexpr int ND = 2 * L - 1, PER_WAVE = 3 * L;      // 12L 1-KB blocks per stage over 4 waves
    __shared__ __attribute__((aligned(1024))) uint8_t s0[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s1[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s2[STAGE];
    const uint32_t m = a.m_dev ? min(*a.m_dev, a.m) : a.m;
    const uint32_t m0 = tm * LM, n0 = tn * LN;
    if (m0 >= m) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;

    v16i acc[2][ND];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[cb][d][r] = 0;

    // lane -> (row within a 16-row block, physical chunk); it fetches the logical chunk stored there
    const int br = lane >> 2, bq = (lane & 3) ^ swz(br);
    const int8_t* Abase = a.A8 + (size_t)(m0 + br) * a.Kp + 16 * bq;
    const int8_t* Bbase = a.B8 + (size_t)(n0 + br) * a.Kp + 16 * bq;
    auto compute = [&](const uint8_t* st) {
        LFrag<L> f;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            stage_frag<L>(st, ks, wm, wn, lane, f);
            frag_mfma<L>(f, acc);
        }
    };
    const uint32_t KT = a.Kp / LK;
    if (KT > 0) stage_issue<L>(a, Abase, Bbase, wave, 0, s0);
    if (KT > 1) stage_issue<L>(a, Abase, Bbase, wave, 1, s1);
    // stage kt: wait for this wave's DMAs of stage kt (stage kt+1 may stay in flight), barrier so
    // every wave's DMAs of stage kt are in LDS and every wave is done reading stage kt-1, refill
    // the slot of stage kt-1 with stage kt+2, then consume stage kt.
    auto step = [&](uint32_t kt, const uint8_t* cur, uint8_t* next2) {
        if (kt + 1 < KT)
            __builtin_amdgcn_s_waitcnt(wait_imm(PER_WAVE, 15));
        else
            __builtin_amdgcn_s_waitcnt(wait_imm(0, 15));
        wg_barrier();
        if (kt + 2 < KT) stage_issue<L>(a, Abase, Bbase, wave, kt + 2, next2);
        compute(cur);
    };
    for (uint32_t kt = 0; kt < KT; kt += 3) {
        step(kt, s0, s2);
        if (kt + 1 < KT) step(kt + 1, s1, s0);
        if (kt + 2 < KT) step(kt + 2, s2, s1);
    }

    limb_store<L>(a, acc, m, m0 + wm * 32, n0 + wn * 64, lane);
}

// limb_tile2 with the LDS reads software-pipelined against the MFMAs: the fragments of the next
// 32-deep k-step are read while the MFMAs of the current one run, across stage boundaries too.
// Per stage kt: read (kt, ks1) | MFMAs (kt, ks0) | wait until this wave's stage kt+1 DMAs and
// (kt, ks1) reads are done, barrier (every wave holds its (kt, ks1) fragments, so slot kt is free,
// and stage kt+1 is in LDS), DMA of stage kt+3 into slot kt | read (kt+1, ks0) | MFMAs (kt, ks1).
// Three stages are in flight.
template <int L>
__device__ __forceinline__ void limb_tile3(const LimbArgs& a, uint32_t tm, uint32_t tn) {
    constexpr int SA = L * LM * 64, SB = L * LN * 64, STAGE = SA + SB;
    constexpr int ND = 2 * L - 1, PER_WAVE = 3 * L;
    __shared__ __attribute__((aligned(1024))) uint8_t s0[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s1[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s2[STAGE];
    const uint32_t m = a.m_dev ? min(*a.m_dev, a.m) : a.m;
    const uint32_t m0 = tm * LM, n0 = tn * LN;
    if (m0 >= m) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;

    v16i acc[2][ND];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[cb][d][r] = 0;

    const int br = lane >> 2, bq = (lane & 3) ^ swz(br);
    const int8_t* Abase = a.A8 + (size_t)(m0 + br) * a.Kp + 16 * bq;
    const int8_t* Bbase = a.B8 + (size_t)(n0 + br) * a.Kp + 16 * bq;
    const uint32_t KT = a.Kp / LK;
    if (KT > 0) {
        stage_issue<L>(a, Abase, Bbase, wave, 0, s0);
        if (KT > 1) stage_issue<L>(a, Abase, Bbase, wave, 1, s1);
        if (KT > 2) stage_issue<L>(a, Abase, Bbase, wave, 2, s2);
        if (KT > 2)
            __builtin_amdgcn_s_waitcnt(wait_imm(2 * PER_WAVE, 15));
        else if (KT > 1)
            __builtin_amdgcn_s_waitcnt(wait_imm(PER_WAVE, 15));
        else
            __builtin_amdgcn_s_waitcnt(wait_imm(0, 15));
        wg_barrier();
        LFrag<L> fa, fb;
        stage_frag<L>(s0, 0, wm, wn, lane, fa);
        __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
        // a stage that has a successor (kt + 1 < KT)
        // (the scheduling barriers keep each group of LDS reads ahead of the MFMAs they overlap)
        auto body = [&](uint32_t kt, uint8_t* cur, const uint8_t* nxt) {
            stage_frag<L>(cur, 1, wm, wn, lane, fb);
            __builtin_amdgcn_sched_barrier(0);
            frag_mfma<L>(fa, acc);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));             // (kt, ks1) in registers
            if (kt + 2 < KT)
                __builtin_amdgcn_s_waitcnt(wait_imm(PER_WAVE, 15));  // stage kt+1 copied
            else
                __builtin_amdgcn_s_waitcnt(wait_imm(0, 15));
            wg_barrier();
            if (kt + 3 < KT) stage_issue<L>(a, Abase, Bbase, wave, kt + 3, cur);
            stage_frag<L>(nxt, 0, wm, wn, lane, fa);
            __builtin_amdgcn_sched_barrier(0);
            frag_mfma<L>(fb, acc);
            // the (kt+1, ks0) reads are long done here; saying so keeps the next group of reads
            // from being followed by a full wait before the MFMAs that do not use them
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
        };
        for (uint32_t kt = 0; kt + 1 < KT; kt += 3) {
            body(kt, s0, s1);
            if (kt + 2 < KT) body(kt + 1, s1, s2);
            if (kt + 3 < KT) body(kt + 2, s2, s0);
        }
        // the last stage (all copies are complete: nothing left in flight)
        const uint32_t ls = (KT - 1) % 3;
        const uint8_t* cur = ls == 0 ? s0 : (ls == 1 ? s1 : s2);
        stage_frag<L>(cur, 1, wm, wn, lane, fb);
        frag_mfma<L>(fa, acc);
        frag_mfma<L>(fb, acc);
    }
    limb_store<L>(a, acc, m, m0 + wm * 32, n0 + wn * 64, lane);
}

// limb_tile3's pipeline with eight waves per workgroup: the same 64 x 128 tile and LDS ring, each
// wave owning one 32 x 32 block (wm = wave & 1: rows, wn = wave >> 1: columns), so the
// accumulators (2L-1 x 16 registers) leave room for two waves per SIMD and one wave's barrier
// waits and epilogue overlap the other's MFMAs.  The 12L copy blocks of a stage are dealt
// round-robin over the eight waves (blk = wave + 8 i).
template <int L>
__device__ __forceinline__ void stage_issue8(const LimbArgs& a, const int8_t* Abase, const int8_t* Bbase, int wave,
                                             uint32_t kt, uint8_t* st) {
    constexpr int SA = L * LM * 64, NBLK = 12 * L, PER = (NBLK + 7) / 8;
    const size_t k0 = (size_t)kt * LK;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int blk = wave + 8 * i;
        if (blk >= NBLK) break;
        const int8_t* src;
        uint8_t* dst;
        if (blk < 4 * L) {
            const int s = blk >> 2, g = blk & 3;
            src = Abase + s * a.a_plane + (size_t)(16 * g) * a.Kp + k0;
            dst = st + s * (LM * 64) + g * 1024;
        } else {
            const int s = (blk - 4 * L) >> 3, g = (blk - 4 * L) & 7;
            src = Bbase + s * a.b_plane + (size_t)(16 * g) * a.Kp + k0;
            dst = st + SA + s * (LN * 64) + g * 1024;
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
}

template <int L>
struct LFrag1 {
    v4i a[L];
    v4i b[L];
};

template <int L>
__device__ __forceinline__ void stage_frag1(const uint8_t* st, int ks, int wm, int wn, int lane, LFrag1<L>& f) {
    constexpr int SA = L * LM * 64;
    const int lr = lane & 31, q = ks * 2 + (lane >> 5);
    const int ra = wm * 32 + lr, rb = wn * 32 + lr;
#pragma unroll
    for (int s = 0; s < L; ++s) f.a[s] = *(const v4i*)(st + s * (LM * 64) + ra * 64 + 16 * (q ^ swz(ra)));
#pragma unroll
    for (int s = 0; s < L; ++s) f.b[s] = *(const v4i*)(st + SA + s * (LN * 64) + rb * 64 + 16 * (q ^ swz(rb)));
}

template <int L>
__device__ __forceinline__ void frag_mfma1(const LFrag1<L>& f, v16i (&acc)[1][2 * L - 1]) {
#pragma unroll
    for (int s = 0; s < L; ++s)
#pragma unroll
        for (int t = 0; t < L; ++t)
            acc[0][s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.a[s], f.b[t], acc[0][s + t], 0, 0, 0);
}

template <int L>
__device__ __forceinline__ void limb_tile4(const LimbArgs& a, uint32_t tm, uint32_t tn) {
    constexpr int SA = L * LM * 64, SB = L * LN * 64, STAGE = SA + SB;
    constexpr int ND = 2 * L - 1, NBLK = 12 * L;
    // DMA instructions per wave per stage: waves below NBLK % 8 issue one more
    constexpr int HI = (NBLK + 7) / 8, LO = NBLK / 8, NHI = NBLK % 8 == 0 ? 8 : NBLK % 8;
    __shared__ __attribute__((aligned(1024))) uint8_t s0[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s1[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s2[STAGE];
    const uint32_t m = a.m_dev ? min(*a.m_dev, a.m) : a.m;
    const uint32_t m0 = tm * LM, n0 = tn * LN;
    if (m0 >= m) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const bool hi = wave < NHI;

    v16i acc[1][ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[0][d][r] = 0;

    const int br = lane >> 2, bq = (lane & 3) ^ swz(br);
    const int8_t* Abase = a.A8 + (size_t)(m0 + br) * a.Kp + 16 * bq;
    const int8_t* Bbase = a.B8 + (size_t)(n0 + br) * a.Kp + 16 * bq;
    // wait until at most `stages` stages of this wave's copies are in flight
    auto wait_stages = [&](int stages) {
        if (hi) {
            if (stages == 2) __builtin_amdgcn_s_waitcnt(wait_imm(2 * HI, 15));
            else if (stages == 1) __builtin_amdgcn_s_waitcnt(wait_imm(HI, 15));
            else __builtin_amdgcn_s_waitcnt(wait_imm(0, 15));
        } else {
            if (stages == 2) __builtin_amdgcn_s_waitcnt(wait_imm(2 * LO, 15));
            else if (stages == 1) __builtin_amdgcn_s_waitcnt(wait_imm(LO, 15));
            else __builtin_amdgcn_s_waitcnt(wait_imm(0, 15));
        }
    };
    const uint32_t KT = a.Kp / LK;
    if (KT > 0) {
        stage_issue8<L>(a, Abase, Bbase, wave, 0, s0);
        if (KT > 1) stage_issue8<L>(a, Abase, Bbase, wave, 1, s1);
        if (KT > 2) stage_issue8<L>(a, Abase, Bbase, wave, 2, s2);
        wait_stages(KT > 2 ? 2 : (KT > 1 ? 1 : 0));
        wg_barrier();
        LFrag1<L> fa, fb;
        stage_frag1<L>(s0, 0, wm, wn, lane, fa);
        __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
        auto body = [&](uint32_t kt, uint8_t* cur, const uint8_t* nxt) {
            stage_frag1<L>(cur, 1, wm, wn, lane, fb);
            __builtin_amdgcn_sched_barrier(0);
            frag_mfma1<L>(fa, acc);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
            wait_stages(kt + 2 < KT ? 1 : 0);
            wg_barrier();
            if (kt + 3 < KT) stage_issue8<L>(a, Abase, Bbase, wave, kt + 3, cur);
            stage_frag1<L>(nxt, 0, wm, wn, lane, fa);
            __builtin_amdgcn_sched_barrier(0);
            frag_mfma1<L>(fb, acc);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
        };
        for (uint32_t kt = 0; kt + 1 < KT; kt += 3) {
            body(kt, s0, s1);
            if (kt + 2 < KT) body(kt + 1, s1, s2);
            if (kt + 3 < KT) body(kt + 2, s2, s0);
        }
        const uint32_t ls = (KT - 1) % 3;
        const uint8_t* cur = ls == 0 ? s0 : (ls == 1 ? s1 : s2);
        stage_frag1<L>(cur, 1, wm, wn, lane, fb);
        frag_mfma1<L>(fa, acc);
        frag_mfma1<L>(fb, acc);
    }
    limb_store<L, 1>(a, acc, m, m0 + wm * 32, n0 + wn * 32, lane);
}

// A 96 x 128 tile on twelve waves (three per SIMD), each wave one 32 x 32 block (wm = wave % 3: rows,
// wn = wave / 3: columns).  Per MAC the tile streams (96 + 128) / (96 * 128) bytes of planes from L2
// against (64 + 128) / (64 * 128) for the 64 x 128 tile: 22 % less, and each XCD's 32 concurrent tiles
// cover 1.5 times the outputs.  The accumulators (2L-1 x 16 registers) plus one k-step's fragments fit
// the third of a SIMD's registers a wave gets, so the fragments are not double-buffered: the other two
// waves of the SIMD cover a wave's LDS reads.  Two stages of 64 k-bytes (112 KB at L = 4): the copy of
// stage kt + 1 runs while stage kt is consumed.  The 14L copy blocks of a stage are dealt round-robin over
// the twelve waves; A rows past the plane's last row (the last row tile of an m that is not a multiple
// of 96) are read from the plane's last row and never stored.
constexpr int LM5 = 96;

template <int L>
__device__ __forceinline__ void stage_issue12(const LimbArgs& a, const int8_t* Abase, const int8_t* Bbase, int wave,
                                              int br, uint32_t m0, uint32_t mp, uint32_t kt, uint8_t* st) {
    constexpr int SA = L * LM5 * 64, NBLK = 14 * L, PER = (NBLK + 11) / 12;
    const size_t k0 = (size_t)kt * LK;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int blk = wave + 12 * i;
        if (blk >= NBLK) break;
        const int8_t* src;
        uint8_t* dst;
        if (blk < 6 * L) {
            const int s = blk / 6, g = blk % 6;
            const uint32_t row = min(m0 + (uint32_t)(16 * g + br), mp - 1) - m0 - (uint32_t)br;
            src = Abase + s * a.a_plane + (size_t)row * a.Kp + k0;
            dst = st + s * (LM5 * 64) + g * 1024;
        } else {
            const int s = (blk - 6 * L) >> 3, g = (blk - 6 * L) & 7;
            src = Bbase + s * a.b_plane + (size_t)(16 * g) * a.Kp + k0;
            dst = st + SA + s * (LN * 64) + g * 1024;
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
}

template <int L>
__device__ __forceinline__ void stage_frag12(const uint8_t* st, int ks, int wm, int wn, int lane, LFrag1<L>& f) {
    constexpr int SA = L * LM5 * 64;
    const int lr = lane & 31, q = ks * 2 + (lane >> 5);
    const int ra = wm * 32 + lr, rb = wn * 32 + lr;
#pragma unroll
    for (int s = 0; s < L; ++s) f.a[s] = *(const v4i*)(st + s * (LM5 * 64) + ra * 64 + 16 * (q ^ swz(ra)));
#pragma unroll
    for (int s = 0; s < L; ++s) f.b[s] = *(const v4i*)(st + SA + s * (LN * 64) + rb * 64 + 16 * (q ^ swz(rb)));
}

template <int L>
__device__ __forceinline__ void limb_tile5(const LimbArgs& a, uint32_t tm, uint32_t tn) {
    constexpr int SA = L * LM5 * 64, SB = L * LN * 64, STAGE = SA + SB;
    constexpr int ND = 2 * L - 1;
    __shared__ __attribute__((aligned(1024))) uint8_t s0[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s1[STAGE];
    const uint32_t m = a.m_dev ? min(*a.m_dev, a.m) : a.m;
    const uint32_t m0 = tm * LM5, n0 = tn * LN;
    if (m0 >= m) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % 3, wn = wave / 3;

    v16i acc[1][ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[0][d][r] = 0;

    const int br = lane >> 2, bq = (lane & 3) ^ swz(br);
    const int8_t* Abase = a.A8 + (size_t)(m0 + br) * a.Kp + 16 * bq;
    const int8_t* Bbase = a.B8 + (size_t)(n0 + br) * a.Kp + 16 * bq;
    const uint32_t KT = a.Kp / LK;
    const uint32_t mp = (uint32_t)(a.a_plane / a.Kp);    // rows per A plane
    auto body = [&](uint32_t kt, const uint8_t* cur, uint8_t* nxt) {
        __builtin_amdgcn_s_waitcnt(wait_imm(0, 15));      // this wave's copies of stage kt have landed
        wg_barrier();                                     // everyone's; and stage kt - 1 is read out
        if (kt + 1 < KT) stage_issue12<L>(a, Abase, Bbase, wave, br, m0, mp, kt + 1, nxt);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            LFrag1<L> f;
            stage_frag12<L>(cur, ks, wm, wn, lane, f);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
            __builtin_amdgcn_sched_barrier(0);
            frag_mfma1<L>(f, acc);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    if (KT > 0) {
        stage_issue12<L>(a, Abase, Bbase, wave, br, m0, mp, 0, s0);
        for (uint32_t kt = 0; kt < KT; kt += 2) {
            body(kt, s0, s1);
            if (kt + 1 < KT) body(kt + 1, s1, s0);
        }
    }
    limb_store<L, 1>(a, acc, m, m0 + wm * 32, n0 + wn * 32, lane);
}

template <int L>
__global__ __launch_bounds__(768, 1) void limb5_kernel(LimbArgs a) {
    uint32_t tm, tn;
    tile_of(blockIdx.x, a.ntm, a.ntn, tm, tn);
    limb_tile5<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(512, 1) void limb4_kernel(LimbArgs a) {
    uint32_t tm, tn;
    tile_of(blockIdx.x, a.ntm, a.ntn, tm, tn);
    limb_tile4<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(512, 1) void limb4_batch_kernel(const LimbArgs* jobs, const uint32_t* prefix,
                                                             uint32_t njobs) {
    const uint32_t j = job_of(prefix, njobs, blockIdx.x);
    const LimbArgs a = jobs[j];
    const uint32_t local = blockIdx.x - prefix[j];
    if (local >= a.ntm * a.ntn) return;          // alignment padding of the job's block range
    uint32_t tm, tn;
    tile_of(local, a.ntm, a.ntn, tm, tn);
    limb_tile4<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(256, 1) void limb2_kernel(LimbArgs a) {
    uint32_t tm, tn;
    tile_of(blockIdx.x, a.ntm, a.ntn, tm, tn);
    limb_tile2<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(256, 1) void limb2_batch_kernel(const LimbArgs* jobs, const uint32_t* prefix,
                                                             uint32_t njobs) {
    const uint32_t j = job_of(prefix, njobs, blockIdx.x);
    const LimbArgs a = jobs[j];
    const uint32_t local = blockIdx.x - prefix[j];
    if (local >= a.ntm * a.ntn) return;          // alignment padding of the job's block range
    uint32_t tm, tn;
    tile_of(local, a.ntm, a.ntn, tm, tn);
    limb_tile2<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(256, 1) void limb3_kernel(LimbArgs a) {
    uint32_t tm, tn;
    tile_of(blockIdx.x, a.ntm, a.ntn, tm, tn);
    limb_tile3<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(256, 1) void limb3_batch_kernel(const LimbArgs* jobs, const uint32_t* prefix,
                                                             uint32_t njobs) {
    const uint32_t j = job_of(prefix, njobs, blockIdx.x);
    const LimbArgs a = jobs[j];
    const uint32_t local = blockIdx.x - prefix[j];
    if (local >= a.ntm * a.ntn) return;          // alignment padding of the job's block range
    uint32_t tm, tn;
    tile_of(local, a.ntm, a.ntn, tm, tn);
    limb_tile3<L>(a, tm, tn);
}

template <int L>
__global__ __launch_bounds__(256, 1) void limb_kernel(LimbArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    uint32_t tm, tn;
    tile_of(blockIdx.x, a.ntm, a.ntn, tm, tn);
    limb_tile<L>(a, tm, tn, smem);
}

template <int L>
__global__ __launch_bounds__(256, 1) void limb_batch_kernel(const LimbArgs* jobs, const uint32_t* prefix,
                                                            uint32_t njobs) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    const uint32_t j = job_of(prefix, njobs, blockIdx.x);
    const LimbArgs a = jobs[j];
    const uint32_t local = blockIdx.x - prefix[j];
    if (local >= a.ntm * a.ntn) return;          // alignment padding of the job's block range
    uint32_t tm, tn;
    tile_of(local, a.ntm, a.ntn, tm, tn);
    limb_tile<L>(a, tm, tn, smem);
}

template <int L>
size_t limb_lds() { return 2 * (size_t)(L * (LM + LN) * LROW); }

template <int L>
void limb_attrs() {
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void*)limb_kernel<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)limb_lds<L>());
        (void)hipFuncSetAttribute((const void*)limb_batch_kernel<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)limb_lds<L>());
    });
}

// ------------------------------------------------------------------------------------------
// residue route: the product mod p through N small moduli (multi-modular int8 GEMM + CRT)
// ------------------------------------------------------------------------------------------
// The operands' balanced residues a_b in [-p/2, p/2] are reduced modulo N pairwise coprime moduli
// q_i <= 255 into balanced bytes (|r| <= 127).  Each modulus is one plain int8 product on the MFMAs
// whose int32 sums (|sum| <= k * 127^2 < 2^31 for k <= 133144) are reduced to bytes in [0, q_i); a
// CRT pass then turns the N bytes into the exact sum's residue mod p.  The exact sum x satisfies
// |x| <= k (p/2)^2, and N is the least count with 20 prod q_i > 41 k (p/2)^2: then |x / prod q_i| < 20/41
// = 0.488, and the CRT's quotient estimate (error below 0.0015) rounds exactly.  For p < 2^32 and k <= 8192 that is
// N = 10 int8 products against the limb route's L^2 = 16 (the same decomposition as Ozaki, Uchino and
// Imamura, "Ozaki Scheme II: a GEMM-oriented emulation of floating-point matrix multiplication using
// an integer modular technique", 2025, here for exact residues).
constexpr int RES_MAXN = 12;
constexpr uint32_t RES_KMAX = 133144;            // k * 127^2 <= 2^31 - 1
__host__ __device__ constexpr uint32_t res_modulus(int i) {
    // pairwise coprime: 3*5*17, 2*127, 11*23, 251, 13*19, then primes
    return i == 0 ? 255u : i == 1 ? 254u : i == 2 ? 253u : i == 3 ? 251u : i == 4 ? 247u : i == 5 ? 241u
         : i == 6 ? 239u : i == 7 ? 233u : i == 8 ? 229u : i == 9 ? 227u : i == 10 ? 223u : 211u;
}

// Per-modulus constants of the operand split, computed on the host (the modulus loop is not unrolled,
// so constants derived in the kernel would cost runtime divisions): q, the bytes (1, 2^8, 2^16, 2^24)
// mod q packed into one word, q - (2^31 mod q), fl(1/q).
struct ResSplit {
    uint32_t nmod;
    uint32_t q[RES_MAXN], tpack[RES_MAXN], kq[RES_MAXN];
    float iq[RES_MAXN];
};

// The balanced residues modulo q_i of four entries given as w = a_b + 2^31 (a_b the balanced residue mod
// p, |a_b| <= p/2 < 2^31), packed as bytes.  One byte dot product t = sum_j byte_j(w) (2^8j mod q) +
// q - (2^31 mod q) is congruent to a_b and below 4 * 255^2 + 255 < 2^18, so fp32 holds it exactly and
// t * fl(1/q) + 1/2 is within 2^18 / 211 * 2^-23 < 2e-4 of t / q + 1/2, whose fractional part is a
// multiple of 1/q at least 1/(2q) > 0.0019 from the next integer when q is odd: the truncation is
// round(t / q) and r = t - q round(t / q) lies in [-(q-1)/2, (q-1)/2].  For q = 254 a tie (t = 127 mod
// 254) may round either way, giving +127 or -127: both are that residue and within a byte.
__device__ __forceinline__ uint32_t res_bytes4(const uint32_t (&w)[4], const ResSplit& c, int i) {
    const uint32_t Q = c.q[i], TP = c.tpack[i], KQ = c.kq[i];
    const float IQ = c.iq[i];
    uint32_t out = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t t = __builtin_amdgcn_udot4(w[b], TP, KQ, false);
        const uint32_t qt = (uint32_t)fmaf((float)t, IQ, 0.5f);
        const int32_t rr = (int32_t)t - (int32_t)__umul24(qt, Q);
        out |= ((uint32_t)rr & 0xffu) << (8 * b);
    }
    return out;
}

// One 64 x 64 (rows x k) tile of the N residue planes per workgroup (any strides, gathers and
// device-side sizes; the transposing split of B): every thread turns 16 k-entries of one row, staged
// through LDS, into N 16-byte strips of balanced residues (res_bytes4).
__global__ __launch_bounds__(256) void res_prep_kernel(PrepArgs a, ResSplit c) {
    __shared__ uint32_t T[64][65];
    const uint32_t r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    prep_load(a, r0, c0, T);
    __syncthreads();
    const int tid = threadIdx.x, r = tid / 4, qd = tid % 4;
    uint32_t w[16];
    const uint32_t p = a.p;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t v = T[r][16 * qd + j];
        w[j] = (v > a.half ? v - p : v) + 0x80000000u;      // a_b + 2^31, a_b in [-p/2, p/2]
    }
    int8_t* out = a.out + (size_t)(r0 + r) * a.Kp + c0 + 16 * qd;
#pragma unroll
    for (int i = 0; i < RES_MAXN; ++i) {
        if (i >= (int)c.nmod) break;
        v4i vw;
#pragma unroll
        for (int wd = 0; wd < 4; ++wd) {
            const uint32_t w4[4] = {w[4 * wd], w[4 * wd + 1], w[4 * wd + 2], w[4 * wd + 3]};
            vw[wd] = (int)res_bytes4(w4, c, i);
        }
        *(v4i*)(out + (size_t)i * a.plane) = vw;
    }
}

// Residue planes of a row-major operand whose rows are 16-byte aligned (no gathers, host-side sizes):
// one wave per row and 1024 k, lane-contiguous 16-byte loads (four entries per lane and instruction,
// 1 KB per wave instruction) and one dword per lane and plane for every four entries (256 contiguous
// bytes per wave store), no LDS.  Entries past K and rows past R are written as zero up to Kp and the
// plane's last row.  Grid: (ceil(Kp / 1024), rows / 4), 256 threads.
__global__ __launch_bounds__(256) void res_prep_rows_kernel(PrepArgs a, ResSplit c) {
    const uint32_t row = blockIdx.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t p = a.p;
    const uint32_t* src = a.src + (size_t)row * a.s_r;
    int8_t* out = a.out + (size_t)row * a.Kp;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t k4 = blockIdx.x * 1024 + (uint32_t)j * 256 + lane * 4;
        if (k4 >= a.Kp) break;
        uint32_t v[4] = {0, 0, 0, 0};
        if (row < a.R) {
            if (k4 + 3 < a.K) {
                const uint4 q4 = *(const uint4*)(src + k4);
                v[0] = q4.x;
                v[1] = q4.y;
                v[2] = q4.z;
                v[3] = q4.w;
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (k4 + b < a.K) v[b] = src[k4 + b];
            }
        }
        uint32_t w[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) w[b] = (v[b] > a.half ? v[b] - p : v[b]) + 0x80000000u;
#pragma unroll
        for (int i = 0; i < RES_MAXN; ++i) {
            if (i >= (int)c.nmod) break;
            *(uint32_t*)(out + (size_t)i * a.plane + k4) = res_bytes4(w, c, i);
        }
    }
}

// Residue planes of an operand whose output rows are contiguous in the source and whose k is strided (a
// row-major B: entry (row r, k c) at src[r s_r + c s_c] with s_r = 1; no gathers, host-side sizes).  One
// 32-row x 128-k tile per workgroup: the load reads 128-byte runs of rows (two k per wave instruction) into
// LDS as T[c][r ^ swizzle(c)]; each thread then takes 16 k of one row (8 threads per row), so a wave stores
// 8 rows x 128 contiguous bytes per plane.  The swizzle (r XOR 4 ((c / 16) mod 8)) makes both the stores
// into T (32 consecutive rows, one k) and the reads out of it (rows 0..3 x eight 16-k chunks, one k offset)
// hit 32 different banks.  Chunks past Kp are not written; entries past K or R read as zero.
// Grid: (ceil(Kp / 128), rows / 32), 256 threads.
__global__ __launch_bounds__(256) void res_prep_cols_kernel(PrepArgs a, ResSplit c) {
    __shared__ uint32_t T[128][33];
    const uint32_t r0 = blockIdx.y * 32, c0 = blockIdx.x * 128;
    const int tid = threadIdx.x;
    {
        const uint32_t r = tid & 31, cb = tid >> 5;              // row along the contiguous direction
        const uint32_t* src = a.src + (int64_t)(r0 + r) * a.s_r + (int64_t)(c0 + cb) * a.s_c;
        const bool in_r = r0 + r < a.R;
        const uint32_t lim = a.K > c0 + cb ? a.K - c0 - cb : 0;
        const int64_t step = 8 * a.s_c;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const uint32_t cc = cb + 8 * t;
            T[cc][r ^ (((cc >> 4) & 7) << 2)] = in_r && 8u * t < lim ? src[t * step] : 0u;
        }
    }
    __syncthreads();
    const uint32_t r = tid >> 3, q = tid & 7;
    if (c0 + 16 * q >= a.Kp) return;
    uint32_t w[16];
    const uint32_t p = a.p;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t v = T[16 * q + j][r ^ (q << 2)];
        w[j] = (v > a.half ? v - p : v) + 0x80000000u;
    }
    int8_t* out = a.out + (size_t)(r0 + r) * a.Kp + c0 + 16 * q;
#pragma unroll
    for (int i = 0; i < RES_MAXN; ++i) {
        if (i >= (int)c.nmod) break;
        v4i vw;
#pragma unroll
        for (int wd = 0; wd < 4; ++wd) {
            const uint32_t w4[4] = {w[4 * wd], w[4 * wd + 1], w[4 * wd + 2], w[4 * wd + 3]};
            vw[wd] = (int)res_bytes4(w4, c, i);
        }
        *(v4i*)(out + (size_t)i * a.plane) = vw;
    }
}

// One modulus' int8 product on a 256 x 256 tile: the limb_tile4 pipeline (eight waves, three-stage
// direct-to-LDS ring of 64-deep k-strips, fragments double-buffered against the MFMAs) with the
// operand "planes" being four 64-row blocks of A and two 128-row blocks of B of one residue plane,
// so each wave owns 4 x 2 separate 32 x 32 accumulator blocks (rows 64 s + 32 wm, columns
// 128 t + 32 wn) and reads 6 fragments per 8 MFMAs.  The epilogue reduces every int32 sum to its
// residue in [0, q) and stores it as a byte of the modulus' residue plane.
// The instance the route runs: eight waves (two per SIMD, 4 x 2 blocks each), a ring of three stages.
// (Measured at 8192^3, 10 moduli: a ring of 3, 4 or 5 stages 5.37 .. 5.43 ms; four waves with 4 x 4 blocks
// each, a third less LDS reading per MFMA, 5.55 .. 5.93 ms; two workgroups per CU of four waves with 4 x 2
// blocks each on 256 x 128 tiles, whose barriers are independent, 6.10 .. 6.18 ms.)
constexpr int RES_WAVES = 8, RES_STAGES = 3;
constexpr int RLA = 4, RLB = 2, RTM = RLA * LM, RTN = RLB * 16 * RES_WAVES;

struct ResArgs {
    const int8_t* A8;            // residue plane i of A at A8 + i * a_plane: Mp rows of Kp bytes
    const int8_t* B8;            // residue plane i of B at B8 + i * b_plane: Np rows of Kp bytes
    size_t a_plane, b_plane;
    uint32_t Kp;
    uint32_t ntm, ntn;
    uint8_t* R;                  // product residues of modulus i at R + i * r_plane: Mp rows of ldr bytes
    size_t r_plane;
    uint32_t ldr;
    uint32_t q[RES_MAXN];
    uint32_t y[RES_MAXN];        // (M / q_i)^-1 mod q_i, the CRT factor of modulus i
    double invq[RES_MAXN];
};

// B blocks hold 16 * WAVES rows (WAVES / 2 waves across, 32 columns each); A blocks 64 rows (two waves down).
template <int LA, int LB, int WAVES>
__device__ __forceinline__ void res_stage_issue(const ResArgs& a, const int8_t* Abase, const int8_t* Bbase, int wave,
                                                uint32_t kt, uint8_t* st) {
    constexpr int BN = 16 * WAVES, BG = BN / 16;
    constexpr int SA = LA * LM * 64, NBLK = 4 * LA + BG * LB, PER = (NBLK + WAVES - 1) / WAVES;
    const size_t k0 = (size_t)kt * LK;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int blk = wave + WAVES * i;
        if (blk >= NBLK) break;
        const int8_t* src;
        uint8_t* dst;
        if (blk < 4 * LA) {
            const int s = blk >> 2, g = blk & 3;
            src = Abase + (size_t)(s * LM + 16 * g) * a.Kp + k0;
            dst = st + s * (LM * 64) + g * 1024;
        } else {
            const int s = (blk - 4 * LA) / BG, g = (blk - 4 * LA) % BG;
            src = Bbase + (size_t)(s * BN + 16 * g) * a.Kp + k0;
            dst = st + SA + s * (BN * 64) + g * 1024;
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
}

template <int LA, int LB>
struct RFrag {
    v4i a[LA];
    v4i b[LB];
};

template <int LA, int LB, int WAVES>
__device__ __forceinline__ void res_stage_frag(const uint8_t* st, int ks, int wm, int wn, int lane, RFrag<LA, LB>& f) {
    constexpr int SA = LA * LM * 64, BN = 16 * WAVES;
    const int lr = lane & 31, q = ks * 2 + (lane >> 5);
    const int ra = wm * 32 + lr, rb = wn * 32 + lr;
#pragma unroll
    for (int s = 0; s < LA; ++s) f.a[s] = *(const v4i*)(st + s * (LM * 64) + ra * 64 + 16 * (q ^ swz(ra)));
#pragma unroll
    for (int t = 0; t < LB; ++t) f.b[t] = *(const v4i*)(st + SA + t * (BN * 64) + rb * 64 + 16 * (q ^ swz(rb)));
}

template <int LA, int LB>
__device__ __forceinline__ void res_mfma(const RFrag<LA, LB>& f, v16i (&acc)[LA][LB]) {
#pragma unroll
    for (int s = 0; s < LA; ++s)
#pragma unroll
        for (int t = 0; t < LB; ++t)
            acc[s][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.a[s], f.b[t], acc[s][t], 0, 0, 0);
}

template <int LA, int LB, int NS, int WAVES>
__device__ __forceinline__ void res_tile(const ResArgs& a, uint32_t tm, uint32_t tn, uint32_t mi) {
    constexpr int BN = 16 * WAVES, SA = LA * LM * 64, SB = LB * BN * 64, STAGE = SA + SB;
    constexpr int NBLK = 4 * LA + (BN / 16) * LB;
    constexpr int HI = (NBLK + WAVES - 1) / WAVES, LO = NBLK / WAVES, NHI = NBLK % WAVES == 0 ? WAVES : NBLK % WAVES;
    static_assert(NS >= 3 && NS <= 5 && NS * STAGE <= 160 * 1024, "ring of 3..5 stages within the LDS");
    __shared__ __attribute__((aligned(1024))) uint8_t s0[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s1[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s2[STAGE];
    __shared__ __attribute__((aligned(1024))) uint8_t s3[NS > 3 ? STAGE : 1];
    __shared__ __attribute__((aligned(1024))) uint8_t s4[NS > 4 ? STAGE : 1];
    const uint32_t m0 = tm * (LA * LM), n0 = tn * (LB * BN);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const bool hi = wave < NHI;

    v16i acc[LA][LB];
#pragma unroll
    for (int s = 0; s < LA; ++s)
#pragma unroll
        for (int t = 0; t < LB; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[s][t][r] = 0;

    const int br = lane >> 2, bq = (lane & 3) ^ swz(br);
    const int8_t* Abase = a.A8 + mi * a.a_plane + (size_t)(m0 + br) * a.Kp + 16 * bq;
    const int8_t* Bbase = a.B8 + mi * a.b_plane + (size_t)(n0 + br) * a.Kp + 16 * bq;
    // wait until at most `stages` stages of this wave's copies are in flight (0 .. NS - 1)
    auto wait_stages = [&](int stages) {
        const int per = hi ? HI : LO;
        switch (stages * per) {
#define HX_RES_WAIT(c) case c: __builtin_amdgcn_s_waitcnt(wait_imm(c, 15)); break;
            HX_RES_WAIT(0) HX_RES_WAIT(1) HX_RES_WAIT(2) HX_RES_WAIT(3) HX_RES_WAIT(4) HX_RES_WAIT(5)
            HX_RES_WAIT(6) HX_RES_WAIT(7) HX_RES_WAIT(8) HX_RES_WAIT(9) HX_RES_WAIT(10) HX_RES_WAIT(11)
            HX_RES_WAIT(12) HX_RES_WAIT(13) HX_RES_WAIT(14) HX_RES_WAIT(15) HX_RES_WAIT(16)
#undef HX_RES_WAIT
            default: __builtin_amdgcn_s_waitcnt(wait_imm(0, 15)); break;
        }
    };
    const uint32_t KT = a.Kp / LK;
    if (KT > 0) {
        // prologue: stages 0 .. NS-1 in flight, stage 0 landed everywhere
        res_stage_issue<LA, LB, WAVES>(a, Abase, Bbase, wave, 0, s0);
        if (KT > 1) res_stage_issue<LA, LB, WAVES>(a, Abase, Bbase, wave, 1, s1);
        if (KT > 2) res_stage_issue<LA, LB, WAVES>(a, Abase, Bbase, wave, 2, s2);
        if (NS > 3 && KT > 3) res_stage_issue<LA, LB, WAVES>(a, Abase, Bbase, wave, 3, s3);
        if (NS > 4 && KT > 4) res_stage_issue<LA, LB, WAVES>(a, Abase, Bbase, wave, 4, s4);
        wait_stages((int)min(KT - 1, (uint32_t)NS - 1));
        wg_barrier();
        RFrag<LA, LB> fa, fb;
        res_stage_frag<LA, LB, WAVES>(s0, 0, wm, wn, lane, fa);
        __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
        // step kt: MFMAs of (kt, ks0) while (kt, ks1) is read; stage kt+1 landed and every wave done
        // with stage kt (barrier); stage kt+NS copied into kt's slot; MFMAs of (kt, ks1) while
        // (kt+1, ks0) is read.  Before the barrier stages kt+2 .. min(KT-1, kt+NS-1) may be in flight.
        auto body = [&](uint32_t kt, uint8_t* cur, const uint8_t* nxt) {
            res_stage_frag<LA, LB, WAVES>(cur, 1, wm, wn, lane, fb);
            __builtin_amdgcn_sched_barrier(0);
            res_mfma<LA, LB>(fa, acc);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
            wait_stages((int)min(KT - kt - 2, (uint32_t)NS - 2));
            wg_barrier();
            if (kt + NS < KT) res_stage_issue<LA, LB, WAVES>(a, Abase, Bbase, wave, kt + NS, cur);
            res_stage_frag<LA, LB, WAVES>(nxt, 0, wm, wn, lane, fa);
            __builtin_amdgcn_sched_barrier(0);
            res_mfma<LA, LB>(fb, acc);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(wait_imm(63, 0));
        };
        uint8_t* const ring[5] = {s0, s1, s2, s3, s4};
        for (uint32_t kt = 0; kt + 1 < KT; kt += NS) {
#pragma unroll
            for (int j = 0; j < NS; ++j)
                if (kt + j + 1 < KT) body(kt + j, ring[j], ring[(j + 1) % NS]);
        }
        const uint32_t ls = (KT - 1) % NS;
        const uint8_t* cur = ls == 0 ? s0 : ls == 1 ? s1 : ls == 2 ? s2 : ls == 3 ? s3 : s4;
        res_stage_frag<LA, LB, WAVES>(cur, 1, wm, wn, lane, fb);
        res_mfma<LA, LB>(fa, acc);
        res_mfma<LA, LB>(fb, acc);
    }
    // s = v y mod q (y = (M / q)^-1 mod q, the CRT's factor, folded in here): t = floor(v y / q) from a
    // double (v y < 2^39 is exact there; the quotient's rounding error stays far below the 1/q spacing of
    // the fractions, so t is exact but for a true multiple of q, where it may come out one low: then
    // v y - t q = q, folded back)
    const int32_t Q = (int32_t)a.q[mi], Y = (int32_t)a.y[mi];
    const double iq = a.invq[mi];
    uint8_t* R = a.R + mi * a.r_plane;
    const int lr = lane & 31, lg = lane >> 5;
#pragma unroll
    for (int s = 0; s < LA; ++s)
#pragma unroll
        for (int t = 0; t < LB; ++t) {
            const uint32_t col = n0 + t * BN + wn * 32 + lr;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t row = m0 + s * LM + wm * 32 + (r / 4) * 8 + lg * 4 + (r % 4);
                const int64_t vy = (int64_t)acc[s][t][r] * Y;
                const int64_t qt = (int64_t)floor((double)vy * iq);
                int32_t u = (int32_t)(vy - qt * Q);
                u = u >= Q ? u - Q : u;
                u = u < 0 ? u + Q : u;
                R[(size_t)row * a.ldr + col] = (uint8_t)u;
            }
        }
}

template <int LA, int LB, int NS, int WAVES>
__global__ __launch_bounds__(WAVES * 64, 8 / WAVES) void res_gemm_kernel(ResArgs a) {
    uint32_t tm, tn;
    tile_of(blockIdx.x, a.ntm, a.ntn, tm, tn);
    res_tile<LA, LB, NS, WAVES>(a, tm, tn, blockIdx.y);
}

// CRT of the N product residues of four adjacent outputs per thread into x mod p, then op into C.
// The product kernel stored s_i = v y_i mod q_i (y_i = (M / q_i)^-1 mod q_i, M = prod q_i), so
// sum_i s_i M / q_i = x + T M with T = round(sum_i s_i / q_i) (that sum is T + x / M, |x / M| < 0.488),
// and x = sum_i s_i (M / q_i mod p) - T (M mod p) (mod p).  All in 24-bit products: c_i = M / q_i mod p
// split into 16-bit halves (the two partial sums stay below 12 * 255 * 2^16 < 2^28), the fraction sum in
// fixed point with f_i = round(2^20 / q_i) (total error <= 12 * 254 / 2 units of 2^-20 < 0.0015, inside
// the 0.5 - 0.488 margin).
struct CrtArgs {
    const uint8_t* R;
    size_t r_plane;
    uint32_t ldr;
    uint32_t nmod;
    uint32_t ch[RES_MAXN], cl[RES_MAXN];   // (M / q_i) mod p = ch 2^16 + cl
    uint32_t f[RES_MAXN];                  // round(2^20 / q_i)
    uint32_t pmm;                          // p - (M mod p)
    uint32_t m, n;
    const uint32_t* Cin;
    int64_t cirs, cics;
    const uint32_t* c_rows;
    uint32_t* Cout;
    int64_t cors, cocs;
    int op;
    DevMod M;
};

__global__ __launch_bounds__(256) void crt_kernel(CrtArgs a) {
    const uint32_t col0 = (blockIdx.x * 64 + (threadIdx.x & 63)) * 4;
    const uint32_t row = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (row >= a.m || col0 >= a.n) return;
    uint32_t hi[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0}, fr[4] = {0, 0, 0, 0};
    const uint8_t* src = a.R + (size_t)row * a.ldr + col0;
#pragma unroll
    for (int i = 0; i < RES_MAXN; ++i) {
        if (i >= (int)a.nmod) break;
        const uint32_t w = *(const uint32_t*)(src + (size_t)i * a.r_plane);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t sb = (w >> (8 * b)) & 0xffu;
            hi[b] += __umul24(sb, a.ch[i]);
            lo[b] += __umul24(sb, a.cl[i]);
            fr[b] += __umul24(sb, a.f[i]);
        }
    }
    const int64_t rr = a.c_rows ? (int64_t)a.c_rows[row] : (int64_t)row;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t col = col0 + b;
        if (col >= a.n) break;
        const uint32_t T = (fr[b] + (1u << 19)) >> 20;
        const uint32_t res = red64(((uint64_t)hi[b] << 16) + lo[b] + (uint64_t)T * a.pmm, a.M);
        uint32_t c = 0;
        if (a.op != HX_OP_SET) c = a.Cin[rr * a.cirs + (int64_t)col * a.cics];
        a.Cout[(int64_t)row * a.cors + (int64_t)col * a.cocs] = apply_op(c, res, a.op, a.M.p);
    }
}

// Limb kernel choice (hx_gemm_opts.tile for the limb variant): 1 = register-staged copy
// (limb_kernel), 2 = direct-to-LDS copy ring (limb2_kernel), 3 = the ring with the LDS reads
// pipelined against the MFMAs (limb3_kernel), 4 = limb3's pipeline on eight waves
// (limb4_kernel), 5 = the 96 x 128 tile on twelve waves (limb5_kernel); 0 = the default.
constexpr int kLimbDefault = 4;
inline int limb_version(int tile) { return tile >= 1 && tile <= 5 ? tile : kLimbDefault; }
// tile 6 of the limb variant: the residue route (res_gemm_kernel) in place of the limb kernels
constexpr int kResidueTile = 6;

template <int L>
void launch_limb(const LimbArgs& la, hipStream_t s, int ver) {
    const uint32_t blocks = la.ntm * la.ntn;
    if (ver == 2) {
        limb2_kernel<L><<<blocks, 256, 0, s>>>(la);
        after_launch("limb2_kernel");
        return;
    }
    if (ver == 3) {
        limb3_kernel<L><<<blocks, 256, 0, s>>>(la);
        after_launch("limb3_kernel");
        return;
    }
    if (ver == 4) {
        limb4_kernel<L><<<blocks, 512, 0, s>>>(la);
        after_launch("limb4_kernel");
        return;
    }
    if (ver == 5) {
        LimbArgs l5 = la;
        l5.ntm = (la.m + LM5 - 1) / LM5;
        limb5_kernel<L><<<l5.ntm * l5.ntn, 768, 0, s>>>(l5);
        after_launch("limb5_kernel");
        return;
    }
    limb_attrs<L>();
    limb_kernel<L><<<blocks, 256, limb_lds<L>(), s>>>(la);
    after_launch("limb_kernel");
}

inline uint32_t round_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

}  // namespace

uint32_t limb_k_chunk(uint32_t limbs) {
    const uint64_t kmax = 2147483647ull / ((uint64_t)limbs * 16384ull);
    return (uint32_t)(kmax / 64 * 64);
}

size_t limb_workspace_bytes(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k) {
    if (ctx.limbs == 0) return 0;
    const uint32_t kc = round_up(std::max<uint32_t>(1, std::min(k, limb_k_chunk(ctx.limbs))), 64);
    return (size_t)ctx.limbs * ((size_t)round_up(std::max<uint32_t>(m, 1), LM) + round_up(std::max<uint32_t>(n, 1), LN)) * kc;
}

// The K chunk one limb pass runs in a workspace of `bytes` (imma_k_chunk's role, CudaKernels.h:100-110):
// min(k, limb_k_chunk) when its planes fit, else the largest multiple of 64 whose planes fit; 0 when
// not even one 64-deep chunk fits.  Chunks after the first accumulate with Add / Sub: same bytes.
uint32_t limb_k_chunk_for(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k, size_t bytes) {
    if (ctx.limbs == 0) return 0;
    const uint32_t full = std::max<uint32_t>(1, std::min(k, limb_k_chunk(ctx.limbs)));
    if (limb_workspace_bytes(ctx, m, n, full) <= bytes) return full;
    const size_t per_k = (size_t)ctx.limbs * ((size_t)round_up(std::max<uint32_t>(m, 1), LM) +
                                              round_up(std::max<uint32_t>(n, 1), LN));
    const size_t kc = bytes / per_k / 64 * 64;
    return (uint32_t)std::min<size_t>(kc, full);
}

// The K slice of one split: a multiple of the kernel's K step, so every split but the last is full.
static uint32_t split_kc(uint32_t k, uint32_t splits) {
    const uint64_t per = ((uint64_t)k + splits - 1) / splits;
    return (uint32_t)((per + PK - 1) / PK * PK);
}

uint32_t plain_splits_clamp(uint32_t k, uint32_t splits) {
    if (splits <= 1 || k == 0) return 1;
    splits = std::min<uint32_t>(splits, HX_MAX_SPLITS);
    const uint64_t kc = split_kc(k, splits);
    return (uint32_t)(((uint64_t)k + kc - 1) / kc);                  // no empty split
}

size_t plain_split_workspace_bytes(uint32_t m, uint32_t n, uint32_t splits) {
    if (splits <= 1 || m == 0 || n == 0) return 0;
    const size_t tiles = (size_t)((m + PM - 1) / PM) * ((n + PN - 1) / PN);
    return (tiles * 4 + 255) / 256 * 256 + tiles * splits * (size_t)(PM * PN) * 4;
}

// The split count the automatic policy takes for the plain kernel (suggest_splits' role,
// cuda_kernels.cu:770-777, sized for gfx950): when the output tiles alone give fewer than 256
// workgroups (one per CU) and K is deep, enough K slices of at least 256 for about 1024 workgroups.
uint32_t plain_suggest_splits(uint32_t m, uint32_t n, uint32_t k) {
    const uint64_t tiles = (uint64_t)((m + PM - 1) / PM) * ((n + PN - 1) / PN);
    if (tiles == 0 || tiles >= 256 || k < 1024) return 1;
    uint64_t sp = (1024 + tiles - 1) / tiles;
    sp = std::min<uint64_t>(sp, k / 256);
    return plain_splits_clamp(k, (uint32_t)std::max<uint64_t>(1, sp));
}

void gemm_plain(const hx_modctx& ctx, const GemmArgs& g, hipStream_t s, int /*tile*/, hipEvent_t ev0, hipEvent_t ev1,
                uint32_t splits, void* ws, size_t ws_bytes) {
    if (g.m == 0 || g.n == 0) return;
    const uint32_t ntm = (g.m + PM - 1) / PM, ntn = (g.n + PN - 1) / PN;
    splits = plain_splits_clamp(g.k, splits);
    if (splits > 1) {
        HX_CHECK(g.m_dev == nullptr && g.k_dev == nullptr, "split-K needs host-side sizes");
        HX_CHECK(ws != nullptr && ws_bytes >= plain_split_workspace_bytes(g.m, g.n, splits),
                 "split-K workspace too small (hx_gemm_split_workspace_bytes)");
        const uint32_t tiles = ntm * ntn;
        SplitK sk;
        sk.arrive = (uint32_t*)ws;
        sk.part = (uint32_t*)((char*)ws + ((size_t)tiles * 4 + 255) / 256 * 256);
        sk.splits = splits;
        sk.kc = split_kc(g.k, splits);
        clear_words_kernel<<<(tiles + 255) / 256, 256, 0, s>>>(sk.arrive, tiles);
        after_launch("clear_words_kernel");
        if (ev0) HX_HIP(hipEventRecord(ev0, s));
        plain_split_kernel<<<tiles * splits, 256, 0, s>>>(g, devmod(ctx), sk);
        after_launch("plain_split_kernel");
        if (ev1) HX_HIP(hipEventRecord(ev1, s));
        return;
    }
    if (ev0) HX_HIP(hipEventRecord(ev0, s));
    plain_kernel<<<ntm * ntn, 256, 0, s>>>(g, devmod(ctx));
    after_launch("plain_kernel");
    if (ev1) HX_HIP(hipEventRecord(ev1, s));
}

// ws: the limb planes' workspace (ws_bytes of it); K runs in the chunks that fit it.  A NULL
// workspace has been resolved by the caller (gemm_dispatch).
void gemm_limb(const hx_modctx& ctx, const GemmArgs& g0, hipStream_t s, void* ws, size_t ws_bytes, int tile,
               hipEvent_t ev0, hipEvent_t ev1) {
    const int L = (int)ctx.limbs;
    HX_CHECK(L >= 1 && L <= 4, "the limb route needs p <= 4278124287 (ctx.limbs > 0)");
    if (g0.m == 0 || g0.n == 0) return;
    const uint32_t kchunk = limb_k_chunk_for(ctx, g0.m, g0.n, g0.k, ws ? ws_bytes : 0);
    HX_CHECK(ws != nullptr && kchunk > 0, "limb workspace too small for one 64-deep K chunk of this shape");
    HX_CHECK(g0.k_dev == nullptr || g0.k <= kchunk, "device-side k needs k <= one limb chunk");
    const uint32_t Mp = round_up(g0.m, LM), Np = round_up(g0.n, LN);
    const uint32_t p = ctx.p, half = (p - 1) / 2;
    GemmArgs g = g0;
    uint32_t k0 = 0;
    do {
        const uint32_t kc = std::min(kchunk, g.k - k0);
        const uint32_t Kp = round_up(kc, 64);
        int8_t* A8 = (int8_t*)ws;
        int8_t* B8 = A8 + (size_t)L * Mp * Kp;
        if (Kp > 0) {
            PrepArgs pa{};
            pa.src = g.a_cols ? g.A : g.A + (int64_t)k0 * g.acs;
            pa.s_r = g.ars;
            pa.s_c = g.acs;
            pa.rlist = g.a_rows;
            pa.clist = g.a_cols ? g.a_cols + k0 : nullptr;
            pa.R = g.m;
            pa.K = kc;
            pa.r_dev = g.m_dev;
            pa.k_dev = g.k_dev;
            pa.Kp = Kp;
            pa.out = A8;
            pa.plane = (size_t)Mp * Kp;
            pa.p = p;
            pa.half = half;
            pa.L = L;
            limb_prep_kernel<<<dim3(Kp / 64, Mp / 64), 256, 0, s>>>(pa);
            after_launch("limb_prep_kernel(A)");
            PrepArgs pb = pa;
            pb.src = g.B + (int64_t)k0 * g.brs;
            pb.s_r = g.bcs;
            pb.s_c = g.brs;
            pb.rlist = nullptr;
            pb.clist = nullptr;
            pb.R = g.n;
            pb.r_dev = nullptr;
            pb.out = B8;
            pb.plane = (size_t)Np * Kp;
            limb_prep_kernel<<<dim3(Kp / 64, Np / 64), 256, 0, s>>>(pb);
            after_launch("limb_prep_kernel(B)");
        }
        LimbArgs la{};
        la.A8 = A8;
        la.B8 = B8;
        la.a_plane = (size_t)Mp * Kp;
        la.b_plane = (size_t)Np * Kp;
        la.Kp = Kp;
        la.m = g.m;
        la.n = g.n;
        la.m_dev = g.m_dev;
        la.ntm = Mp / LM;
        la.ntn = Np / LN;
        la.Cin = g.Cin;
        la.cirs = g.cirs;
        la.cics = g.cics;
        la.c_rows = g.c_rows;
        la.Cout = g.Cout;
        la.cors = g.cors;
        la.cocs = g.cocs;
        la.op = g.op;
        limb_epilogue_consts(p, la);
        la.M = devmod(ctx);
        const bool last = k0 + kc >= g.k;
        const int ver = limb_version(tile);
        if (ev0 && k0 == 0) HX_HIP(hipEventRecord(ev0, s));
        switch (L) {
            case 1: launch_limb<1>(la, s, ver); break;
            case 2: launch_limb<2>(la, s, ver); break;
            case 3: launch_limb<3>(la, s, ver); break;
            default: launch_limb<4>(la, s, ver); break;
        }
        if (ev1 && last) HX_HIP(hipEventRecord(ev1, s));
        // later chunks accumulate into the C just written
        k0 += kc;
        g.op = g.op == HX_OP_SUB ? HX_OP_SUB : HX_OP_ADD;
        g.Cin = g.Cout;
        g.cirs = g.cors;
        g.cics = g.cocs;
        g.c_rows = nullptr;
    } while (k0 < g.k);
}

// ------------------------------------------------------------------------------------------
// residue route (host side)
// ------------------------------------------------------------------------------------------
// The least N with 20 prod_{i<N} q_i > 41 k (p/2)^2; 0 when k is 0 or past RES_KMAX (int32 sums) or when
// RES_MAXN moduli do not reach the bound.
uint32_t residue_moduli(uint32_t p, uint32_t k) {
    if (k == 0 || k > RES_KMAX) return 0;
    const unsigned __int128 h = p / 2, bound = (unsigned __int128)41 * k * h * h;     // < 2^86
    unsigned __int128 M = 1;
    for (int i = 0; i < RES_MAXN; ++i) {
        M *= res_modulus(i);
        if (M * 20 > bound) return (uint32_t)i + 1;
    }
    return 0;
}

size_t residue_workspace_bytes(uint32_t m, uint32_t n, uint32_t k, uint32_t nmod) {
    const size_t Mp = round_up(std::max<uint32_t>(m, 1), RTM), Np = round_up(std::max<uint32_t>(n, 1), RTN);
    const size_t Kp = round_up(std::max<uint32_t>(k, 1), 64);
    return (size_t)nmod * (Mp * Kp + Np * Kp + Mp * Np);
}

// The moduli count of the residue route when the limb variant takes it for this product, else 0.  It
// needs host-side sizes, 1 <= k <= RES_KMAX and enough moduli; tile kResidueTile asks for it wherever it
// applies; the default (tile 0) takes it where it runs at most L^2 - 4 int8 products (four-limb moduli:
// 10 or 11 for 16) and m, n, k >= 4096, where the saved MFMA work pays for the operand splits and the CRT
// pass (mod 2147483629, wall ms limb kernel 4 / residue route: 8192^3 7.76 / 6.26, 4096^3 1.13 / 1.01,
// 8192^2 x 2048 2.05 / 2.11, 2048^2 x 16384 1.14 / 1.18, 2048^3 0.17 / 0.24).
uint32_t residue_choice(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k, bool device_sizes, int tile) {
    if (device_sizes || (tile != 0 && tile != kResidueTile)) return 0;
    const uint32_t N = residue_moduli(ctx.p, k);
    if (N == 0 || tile == kResidueTile) return N;
    const bool large = m >= 4096 && n >= 4096 && k >= 4096;
    return large && ctx.limbs > 0 && N + 4 <= ctx.limbs * ctx.limbs ? N : 0;
}

// C op= A B mod p by the residue route with N moduli: residue planes of A and B (res_prep_kernel),
// one int8 product per modulus (res_gemm_kernel, grid.y = modulus), CRT into C (crt_kernel).
// ws holds N (Mp Kp + Np Kp + Mp Np) bytes (Mp, Np: m, n rounded up to 256; Kp: k to 64).
void gemm_residue(const hx_modctx& ctx, const GemmArgs& g, hipStream_t s, uint32_t N, void* ws, size_t ws_bytes,
                  hipEvent_t ev0, hipEvent_t ev1) {
    if (g.m == 0 || g.n == 0) return;
    HX_CHECK(N >= 1 && N <= (uint32_t)RES_MAXN && g.m_dev == nullptr && g.k_dev == nullptr && g.k >= 1,
             "residue route: no plan for this product");
    HX_CHECK(ws != nullptr && ws_bytes >= residue_workspace_bytes(g.m, g.n, g.k, N),
             "residue route: workspace too small");
    const uint32_t p = ctx.p;
    const uint32_t Mp = round_up(g.m, RTM), Np = round_up(g.n, RTN), Kp = round_up(g.k, 64);
    int8_t* A8 = (int8_t*)ws;
    int8_t* B8 = A8 + (size_t)N * Mp * Kp;
    uint8_t* R = (uint8_t*)(B8 + (size_t)N * Np * Kp);

    PrepArgs pa{};
    pa.src = g.A;
    pa.s_r = g.ars;
    pa.s_c = g.acs;
    pa.rlist = g.a_rows;
    pa.clist = g.a_cols;
    pa.R = g.m;
    pa.K = g.k;
    pa.Kp = Kp;
    pa.out = A8;
    pa.plane = (size_t)Mp * Kp;
    pa.p = p;
    pa.half = (p - 1) / 2;
    ResSplit rs{};
    rs.nmod = N;
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t q = res_modulus((int)i);
        rs.q[i] = q;
        rs.tpack[i] = 1u | ((256u % q) << 8) | ((65536u % q) << 16) | ((uint32_t)(16777216ull % q) << 24);
        rs.kq[i] = q - (0x80000000u % q);
        rs.iq[i] = 1.0f / (float)q;
    }
    // the direct split where the operand's k runs contiguously along 16-byte aligned rows (A row-major;
    // B given transposed, as gemm_nt does), the LDS-staged one otherwise
    auto split = [&](const PrepArgs& pr, uint32_t rows, const char* what) {
        const bool direct = pr.rlist == nullptr && pr.clist == nullptr && pr.s_c == 1 && pr.s_r % 4 == 0 &&
                            ((uintptr_t)pr.src & 15) == 0;
        const bool cols = pr.rlist == nullptr && pr.clist == nullptr && pr.s_r == 1;
        if (direct) res_prep_rows_kernel<<<dim3((Kp + 1023) / 1024, rows / 4), 256, 0, s>>>(pr, rs);
        else if (cols) res_prep_cols_kernel<<<dim3((Kp + 127) / 128, rows / 32), 256, 0, s>>>(pr, rs);
        else res_prep_kernel<<<dim3(Kp / 64, rows / 64), 256, 0, s>>>(pr, rs);
        after_launch(what);
    };
    split(pa, Mp, "res_prep(A)");
    PrepArgs pb = pa;
    pb.src = g.B;
    pb.s_r = g.bcs;
    pb.s_c = g.brs;
    pb.rlist = nullptr;
    pb.clist = nullptr;
    pb.R = g.n;
    pb.out = B8;
    pb.plane = (size_t)Np * Kp;
    split(pb, Np, "res_prep(B)");

    ResArgs ra{};
    ra.A8 = A8;
    ra.B8 = B8;
    ra.a_plane = (size_t)Mp * Kp;
    ra.b_plane = (size_t)Np * Kp;
    ra.Kp = Kp;
    ra.ntm = Mp / RTM;
    ra.ntn = Np / RTN;
    ra.R = R;
    ra.r_plane = (size_t)Mp * Np;
    ra.ldr = Np;
    CrtArgs ca{};
    unsigned __int128 M = 1;
    for (uint32_t i = 0; i < N; ++i) M *= res_modulus((int)i);
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t q = res_modulus((int)i);
        ra.q[i] = q;
        ra.invq[i] = 1.0 / (double)q;
        const unsigned __int128 Mi = M / q;
        const uint32_t mq = (uint32_t)(Mi % q);
        uint32_t y = 1;
        while (y < q && (mq * y) % q != 1) ++y;              // q <= 255 and gcd(M / q, q) = 1
        HX_CHECK(y < q, "residue route: moduli not coprime");
        ra.y[i] = y;
        const uint32_t c = (uint32_t)(Mi % p);
        ca.ch[i] = c >> 16;
        ca.cl[i] = c & 0xffffu;
        ca.f[i] = (uint32_t)(((1u << 20) + q / 2) / q);
    }
    ca.pmm = p - (uint32_t)(M % p);
    if (ev0) HX_HIP(hipEventRecord(ev0, s));
    res_gemm_kernel<RLA, RLB, RES_STAGES, RES_WAVES><<<dim3(ra.ntm * ra.ntn, N), RES_WAVES * 64, 0, s>>>(ra);
    after_launch("res_gemm_kernel");
    if (ev1) HX_HIP(hipEventRecord(ev1, s));

    ca.R = R;
    ca.r_plane = ra.r_plane;
    ca.ldr = Np;
    ca.nmod = N;
    ca.m = g.m;
    ca.n = g.n;
    ca.Cin = g.Cin;
    ca.cirs = g.cirs;
    ca.cics = g.cics;
    ca.c_rows = g.c_rows;
    ca.Cout = g.Cout;
    ca.cors = g.cors;
    ca.cocs = g.cocs;
    ca.op = g.op;
    ca.M = devmod(ctx);
    crt_kernel<<<dim3((g.n + 255) / 256, (g.m + 3) / 4), 256, 0, s>>>(ca);
    after_launch("crt_kernel");
}

// The automatic choice: the limb route once the product is big enough to pay for the two
// operand splits, and deep enough (k >= 64, one full LDS stage) for the MFMA work to outweigh the
// recombination epilogue; thinner products (the k = 32 Schur updates of the eliminations) run
// faster on the plain kernel.
int gemm_auto_variant(const hx_modctx& ctx, uint32_t m, uint32_t n, uint32_t k) {
    if (ctx.limbs == 0 || k < 64) return HX_GEMM_PLAIN;
    const double work = (double)m * n * k;
    return work >= 4.0e6 ? HX_GEMM_LIMB : HX_GEMM_PLAIN;
}

void gemm_dispatch(const hx_modctx& ctx, const GemmArgs& g, hipStream_t s, int variant, void* ws, size_t ws_bytes,
                   int tile, hipEvent_t ev0, hipEvent_t ev1, uint32_t splits, void* split_ws, size_t split_ws_bytes) {
    const bool automatic = variant == HX_GEMM_AUTO;
    if (automatic) variant = gemm_auto_variant(ctx, g.m, g.n, g.k);
    // a thin deep product whose limb tiles cannot fill the GPU runs faster on the plain kernel split
    // over K (256 x 256 x 65536: plain split 64 ways 0.87 ms, limb route 1.50 ms, plain unsplit 11.7 ms)
    if (automatic && variant == HX_GEMM_LIMB && splits != 1 && g.m_dev == nullptr && g.k_dev == nullptr) {
        const uint64_t limb_tiles = (uint64_t)((g.m + LM - 1) / LM) * ((g.n + LN - 1) / LN);
        if (limb_tiles < 32 && plain_suggest_splits(g.m, g.n, g.k) > 1) variant = HX_GEMM_PLAIN;
    }
    uint32_t rn = 0;            // moduli of the residue route, 0 = the limb kernels
    if (variant == HX_GEMM_LIMB && g.m != 0 && g.n != 0)
        rn = residue_choice(ctx, g.m, g.n, g.k, g.m_dev != nullptr || g.k_dev != nullptr, tile);
    if (rn != 0) {
        // the default route falls back to the limb kernels when its planes do not fit (a caller workspace
        // sized for the limb planes, or the library's own that cannot grow inside a capture)
        const bool asked = tile == kResidueTile;
        const size_t need = residue_workspace_bytes(g.m, g.n, g.k, rn);
        if (ws == nullptr) {
            void* own = stream_scratch(s, SCRATCH_LIMB, need);
            if (own != nullptr) {
                ws = own;
                ws_bytes = need;
            } else if (asked) {
                throw Error{HX_E_CAPTURE, "the residue route needs a caller workspace inside a stream capture "
                                          "(hx_gemm_opts.workspace; the library's own cannot grow there)"};
            } else {
                rn = 0;
            }
        } else if (ws_bytes < need) {
            HX_CHECK(!asked, "residue route: workspace too small (hx_gemm_workspace_bytes_tile)");
            rn = 0;
        }
    }
    if (rn == 0 && variant == HX_GEMM_LIMB && g.m != 0 && g.n != 0) {
        if (ws == nullptr) {
            // library-owned planes private to this stream (never shared with another stream); inside a
            // capture they can only be used when they are already large enough
            const size_t need = limb_workspace_bytes(ctx, g.m, g.n, g.k);
            ws = stream_scratch(s, SCRATCH_LIMB, need);
            ws_bytes = ws ? need : 0;
            if (ws == nullptr && !automatic)
                throw Error{HX_E_CAPTURE, "the limb route needs a caller workspace inside a stream capture "
                                          "(hx_gemm_opts.workspace; the library's own cannot grow there)"};
        }
        // the limb route runs when one 64-deep K chunk fits the workspace (the reference takes its IMMA
        // route only with a large enough imma_workspace); otherwise AUTO runs the plain kernel
        if (limb_k_chunk_for(ctx, g.m, g.n, g.k, ws == nullptr ? 0 : ws_bytes) == 0) {
            HX_CHECK(automatic, "limb workspace too small for one 64-deep K chunk of this shape");
            variant = HX_GEMM_PLAIN;
        }
    }
    // split-K (plain kernel only, host-side sizes only): splits == 0 is the automatic policy
    uint32_t sp = 1;
    if (variant != HX_GEMM_LIMB && g.m != 0 && g.n != 0 && g.m_dev == nullptr && g.k_dev == nullptr) {
        const bool auto_split = splits == 0;
        sp = auto_split ? plain_suggest_splits(g.m, g.n, g.k) : plain_splits_clamp(g.k, splits);
        if (sp > 1) {
            const size_t need = plain_split_workspace_bytes(g.m, g.n, sp);
            if (split_ws == nullptr) {
                // library-owned partials private to this stream; inside a capture only when already large enough
                split_ws = stream_scratch(s, SCRATCH_SPLIT, need);
                split_ws_bytes = split_ws ? need : 0;
                if (split_ws == nullptr && !auto_split)
                    throw Error{HX_E_CAPTURE, "split-K needs a caller workspace inside a stream capture "
                                              "(hx_gemm_opts.split_workspace; the library's own cannot grow there)"};
            } else {
                HX_CHECK(auto_split || split_ws_bytes >= need,
                         "split-K workspace too small (hx_gemm_split_workspace_bytes)");
            }
            if (split_ws == nullptr || split_ws_bytes < need) sp = 1;      // automatic: no split
        }
    }
    if (rn != 0)
        gemm_residue(ctx, g, s, rn, ws, ws_bytes, ev0, ev1);
    else if (variant == HX_GEMM_LIMB)
        gemm_limb(ctx, g, s, ws, ws_bytes, tile, ev0, ev1);
    else
        gemm_plain(ctx, g, s, tile, ev0, ev1, sp, split_ws, split_ws_bytes);
    if (g.m != 0 && g.n != 0) {
        stat_add(ST_GEMMS, 1);
        if (variant == HX_GEMM_LIMB) stat_add(ST_LIMB, 1);
        if (sp > 1) stat_add(ST_SPLITK, 1);
    }
}

// ---- batched uneven GEMM -------------------------------------------------------------------
namespace {

// The limb jobs of one kernel choice (hx_gemm_job.tile 1..4): one launch each.
struct LimbGroup {
    uint32_t n = 0, tiles = 0;
    LimbArgs* d_args = nullptr;
    uint32_t* d_prefix = nullptr;         // n + 1 entries
};

struct GBatch {
    hx_modctx ctx;
    uint32_t n_plain = 0, n_limb = 0, n_prep = 0;
    uint32_t tiles_plain = 0, tiles_prep = 0;
    GemmArgs* d_plain = nullptr;
    PrepArgs* d_prep = nullptr;
    uint32_t* d_prefix = nullptr;        // [plain prefix (n_plain+1) | prep prefix (n_prep+1)]
    LimbGroup group[5];                  // index = kernel choice 1..4
    int8_t* d_ws = nullptr;
};

void gbatch_free(GBatch* b) {
    if (!b) return;
    (void)hipFree(b->d_plain);
    (void)hipFree(b->d_prep);
    (void)hipFree(b->d_prefix);
    for (LimbGroup& g : b->group) {
        (void)hipFree(g.d_args);
        (void)hipFree(g.d_prefix);
    }
    (void)hipFree(b->d_ws);
    delete b;
}

template <class T>
T* upload_vec(const std::vector<T>& v) {
    T* d = nullptr;
    if (v.empty()) return nullptr;
    HX_HIP(hipMalloc(&d, v.size() * sizeof(T)));
    HX_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

void* gemm_batch_create(const hx_modctx& ctx, const hx_gemm_job* jobs, uint32_t njobs, int variant) {
    GBatch* b = new GBatch();
    try {
        b->ctx = ctx;
        std::vector<GemmArgs> plain;
        std::vector<LimbArgs> limb;
        std::vector<int> limb_ver;          // each limb job's kernel choice (1..4)
        std::vector<PrepArgs> prep;
        std::vector<uint32_t> pp{0}, prp{0};
        size_t ws = 0;
        const int L = (int)ctx.limbs;
        const uint32_t kchunk = L ? limb_k_chunk(L) : 0;
        struct Pending { size_t a_off, b_off; uint32_t Mp, Np, Kp; };
        std::vector<Pending> pend;
        const uint32_t p = ctx.p, half = (p - 1) / 2;
        // largest products first, so the big jobs' tiles are dispatched first and the small ones fill
        // the tail (the order of the job table does not change any value)
        std::vector<uint32_t> ord(njobs);
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
            return (uint64_t)jobs[x].m * jobs[x].n * jobs[x].k > (uint64_t)jobs[y].m * jobs[y].n * jobs[y].k;
        });
        for (uint32_t oi = 0; oi < njobs; ++oi) {
            const hx_gemm_job& j = jobs[ord[oi]];
            HX_CHECK(j.op == HX_OP_SET || j.op == HX_OP_ADD || j.op == HX_OP_SUB, "gemm batch: bad op");
            if (j.m == 0 || j.n == 0) continue;
            HX_CHECK(j.C != nullptr && (j.k == 0 || (j.A && j.B)), "gemm batch: NULL operand");
            HX_CHECK(j.lda >= (int64_t)j.k && j.ldb >= (int64_t)j.n && j.ldc >= (int64_t)j.n, "gemm batch: bad ld");
            GemmArgs g;
            g.m = j.m; g.n = j.n; g.k = j.k;
            g.A = j.A; g.ars = j.lda; g.acs = 1;
            g.B = j.B; g.brs = j.ldb; g.bcs = 1;
            g.Cin = j.C; g.cirs = j.ldc; g.cics = 1;
            g.Cout = j.C; g.cors = j.ldc; g.cocs = 1;
            g.op = j.op;
            int v = j.variant != HX_GEMM_AUTO ? j.variant : variant;
            if (v == HX_GEMM_AUTO) v = gemm_auto_variant(ctx, j.m, j.n, j.k);
            if (v == HX_GEMM_LIMB && (L == 0 || j.k > kchunk)) v = HX_GEMM_PLAIN;
            if (v != HX_GEMM_LIMB) {
                plain.push_back(g);
                pp.push_back(pp.back() + ((j.m + PM - 1) / PM) * ((j.n + PN - 1) / PN));
                continue;
            }
            const uint32_t Mp = round_up(j.m, LM), Np = round_up(j.n, LN), Kp = round_up(j.k, 64);
            pend.push_back({ws, ws + (size_t)L * Mp * Kp, Mp, Np, Kp});
            ws += (size_t)L * (Mp + Np) * Kp;
            LimbArgs la{};
            la.a_plane = (size_t)Mp * Kp;
            la.b_plane = (size_t)Np * Kp;
            la.Kp = Kp;
            la.m = j.m; la.n = j.n;
            la.ntm = Mp / LM; la.ntn = Np / LN;
            la.Cin = j.C; la.cirs = j.ldc; la.cics = 1;
            la.Cout = j.C; la.cors = j.ldc; la.cocs = 1;
            la.op = j.op; la.M = devmod(ctx);
            limb_epilogue_consts(p, la);
            HX_CHECK(j.tile >= 0 && j.tile <= 4, "gemm batch: tile must be 0 .. 4");
            limb.push_back(la);
            limb_ver.push_back(limb_version(j.tile));
            if (Kp > 0) {
                PrepArgs pa{};
                pa.src = j.A; pa.s_r = j.lda; pa.s_c = 1;
                pa.R = j.m; pa.K = j.k; pa.Kp = Kp; pa.plane = (size_t)Mp * Kp;
                pa.p = p; pa.half = half; pa.L = L;
                prep.push_back(pa);
                prp.push_back(prp.back() + (Kp / 64) * (Mp / 64));
                PrepArgs pb = pa;
                pb.src = j.B; pb.s_r = 1; pb.s_c = j.ldb; pb.R = j.n; pb.plane = (size_t)Np * Kp;
                prep.push_back(pb);
                prp.push_back(prp.back() + (Kp / 64) * (Np / 64));
            }
        }
        if (ws) HX_HIP(hipMalloc(&b->d_ws, ws));
        size_t pi = 0;
        for (size_t i = 0; i < limb.size(); ++i) {
            limb[i].A8 = b->d_ws + pend[i].a_off;
            limb[i].B8 = b->d_ws + pend[i].b_off;
            if (pend[i].Kp > 0) {
                prep[pi++].out = b->d_ws + pend[i].a_off;
                prep[pi++].out = b->d_ws + pend[i].b_off;
            }
        }
        b->n_plain = (uint32_t)plain.size();
        b->n_limb = (uint32_t)limb.size();
        b->n_prep = (uint32_t)prep.size();
        b->tiles_plain = pp.back();
        b->tiles_prep = prp.back();
        b->d_plain = upload_vec(plain);
        b->d_prep = upload_vec(prep);
        std::vector<uint32_t> pre(pp);
        pre.insert(pre.end(), prp.begin(), prp.end());
        b->d_prefix = upload_vec(pre);
        // the limb jobs by kernel choice, largest first within a group (their order above); a job of 64+
        // tiles starts at a multiple of 8 blocks, so block b runs on XCD b mod 8 with the job-local
        // XCD-aware tile order (tile_of) intact; the skipped blocks exit at once
        for (int v = 1; v <= 4; ++v) {
            std::vector<LimbArgs> ga;
            std::vector<uint32_t> gp{0};
            for (size_t i = 0; i < limb.size(); ++i) {
                if (limb_ver[i] != v) continue;
                ga.push_back(limb[i]);
                const uint32_t tiles = limb[i].ntm * limb[i].ntn;
                if (tiles >= 64) gp.back() = (gp.back() + 7) / 8 * 8;
                gp.push_back(gp.back() + tiles);
            }
            LimbGroup& g = b->group[v];
            g.n = (uint32_t)ga.size();
            g.tiles = gp.back();
            if (g.n) {
                g.d_args = upload_vec(ga);
                g.d_prefix = upload_vec(gp);
            }
        }
        switch (L) {
            case 1: limb_attrs<1>(); break;
            case 2: limb_attrs<2>(); break;
            case 3: limb_attrs<3>(); break;
            case 4: limb_attrs<4>(); break;
            default: break;
        }
    } catch (...) {
        gbatch_free(b);
        throw;
    }
    return b;
}

template <int L>
void launch_limb_group(const LimbGroup& g, int v, hipStream_t s) {
    if (g.tiles == 0) return;
    switch (v) {
        case 1: limb_batch_kernel<L><<<g.tiles, 256, limb_lds<L>(), s>>>(g.d_args, g.d_prefix, g.n); break;
        case 2: limb2_batch_kernel<L><<<g.tiles, 256, 0, s>>>(g.d_args, g.d_prefix, g.n); break;
        case 3: limb3_batch_kernel<L><<<g.tiles, 256, 0, s>>>(g.d_args, g.d_prefix, g.n); break;
        default: limb4_batch_kernel<L><<<g.tiles, 512, 0, s>>>(g.d_args, g.d_prefix, g.n); break;
    }
    after_launch("limb_batch_kernel");
}

void gemm_batch_run(void* vb, hipStream_t s) {
    GBatch* b = (GBatch*)vb;
    const uint32_t* pre_plain = b->d_prefix;
    const uint32_t* pre_prep = pre_plain + b->n_plain + 1;
    stat_add(ST_BATCHED, (uint64_t)b->n_plain + b->n_limb);
    if (b->tiles_prep) {
        prep_batch_kernel<<<b->tiles_prep, 256, 0, s>>>(b->d_prep, pre_prep, b->n_prep);
        after_launch("prep_batch_kernel");
    }
    for (int v = 1; v <= 4; ++v) {
        const LimbGroup& g = b->group[v];
        if (g.tiles == 0) continue;
        stat_add(ST_GEMMS, 1);
        stat_add(ST_LIMB, 1);
        switch (b->ctx.limbs) {
            case 1: launch_limb_group<1>(g, v, s); break;
            case 2: launch_limb_group<2>(g, v, s); break;
            case 3: launch_limb_group<3>(g, v, s); break;
            default: launch_limb_group<4>(g, v, s); break;
        }
    }
    if (b->tiles_plain) {
        stat_add(ST_GEMMS, 1);
        plain_batch_kernel<<<b->tiles_plain, 256, 0, s>>>(b->d_plain, pre_plain, b->n_plain, devmod(b->ctx));
        after_launch("plain_batch_kernel");
    }
}

}  // namespace hx

using namespace hx;

namespace {

GemmArgs args_of(hx_view A, hx_view B, hx_view C, int op) {
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

// Calibration of the limb kernel's instruction: each wave runs `iters` rounds of the 16 products of a
// 4-limb k-step (a[s] x b[t] onto acc[s + t], the accumulation pattern of limb_tile4) from registers only,
// then folds its accumulators into one word per lane (out[blockIdx * 256 + tid]).  With `random` set the
// operands are SplitMix64 bits of (salt, lane, slot), as uniform as the limb planes of random residues;
// else they are small constants that barely toggle.  The ratio of the two rates is what operand data
// costs the chip's clock.
__global__ __launch_bounds__(256) void mfma_rate_kernel(int* out, uint32_t iters, uint64_t salt, int random) {
    const uint64_t id = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    v4i a[4], b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (random) {
            uint64_t x = salt ^ (id * 8 + s), y = salt ^ (id * 8 + 4 + s);
            auto mix = [](uint64_t z) {
                z += 0x9E3779B97F4A7C15ull;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                return z ^ (z >> 31);
            };
            const uint64_t x0 = mix(x), x1 = mix(x0), y0 = mix(y), y1 = mix(y0);
            a[s] = v4i{(int)x0, (int)(x0 >> 32), (int)x1, (int)(x1 >> 32)};
            b[s] = v4i{(int)y0, (int)(y0 >> 32), (int)y1, (int)(y1 >> 32)};
        } else {
            a[s] = v4i{(int)salt + s, (int)threadIdx.x, s, 1};
            b[s] = v4i{(int)salt - s, 3, (int)threadIdx.x, 2};
        }
    }
    v16i acc[7];
#pragma unroll
    for (int d = 0; d < 7; ++d)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[d][r] = 0;
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[s], b[t], acc[s + t], 0, 0, 0);
    }
    int sum = 0;
#pragma unroll
    for (int d = 0; d < 7; ++d)
#pragma unroll
        for (int r = 0; r < 16; ++r) sum += acc[d][r];
    out[id] = sum;
}

}  // namespace

extern "C" {

HX_API size_t hx_gemm_workspace_bytes_tile(const hx_modctx* ctx, uint32_t m, uint32_t n, uint32_t k, int variant,
                                           int tile) {
    if (ctx == nullptr) return 0;
    if (variant == HX_GEMM_AUTO) variant = gemm_auto_variant(*ctx, m, n, k);
    if (variant != HX_GEMM_LIMB) return 0;
    const uint32_t rn = residue_choice(*ctx, m, n, k, false, tile);
    return rn ? residue_workspace_bytes(m, n, k, rn) : limb_workspace_bytes(*ctx, m, n, k);
}

HX_API size_t hx_gemm_workspace_bytes(const hx_modctx* ctx, uint32_t m, uint32_t n, uint32_t k, int variant) {
    return hx_gemm_workspace_bytes_tile(ctx, m, n, k, variant, 0);
}

HX_API uint32_t hx_gemm_residue_moduli(uint32_t p, uint32_t k) { return p >= 2 ? residue_moduli(p, k) : 0; }

HX_API int hx_gemm_route(const hx_modctx* ctx, uint32_t m, uint32_t n, uint32_t k, int variant, int tile) {
    if (ctx == nullptr) return -1;
    if (variant == HX_GEMM_AUTO) variant = gemm_auto_variant(*ctx, m, n, k);
    if (variant != HX_GEMM_LIMB || ctx->limbs == 0) return 0;
    if (residue_choice(*ctx, m, n, k, false, tile) != 0) return kResidueTile;
    return limb_version(tile);
}

HX_API int hx_gemm_u32(const hx_modctx* ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s,
                       const hx_gemm_opts* opts) {
    return guard("hx_gemm_u32", [&] {
        HX_CHECK(ctx != nullptr && ctx->p >= 2, "bad modulus context");
        HX_CHECK(A.cols == B.rows && C.rows == A.rows && C.cols == B.cols, "gemm: shape mismatch");
        HX_CHECK(op == HX_OP_SET || op == HX_OP_ADD || op == HX_OP_SUB, "gemm: bad op");
        HX_CHECK(!views_overlap(C, A) && !views_overlap(C, B), "gemm: C aliases an operand");
        HX_CHECK(C.rows == 0 || C.cols == 0 || C.base != nullptr, "gemm: C is NULL");
        const int variant = opts ? opts->variant : HX_GEMM_AUTO;
        HX_CHECK(variant != HX_GEMM_LIMB || ctx->limbs > 0, "gemm: the limb route needs p <= 4278124287");
        gemm_dispatch(*ctx, args_of(A, B, C, op), (hipStream_t)s, variant, opts ? opts->workspace : nullptr,
                      opts ? opts->workspace_bytes : 0, opts ? opts->tile : 0,
                      opts ? (hipEvent_t)opts->main_begin : nullptr, opts ? (hipEvent_t)opts->main_end : nullptr,
                      opts ? opts->splits : 0u, opts ? opts->split_workspace : nullptr,
                      opts ? opts->split_workspace_bytes : 0);
    });
}

HX_API uint32_t hx_gemm_suggest_splits(uint32_t m, uint32_t n, uint32_t k) { return plain_suggest_splits(m, n, k); }

HX_API size_t hx_gemm_split_workspace_bytes(uint32_t m, uint32_t n, uint32_t k, uint32_t splits) {
    return plain_split_workspace_bytes(m, n, plain_splits_clamp(k, splits));
}

HX_API uint32_t hx_gemm_limb_k_chunk(const hx_modctx* ctx, uint32_t m, uint32_t n, uint32_t k, size_t workspace_bytes) {
    if (ctx == nullptr) return 0;
    return limb_k_chunk_for(*ctx, m, n, k, workspace_bytes);
}

HX_API int hx_gemm_batch_create(const hx_modctx* ctx, const hx_gemm_job* jobs, uint32_t njobs, int variant,
                                hx_gemm_batch* out) {
    return guard("hx_gemm_batch_create", [&] {
        HX_CHECK(ctx != nullptr && out != nullptr && (jobs != nullptr || njobs == 0), "NULL argument");
        HX_CHECK(variant != HX_GEMM_LIMB || ctx->limbs > 0, "gemm batch: the limb route needs p <= 4278124287");
        *out = gemm_batch_create(*ctx, jobs, njobs, variant);
    });
}
HX_API int hx_gemm_batch_run(hx_gemm_batch b, hx_stream s) {
    return guard("hx_gemm_batch_run", [&] {
        HX_CHECK(b != nullptr, "NULL batch");
        gemm_batch_run(b, (hipStream_t)s);
    });
}
HX_API int hx_gemm_batch_destroy(hx_gemm_batch b) {
    return guard("hx_gemm_batch_destroy", [&] { gbatch_free((GBatch*)b); });
}

HX_API int hx_schur_update_u32(const hx_modctx* ctx, hx_view L, hx_view U, hx_view C, hx_stream s,
                               const hx_gemm_opts* opts) {
    return hx_gemm_u32(ctx, L, U, C, HX_OP_SUB, s, opts);
}

HX_API int hx_mfma_rate(uint32_t blocks, uint32_t iters, int random_operands, hx_stream s, double* tops) {
    return guard("hx_mfma_rate", [&] {
        HX_CHECK(tops != nullptr, "NULL argument");
        HX_CHECK(blocks >= 1 && blocks <= 65536 && iters >= 1 && iters <= 1000000, "mfma rate: bad size");
        const hipStream_t st = (hipStream_t)s;
        struct Res {       // released on every exit, the stream drained first
            hipStream_t st;
            int* out = nullptr;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            ~Res() {
                (void)hipStreamSynchronize(st);
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                if (out) arena_free(out);
            }
        } r{st};
        r.out = (int*)arena_alloc(sizeof(int) * 256 * (size_t)blocks);
        HX_HIP(hipEventCreate(&r.e0));
        HX_HIP(hipEventCreate(&r.e1));
        mfma_rate_kernel<<<blocks, 256, 0, st>>>(r.out, 4, 1, random_operands);     // first launch, untimed
        HX_HIP(hipGetLastError());
        HX_HIP(hipEventRecord(r.e0, st));
        mfma_rate_kernel<<<blocks, 256, 0, st>>>(r.out, iters, 0x6D66ull, random_operands);
        HX_HIP(hipGetLastError());
        HX_HIP(hipEventRecord(r.e1, st));
        HX_HIP(hipEventSynchronize(r.e1));
        float ms = 0.f;
        HX_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
        *tops = (double)blocks * 4.0 * iters * 16.0 * 65536.0 / (ms * 1e-3) / 1e12;
    });
}

}  // extern "C"
