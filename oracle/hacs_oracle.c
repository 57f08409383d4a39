/*
 * hacs_oracle.c -- the CPU oracle of the exact mod-p matrix path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * check of __graft_entry__.py and the cpu_baseline leg of bench.py load it;
 * the product (hacs_amd/, libhacs_mi355x.so) never links, loads or calls it.
 *
 * It restates, in plain C, the reference CPU algorithms of HACS's matrix
 * engine.  Every function cites the reference file:line it follows (paths are
 * relative to the reference's src/ModArithmetic/ unless stated).  All
 * arithmetic is exact integer arithmetic: every output is the unique canonical
 * value the reference produces, so results are compared with memcmp.
 *
 * Pinning (what checks each function independently of its own formula) is
 * listed per function and exercised by tests/test_oracle.py.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define HXO_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Fixed-seed generators: PerformanceSmokeValidation.cpp:64-89              */
/* Pinned by: tests/golden vectors written from the compiled reference, and */
/* by the published SplitMix64 test vector (seed 0 -> 0xE220A8397B1DCDAF).  */
/* ------------------------------------------------------------------------ */
HXO_API uint64_t hxo_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* SeedFor: PerformanceSmokeValidation.cpp:74-80 */
HXO_API uint64_t hxo_seed_for(uint32_t r, uint32_t c, uint32_t ell, uint64_t salt) {
    uint64_t s = 0xC0FFEE123456789ull ^ salt;
    s = s * 0x100000001B3ull + r;
    s = s * 0x100000001B3ull + c;
    s = s * 0x100000001B3ull + ell;
    return s;
}

/* RandomEntries: PerformanceSmokeValidation.cpp:82-89 */
HXO_API void hxo_random_entries(uint32_t *out, size_t count, uint32_t ell, uint64_t seed) {
    uint64_t st = seed;
    for (size_t i = 0; i < count; ++i) out[i] = (uint32_t)(hxo_splitmix64(&st) % ell);
}

/* ------------------------------------------------------------------------ */
/* Modular primitives: ModCtx.h:202-250 (init), :323-329 (reduce_u64),      */
/* ModArithmetic.cpp:54-78 (mod_inv, extended Euclid).                      */
/* Pinned by: closed forms (k_max tightness, reduce == x % p on boundary    */
/* values, a * inv(a) == 1) in tests/test_oracle.py.                        */
/* ------------------------------------------------------------------------ */
HXO_API uint64_t hxo_barrett_m(uint32_t p) { return (uint64_t)(((u128)1 << 64) / p); }

HXO_API uint64_t hxo_k_max(uint32_t p) {
    const uint64_t d = (uint64_t)(p - 1) * (p - 1);
    return d == 0 ? UINT64_MAX : UINT64_MAX / d;
}

static inline uint32_t red64(uint64_t x, uint32_t p, uint64_t bm) {
    const uint64_t q = (uint64_t)(((u128)x * bm) >> 64);
    uint64_t r = x - q * (uint64_t)p;
    if (r >= p) r -= p;
    return (uint32_t)r;
}

HXO_API uint32_t hxo_reduce_u64(uint32_t p, uint64_t x) { return red64(x, p, hxo_barrett_m(p)); }

static inline uint32_t madd(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t r = a + b;
    if (r < a || r >= p) r -= p;
    return r;
}
static inline uint32_t msub(uint32_t a, uint32_t b, uint32_t p) { return a >= b ? a - b : a + (p - b); }
static inline uint32_t mmul(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)(((uint64_t)a * b) % p); }

/* Extended Euclid, signed coefficients; 0 when a has no inverse. */
HXO_API uint32_t hxo_inv(uint32_t a, uint32_t p) {
    int64_t r0 = p, r1 = a % p, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const int64_t q = r0 / r1, r2 = r0 - q * r1, s2 = s0 - q * s1;
        r0 = r1; r1 = r2; s0 = s1; s1 = s2;
    }
    if (r0 != 1) return 0;
    int64_t s = s0 % (int64_t)p;
    if (s < 0) s += p;
    return (uint32_t)s;
}

/* ------------------------------------------------------------------------ */
/* FNV-1a-64 over (rows, cols, entries row-major) fed as u64 words, least   */
/* significant byte first: Fnv.h:49-55 (FnvU64) and                         */
/* PerformanceSmokeValidation.cpp:117-125 (feed_buf).                       */
/* Pinned by: the published FNV-1a vectors (through hxo_fnv1a_bytes over   */
/* the same bytes), the reference's engine::FnvHasher on fixed word lists   */
/* and reference-written fixtures.                                          */
/* ------------------------------------------------------------------------ */
static inline uint64_t fnv_u64(uint64_t h, uint64_t v) {
    for (int i = 0; i < 8; ++i) {
        h ^= (v >> (8 * i)) & 0xffu;
        h *= 0x100000001b3ull;
    }
    return h;
}
/* FNV-1a-64 of a byte string (the byte step of Fnv.h:58-64).  Pinned by the published FNV-1a
 * test vectors and by the reference's FnvBytes (a fixture written with oracle/_ref). */
HXO_API uint64_t hxo_fnv1a_bytes(const uint8_t *b, size_t n) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
HXO_API uint64_t hxo_fnv_matrix(const uint32_t *M, uint32_t rows, uint32_t cols, size_t ld) {
    uint64_t h = 0xcbf29ce484222325ull;
    h = fnv_u64(h, rows);
    h = fnv_u64(h, cols);
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t j = 0; j < cols; ++j) h = fnv_u64(h, M[(size_t)i * ld + j]);
    return h;
}

/* ElimWork: PerformanceSmokeValidation.cpp:202-208 */
HXO_API double hxo_elim_work(uint32_t rows, uint32_t cols, uint32_t rank) {
    double w = 0.0;
    for (uint32_t s = 0; s < rank; ++s) w += (double)(rows - 1) * (double)(cols - s);
    return w;
}

/* ------------------------------------------------------------------------ */
/* A tiny persistent pthread team: run fn(tid, T, arg) on T threads.        */
/* ------------------------------------------------------------------------ */
typedef void (*team_fn)(int tid, int nthreads, void *arg);
struct team_job { team_fn fn; void *arg; int tid, nthreads; };
static void *team_entry(void *p) {
    struct team_job *j = (struct team_job *)p;
    j->fn(j->tid, j->nthreads, j->arg);
    return NULL;
}
static void team_run(int nthreads, team_fn fn, void *arg) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    struct team_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].fn = fn; jobs[t].arg = arg; jobs[t].tid = t; jobs[t].nthreads = nthreads;
        if (t > 0) pthread_create(&th[t], NULL, team_entry, &jobs[t]);
    }
    fn(0, nthreads, arg);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* RankedMatrix: PerformanceSmokeValidation.cpp:144-170.  L (rows x r, salt */
/* 0x21) times R (r x cols, salt 0x22) mod ell; a plain random matrix (salt */
/* 0x11) when the rank target is 0 or >= min(rows, cols).  The product is   */
/* the unique canonical residue of the exact integer product.               */
/* ------------------------------------------------------------------------ */
HXO_API void hxo_gemm_fast(uint32_t p, uint32_t m, uint32_t n, uint32_t k,
                           const uint32_t *A, int64_t ars, int64_t acs,
                           const uint32_t *B, int64_t brs, int64_t bcs,
                           uint32_t *C, int64_t crs, int64_t ccs, int op, int nthreads);

HXO_API void hxo_ranked_matrix(uint32_t *out, uint32_t rows, uint32_t cols, uint32_t rank_target,
                               uint32_t ell, int nthreads) {
    const uint32_t full = rows < cols ? rows : cols;
    if (rank_target == 0 || rank_target >= full) {
        hxo_random_entries(out, (size_t)rows * cols, ell, hxo_seed_for(rows, cols, ell, 0x11));
        return;
    }
    const uint32_t r = rank_target;
    uint32_t *L = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)rows * r);
    uint32_t *R = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)r * cols);
    hxo_random_entries(L, (size_t)rows * r, ell, hxo_seed_for(rows, r, ell, 0x21));
    hxo_random_entries(R, (size_t)r * cols, ell, hxo_seed_for(r, cols, ell, 0x22));
    hxo_gemm_fast(ell, rows, cols, r, L, r, 1, R, cols, 1, out, cols, 1, 0, nthreads);
    free(L);
    free(R);
}

/* The rank-deficient LU input with forced swaps: RankedMatrix, then `nover`
 * rows overwritten by random linear combinations of four other rows.  The
 * rule is this build's own (the reference has no such generator): a SplitMix64
 * stream seeded SeedFor(rows, cols, ell, 0x31) draws, per overwritten row, the
 * target row (next % rows), then four (source row, coefficient) pairs
 * (next % rows, skipped to the next row when equal to the target; next % ell).
 * Rows are overwritten in order, so a later combination may read an earlier
 * overwritten row. */
HXO_API void hxo_overwrite_rows(uint32_t *A, uint32_t rows, uint32_t cols, uint32_t ell, uint32_t nover) {
    uint64_t st = hxo_seed_for(rows, cols, ell, 0x31);
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * cols);
    for (uint32_t t = 0; t < nover; ++t) {
        const uint32_t i = (uint32_t)(hxo_splitmix64(&st) % rows);
        memset(tmp, 0, sizeof(uint32_t) * cols);
        for (int s = 0; s < 4; ++s) {
            uint32_t j = (uint32_t)(hxo_splitmix64(&st) % rows);
            if (j == i) j = (j + 1) % rows;
            const uint32_t c = (uint32_t)(hxo_splitmix64(&st) % ell);
            const uint32_t *src = A + (size_t)j * cols;
            for (uint32_t q = 0; q < cols; ++q) tmp[q] = madd(tmp[q], mmul(c, src[q], ell), ell);
        }
        memcpy(A + (size_t)i * cols, tmp, sizeof(uint32_t) * cols);
    }
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* GEMM                                                                     */
/* ------------------------------------------------------------------------ */
/* Naive: RefLinAlg.cpp:94-106 (canonical after every product), combine     */
/* :84-90 (op 0 = Set, 1 = Add, 2 = Sub).  Pinned by: exact u128 sums on    */
/* small cases and the "rank-1 x all-(p-1)" closed form in the tests.       */
HXO_API void hxo_gemm_naive(uint32_t p, uint32_t m, uint32_t n, uint32_t k,
                            const uint32_t *A, int64_t lda, const uint32_t *B, int64_t ldb,
                            uint32_t *C, int64_t ldc, int op) {
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            uint32_t s = 0;
            for (uint32_t t = 0; t < k; ++t) s = madd(s, mmul(A[i * lda + t], B[t * ldb + j], p), p);
            uint32_t *c = &C[i * ldc + j];
            *c = op == 0 ? s : op == 1 ? madd(*c, s, p) : msub(*c, s, p);
        }
}

/* Fast: the reference's u32 route, Kernels.cpp:680-710 (k-blocks of
 * kc = min(256, k_max), exact u64 sums per block, one Barrett reduction per
 * block and output cell, blocks combined with mod_add, op into C), threaded
 * over row tiles of C as the tape executor splits output tiles
 * (Engine/Executor.cpp:69-100). */
struct gemm_arg {
    uint32_t p, m, n, k;
    const uint32_t *A; int64_t ars, acs;
    const uint32_t *B; int64_t brs, bcs;
    uint32_t *C; int64_t crs, ccs;
    int op;
};

#define GF_NB 256
static void gemm_fast_worker(int tid, int T, void *va) {
    const struct gemm_arg *g = (const struct gemm_arg *)va;
    const uint32_t p = g->p;
    const uint64_t bm = hxo_barrett_m(p);
    uint64_t km = hxo_k_max(p);
    const uint32_t kc = (uint32_t)(km < 256 ? km : 256);
    uint64_t acc[4][GF_NB];
    uint32_t res[4][GF_NB];
    uint32_t *brow = (uint32_t *)malloc(sizeof(uint32_t) * GF_NB);
    const uint32_t tiles = (g->m + 3) / 4;
    for (uint32_t tile = (uint32_t)tid; tile < tiles; tile += (uint32_t)T) {
        const uint32_t i0 = tile * 4, mr = g->m - i0 < 4 ? g->m - i0 : 4;
        for (uint32_t j0 = 0; j0 < g->n; j0 += GF_NB) {
            const uint32_t nb = g->n - j0 < GF_NB ? g->n - j0 : GF_NB;
            for (uint32_t r = 0; r < mr; ++r)
                for (uint32_t j = 0; j < nb; ++j) res[r][j] = 0;
            for (uint32_t k0 = 0; k0 < g->k; k0 += kc) {
                const uint32_t kb = g->k - k0 < kc ? g->k - k0 : kc;
                for (uint32_t r = 0; r < mr; ++r)
                    for (uint32_t j = 0; j < nb; ++j) acc[r][j] = 0;
                for (uint32_t t = k0; t < k0 + kb; ++t) {
                    const uint32_t *bp = g->B + (int64_t)t * g->brs + (int64_t)j0 * g->bcs;
                    if (g->bcs == 1) {
                        memcpy(brow, bp, sizeof(uint32_t) * nb);
                    } else {
                        for (uint32_t j = 0; j < nb; ++j) brow[j] = bp[(int64_t)j * g->bcs];
                    }
                    for (uint32_t r = 0; r < mr; ++r) {
                        const uint64_t a = g->A[(int64_t)(i0 + r) * g->ars + (int64_t)t * g->acs];
                        uint64_t *ac = acc[r];
                        for (uint32_t j = 0; j < nb; ++j) ac[j] += a * brow[j];
                    }
                }
                for (uint32_t r = 0; r < mr; ++r)
                    for (uint32_t j = 0; j < nb; ++j) res[r][j] = madd(res[r][j], red64(acc[r][j], p, bm), p);
            }
            for (uint32_t r = 0; r < mr; ++r)
                for (uint32_t j = 0; j < nb; ++j) {
                    uint32_t *c = g->C + (int64_t)(i0 + r) * g->crs + (int64_t)(j0 + j) * g->ccs;
                    *c = g->op == 0 ? res[r][j] : g->op == 1 ? madd(*c, res[r][j], p) : msub(*c, res[r][j], p);
                }
        }
    }
    free(brow);
}

HXO_API void hxo_gemm_fast(uint32_t p, uint32_t m, uint32_t n, uint32_t k,
                           const uint32_t *A, int64_t ars, int64_t acs,
                           const uint32_t *B, int64_t brs, int64_t bcs,
                           uint32_t *C, int64_t crs, int64_t ccs, int op, int nthreads) {
    struct gemm_arg g = {p, m, n, k, A, ars, acs, B, brs, bcs, C, crs, ccs, op};
    if (m == 0 || n == 0) return;
    team_run(nthreads, gemm_fast_worker, &g);
}

/* ------------------------------------------------------------------------ */
/* RREF mod p                                                               */
/* ------------------------------------------------------------------------ */
/* Legacy Gauss-Jordan: RefLinAlg.cpp:140-175 (leftmost column, topmost     */
/* non-zero row from r downward with a swap, leading 1, eliminate above and */
/* below).  In place on M (m x n, row stride ld); returns the rank and the  */
/* pivot columns.  Pinned by: identity / zero / planted-rank cases, the     */
/* RREF uniqueness property (RREF of a row-permuted input is equal), and    */
/* reference-written fixtures.                                              */
HXO_API uint32_t hxo_rref_naive(uint32_t p, uint32_t m, uint32_t n, uint32_t *M, int64_t ld, uint32_t *pivots) {
    uint32_t r = 0;
    for (uint32_t c = 0; c < n && r < m; ++c) {
        uint32_t piv = r;
        while (piv < m && M[piv * ld + c] == 0) ++piv;
        if (piv == m) continue;
        if (piv != r)
            for (uint32_t j = 0; j < n; ++j) {
                const uint32_t t = M[r * ld + j];
                M[r * ld + j] = M[piv * ld + j];
                M[piv * ld + j] = t;
            }
        const uint32_t inv = hxo_inv(M[r * ld + c], p);
        for (uint32_t j = 0; j < n; ++j) M[r * ld + j] = mmul(M[r * ld + j], inv, p);
        for (uint32_t i = 0; i < m; ++i) {
            if (i == r) continue;
            const uint32_t f = M[i * ld + c];
            if (f == 0) continue;
            for (uint32_t j = 0; j < n; ++j) M[i * ld + j] = msub(M[i * ld + j], mmul(f, M[r * ld + j], p), p);
        }
        pivots[r] = c;
        ++r;
    }
    return r;
}

/* The smallest-original-row rule: Cuda/CudaImmaElimValidation.cpp:146-173
 * (rule_reference) -- per column, the unused row of smallest index with a
 * non-zero Schur-complement entry; rows never move.  pivot_rows[t] is the row
 * chosen for pivot t (the row rank profile in pivot order, CudaKernels.h:
 * 255-265).  Then the RREF rows from those U rows (DenseElim.h:43-47: RREF =
 * U1^{-1} U): each U row is scaled to a leading 1 and the rows are cleared
 * above from the bottom.  R receives rank x n rows (row stride ldr); it may
 * be NULL (rank, pivots, pivot_rows only).  Threaded over the rows of each
 * column step.  Pinned by: pivot_rows set == the column rank profile of A^T
 * (computed with hxo_rref_naive on the transpose), RREF == hxo_rref_naive,
 * and a planted-order case where the largest-row rule must disagree. */
struct rule_arg {
    uint32_t p, m, n;
    uint32_t *W;             /* m x n working copy, row stride n */
    char *used;
    pthread_barrier_t *bar;
    volatile int32_t *chosen;      /* per column: the chosen row (-1 none), written by thread 0 */
    int32_t *cand;           /* per thread candidate */
    uint32_t *pivots, *pivot_rows;
    uint32_t rank;
};

static inline uint32_t shoup_pre(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) / p); }
static inline uint32_t shoup_mul(uint32_t a, uint32_t w, uint32_t ws, uint32_t p) {
    const uint32_t q = (uint32_t)(((uint64_t)a * ws) >> 32);
    uint32_t r = a * w - q * p;
    return r >= p ? r - p : r;
}

/* row[c:n] -= f * prow[c:n] mod p, p <= 2^31 (Shoup multiplication by the fixed f; fs =
 * shoup_pre(f, p)).  Built for AVX-512F / AVX2 / baseline x86-64, picked at run time: the
 * arithmetic is the same scalar formula in every clone. */
__attribute__((target_clones("avx512f", "avx2", "default")))
static void row_sub_shoup(uint32_t *restrict row, const uint32_t *restrict prow, uint32_t f, uint32_t fs,
                          uint32_t p, uint32_t c, uint32_t n) {
    for (uint32_t j = c; j < n; ++j) row[j] = msub(row[j], shoup_mul(prow[j], f, fs, p), p);
}

static void rule_worker(int tid, int T, void *va) {
    struct rule_arg *g = (struct rule_arg *)va;
    const uint32_t p = g->p, m = g->m, n = g->n;
    const int shoup = p <= 0x80000000u;
    uint32_t rank = 0;
    for (uint32_t c = 0; c < n; ++c) {
        /* each thread: the smallest unused row of its stripe with a non-zero at c */
        int32_t best = -1;
        for (uint32_t i = (uint32_t)tid; i < m; i += (uint32_t)T)
            if (!g->used[i] && g->W[(size_t)i * n + c] != 0) { best = (int32_t)i; break; }
        g->cand[tid] = best;
        pthread_barrier_wait(g->bar);
        if (tid == 0) {
            int32_t pr = -1;
            for (int t = 0; t < T; ++t)
                if (g->cand[t] >= 0 && (pr < 0 || g->cand[t] < pr)) pr = g->cand[t];
            g->chosen[0] = pr;
            if (pr >= 0) {
                g->used[pr] = 1;
                g->pivots[rank] = c;
                g->pivot_rows[rank] = (uint32_t)pr;
            }
        }
        pthread_barrier_wait(g->bar);
        const int32_t pr = g->chosen[0];
        if (pr < 0) continue;
        const uint32_t *prow = g->W + (size_t)pr * n;
        const uint32_t inv = hxo_inv(prow[c], p);
        for (uint32_t i = (uint32_t)tid; i < m; i += (uint32_t)T) {
            if (g->used[i]) continue;
            uint32_t *row = g->W + (size_t)i * n;
            if (row[c] == 0) continue;
            const uint32_t f = mmul(row[c], inv, p);
            if (shoup) {
                row_sub_shoup(row, prow, f, shoup_pre(f, p), p, c, n);
            } else {
                for (uint32_t j = c; j < n; ++j) row[j] = msub(row[j], mmul(f, prow[j], p), p);
            }
        }
        ++rank;
        pthread_barrier_wait(g->bar);
    }
    if (tid == 0) g->rank = rank;
}

HXO_API uint32_t hxo_rref_rule(uint32_t p, uint32_t m, uint32_t n, const uint32_t *A, int64_t lda,
                               uint32_t *pivots, uint32_t *pivot_rows, uint32_t *R, int64_t ldr, int nthreads) {
    if (m == 0 || n == 0) return 0;
    if (nthreads < 1) nthreads = 1;
    uint32_t *W = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)m * n);
    for (uint32_t i = 0; i < m; ++i) memcpy(W + (size_t)i * n, A + (size_t)i * lda, sizeof(uint32_t) * n);
    char *used = (char *)calloc(m, 1);
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nthreads);
    int32_t chosen = -1;
    int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)nthreads);
    struct rule_arg g = {p, m, n, W, used, &bar, &chosen, cand, pivots, pivot_rows, 0};
    team_run(nthreads, rule_worker, &g);
    pthread_barrier_destroy(&bar);
    const uint32_t rank = g.rank;
    if (R != NULL && rank > 0) {
        /* U rows in pivot order, scaled to a leading 1 */
        for (uint32_t t = 0; t < rank; ++t) {
            const uint32_t *u = W + (size_t)pivot_rows[t] * n;
            uint32_t *dst = R + (size_t)t * ldr;
            const uint32_t inv = hxo_inv(u[pivots[t]], p);
            for (uint32_t j = 0; j < n; ++j) dst[j] = j < pivots[t] ? 0 : mmul(u[j], inv, p);
        }
        /* clear above every pivot, bottom up */
        for (uint32_t t = rank; t-- > 0;) {
            const uint32_t c = pivots[t];
            const uint32_t *rt = R + (size_t)t * ldr;
            for (uint32_t s = 0; s < t; ++s) {
                uint32_t *rs = R + (size_t)s * ldr;
                const uint32_t f = rs[c];
                if (f == 0) continue;
                if (p <= 0x80000000u)
                    row_sub_shoup(rs, rt, f, shoup_pre(f, p), p, c, n);
                else
                    for (uint32_t j = c; j < n; ++j) rs[j] = msub(rs[j], mmul(f, rt[j], p), p);
            }
        }
    }
    free(cand);
    free(used);
    free(W);
    return rank;
}

/* ------------------------------------------------------------------------ */
/* Howell form over Z/2^k, 1 <= k <= 32                                     */
/* ------------------------------------------------------------------------ */
/* Naive: RefLinAlg.cpp:231-290 (ENGINE_API.md Sec. 13.5's pass on A        */
/* itself: minimal valuation, ties to the topmost active row; scale by the  */
/* inverse of the unit part; exact elimination; closure row 2^(k-v) * P at  */
/* the bottom of the active list when v > 0; then reduce above, pivot       */
/* columns left to right).  Arithmetic in u64 modulo 2^k.  H receives at    */
/* most n rows of n entries (row stride n); returns h.  Pinned by: k = 1    */
/* equals the RREF without zero rows, howell([[2,1]]) over Z/4 =            */
/* [[2,1],[0,2]] (ENGINE_API.md:1939-1942), row-permutation invariance and  */
/* reference-written fixtures.                                              */
static inline uint32_t val2(uint64_t x, uint32_t k) {
    if (x == 0) return k;
    uint32_t v = (uint32_t)__builtin_ctzll(x);
    return v < k ? v : k;
}
static inline uint64_t inv_odd64(uint64_t u) {    /* u^-1 mod 2^64, Newton from 3 correct bits */
    uint64_t x = u;
    for (int i = 0; i < 5; ++i) x *= 2 - u * x;
    return x;
}

HXO_API uint32_t hxo_howell_naive(uint32_t k, uint32_t m, uint32_t n, const uint32_t *A, int64_t lda,
                                  uint32_t *H, uint32_t *pivots, uint32_t *pivot_v) {
    const uint64_t mask = k >= 64 ? ~0ull : ((1ull << k) - 1);
    const uint32_t cap = m + n + 1;
    uint64_t *rows = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cap * (n ? n : 1));
    uint32_t *act = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    uint32_t nact = 0, nrows = 0;
    for (uint32_t i = 0; i < m; ++i) {
        for (uint32_t j = 0; j < n; ++j) rows[(size_t)nrows * n + j] = A[i * lda + j] & mask;
        act[nact++] = nrows++;
    }
    uint32_t *prow = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t h = 0;
    for (uint32_t c = 0; c < n; ++c) {
        uint32_t best = nact, bv = k;
        for (uint32_t t = 0; t < nact; ++t) {
            const uint64_t x = rows[(size_t)act[t] * n + c];
            if (x == 0) continue;
            const uint32_t v = val2(x, k);
            if (v < bv) { bv = v; best = t; }
        }
        if (best == nact) continue;
        const uint32_t pid = act[best];
        for (uint32_t t = best; t + 1 < nact; ++t) act[t] = act[t + 1];
        --nact;
        uint64_t *P = rows + (size_t)pid * n;
        const uint64_t ui = inv_odd64(P[c] >> bv);
        for (uint32_t j = 0; j < n; ++j) P[j] = (P[j] * ui) & mask;
        const uint64_t lead = 1ull << bv;
        for (uint32_t t = 0; t < nact; ++t) {
            uint64_t *r = rows + (size_t)act[t] * n;
            if (r[c] == 0) continue;
            const uint64_t q = r[c] / lead;
            for (uint32_t j = 0; j < n; ++j) r[j] = (r[j] - q * P[j]) & mask;
        }
        if (bv > 0) {
            uint64_t *cl = rows + (size_t)nrows * n;
            const uint64_t f = 1ull << (k - bv);
            for (uint32_t j = 0; j < n; ++j) cl[j] = (f * P[j]) & mask;
            act[nact++] = nrows++;
        }
        prow[h] = pid;
        pivots[h] = c;
        pivot_v[h] = bv;
        ++h;
    }
    /* reduce above, pivot columns left to right */
    for (uint32_t i = 0; i < h; ++i) {
        const uint64_t *ri = rows + (size_t)prow[i] * n;
        const uint64_t d = 1ull << pivot_v[i];
        for (uint32_t ip = 0; ip < i; ++ip) {
            uint64_t *rp = rows + (size_t)prow[ip] * n;
            const uint64_t q = rp[pivots[i]] / d;
            if (q == 0) continue;
            for (uint32_t j = 0; j < n; ++j) rp[j] = (rp[j] - q * ri[j]) & mask;
        }
    }
    for (uint32_t i = 0; i < h; ++i)
        for (uint32_t j = 0; j < n; ++j) H[(size_t)i * n + j] = (uint32_t)rows[(size_t)prow[i] * n + j];
    free(prow);
    free(act);
    free(rows);
    return h;
}

/* Fast: Binary/Ring2k.h:959-969 (howell) = echelon_blocked with the
 * MinValuation rule (:538-613, the serial T <= 1 path: per column the row of
 * minimal valuation from the current row down, ties to the topmost, early
 * stop at v = 0; swap; eliminate below with f = (x >> v) * u^-1 in
 * wraparound) followed by howell_pass (:852-956) on the masked echelon rows.
 * Native u32 wraparound throughout; the output is masked to k bits.  The
 * panel / reduce blocking of the reference does not change a value (the
 * Howell form is unique, ENGINE_API.md Sec. 13.5), so it is run unblocked. */
static inline uint32_t inv_odd32(uint32_t u) {
    uint32_t x = u;
    for (uint32_t b = 3; b < 32; b *= 2) x = x * (2u - u * x);
    return x;
}
static inline uint32_t val32(uint32_t x, uint32_t k) {
    const uint32_t mk = k >= 32 ? 0xffffffffu : ((1u << k) - 1);
    x &= mk;
    if (x == 0) return k;
    const uint32_t c = (uint32_t)__builtin_ctz(x);
    return c < k ? c : k;
}

HXO_API uint32_t hxo_howell_fast(uint32_t k, uint32_t m, uint32_t n, const uint32_t *A, int64_t lda,
                                 uint32_t *H, uint32_t *pivots, uint32_t *pivot_v) {
    const uint32_t mk = k >= 32 ? 0xffffffffu : ((1u << k) - 1);
    if (n == 0) return 0;
    uint32_t *M = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(m ? m : 1) * n);
    for (uint32_t i = 0; i < m; ++i) memcpy(M + (size_t)i * n, A + i * lda, sizeof(uint32_t) * n);
    /* 1. minimal-valuation LU echelon (rowspan unchanged) */
    uint32_t r = 0;
    for (uint32_t j = 0; j < n && r < m; ++j) {
        uint32_t best = m, bv = k;
        for (uint32_t i = r; i < m; ++i) {
            const uint32_t v = val32(M[(size_t)i * n + j], k);
            if (v < bv) { bv = v; best = i; if (v == 0) break; }
        }
        if (best == m) continue;
        if (best != r)
            for (uint32_t q = 0; q < n; ++q) {
                const uint32_t t = M[(size_t)best * n + q];
                M[(size_t)best * n + q] = M[(size_t)r * n + q];
                M[(size_t)r * n + q] = t;
            }
        const uint32_t *P = M + (size_t)r * n;
        const uint32_t ui = inv_odd32(P[j] >> bv);
        for (uint32_t i = r + 1; i < m; ++i) {
            uint32_t *R = M + (size_t)i * n;
            if ((R[j] & mk) == 0) continue;
            const uint32_t f = (R[j] >> bv) * ui;
            for (uint32_t q = j; q < n; ++q) R[q] -= f * P[q];
        }
        ++r;
    }
    /* 2. the Sec. 13.5 pass on the r echelon rows (closure rows on demand) */
    const uint32_t cap = r + n + 1;
    uint32_t *rows = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cap * n);
    for (uint32_t i = 0; i < r; ++i)
        for (uint32_t q = 0; q < n; ++q) rows[(size_t)i * n + q] = M[(size_t)i * n + q] & mk;
    free(M);
    uint32_t *act = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    uint32_t *prow = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t nact = 0, nrows = r, h = 0;
    for (uint32_t i = 0; i < r; ++i) act[nact++] = i;
    for (uint32_t c = 0; c < n && nact > 0; ++c) {
        uint32_t best = nact, bv = k;
        for (uint32_t t = 0; t < nact; ++t) {
            const uint32_t v = val32(rows[(size_t)act[t] * n + c], k);
            if (v < bv) { bv = v; best = t; if (v == 0) break; }
        }
        if (best == nact) continue;
        const uint32_t pid = act[best];
        for (uint32_t t = best; t + 1 < nact; ++t) act[t] = act[t + 1];
        --nact;
        uint32_t *P = rows + (size_t)pid * n;
        const uint32_t ui = inv_odd32(P[c] >> bv);
        for (uint32_t q = c; q < n; ++q) P[q] = (P[q] * ui) & mk;
        for (uint32_t t = 0; t < nact; ++t) {
            uint32_t *R = rows + (size_t)act[t] * n;
            if ((R[c] & mk) == 0) continue;
            const uint32_t f = (R[c] & mk) >> bv;
            for (uint32_t q = c; q < n; ++q) R[q] -= f * P[q];
        }
        if (bv > 0) {
            uint32_t *Z = rows + (size_t)nrows * n;
            for (uint32_t q = 0; q <= c; ++q) Z[q] = 0;
            for (uint32_t q = c + 1; q < n; ++q) Z[q] = (k - bv) >= 32 ? 0 : (P[q] << (k - bv));
            act[nact++] = nrows++;
        }
        prow[h] = pid;
        pivots[h] = c;
        pivot_v[h] = bv;
        ++h;
    }
    for (uint32_t i = 0; i < h; ++i)
        for (uint32_t q = 0; q < n; ++q) H[(size_t)i * n + q] = rows[(size_t)prow[i] * n + q] & mk;
    /* 3. reduce above, pivots left to right (the unblocked form of :915-948) */
    for (uint32_t t = 1; t < h; ++t) {
        const uint32_t ct = pivots[t];
        const uint32_t *Ht = H + (size_t)t * n;
        for (uint32_t i = 0; i < t; ++i) {
            uint32_t *Hi = H + (size_t)i * n;
            const uint32_t q = (Hi[ct] & mk) >> pivot_v[t];
            if (q == 0) continue;
            for (uint32_t j = ct; j < n; ++j) Hi[j] -= q * Ht[j];
        }
    }
    for (size_t i = 0; i < (size_t)h * n; ++i) H[i] &= mk;
    free(prow);
    free(act);
    free(rows);
    return h;
}

/* ------------------------------------------------------------------------ */
/* Blocked LU over Z/2^k on the native u32 / u64 rungs                      */
/* ------------------------------------------------------------------------ */
/* Restates ring2k::echelon_blocked's serial route (Binary/Ring2k.h:538-612 */
/* and :667-683, the MinValuation rule) with NativeTraits' arithmetic      */
/* (:160-238): entries are native wraparound words understood mod 2^k;      */
/* per panel [c, c1) and column j the pivot is the row of least valuation  */
/* mod 2^k from row r + np down (the first such row; a unit stops the       */
/* search), swapped into place with its L row; every row below it gets      */
/* slot f = (R[j] >> v) u^-1 (u the pivot's odd part; slot 0 when R[j] is   */
/* 0 mod 2^k) and R[j, c1) -= f P[j, c1); then the unit-lower TRSM on the   */
/* panel's pivot rows and C -= L U on the rows below, over columns [c1, n). */
/* The device session's host driver must leave A, rank, pivots and          */
/* pivot_v exactly so (Ring2kDeviceLU.h's contract).  Returns the rank;     */
/* pivots / pivot_v receive min(m, n) entries at most.  Pinned by the       */
/* reference's own echelon_blocked (fixture written with oracle/_ref), and  */
/* by the invariant that the Howell form of A mod 2^k is unchanged.          */
#define HXO_R2_ECHELON(NAME, E, BITS)                                                                         \
    static inline E NAME##_inv_odd(E u) {                                                                     \
        E x = u;                                                                                              \
        for (uint32_t b = 3; b < BITS; b *= 2) x = (E)(x * (E)((E)2 - (E)(u * x)));                          \
        return x;                                                                                             \
    }                                                                                                         \
    static inline uint32_t NAME##_val(E a, uint32_t k) {                                                      \
        const E mk = k >= BITS ? (E)~(E)0 : (E)(((E)1 << k) - 1);                                             \
        const E x = a & mk;                                                                                   \
        if (x == 0) return k;                                                                                 \
        const uint32_t c = (uint32_t)__builtin_ctzll((uint64_t)x);                                           \
        return c < k ? c : k;                                                                                 \
    }                                                                                                         \
    HXO_API uint32_t NAME(uint32_t k, uint32_t m, uint32_t n, E *A, int64_t lda, uint32_t panel, uint32_t climit, \
                          uint32_t *pivots, uint32_t *pivot_v) {                                              \
        if (panel == 0) panel = 1;                                                                            \
        if (m == 0 || n == 0) return 0;                                                                       \
        if (climit > n) climit = n;                                                                           \
        E *L = (E *)calloc((size_t)m * panel, sizeof(E));                                                     \
        uint32_t r = 0, c = 0, npiv = 0;                                                                      \
        while (c < climit && r < m) {                                                                         \
            const uint32_t c1 = c + panel < climit ? c + panel : climit;                                      \
            uint32_t np = 0;                                                                                  \
            for (uint32_t j = c; j < c1 && r + np < m; ++j) {                                                 \
                const uint32_t prow = r + np;                                                                 \
                uint32_t best = m, bv = k;                                                                    \
                for (uint32_t i = prow; i < m; ++i) {                                                         \
                    const uint32_t v = NAME##_val(A[(int64_t)i * lda + j], k);                                \
                    if (v < bv) { bv = v; best = i; if (v == 0) break; }                                      \
                }                                                                                             \
                if (best == m) continue;                                                                      \
                if (best != prow) {                                                                           \
                    for (uint32_t q = 0; q < n; ++q) {                                                        \
                        E t = A[(int64_t)best * lda + q];                                                     \
                        A[(int64_t)best * lda + q] = A[(int64_t)prow * lda + q];                              \
                        A[(int64_t)prow * lda + q] = t;                                                       \
                    }                                                                                         \
                    for (uint32_t t = 0; t < panel; ++t) {                                                    \
                        E x = L[(size_t)best * panel + t];                                                    \
                        L[(size_t)best * panel + t] = L[(size_t)prow * panel + t];                            \
                        L[(size_t)prow * panel + t] = x;                                                      \
                    }                                                                                         \
                }                                                                                             \
                const E *P = A + (int64_t)prow * lda;                                                         \
                const E ui = NAME##_inv_odd((E)(P[j] >> bv));                                                 \
                for (uint32_t i = prow + 1; i < m; ++i) {                                                     \
                    E *R = A + (int64_t)i * lda;                                                              \
                    E *slot = L + (size_t)i * panel + np;                                                     \
                    if (NAME##_val(R[j], k) == k) { *slot = 0; continue; }                                    \
                    const E f = (E)((E)(R[j] >> bv) * ui);                                                    \
                    *slot = f;                                                                                \
                    for (uint32_t q = j; q < c1; ++q) R[q] = (E)(R[q] - (E)(f * P[q]));                       \
                }                                                                                             \
                pivots[npiv] = j;                                                                             \
                pivot_v[npiv] = bv;                                                                           \
                ++npiv;                                                                                       \
                ++np;                                                                                         \
            }                                                                                                 \
            if (np > 0 && c1 < n) {                                                                           \
                for (uint32_t t = 1; t < np; ++t) {                                                           \
                    const E *lt = L + (size_t)(r + t) * panel;                                                \
                    E *Rt = A + (int64_t)(r + t) * lda;                                                       \
                    for (uint32_t t2 = 0; t2 < t; ++t2) {                                                     \
                        const E f = lt[t2];                                                                   \
                        if (f == 0) continue;                                                                 \
                        const E *Rs = A + (int64_t)(r + t2) * lda;                                            \
                        for (uint32_t q = c1; q < n; ++q) Rt[q] = (E)(Rt[q] - (E)(f * Rs[q]));               \
                    }                                                                                         \
                }                                                                                             \
                for (uint32_t i = r + np; i < m; ++i) {                                                       \
                    E *Ci = A + (int64_t)i * lda;                                                             \
                    const E *li = L + (size_t)i * panel;                                                      \
                    for (uint32_t t = 0; t < np; ++t) {                                                       \
                        const E f = li[t];                                                                    \
                        if (f == 0) continue;                                                                 \
                        const E *Ut = A + (int64_t)(r + t) * lda;                                             \
                        for (uint32_t q = c1; q < n; ++q) Ci[q] = (E)(Ci[q] - (E)(f * Ut[q]));               \
                    }                                                                                         \
                }                                                                                             \
            }                                                                                                 \
            r += np;                                                                                          \
            c = c1;                                                                                           \
        }                                                                                                     \
        free(L);                                                                                              \
        return r;                                                                                             \
    }
HXO_R2_ECHELON(hxo_r2_echelon32, uint32_t, 32u)
HXO_R2_ECHELON(hxo_r2_echelon64, uint64_t, 64u)

/* ------------------------------------------------------------------------ */
/* GEMM mod p on 64-bit residues, 2 <= p < 2^64                             */
/* ------------------------------------------------------------------------ */
/* C op= A B mod p with every product and sum exact in unsigned __int128     */
/* (each product reduced before the sum, as RefLinAlg.cpp:94-106 does for   */
/* the u32 naive GEMM).  Threaded over rows.  Pinned by exact Python        */
/* integers, the rank-1 x all-(p-1) closed form, and the reference's own    */
/* kernels::gemm<uint64_t> (fixture written with oracle/_ref).               */
struct gemm64_arg {
    uint64_t p;
    uint32_t m, n, k;
    const uint64_t *A, *B;
    uint64_t *C;
    int op;
};
static void gemm64_worker(int tid, int T, void *va) {
    const struct gemm64_arg *g = (const struct gemm64_arg *)va;
    const uint64_t p = g->p;
    for (uint32_t i = (uint32_t)tid; i < g->m; i += (uint32_t)T)
        for (uint32_t j = 0; j < g->n; ++j) {
            uint64_t s = 0;
            for (uint32_t t = 0; t < g->k; ++t) {
                const uint64_t x = (uint64_t)(((unsigned __int128)g->A[(size_t)i * g->k + t] * g->B[(size_t)t * g->n + j]) % p);
                s = (uint64_t)(((unsigned __int128)s + x) % p);
            }
            uint64_t *c = g->C + (size_t)i * g->n + j;
            if (g->op == 0) *c = s;
            else if (g->op == 1) *c = (uint64_t)(((unsigned __int128)*c + s) % p);
            else *c = (uint64_t)(((unsigned __int128)*c + (p - s)) % p);
        }
}
HXO_API void hxo_gemm_u64(uint64_t p, uint32_t m, uint32_t n, uint32_t k, const uint64_t *A, const uint64_t *B,
                          uint64_t *C, int op, int nthreads) {
    struct gemm64_arg g = {p, m, n, k, A, B, C, op};
    if (m == 0 || n == 0) return;
    team_run(nthreads, gemm64_worker, &g);
}

/* ------------------------------------------------------------------------- */
/* Smith valuations over Z/2^k by global minimal-valuation pivoting: the     */
/* reference's smith_naive (Binary/Ring2k.h:1004-1035, "the prototype's      */
/* smith mode").  A: m x n words (row-major, consumed), values read mod 2^k   */
/* (k <= 64; the reference's u32 / u64 rungs agree mod 2^k).  sigma: the     */
/* valuations in pivot order (ascending).  Returns their count.               */
static uint64_t sm_inv_odd(uint64_t u) { /* u^-1 mod 2^64, Newton from 3 bits */
    uint64_t x = u;
    for (int i = 0; i < 6; ++i) x *= 2u - u * x;
    return x;
}
HXO_API int64_t hxo_smith_naive(uint32_t k, uint32_t m, uint32_t n, uint64_t *A, uint32_t *sigma) {
    const uint64_t mask = k >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << k) - 1);
#define SM_AT(i, j) A[(size_t)(i) * n + (j)]
    int64_t cnt = 0;
    for (uint32_t r = 0; r < m && r < n; ++r) {
        uint32_t bi = m, bj = n, bv = k;
        for (uint32_t i = r; i < m && bv > 0; ++i)
            for (uint32_t j = r; j < n; ++j) {
                const uint64_t x = SM_AT(i, j) & mask;
                const uint32_t v = x == 0 ? k : (uint32_t)__builtin_ctzll(x);
                if (v < bv) {
                    bv = v;
                    bi = i;
                    bj = j;
                    if (v == 0) break;
                }
            }
        if (bi == m) break;
        if (bi != r)
            for (uint32_t j = 0; j < n; ++j) {
                const uint64_t t = SM_AT(bi, j);
                SM_AT(bi, j) = SM_AT(r, j);
                SM_AT(r, j) = t;
            }
        if (bj != r)
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t t = SM_AT(i, bj);
                SM_AT(i, bj) = SM_AT(i, r);
                SM_AT(i, r) = t;
            }
        const uint64_t ui = sm_inv_odd((SM_AT(r, r) & mask) >> bv);
        for (uint32_t i = r + 1; i < m; ++i) {
            const uint64_t x = SM_AT(i, r) & mask;
            if (x == 0) continue;
            const uint64_t f = (x >> bv) * ui;
            for (uint32_t j = r; j < n; ++j) SM_AT(i, j) -= f * SM_AT(r, j);
        }
        sigma[cnt++] = bv;
    }
#undef SM_AT
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Howell form over Z/ell^k, ell prime, q = ell^k < 2^32: elim::howell      */
/* (DenseElim.cpp:1450 -> howell_rows :1130-1300; ModCtx's valuation,       */
/* unit_part, divide_exact, ell_pow, ModCtx.h:283-312).  The Sec. 13.5      */
/* pass on the rows of A: columns left to right; among the active rows the  */
/* entry of minimal ell-adic valuation, ties to the topmost (active order = */
/* the input rows, then closure rows as created; the chosen row leaves the  */
/* list); the pivot row scaled by the inverse of its unit part (it leads    */
/* with ell^v); every other active row r -= (r[c] / ell^v) P (exact         */
/* division); for v > 0 the closure row ell^(k-v) P, zero up to column c,   */
/* joins the active list at the bottom.  Then reduce above, pivots left to  */
/* right: H[i'] -= floor(H[i'][c_i] / ell^v_i) H[i].  Entries canonical in  */
/* [0, q).  Returns h; H is h x n (ld n), pivots / pivot_v h entries.       */
/* Pinned by: the reference's own elim::howell on the committed fixture     */
/* (tests/golden/howell_ell_reference.npz); at ell = 2 the Z/2^k oracles;   */
/* at k = 1 the RREF over F_ell (hxo_rref_rule); [[ell, 1]] over Z/ell^2 =  */
/* [[ell, 1], [0, ell]]; on tiny matrices over Z/9 the row span of H equals */
/* the row span of A, enumerated.                                           */
/* ------------------------------------------------------------------------ */
static uint32_t ell_val(uint64_t x, uint32_t ell, uint32_t k) {
    if (x == 0) return k;
    uint32_t v = 0;
    while (v < k && x % ell == 0) { x /= ell; ++v; }
    return v;
}

static uint64_t ell_inv(uint64_t a, uint64_t q) {   /* a^-1 mod q for gcd(a, q) = 1 (extended Euclid) */
    int64_t r0 = (int64_t)q, r1 = (int64_t)(a % q), s0 = 0, s1 = 1;
    while (r1 != 0) {
        const int64_t t = r0 / r1, r2 = r0 - t * r1, s2 = s0 - t * s1;
        r0 = r1; r1 = r2; s0 = s1; s1 = s2;
    }
    int64_t s = s0 % (int64_t)q;
    if (s < 0) s += (int64_t)q;
    return (uint64_t)s;
}

/* row[c0..n) -= f * src[c0..n) mod q */
static void ell_row_sub(uint64_t *row, const uint64_t *src, uint64_t f, uint32_t c0, uint32_t n, uint64_t q) {
    if (f == 0) return;
    for (uint32_t j = c0; j < n; ++j)
        if (src[j] != 0) row[j] = (row[j] + q - (f * src[j]) % q) % q;
}

HXO_API uint32_t hxo_howell_ell(uint32_t ell, uint32_t k, uint32_t m, uint32_t n, const uint32_t *A, int64_t lda,
                                uint32_t *H, uint32_t *pivots, uint32_t *pivot_v) {
    uint64_t pw[65];
    pw[0] = 1;
    for (uint32_t v = 1; v <= k && v < 65; ++v) pw[v] = pw[v - 1] * ell;
    const uint64_t q = pw[k];
    const uint32_t cap = m + n + 1;
    uint64_t *rows = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cap * (n ? n : 1));
    uint32_t *act = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    uint32_t *prow = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t nact = 0, nrows = 0, h = 0;
    for (uint32_t i = 0; i < m; ++i) {
        for (uint32_t j = 0; j < n; ++j) rows[(size_t)nrows * n + j] = A[i * lda + j] % q;
        act[nact++] = nrows++;
    }
    for (uint32_t c = 0; c < n && nact > 0; ++c) {
        uint32_t best = nact, bv = k;
        for (uint32_t t = 0; t < nact; ++t) {
            const uint64_t e = rows[(size_t)act[t] * n + c];
            if (e == 0) continue;
            const uint32_t v = ell_val(e, ell, k);
            if (v < bv) { bv = v; best = t; if (v == 0) break; }
        }
        if (best == nact) continue;
        const uint32_t pid = act[best];
        memmove(act + best, act + best + 1, sizeof(uint32_t) * (nact - best - 1));
        --nact;
        uint64_t *P = rows + (size_t)pid * n;
        uint64_t u = P[c];
        while (u % ell == 0) u /= ell;
        if (u != 1) {
            const uint64_t ui = ell_inv(u, q);
            for (uint32_t j = c; j < n; ++j) P[j] = (P[j] * ui) % q;
        }
        const uint64_t lead = pw[bv];
        for (uint32_t t = 0; t < nact; ++t) {
            uint64_t *R = rows + (size_t)act[t] * n;
            if (R[c] == 0) continue;
            ell_row_sub(R, P, R[c] / lead, c, n, q);
        }
        if (bv > 0) {
            const uint64_t f = pw[k - bv];
            uint64_t *Z = rows + (size_t)nrows * n;
            P = rows + (size_t)pid * n;
            for (uint32_t j = 0; j < n; ++j) Z[j] = j <= c ? 0 : (P[j] * f) % q;
            act[nact++] = nrows++;
        }
        prow[h] = pid;
        pivots[h] = c;
        pivot_v[h] = bv;
        ++h;
    }
    for (uint32_t i = 0; i < h; ++i)
        for (uint32_t j = 0; j < n; ++j) H[(size_t)i * n + j] = (uint32_t)rows[(size_t)prow[i] * n + j];
    uint64_t *hi = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint64_t *hp = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < h; ++i) {
        const uint32_t ji = pivots[i];
        const uint64_t d = pw[pivot_v[i]];
        for (uint32_t j = 0; j < n; ++j) hi[j] = H[(size_t)i * n + j];
        for (uint32_t ip = 0; ip < i; ++ip) {
            for (uint32_t j = 0; j < n; ++j) hp[j] = H[(size_t)ip * n + j];
            ell_row_sub(hp, hi, hp[ji] / d, ji, n, q);
            for (uint32_t j = 0; j < n; ++j) H[(size_t)ip * n + j] = (uint32_t)hp[j];
        }
    }
    free(hi);
    free(hp);
    free(rows);
    free(act);
    free(prow);
    return h;
}
