/*
 * hacs_mi355x.h -- C-ABI of the MI355X (gfx950) device backend of HACS's
 * exact matrix engine.
 *
 * The library (libhacs_mi355x.so) replaces the device side that the reference
 * reaches through src/Cuda (namespace engine::cuda): the runtime context and
 * buffers (Cuda/CudaCtx.h, Cuda/DeviceBuf.h), the dense kernels
 * (Cuda/CudaKernels.h), the device elimination (Cuda/ImmaElim.cuh through
 * CudaKernels.h:249-298) and the Z/2^k device work (Cuda/CudaRing2k.h).  Each
 * declaration below cites the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative HX_E* code; the text
 *     of the last error of the calling thread is hx_last_error().  This is the
 *     C form of the reference's engine::cuda::CudaError / ENGINE_CHECK
 *     (CudaCtx.h:40-54, cuda_kernels.cu:384-394); a C++ facade rethrows it.
 *   - Device pointers are plain pointers into HBM; hx_stream is a hipStream_t
 *     (NULL = the null stream).  Launches are asynchronous on the stream
 *     unless a function says it synchronises.
 *   - All arithmetic is exact integer arithmetic on canonical residues in
 *     [0, p); results are bit-identical to the reference CPU engine.
 *   - No torch type appears in any signature.
 */
#ifndef HACS_MI355X_H
#define HACS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HX_API __attribute__((visibility("default")))

/* ---- status codes -------------------------------------------------------- */
enum {
    HX_OK = 0,
    HX_E_RUNTIME = -1,     /* a HIP runtime / launch error (CudaError) */
    HX_E_ARGUMENT = -2,    /* a contract breach: shape, stride, range (ENGINE_CHECK) */
    HX_E_UNAVAILABLE = -3, /* no usable device (available() == false) */
    HX_E_NOMEM = -4,       /* device allocation failed */
    HX_E_NOT_INVERTIBLE = -5, /* a pivot had no inverse mod p (p not prime) */
    HX_E_CAPTURE = -6      /* the call is not legal inside a stream capture */
};

/* ---- opaque handles (CudaCtx.h:63-64 Stream / Event) --------------------- */
typedef void *hx_stream;   /* hipStream_t */
typedef void *hx_event;    /* hipEvent_t */
typedef void *hx_graph;    /* an instantiated hipGraphExec_t */

/* ---- the modulus (ModCtx.h:123-250, the u32 fields the kernels read) ----- */
typedef struct hx_modctx {
    uint32_t p;          /* modulus, 2 <= p < 2^32 */
    uint32_t limbs;      /* signed 8-bit limbs of a balanced residue (1..4); 0 = too wide for the limb route */
    uint64_t barrett_m;  /* floor(2^64 / p)                          (ModCtx.h:219) */
    uint64_t k_max;      /* floor((2^64 - 1) / (p - 1)^2)            (ModCtx.h:220-222) */
    uint64_t two64_mod;  /* 2^64 mod p */
} hx_modctx;

/* ---- a strided device matrix view (DeviceBuf.h:50-91 DeviceView) -------- */
/* element (i, j) lives at base[i * rs + j * cs]; rs, cs in elements. */
typedef struct hx_view {
    void *base;
    int64_t rs, cs;
    uint32_t rows, cols;
} hx_view;

/* GemmOp (ModArithmetic/Kernels.h, reference ref::gemm encoding RefLinAlg.h:37-39) */
enum { HX_OP_SET = 0, HX_OP_ADD = 1, HX_OP_SUB = 2 };

/* GEMM kernel variants (no counterpart: the reference has one u32 kernel and the p <= 127 IMMA route) */
enum {
    HX_GEMM_AUTO = 0,   /* the library picks: LIMB when m*n*k >= 4e6 and k >= 64, else PLAIN (never changes a value) */
    HX_GEMM_PLAIN = 1,  /* u32 x u32 -> exact u64 (+carry word) accumulation on the VALU, Barrett epilogue */
    HX_GEMM_LIMB = 2    /* balanced residues as signed 8-bit limbs on the INT8 MFMA cores, exact recombination */
};

typedef struct hx_gemm_opts {   /* GemmOpts, CudaKernels.h:64-79 */
    int variant;            /* HX_GEMM_* */
    int tile;               /* 0 = automatic; a variant-specific choice (tuning only, same bytes): for LIMB,
                               1 = register-staged LDS copy, 2 = direct-to-LDS copy ring,
                               3 = the ring with LDS reads pipelined against the MFMAs, 4 = the ring on eight
                               waves, one 32 x 32 block each (the default), 5 = 96 x 128 workgroup tiles on
                               twelve waves, 6 = the residue route: N <= 12 int8 products modulo small
                               coprime moduli and a CRT recombination (N = 10 for p < 2^32, k <= 8192;
                               host-side sizes and k <= 133144 only, else the limb kernels; its workspace
                               is hx_gemm_workspace_bytes_tile's, all of it: K is not chunked).  0 takes
                               the residue route for large four-limb products when its workspace fits
                               (hx_gemm_route), else kernel 4 */
    void *workspace;        /* the limb route's planes (the reference's imma_workspace, CudaKernels.h:72-75):
                               hx_gemm_workspace_bytes() runs K in one piece, less (at least the bytes of one
                               64-deep K chunk, hx_gemm_limb_k_chunk) runs K in chunks, same bytes; smaller: AUTO
                               runs the plain kernel, LIMB fails (HX_E_ARGUMENT).  NULL = library-owned planes
                               private to the stream (grown outside a capture only: inside one, AUTO falls back
                               to the plain kernel when they are too small and LIMB fails with HX_E_CAPTURE) */
    size_t workspace_bytes;
    hx_event main_begin;    /* optional timing events recorded on the stream right before / after the */
    hx_event main_end;      /* main product kernel (instrumentation only; NULL = none) */
    uint32_t splits;        /* split-K of the plain kernel (GemmOpts::splits, CudaKernels.h:64-66): 0 = automatic
                               (hx_gemm_suggest_splits), 1 = never, 2 .. HX_MAX_SPLITS = that many K slices (fewer
                               when k is short).  The limb route never splits.  Same bytes either way. */
    void *split_workspace;  /* the partial sums (GemmOpts::workspace): >= hx_gemm_split_workspace_bytes(); NULL =
                               library-owned partials private to the stream (grown outside a capture only: inside
                               one, automatic splitting is skipped when they are too small and an explicit
                               count fails with HX_E_CAPTURE) */
    size_t split_workspace_bytes;
} hx_gemm_opts;

#define HX_MAX_SPLITS 64

typedef struct hx_device_info {   /* CudaCtx.h:55-64 DeviceInfo */
    char name[128];
    char arch[64];
    int index;
    int compute_units;
    int clock_khz;
    size_t total_mem, free_mem;
    int runtime_version, driver_version;
} hx_device_info;

/* ---- availability, errors (CudaCtx.h:56-62) ------------------------------ */
HX_API int hx_available(void);
HX_API const char *hx_unavailable_reason(void);
HX_API int hx_device_info_get(hx_device_info *out);
HX_API const char *hx_last_error(void);
HX_API const char *hx_version(void);   /* kernel_layer_version(), CudaKernels.h:307 */

/* ModCtx<uint32_t>::init (ModCtx.h:202-250), device-relevant fields only. */
HX_API int hx_modctx_init(uint32_t p, hx_modctx *out);

/* ---- streams, events, memory, transfers (CudaCtx.h:120-214) -------------- */
HX_API int hx_stream_create(hx_stream *s);
HX_API int hx_stream_destroy(hx_stream s);
HX_API int hx_stream_sync(hx_stream s);
HX_API int hx_device_sync(void);
HX_API int hx_event_create(hx_event *e, int timing);
HX_API int hx_event_record(hx_event e, hx_stream s);
HX_API int hx_event_wait(hx_stream s, hx_event e);
HX_API int hx_event_sync(hx_event e);
HX_API int hx_event_elapsed_ms(hx_event a, hx_event b, float *ms);
HX_API int hx_event_destroy(hx_event e);
/* Device memory through the library's arena (CudaCtx::dalloc / dfree / release_cached,
 * CudaCtx.h:104-168): blocks are cached in exact-size free lists (sizes rounded up to 512 B below
 * 1 MiB, 2 MiB above) and reused by later requests of the same rounded size; a failed device
 * allocation is answered by a trim and one retry.  hx_free returns the block to the cache at once:
 * no queued work may still use it (stream-ordered reuse is the caller's, as with dfree).  None of
 * these is legal inside a stream capture. */
HX_API int hx_malloc(void **ptr, size_t bytes);        /* dalloc */
HX_API int hx_free(void *ptr);                          /* dfree: HX_E_ARGUMENT for a pointer not from hx_malloc */
typedef struct hx_arena_stats {   /* ArenaStats, CudaCtx.h:107-117 */
    size_t reserved_bytes;        /* held from the driver (live + cached) */
    size_t live_bytes;            /* handed out and not yet freed (rounded sizes) */
    size_t peak_live_bytes;
    size_t peak_reserved_bytes;
    uint64_t driver_mallocs;      /* device allocations made */
    uint64_t reuses;              /* requests served from a free list */
    uint64_t driver_frees;        /* cached blocks returned to the driver by trims */
    uint64_t oom_retries;         /* failed allocations answered by a trim and a retry */
} hx_arena_stats;
HX_API int hx_arena_stats_get(hx_arena_stats *out);
/* release_cached (CudaCtx.h:157-168): waits for the device, returns every cached (not live) block
 * to the driver; *freed (may be NULL) receives the bytes returned. */
HX_API int hx_arena_trim(size_t *freed);
/* upload / download (CudaCtx::upload / download, CudaCtx.h:171-177): pageable host memory moves through two
 * pinned 16 MiB staging chunks (the memcpy of one overlapping the DMA of the other); a pinned host buffer
 * (hx_host_alloc) goes straight to the DMA.  Both return when the host buffer may be reused / holds the
 * bytes.  Not legal inside a stream capture (HX_E_ARGUMENT). */
HX_API int hx_upload(void *dev, const void *host, size_t bytes, hx_stream s);
HX_API int hx_download(void *host, const void *dev, size_t bytes, hx_stream s);
/* Pinned host memory (CudaCtx::pinned_alloc / pinned_free, CudaCtx.h:182-184). */
HX_API int hx_host_alloc(void **ptr, size_t bytes);
HX_API int hx_host_free(void *ptr);
typedef struct hx_staging_stats {   /* PinnedPoolStats, CudaCtx.h:82-91 */
    size_t chunk_bytes, cap_bytes, allocated_bytes, peak_bytes;
    uint64_t acquisitions;           /* chunks taken */
    uint64_t waits;                  /* takes that had to wait for the chunk's previous DMA */
} hx_staging_stats;
HX_API int hx_staging_stats_get(hx_staging_stats *out);
HX_API int hx_copy_d2d(void *dst, const void *src, size_t bytes, hx_stream s);
HX_API int hx_memset(void *dst, int value, size_t bytes, hx_stream s);

/* ---- diagnostic counters (launch_stats / reset_launch_stats, CudaKernels.h:299-306;
 *      TransferStats, CudaCtx.h:94-101).  Process-wide, counts only. ------------------- */
typedef struct hx_stats {
    uint64_t kernel_launches;    /* every kernel the library launched */
    uint64_t gemm_launches;      /* GEMM product launches (single or batched; a batch counts per kernel) */
    uint64_t batched_instances;  /* GEMM jobs run through batches */
    uint64_t limb_launches;      /* launches of the INT8-MFMA limb route (the IMMA counter's counterpart) */
    uint64_t h2d_bytes, d2h_bytes, d2d_bytes;   /* hx_upload / hx_download / hx_copy_d2d */
    uint64_t h2d_calls, d2h_calls;
    uint64_t mallocs, frees;     /* hx_malloc / hx_free */
    uint64_t splitk_launches;    /* plain GEMMs run split-K */
} hx_stats;
HX_API int hx_stats_get(hx_stats *out);
HX_API int hx_stats_reset(void);

/* Elimination counters (ImmaElimStats / imma_elim_stats, CudaKernels.h:289-298), over the whole-GPU RREF
 * (hx_rref_u32 on large shapes, hx_rref_u32_device, hx_rref_u32_host). */
typedef struct hx_elim_stats {
    uint64_t calls;              /* eliminations enqueued */
    uint64_t panels;             /* column panels (the reference's bases) */
    uint64_t panel_columns;      /* their columns */
    uint64_t kernel_launches;    /* kernels launched while an elimination was enqueued (process-wide difference) */
    uint64_t gemms;              /* the panels' update products (matrix-core launches) */
    uint64_t h2d_bytes, d2h_bytes;   /* the host-slab form's row transfers */
    size_t peak_device_bytes;    /* the largest single call's device footprint */
    double seconds;              /* wall time inside the synchronous calls, summed */
    uint64_t spec_panels;        /* panels whose pivots the search over their first 64 rows settled (synchronous
                                    calls only; the rest ran the full tournament) */
} hx_elim_stats;
HX_API int hx_elim_stats_get(hx_elim_stats *out);
HX_API int hx_elim_stats_reset(void);

/* ---- graph capture (CudaCtx.h:223-253 GraphCapture) ---------------------- */
HX_API int hx_graph_begin(hx_stream s);
HX_API int hx_graph_end(hx_stream s, hx_graph *g);
HX_API int hx_graph_launch(hx_graph g, hx_stream s);
HX_API int hx_graph_destroy(hx_graph g);

/* ---- GEMM (CudaKernels.h:120-133 gemm / gemm_nt / schur_update) --------- */
/* C op= A * B mod p; A m x k, B k x n, C m x n, any strides; C must not overlap A or B.
 * variant PLAIN: any p; LIMB: p with ctx->limbs > 0 (p <= 4278124287), else HX_E_ARGUMENT. */
HX_API int hx_gemm_u32(const hx_modctx *ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s,
                       const hx_gemm_opts *opts);
/* Device workspace bytes the variant's default route needs for these shapes in one K piece (0 for PLAIN): the
 * limb planes, or the residue planes where the default takes the residue route (see hx_gemm_route). */
HX_API size_t hx_gemm_workspace_bytes(const hx_modctx *ctx, uint32_t m, uint32_t n, uint32_t k, int variant);
/* The same for a given hx_gemm_opts.tile (tile 6, the residue route: N (Mp k' + Np k' + Mp Np) bytes with Mp, Np
 * = m, n rounded up to 256 and k' = k rounded up to 64).  No counterpart in the reference (its IMMA route has
 * one workspace formula). */
HX_API size_t hx_gemm_workspace_bytes_tile(const hx_modctx *ctx, uint32_t m, uint32_t n, uint32_t k, int variant,
                                           int tile);
/* The moduli count N of the residue route (tile 6) for modulus p and depth k: the least N with the product of
 * the first N moduli (255, 254, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211) above 2.05 k (p/2)^2; 0 when
 * the route does not apply (k = 0, k > 133144, p < 2, or 12 moduli too few).  Host-only arithmetic.  No
 * counterpart in the reference. */
HX_API uint32_t hx_gemm_residue_moduli(uint32_t p, uint32_t k);
/* The kernel a product with host-side sizes runs for (variant, tile), given a large enough workspace: 0 = the
 * plain kernel, 1 .. 5 = that limb kernel, 6 = the residue route; -1 for a NULL ctx.  The default (tile 0)
 * takes the residue route where it needs at most limbs^2 - 4 int8 products and m, n, k >= 4096, else limb
 * kernel 4.  No counterpart in the reference. */
HX_API int hx_gemm_route(const hx_modctx *ctx, uint32_t m, uint32_t n, uint32_t k, int variant, int tile);
/* The K chunk a limb workspace of workspace_bytes runs (imma_k_chunk, CudaKernels.h:100-110): min(k, the
 * int32-exact chunk) when its planes fit, else the largest multiple of 64 that fits; 0 when not even a
 * 64-deep chunk fits.  Host-only arithmetic (no device needed). */
HX_API uint32_t hx_gemm_limb_k_chunk(const hx_modctx *ctx, uint32_t m, uint32_t n, uint32_t k, size_t workspace_bytes);
/* Split-K (suggest_splits / gemm_workspace_bytes, CudaKernels.h:80-81): the count the automatic policy takes
 * for an m x n x k plain product (1 when the output tiles alone fill the GPU or k < 1024), and the workspace
 * bytes of a split product (0 for splits <= 1).  Host-only arithmetic. */
HX_API uint32_t hx_gemm_suggest_splits(uint32_t m, uint32_t n, uint32_t k);
HX_API size_t hx_gemm_split_workspace_bytes(uint32_t m, uint32_t n, uint32_t k, uint32_t splits);
/* hx_gemm_u32 without any device allocation.  No counterpart in the reference: for callers that cannot
 * allocate (code under graph capture, tight serving loops, harnesses that forbid it).  The limb route's
 * planes live in a fixed 64 MiB device array per GPU (hx_gemm_noalloc_scratch_bytes), K running in the
 * chunks that fit it; opts->workspace is ignored.  Limit: limbs * (round_up(m, 64) + round_up(n, 128)) * 64
 * bytes (one 64-deep K chunk) must fit the array, else LIMB fails with HX_E_ARGUMENT and AUTO runs the plain
 * kernel (which needs no scratch).  Every call on one GPU shares the array: calls that overlap on a GPU
 * (two streams, two host threads) race and must be serialised by the caller. */
HX_API int hx_gemm_u32_noalloc(const hx_modctx *ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s,
                               const hx_gemm_opts *opts);
HX_API size_t hx_gemm_noalloc_scratch_bytes(void);
/* Calibration, no counterpart in the reference: the sustained rate, in TOP/s (65536 ops per instruction), of
 * v_mfma_i32_32x32x32_i8 issued from registers only in the limb kernel's 16-product pattern, `blocks`
 * workgroups of 256 threads each running `iters` rounds, timed with events on stream s (one untimed launch
 * first).  random_operands = 1 feeds uniformly random bits, as the limb planes of random residues are; 0 feeds
 * small constants.  The first is the ceiling the limb GEMM can reach on this chip; the published peak
 * assumes the clock the second keeps.  Blocks 1..65536, iters 1..1000000. */
HX_API int hx_mfma_rate(uint32_t blocks, uint32_t iters, int random_operands, hx_stream s, double *tops);
/* C -= L * U (schur_update, CudaKernels.h:130-133). */
HX_API int hx_schur_update_u32(const hx_modctx *ctx, hx_view L, hx_view U, hx_view C, hx_stream s,
                               const hx_gemm_opts *opts);

/* ---- F_2 on bit-packed rows (binary::BitMatrix, Binary/BitMatrix.h:49-140) ---------------------------
 * A bit matrix in device memory: row i is `ldw` 64-bit words, bit j of the row is bit j % 64 of word
 * j / 64 (ldw >= ceil(n / 64)); padding bits (columns >= n) are zero and stay zero.  Over F_2 the RREF,
 * its pivot columns and the rank are unique, so the results equal the reference's bit for bit. */
/* In-place reduced row echelon form (rref(BitMatrix&), BitMatrix.h:101-105: leading 1 in each pivot
 * column, zero above and below, zero rows last); rank and the pivot columns ascending (pivots: room for
 * min(m, n), or NULL).  Synchronous on s; its workspace comes from the library's arena
 * (hx_f2_rref_workspace_bytes, per pass).  Up to 16384 rows in one pass; taller matrices in passes of the RREF
 * so far stacked with the next rows (16384 more rows of device memory), which needs the rank to stay below
 * 16384 (HX_E_ARGUMENT when it reaches 16384 with rows left). */
HX_API int hx_f2_rref(uint32_t m, uint32_t n, uint64_t *A, uint32_t ldw, uint32_t *rank, uint32_t *pivots, hx_stream s);
HX_API size_t hx_f2_rref_workspace_bytes(uint32_t m, uint32_t n, uint32_t ldw);
/* C = A B (accumulate = 0) or C += A B (accumulate = 1) over F_2 (multiply_m4rm, BitMatrix.h:129):
 * A m x k, B k x n, C m x n, all bit-packed with the given row strides in words. */
HX_API int hx_f2_gemm(uint32_t m, uint32_t n, uint32_t k, const uint64_t *A, uint32_t lda, const uint64_t *B,
                      uint32_t ldb, uint64_t *C, uint32_t ldc, int accumulate, hx_stream s);
/* ---- F_3 on bit-sliced rows (binary::TritMatrix, Binary/TritMatrix.h:42-125) ---------------------------
 * Row i of an m x n matrix: plane `one` (ldw words: bit j = [x_j == 1]) then plane `two` (ldw words: bit j =
 * [x_j == 2]), row stride 2 ldw words (the reference's layout with ldw = its words_per_plane()); padding
 * bits zero.  hx_f3_rref: rref(TritMatrix&) (TritMatrix.h:96-98) in place on the device -- the RREF (leading
 * 1s, zero above and below, zero rows last), rank and pivot columns (host, min(m, n) entries); up to 8192 rows
 * in one pass, taller matrices in stacked passes while the rank stays below 8192 (HX_E_ARGUMENT when it
 * reaches 8192 with rows left); workspace from the library's arena (hx_f3_rref_workspace_bytes, per pass).  hx_f3_gemm: multiply (TritMatrix.h:117),
 * C = A B or C += A B with A m x k, B k x n, C m x n (lda / ldb / ldc: words per plane). */
HX_API int hx_f3_rref(uint32_t m, uint32_t n, uint64_t *A, uint32_t ldw, uint32_t *rank, uint32_t *pivots, hx_stream s);
HX_API size_t hx_f3_rref_workspace_bytes(uint32_t m, uint32_t n, uint32_t ldw);
HX_API int hx_f3_gemm(uint32_t m, uint32_t n, uint32_t k, const uint64_t *A, uint32_t lda, const uint64_t *B,
                      uint32_t ldb, uint64_t *C, uint32_t ldc, int accumulate, hx_stream s);

/* ---- GEMM mod p for 64-bit residues (Cuda/cuda_kernels.cu:104-164, the u64 kernel; CPU path of record
 *      kernels::gemm<uint64_t>, ModArithmetic/Kernels.h:184) ----------------------------------------- */
typedef struct hx_modctx64 {    /* ModCtx<uint64_t>'s device-relevant constants */
    uint64_t p;                 /* modulus, 2 <= p < 2^64 */
    uint64_t ninv;              /* -p^-1 mod 2^64 (odd p; 0 for even p) */
    uint64_t r2;                /* 2^128 mod p */
    uint64_t two64;             /* 2^64 mod p */
} hx_modctx64;
HX_API int hx_modctx64_init(uint64_t p, hx_modctx64 *out);
/* C op= A * B mod p; views of 8-byte words (strides in elements), operands canonical (< p); k < 2^32.
 * Exact 160-bit sums, one reduction per output (Montgomery for odd p, long division for even p). */
HX_API int hx_gemm_u64(const hx_modctx64 *ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s);

/* ---- batched GEMM of uneven shapes (generalises GemmBatch<Elt>, CudaKernels.h:135-185, whose
 *      instances must share one shape) ------------------------------------------------------ */
typedef struct hx_gemm_job {
    uint32_t m, n, k;
    int32_t op;                 /* HX_OP_* */
    int32_t variant;            /* HX_GEMM_*: AUTO = the batch's variant */
    int32_t tile;               /* the limb kernel of this job (hx_gemm_opts.tile: 0 = default, 1 .. 4; 5 runs
                                   as 4 in a batch); tuning only */
    const uint32_t *A;          /* row-major, leading dimension lda (device) */
    const uint32_t *B;
    uint32_t *C;
    int64_t lda, ldb, ldc;
} hx_gemm_job;
typedef void *hx_gemm_batch;
/* set_shape / push / upload (NOT capturable): the job table goes to device memory, with the limb
 * planes' workspace.  variant: HX_GEMM_PLAIN / HX_GEMM_LIMB for every job, or AUTO (per-job policy);
 * a job's own variant field overrides it.  Jobs of the limb route need k <= one limb chunk. */
HX_API int hx_gemm_batch_create(const hx_modctx *ctx, const hx_gemm_job *jobs, uint32_t njobs, int variant,
                                hx_gemm_batch *out);
/* run (capturable once created): every job in one launch. */
HX_API int hx_gemm_batch_run(hx_gemm_batch b, hx_stream s);
HX_API int hx_gemm_batch_destroy(hx_gemm_batch b);

/* ---- triangular solve (CudaKernels.h:240-247 trsm_unit) ------------------ */
/* B <- T^{-1} B in place, T square unit triangular (only the strict triangle is read). */
HX_API int hx_trsm_unit_u32(const hx_modctx *ctx, hx_view T, hx_view B, int upper, hx_stream s);

/* ---- level-1/2 and tape-support kernels (CudaKernels.h:186-238) --------------------------
 * Views are any-stride windows (hx_view); every operand is canonical (< p) unless stated, and an
 * output must not overlap an input (HX_E_ARGUMENT), as in cuda_kernels.cu:950-1087.  All are
 * asynchronous on s, allocate nothing and are capturable. */
/* C op= A * B^T mod p (gemm_nt, CudaKernels.h:125-128): B is n x k. */
HX_API int hx_gemm_nt_u32(const hx_modctx *ctx, hx_view A, hx_view B, hx_view C, int op, hx_stream s,
                          const hx_gemm_opts *opts);
/* y op= A * x (gemv, CudaKernels.h:187-191): x a k x 1 view, y an m x 1 view. */
HX_API int hx_gemv_u32(const hx_modctx *ctx, hx_view A, hx_view x, hx_view y, int op, hx_stream s);
/* y = y + alpha x, or y - alpha x when subtract != 0 (axpy, CudaKernels.h:193-196). */
HX_API int hx_axpy_u32(const hx_modctx *ctx, uint32_t alpha, hx_view x, hx_view y, int subtract, hx_stream s);
/* x = alpha x (scal, CudaKernels.h:198-200). */
HX_API int hx_scal_u32(const hx_modctx *ctx, uint32_t alpha, hx_view x, hx_stream s);
/* dst(i, j) = src(i, j), any strides (copy, CudaKernels.h:202-204); identical views are a no-op. */
HX_API int hx_copy_u32(hx_view src, hx_view dst, hx_stream s);
/* dst(i, j) = src(j, i) (transpose, CudaKernels.h:206-208). */
HX_API int hx_transpose_u32(hx_view src, hx_view dst, hx_stream s);
/* x(i, j) = value (fill, CudaKernels.h:211-213). */
HX_API int hx_fill_u32(hx_view x, uint32_t value, hx_stream s);
/* y(i, i) = y(i, i) +/- alpha for i < min(rows, cols) (axpy_identity, CudaKernels.h:215-219). */
HX_API int hx_axpy_identity_u32(const hx_modctx *ctx, uint32_t alpha, hx_view y, int subtract, hx_stream s);
/* The tape's GATHER (gather, CudaKernels.h:221-230):
 *   c(i, t) [+]= alpha * [row_off + r(i) == col(t)] + beta * a(r(i), col(t))
 * col(t) = cols[t] (device list of c.cols entries) or t when cols == NULL; r(i) = rows[i] (device
 * list of c.rows entries) or i when rows == NULL.  a may be empty (a.base == NULL: identity source);
 * accumulate == 0 is Set.  a is read only when a.base != NULL and beta != 0. */
HX_API int hx_gather_u32(const hx_modctx *ctx, uint32_t alpha, uint32_t beta, hx_view a, uint32_t row_off,
                         const uint32_t *cols, hx_view c, int accumulate, const uint32_t *rows, hx_stream s);
/* Lazy2p -> canonical in place: v >= p ? v - p : v (canonicalise, CudaKernels.h:232-234). */
HX_API int hx_canonicalise_u32(const hx_modctx *ctx, hx_view v, hx_stream s);
/* dst(i, j) = acc(i, j) mod p for a uint64 accumulator view (strides in 8-byte elements)
 * (canonicalise of an Accum64 view, CudaKernels.h:236-238). */
HX_API int hx_canonicalise_acc64_u32(const hx_modctx *ctx, hx_view acc, hx_view dst, hx_stream s);

/* ---- device RREF over F_p (CudaKernels.h:249-298 imma_rref; DenseElim.h:111-123 hook) ----
 * p prime, 2 <= p < 2^31.  A (m x n, row stride lda, device) is not modified.  R (device, row
 * stride ldr, at least m rows) receives RREF(A) with its zero rows at the bottom (the in-place
 * convention of elim::rref); pass R == NULL for rank / pivots / pivot_rows only.
 * pivots[t]: the pivot columns, ascending (the column rank profile).  pivot_rows[t]: the original
 * row that supplied pivot t under the smallest-original-row rule (CudaKernels.h:255-265), i.e. the
 * row rank profile in pivot order.  pivots / pivot_rows are HOST arrays of min(m, n) entries.
 * Synchronous: returns when rank and pivots are on the host. */
HX_API int hx_rref_u32(const hx_modctx *ctx, uint32_t m, uint32_t n, const uint32_t *A, int64_t lda, uint32_t *R,
                       int64_t ldr, uint32_t *rank, uint32_t *pivots, uint32_t *pivot_rows, hx_stream s);
/* The same elimination entirely on the device, for callers that capture it in a graph (no counterpart in
 * the reference, whose imma_rref synchronises): nothing is allocated and nothing is read back.  d_rank
 * (device, one entry) receives the rank, or 0xffffffff when a pivot had no inverse (p not prime);
 * d_pivots / d_pivot_rows (device, min(m, n) entries, may be NULL) and R (device, may be NULL) as above.
 * workspace: >= hx_rref_workspace_bytes(m, n) bytes of device memory, caller-owned.  Every shape runs the
 * whole-GPU path.  Two calls on one stream may share a workspace; calls on different streams need their own. */
HX_API size_t hx_rref_workspace_bytes(uint32_t m, uint32_t n);
HX_API int hx_rref_u32_device(const hx_modctx *ctx, uint32_t m, uint32_t n, const uint32_t *A, int64_t lda, uint32_t *R,
                              int64_t ldr, uint32_t *d_rank, uint32_t *d_pivots, uint32_t *d_pivot_rows,
                              void *workspace, size_t workspace_bytes, hx_stream s);
/* The reference's row-slab form (imma_rref, CudaKernels.h:276-288): src(user, r0, nr) returns a
 * host pointer to rows [r0, r0 + nr) (row-contiguous, row stride = the returned ld); sink(user, rank,
 * &ld) returns where the rank x n RREF rows go (NULL: no rows wanted). */
typedef const uint32_t *(*hx_row_source)(void *user, uint32_t r0, uint32_t nr, int64_t *ld);
typedef uint32_t *(*hx_row_sink)(void *user, uint32_t rank, int64_t *ld);
HX_API int hx_rref_u32_host(const hx_modctx *ctx, uint32_t m, uint32_t n, hx_row_source src, hx_row_sink sink,
                            void *user, uint32_t slab_rows, uint32_t *rank, uint32_t *pivots, uint32_t *pivot_rows);

/* ---- batched uneven RREF (no counterpart: the reference runs one elimination per call) ---- */
typedef struct hx_rref_job {
    uint32_t m, n;
    uint32_t *A;            /* device, row stride lda; overwritten by RREF(A), zero rows at the bottom */
    int64_t lda;
    uint32_t *pivots;       /* device, min(m, n) entries */
    uint32_t *pivot_rows;   /* device, min(m, n) entries */
    uint32_t *rank;         /* device, one entry */
    uint32_t panel;         /* tuning: widest column panel (8, 16 or 32; 0 = 32); never changes a value */
} hx_rref_job;
typedef void *hx_rref_batch;
/* Launch shape of the persistent batch (tuning only; never changes a value).  AUTO: two workgroups
 * per CU, in half the LDS each, when every job fits there and the batch's total work is at least
 * 512x its largest job's; otherwise one.  The other two force a shape (TWO_PER_CU fails with
 * HX_E_ARGUMENT when a job does not fit half the LDS). */
enum { HX_RREF_BATCH_AUTO = 0, HX_RREF_BATCH_ONE_PER_CU = 1, HX_RREF_BATCH_TWO_PER_CU = 2 };
HX_API int hx_rref_batch_create(const hx_modctx *ctx, const hx_rref_job *jobs, uint32_t njobs, int mode,
                                hx_rref_batch *out);
/* What create chose (for tests and profiles). */
typedef struct hx_rref_batch_info {
    uint32_t njobs;
    uint32_t workgroups_per_cu;       /* 1 or 2 */
    uint32_t grid;                    /* persistent workgroups launched */
    uint32_t region_words;            /* LDS words per workgroup for the panel / X region */
    size_t lds_bytes;                 /* dynamic LDS per workgroup */
    uint32_t jobs_planes_outside_lds; /* jobs whose first panel reads X's limb planes from HBM (region too small) */
} hx_rref_batch_info;
HX_API int hx_rref_batch_info_get(hx_rref_batch b, hx_rref_batch_info *out);
/* One persistent launch over every job (capturable once created). */
HX_API int hx_rref_batch_run(hx_rref_batch b, hx_stream s);
HX_API int hx_rref_batch_destroy(hx_rref_batch b);
/* HX_E_NOT_INVERTIBLE if any job met a pivot without an inverse (synchronises the stream). */
HX_API int hx_rref_batch_status(hx_rref_batch b, hx_stream s);

/* ---- Howell form over Z/2^k, 1 <= k <= 32 (ring2k::howell, Binary/Ring2k.h:959-969; the device
 *      side of the ring2k session, Cuda/CudaRing2k.h:28-46, generalised to whole jobs) --------- */
typedef struct hx_howell_job {
    uint32_t m, n;
    const uint32_t *A;      /* device, row stride lda (entries read mod 2^k) */
    int64_t lda;
    uint32_t *H;            /* device, n x n, row stride n: the h Howell rows, masked to k bits */
    uint32_t *pivots;       /* device, n entries (strictly increasing leading columns) */
    uint32_t *pivot_v;      /* device, n entries (row i leads with exactly 2^pivot_v[i]) */
    uint32_t *h;            /* device, one entry */
    uint32_t *work;         /* device scratch, >= hx_howell_work_words(m, n) u32 */
} hx_howell_job;
HX_API size_t hx_howell_work_words(uint32_t m, uint32_t n);
typedef void *hx_howell_batch;
HX_API int hx_howell_batch_create(uint32_t k, const hx_howell_job *jobs, uint32_t njobs, hx_howell_batch *out);
/* elim::howell over a prime power (DenseElim.cpp:1450 with ModCtx::prime_power(ell, k), ModCtx.h:163-182): the
 * same jobs over Z/ell^k for a prime ell and k >= 1 with ell^k < 2^32 (k = 1: the RREF over F_ell; ell = 2:
 * the Z/2^k batch above).  Entries are read mod ell^k; H holds canonical residues in [0, ell^k), row i
 * leading with exactly ell^pivot_v[i]. */
HX_API int hx_howell_batch_create_ell(uint32_t ell, uint32_t k, const hx_howell_job *jobs, uint32_t njobs,
                                      hx_howell_batch *out);
HX_API int hx_howell_batch_run(hx_howell_batch b, hx_stream s);
HX_API int hx_howell_batch_destroy(hx_howell_batch b);

/* ---- the Z/2^32 / Z/2^64 LU session (Cuda/CudaRing2k.h:28-46; installed through the
 *      ring2k::device::Ops table, Binary/Ring2kDevice.h:64-83; host driver Binary/Ring2kDeviceLU.h) ----
 * A device-resident column window [0, cw) of an m x cw matrix of elem-byte entries (4 or 8); the host
 * does the pivot search and hands each step its factored panel, row swaps and L slots (panel <= 64 columns,
 * at most one swap per panel column).  Wraparound arithmetic: exact.  put / get / fetch are synchronous;
 * step copies its host inputs into pinned staging and returns, the device work runs on the session's
 * streams (the lookahead copy on a second one). */
typedef struct hx_ring2k_step_args {   /* ring2k::device::StepArgs, Binary/Ring2kDevice.h:49-62 */
    uint32_t m;                   /* rows of A */
    uint32_t r, np;               /* the panel's pivot rows r .. r + np - 1 */
    uint32_t pc0, pc1;            /* the panel's columns (the swaps skip them) */
    const void *panel;            /* factored panel, rows r .. m-1, row-major, ld = ldp elements; NULL = none */
    size_t ldp;
    const uint32_t *swaps;        /* nswaps pairs (a, b) of rows, swapped in this order */
    uint32_t nswaps;
    const void *L;                /* L slots from row r on (ld = ldl elements), columns 0 .. np-1 read */
    size_t ldl;
    uint32_t u0;                  /* TRSM / GEMM columns [u0, w) (u0 >= w: no update) */
    uint32_t la0, la1;            /* lookahead: rows r + np .. m-1 of columns [la0, la1), collected by fetch */
    uint32_t w;                   /* window columns in use, [0, w) (0 = cw) */
} hx_ring2k_step_args;
typedef void *hx_ring2k;
HX_API size_t hx_ring2k_session_bytes(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel);
HX_API size_t hx_ring2k_staging_bytes(uint32_t elem, uint32_t m, uint32_t panel);
HX_API int hx_ring2k_open(uint32_t elem, uint32_t m, uint32_t cw, uint32_t panel, hx_ring2k *out);
HX_API int hx_ring2k_put(hx_ring2k s, const void *host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0,
                         uint32_t cols);
HX_API int hx_ring2k_get(hx_ring2k s, void *host, size_t ldh, uint32_t r0, uint32_t rows, uint32_t c0, uint32_t cols);
HX_API int hx_ring2k_step(hx_ring2k s, const hx_ring2k_step_args *a);
HX_API int hx_ring2k_fetch(hx_ring2k s, void *host, size_t ldh);
HX_API int hx_ring2k_close(hx_ring2k s);   /* never fails; frees everything */
/* Ops::plan (Binary/Ring2kDevice.h:65-67): the column window for an m x n LU of elem-byte words with this
 * panel width under a device memory budget (0 = 90 % of the free device memory): n (resident), a multiple
 * of panel below n (tiled), or 0 (declined; the host code runs). */
HX_API uint32_t hx_ring2k_plan(uint32_t elem, uint32_t m, uint32_t n, uint32_t panel, uint64_t budget_bytes);
/* Ops::store_budget (Ring2kDevice.h:80-81): host bytes one phase's L record may hold; from the environment
 * variable HX_RING2K_HOST_MB (MiB), 0 = unlimited. */
HX_API uint64_t hx_ring2k_store_budget(void);
/* Ops::verify (Ring2kDevice.h:78-79): 1 when HX_VERIFY_DEVICE is set (not "0"): the caller re-runs each device
 * LU on the host and compares. */
HX_API int hx_ring2k_verify(void);
/* The device route of ring2k::echelon_blocked (Binary/Ring2kDeviceLU.h; the MinValuation rule) on the session
 * above: A (m x n host words of elem bytes, row stride lda) in place, entries understood mod 2^k.  rank, and
 * pivots / pivot_v (host, min(m, n) entries) exactly as the host code leaves them.  col_limit: pivots are
 * sought in columns [0, col_limit) only.  budget_bytes: as hx_ring2k_plan (a small budget forces the tiled
 * route); store_bytes: the phase record's cap (0 = unlimited).  HX_E_NOMEM when not one panel-wide window
 * fits (nothing was written). */
HX_API int hx_ring2k_echelon(uint32_t elem, uint32_t k, uint32_t m, uint32_t n, void *A, size_t lda, uint32_t panel,
                             uint32_t col_limit, uint64_t budget_bytes, uint64_t store_bytes, uint32_t *rank,
                             uint32_t *pivots, uint32_t *pivot_v);
/* ring2k::smith (Binary/Ring2k.h:701-841; SmithResult, Ring2k.h:408-413): the Smith valuations of A (m x n host
 * words of elem bytes, row stride lda, entries understood mod 2^k; A is not modified) by levels, the matrix
 * resident on the device: per level the F_2 rank profile of the low bits (hx_f2_rref), the odd-pivot LU of the
 * session above on those columns, the even Schur complement halved.  sigma (min(m, n) entries) receives the
 * valuations in ascending order, *nsigma their count (the non-zero Smith diagonal entries 2^sigma_i);
 * level_ranks (k entries) receives #{sigma == s} for each level run, *nlevels their count.  panel: 0 (64) or
 * 1 .. 64.  min(m, n) <= 16384 (a taller matrix is transposed on the device), and the matrix must fit in
 * device memory (HX_E_ARGUMENT otherwise). */
HX_API int hx_ring2k_smith(uint32_t elem, uint32_t k, uint32_t m, uint32_t n, const void *A, size_t lda, uint32_t panel,
                           uint32_t *sigma, uint32_t *nsigma, uint32_t *level_ranks, uint32_t *nlevels);
typedef struct hx_ring2k_stats {   /* ring2k::device::Stats, Binary/Ring2kDevice.h:86-101 */
    uint64_t calls, device, resident, tiled, slabs, replays, fallbacks, panels, lookaheads, verified;
    uint64_t ns_open, ns_fetch, ns_panel, ns_step, ns_host, ns_close;   /* wall ns on the calling thread */
} hx_ring2k_stats;
HX_API int hx_ring2k_stats_get(hx_ring2k_stats *out);
HX_API int hx_ring2k_stats_reset(void);
/* One trailing GEMM C (m x n) -= L (m x k) U (k x n) on the device, all host row-major (gemm_sub_host). */
HX_API int hx_ring2k_gemm_sub_host(uint32_t elem, void *C, size_t ldc, const void *L, size_t ldl, const void *U,
                                   size_t ldu, uint32_t m, uint32_t n, uint32_t k);

/* ---- the learned selector's network (no counterpart in the reference: its cost model is static,
 *      Engine/EngineCostModel.h:22-51) -----------------------------------------------------------
 * A 2-layer MLP, FP32: h = relu(X W1^T + b1), y = h W2^T + b2.  X is B x F, W1 H x F, b1 H,
 * W2 O x H, b2 O, y B x O (all row-major device arrays).  Only the variant / tile CHOICE comes from
 * it; it never touches a value of the exact path. */
HX_API int hx_mlp_forward(const float *X, uint32_t B, uint32_t F, const float *W1, const float *b1, uint32_t H,
                          const float *W2, const float *b2, uint32_t O, float *hidden, float *y, hx_stream s);
/* Mean-squared-error loss L = (1/(B*O)) sum (y - t)^2 with an optional per-entry mask (NULL = all);
 * writes dW1, db1, dW2, db2 (overwritten) and the loss (device scalar). */
HX_API int hx_mlp_backward(const float *X, uint32_t B, uint32_t F, const float *W1, uint32_t H, const float *W2,
                           uint32_t O, const float *hidden, const float *y, const float *target, const float *mask,
                           float *dW1, float *db1, float *dW2, float *db2, float *loss, hx_stream s);
/* Adam step on n parameters (t = 1-based step count). */
HX_API int hx_adam_step(float *param, const float *grad, float *m1, float *m2, uint32_t n, float lr, float beta1,
                        float beta2, float eps, uint32_t t, hx_stream s);

/* ---- fixed-seed device input generation (PerformanceSmokeValidation.cpp:64-89) ----------------
 * out[i * ld + j] = SplitMix64 draw (i * cols + j) of a stream seeded `seed`, mod ell (the
 * RandomEntries(rows * cols, ell, seed) matrix).  Used for synthetic inputs that never leave HBM. */
HX_API int hx_random_entries(uint32_t *out, uint32_t rows, uint32_t cols, int64_t ld, uint32_t ell, uint64_t seed,
                             hx_stream s);
/* The sparse coboundary-like entries: draw e of the stream seeded `seed` decides entry e (row-major):
 * non-zero iff (draw >> 32) % 20 == 0 (density 5 %), value {p-2, p-1, 1, 2}[draw & 3] reduced mod p. */
HX_API int hx_random_sparse(uint32_t *out, uint32_t rows, uint32_t cols, int64_t ld, uint32_t p, uint64_t seed,
                            hx_stream s);

#ifdef __cplusplus
}
#endif

#endif /* HACS_MI355X_H */
