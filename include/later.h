/*
 * later.h -- C-ABI of the MI355X-native recursive Gram-Schmidt QR.
 *
 * Drop-in for the reference's QR entry point
 *   void later_rgsqrf(cudaCtxt ctxt, int m, int n, float* A, int lda,
 *                     float* R, int ldr, float* work, int lwork,
 *                     __half* hwork, int lhwork);          (include/LATER.h:39)
 * with the same argument order and meaning: column-major FP32 A (m x n, lda)
 * overwritten by Q, FP32 R (n x n, ldr) receiving the upper-triangular factor,
 * caller-owned device workspaces `work` (FP32, lwork elements) and `hwork`
 * (16-bit, lhwork elements).  Work is enqueued on ctxt.stream and the call
 * returns without synchronising, as the reference enqueues on its cuBLAS
 * handle's stream (QR/later_rgsqrf.cu:62-79, test/test_qr.cu:80-84).
 *
 * Differences (a strict superset of the reference's behaviour):
 *   - lda is honoured (the reference forces lda := m, QR/later_rgsqrf.cu:70);
 *   - any m >= n >= 1 is accepted (the reference needs n = 128*2^k, m >= 256);
 *   - R's strictly-lower triangle is zeroed (the reference leaves the
 *     off-diagonal lower blocks untouched, QR/panel.cu:321-323);
 *   - errors are reported as status codes by later_rgsqrf_ex (the reference
 *     prints and continues, QR/panel.cu:12-15);
 *   - R, work and hwork must not alias each other or A.
 * No function here allocates device memory except the -check helpers, the
 * vendor-baseline timer and the serving pool, each marked below.
 *
 * Plain C: no HIP, torch or C++ types appear in the signatures.  Streams are
 * hipStream_t values passed as void*.
 */
#ifndef LATER_H_
#define LATER_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces cudaCtxt {cublasHandle_t, cusolverDnHandle_t} (include/LATER.h:19-22).
 * stream: hipStream_t (NULL = the null stream).  handle: reserved, may be NULL. */
typedef struct laterCtxt {
    void *stream;
    void *handle;
} laterCtxt;

/* Status codes. */
enum {
    LATER_OK = 0,
    LATER_ERR_SHAPE = 1,      /* m < n, n < 1, lda < m, ldr < n, ... */
    LATER_ERR_WORKSPACE = 2,  /* lwork / lhwork below the query */
    LATER_ERR_OPTION = 3,     /* unknown option value */
    LATER_ERR_DEVICE = 4,     /* HIP runtime error */
    LATER_ERR_BREAKDOWN = 5,  /* zero or non-finite column norm (rank deficiency, NaN/Inf) */
    LATER_ERR_ALIAS = 6,      /* R / work / hwork overlap */
    LATER_ERR_BUDGET = 7,     /* serving: request exceeds the HBM budget */
    LATER_ERR_CANCELLED = 8   /* serving: request cancelled */
};

/* GEMM operand precision of the recursion nodes. */
enum {
    LATER_GEMM_F32 = 0,   /* FP32 MFMA (the plain-FP32 form) */
    LATER_GEMM_F16 = 1,   /* FP16 operands, FP32 accumulate: the reference's
                             GemmEx TENSOR_OP pair (QR/later_rgsqrf.cu:45-56) */
    LATER_GEMM_BF16 = 2,  /* BF16 operands, FP32 accumulate */
    LATER_GEMM_F16X3 = 3, /* FP16 hi/lo split of both operands, three MFMA
                             products, power-of-two column scaling */
    LATER_GEMM_F16X6 = 4, /* FP16 three-way split (2^-33 representation), six
                             MFMA products; projections only (updates use X3) */
    LATER_GEMM_FP8 = 5,   /* FP8 e4m3 operands, each operand column scaled by a
                             power of two per 128 rows of K (the E8M0 block
                             scales of v_mfma_scale_f32_32x32x64_f8f6f4), FP32
                             accumulation; projections only (updates default
                             to X3); the operands are staged in hwork, which
                             this precision requires */
    LATER_GEMM_I8X4 = 6   /* int8 four-slice split of both operands (each
                             column scaled by a power of two, x 2^e =
                             s0 + s1 2^-7 + s2 2^-14 + s3 2^-21 with |s| <= 64),
                             the thirteen products of slice pairs i + j <= 4
                             on v_mfma_i32_32x32x32_i8, summed exactly in
                             int32; projections only, K <= 131008 rows
                             (longer products use X6), staged in hwork */
};

/* Panel (leaf) algorithm. */
enum {
    LATER_PANEL_MGS = 0,    /* reference form: TSQR tree of block MGS, FP32 sums */
    LATER_PANEL_MGS64 = 1,  /* same tree, FP64 column norms and dot products */
    LATER_PANEL_CHOLQR2 = 2,/* Cholesky QR twice with FP64 Gram factors: three
                               grid-wide launches, no sequential column loop;
                               leaves of up to 64 columns (one pass in the
                               second factorisation of LATER_REORTH_PASS) */
    LATER_PANEL_CHOLQR1 = 3 /* the same leaves with one Cholesky QR pass in
                               every factorisation (orthogonality like
                               kappa(leaf)^2 u_64 plus the FP32 rounding of Q:
                               for a first pass that a second pass repairs) */
};

/* Re-orthogonalisation (bit flags; NODE|PASS = 3 combines both). */
enum {
    LATER_REORTH_NONE = 0,  /* single pass (the reference) */
    LATER_REORTH_NODE = 1,  /* second projection at every node (block CGS2) */
    LATER_REORTH_PASS = 2   /* second full factorisation of Q, R := R2 * R1 */
};

typedef struct later_opts {
    int gemm;         /* LATER_GEMM_* for nodes wider than nmin */
    int gemm_pass2;   /* LATER_GEMM_* for the nodes of the second pass */
    int panel;        /* LATER_PANEL_* */
    int reorth;       /* LATER_REORTH_* */
    int nmin;         /* nodes of width <= nmin use FP32 products (reference: 128) */
    int chunk;        /* rows per FP32 accumulation chunk folded into FP64 (0 = off) */
    int chunk2;       /* the same for the second pass (0 = use chunk) */
    int gemm_update;  /* LATER_GEMM_* of the node updates A2 -= Q1 R12 above
                         nmin (F32 .. F16X3), or -1: paired with the
                         projection precision (F16X6 and FP8 -> F16X3, the
                         others the same) */
    int nmin2;        /* the second pass's projections use FP32 products at
                         nodes of width <= nmin2 (0: nmin) and its updates at
                         width <= max(nmin, nmin2) */
} later_opts;

/* Defaults reproduce the reference form: FP16 GEMMs above 128 columns, FP32
 * below, FP32 block-MGS TSQR panel, no re-orthogonalisation. */
void later_opts_default(later_opts *o);
/* The accuracy-recovering mixed-precision form: FP16X3 first pass (FP32
 * accumulation, chunk 0) above 256 columns, I8X4 second-pass projections
 * above 128 columns (FP16X6 with FP64 folds every 32 rows when hwork cannot
 * stage the slices or m > 131008), FP16X3 second-pass updates above 256
 * columns, FP32 products below,
 * Cholesky QR leaves of up to 64 columns (one pass per leaf,
 * LATER_PANEL_CHOLQR1), second full pass. */
void later_opts_mixed(later_opts *o);
/* The plain-FP32 form: FP32 MFMA everywhere, reference panel, one pass. */
void later_opts_fp32(later_opts *o);

/* Workspace query: *lwork = FP32 elements of `work` that later_rgsqrf_ex
 * needs for (m, n, opts); *lhwork = 16-bit elements of `hwork` with which
 * the node GEMMs of 512+ columns stage their operands as pre-split FP16
 * pieces (the reference's s2h staging, util/util.cu:24-32) instead of
 * converting them in the kernel: faster, same result bits; a smaller or
 * absent hwork selects the in-kernel conversion (0: nothing to stage). */
int later_rgsqrf_workspace(int m, int n, const later_opts *opts, int64_t *lwork, int64_t *lhwork);

/* Reference entry point, same argument list as include/LATER.h:39 (hwork is
 * the reference's __half*).  Runs the default options.  Returns nothing, as
 * the reference; a shape or workspace error leaves A and R untouched.  The
 * reference's own caller sizes (lwork = m/256*32*n floats, lhwork = m*n
 * halves, test/test_qr.cu:73-76) always suffice: what does not fit in `work`
 * goes to `hwork` (see later_read_status).  R may alias `work`, as the
 * reference's polar-decomposition caller passes it (EVD/later_qdwh_polar.cu:79):
 * the whole workspace then lives in `hwork`. */
void later_rgsqrf(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr,
                  float *work, int lwork, uint16_t *hwork, int lhwork);

/* Status-returning form with options.  info (optional) receives per-call
 * counters. */
typedef struct later_info {
    int status;
    int kernels;          /* kernel launches enqueued */
    int tree_levels;      /* recursion depth */
    double flops_ref;     /* 2 n^2 (m - n/3), the reference's formula (test/test_qr.cu:85-86) */
    double flops_mgs;     /* 2 m n^2 */
} later_info;

int later_rgsqrf_ex(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr,
                    float *work, int64_t lwork, uint16_t *hwork, int64_t lhwork,
                    const later_opts *opts, later_info *info);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  later_rgsqrf_ex with its FP32 workspace in a file-scope
 * device array of LATER_NOALLOC_SCRATCH_FLOATS floats (160 MiB; each GPU the
 * entry point runs on holds its own copy until the process ends) and no
 * 16-bit arena (in-kernel operand conversion, the same result bits).  The
 * caller still passes A and R and has the stream's device current.  Shapes
 * and options whose later_rgsqrf_workspace exceeds the array return
 * LATER_ERR_WORKSPACE: every reference- and FP32-form call up to 16384^2 and
 * the tall-skinny 1048576 x 128 and 262144 x 256 fit, the mixed form up to
 * 4096^2.  Nothing is allocated or freed after the first call on a device
 * (which looks up the array's device address).  All calls on one GPU share
 * the array: calls that overlap in time on one GPU race. */
#define LATER_NOALLOC_SCRATCH_FLOATS (40 * 1024 * 1024)
int later_rgsqrf_noalloc(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr,
                         const later_opts *opts, later_info *info);

/* Out-of-core QR, replacing later_oc_qr_rec (include/LATER.h:49,
 * QR/later_oc_qr.cu:29-94): A (m x n, lda) and R (n x n, ldr) are HOST
 * memory (hipHostMalloc'd memory lets copies overlap products).  The
 * reference's Mem_pool argument becomes `device_bytes`, the device memory
 * the call may use (allocated once per call).  The recursion splits columns
 * as the in-core one does until a column block fits, factors it in core with
 * `opts`, and runs each wider node as a block Gram-Schmidt step over row
 * blocks streamed through the device (block CGS2 when opts->reorth != 0).
 * Synchronous: returns when A holds Q and R holds R.  LATER_ERR_WORKSPACE
 * (before anything is copied) when a 32-column leaf or a 256-row node block
 * does not fit the budget. */
int later_oc_qr_rec(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr,
                    int64_t device_bytes, const later_opts *opts, later_info *info);

/* Graph-captured plans (no counterpart in the reference, which launches
 * eagerly): later_plan_create records every kernel of one later_rgsqrf_ex
 * call on these exact buffers and options into a HIP graph (it enqueues
 * nothing and computes nothing); later_plan_execute replays it on `stream`
 * with one graph launch and the same effect as the eager call.  The buffers
 * must stay allocated while the plan exists. */
int later_plan_create(int m, int n, float *A, int lda, float *R, int ldr, float *work, int64_t lwork,
                      uint16_t *hwork, int64_t lhwork, const later_opts *opts, void **plan);
int later_plan_execute(void *plan, void *stream);
int later_plan_kernels(void *plan);
int later_plan_destroy(void *plan);

/* Batched QR of `count` independent matrices A_i (m_i x n_i, lda_i, device
 * pointers) into R_i (n_i x n_i, ldr_i), no counterpart in the reference.
 * later_batch_workspace gives the FP32 workspace; later_batch_plan_create
 * builds the grouped task tables (matrices bucketed by ceil(n/32), one
 * grouped launch per step of the shared recursion tree, one stream per
 * bucket) and captures them into a HIP graph -- it allocates its tables with
 * hipMalloc once; later_batch_plan_execute replays the graph on `stream`.
 * Options as for later_rgsqrf_ex except LATER_REORTH_NODE; the leaves always
 * run the block-MGS tree, with FP64 sums for LATER_PANEL_MGS64 and
 * LATER_PANEL_CHOLQR2.  Matrix i's
 * status word (bit 0: breakdown) is the first int of its workspace block. */
int later_batch_workspace(int count, const int *m, const int *n, const later_opts *opts, int64_t *lwork);
int later_batch_plan_create(int count, const int *m, const int *n, float *const *A, const int *lda,
                            float *const *R, const int *ldr, float *work, int64_t lwork,
                            const later_opts *opts, void **plan);
int later_batch_plan_execute(void *plan, void *stream);
int later_batch_plan_launches(void *plan);
int later_batch_plan_destroy(void *plan);

/* Batched QR in ONE persistent launch (no counterpart in the reference).
 * later_batch_wg_create builds the task table (one task per matrix, ordered
 * by size bucket and decreasing cost, largest first) and the device work
 * queue, allocating them once with hipMalloc; later_batch_wg_execute
 * enqueues one kernel on `stream` whose workgroups each take whole matrices
 * from the queue (an atomic queue head; the last workgroup to finish resets
 * it, so the plan can be executed again).  Per matrix: left-looking block
 * Gram-Schmidt, each panel projected out of all earlier columns of Q twice
 * when opts->reorth != 0 (once otherwise), then factored by Cholesky QR
 * twice with FP64 Gram matrices; products on MFMA in FP32
 * (opts->gemm F32 and the split forms), BF16 or FP16 (opts->gemm BF16 /
 * F16).  A is overwritten by Q and R (n x n, zero below the diagonal).
 * status (optional, count ints): bit 0 set for a matrix whose panel had a
 * bad pivot.  tuning (optional): LATER_BATCH_BUCKETS variant ids, one per
 * size bucket later_batch_wg_bucket(m, n) = 7 (min(ceil(n/32), 8) - 1) +
 * clamp(floor(log2 m) - 6, 0, 6); variant v uses 16 << (v & 1) columns per
 * panel and blocks of 64 << (v >> 1) earlier Q columns; NULL (or -1 entries)
 * selects the built-in table written by later_amd/autotune.py. */
#define LATER_BATCH_BUCKETS 56
int later_batch_wg_create(int count, const int *m, const int *n, float *const *A, const int *lda,
                          float *const *R, const int *ldr, int *status, const later_opts *opts,
                          const int *tuning, void **plan);
int later_batch_wg_execute(void *plan, void *stream);
int later_batch_wg_destroy(void *plan);
int later_batch_wg_bucket(int m, int n);

/* Synchronises ctxt.stream and reads the device status word that the last
 * later_rgsqrf_ex / later_caqr_panel call on this `work` left: bit 0 set
 * means a zero or non-finite column norm (LATER_ERR_BREAKDOWN); the
 * affected Q column is zero and its R diagonal entry 0.  (A non-finite input
 * sets the bit too; Q and R are then unspecified.)  The word is work[0]
 * whenever lwork >= 64; with a smaller `work` (the reference's caller sizes
 * give lwork = 0 for m < 256) it is the first word of hwork: pass hwork. */
int later_read_status(laterCtxt ctxt, const float *work, int *device_status);

/* Panel of the reference alone: mgs_caqr_panel_256x32 (QR/panel.cu:65-134)
 * generalised to w <= 32 columns.  A (m x w) becomes Q, R (w x w).
 * fp64_sums: 0 the reference's FP32 block MGS, 1 block MGS with FP64 sums,
 * 2 Cholesky QR twice with FP64 Gram factors (LATER_PANEL_CHOLQR2).
 * later_caqr_panel_workspace: the FP32 `work` elements it needs. */
int later_caqr_panel_workspace(int m, int w, int fp64_sums, int64_t *lwork);
int later_caqr_panel(laterCtxt ctxt, int m, int w, float *A, int lda, float *R, int ldr,
                     float *work, int64_t lwork, int fp64_sums);

/* The two node GEMMs alone (QR/later_rgsqrf.cu:45-49 and :52-56):
 *   later_gemm_proj:   C (n1 x n2, ldc)  = Q^T B,   Q m x n1, B m x n2
 *   later_gemm_update: B (m x n2, ldb)  -= Q C
 * prec is LATER_GEMM_*; chunk as in later_opts.  hwork (optional, 16-bit,
 * lhwork elements): when at least the query's *lhwork, the operands are
 * first split into FP16/BF16 pieces there (the reference's s2h staging,
 * util/util.cu:24-32) and the products stream the pieces -- the kernels the
 * recursion's wide nodes run; NULL converts in the kernel.
 * later_gemm_workspace: *lwork FP32 elements of `work` (column exponents and
 * split-K partials) and *lhwork 16-bit elements of the pieces, for the
 * projection (update = 0) or the update (update = 1). */
int later_gemm_workspace(int m, int n1, int n2, int prec, int chunk, int update, int64_t *lwork,
                         int64_t *lhwork);
int later_gemm_proj(laterCtxt ctxt, int m, int n1, int n2, const float *Q, int ldq,
                    const float *B, int ldb, float *C, int ldc, int prec, int chunk,
                    float *work, int64_t lwork, uint16_t *hwork, int64_t lhwork);
int later_gemm_update(laterCtxt ctxt, int m, int n1, int n2, const float *Q, int ldq,
                      const float *C, int ldc, float *B, int ldb, int prec, float *work,
                      int64_t lwork, uint16_t *hwork, int64_t lhwork);

/* Least squares (the LS-solve of the serving path): triangular solve
 * B (n x k, ldb) <- R^{-1} B with R upper triangular (n <= 8192), FP64 sums;
 * stands in for the reference's recursive TRSM (BLAS/later_rtrsm.cu:75-187). */
int later_trsm_upper(laterCtxt ctxt, int n, int k, const float *R, int ldr, float *B, int ldb);

/* min ||A X - B||_2: A (m x n) <- Q, R (n x n) <- R, X (n x k) <- R^{-1} Q^T B.
 * B is read only; Q^T B uses FP16 hi/lo products with FP64-folded chunks.
 * later_ls_workspace gives the FP32 workspace for (m, n, k, opts). */
int later_ls_workspace(int m, int n, int k, const later_opts *opts, int64_t *lwork);
int later_gels(laterCtxt ctxt, int m, int n, int k, float *A, int lda, float *R, int ldr,
               const float *B, int ldb, float *X, int ldx, float *work, int64_t lwork,
               const later_opts *opts, later_info *info);

/* Least-squares serving runtime (continuous batching on one GPU).
 * later_server_create allocates the HBM pool of `hbm_budget_bytes` once
 * (the only allocation; requests never allocate) and `slots` streams.
 * later_server_submit queues a request (device pointers A: m x n, B: m x k,
 * owned by the caller until the request is done); it is rejected with
 * LATER_ERR_BUDGET when its working set exceeds the pool.  Requests are
 * admitted in arrival order as slots and pool memory free up and run
 * concurrently, each on its own stream.  later_server_poll advances every
 * request without blocking.  Results (Q, R, X) stay resident in the pool
 * until later_server_release.  States: 0 waiting, 1 QR, 2 solve, 3 done,
 * 4 failed, 5 robust re-factorisation (after an orthogonality sketch above
 * 1e-5), 6 cancelled.  later_server_cancel removes a waiting request from
 * the queue at once; a running one finishes the stage already enqueued on
 * its stream and is then retired without further launches (state 6 at the
 * next poll); a finished one cannot be cancelled.  Cancelled requests are
 * released like finished ones.  Calls on one server are serialised by an
 * internal mutex. */
int later_server_create(int slots, int64_t hbm_budget_bytes, const later_opts *qr_opts, void **srv);
int later_server_submit(void *srv, int m, int n, int k, const float *A, int lda, const float *B, int ldb,
                        uint64_t *id);
int later_server_poll(void *srv, int *active);
int later_server_result(void *srv, uint64_t id, int *state, int *status, const float **X,
                        const float **Q, const float **R, double *latency_s, double *queue_s,
                        double *cond, int *iters);
int later_server_fetch(void *srv, uint64_t id, float *X, int ldx, float *R, int ldr, float *Q, int ldq);
int later_server_cancel(void *srv, uint64_t id);
int later_server_release(void *srv, uint64_t id);
int later_server_stats(void *srv, int64_t *pool_bytes, int64_t *used_bytes, int64_t *peak_bytes,
                       int *completed, int *rejected);
int later_server_destroy(void *srv);

/* Input generator (replaces util/util.cu:93-109, which uses cuRAND):
 * Philox4x32-10 keyed by (seed, mat), counter = column-major index i + j*m.
 * uniform: 2u-1 with u from 24 bits; normal: FP64 Box-Muller. */
int later_generate_uniform(laterCtxt ctxt, int64_t m, int64_t n, float *A, int64_t lda,
                           uint32_t seed, uint32_t mat);
int later_generate_normal(laterCtxt ctxt, int64_t m, int64_t n, float *A, int64_t lda,
                          uint32_t seed, uint32_t mat);

/* -check metrics with FP64 accumulation (test/test_qr.cu:216-268).  These
 * allocate a small device buffer and synchronise the stream.
 *   orth: ||I - Q^T Q||_F (the reference prints this divided by n)
 *   backward: ||A - Q triu(R)||_F / ||A||_F */
int later_check_orth(laterCtxt ctxt, int m, int n, const float *Q, int ldq, double *orth);
int later_check_backward(laterCtxt ctxt, int m, int n, const float *A, int lda, const float *Q,
                         int ldq, const float *R, int ldr, double *backward);

/* Kernel-class timing with HIP events on the launch stream (instrumentation
 * for benchmarks; off by default).  Classes: 0 node projection GEMM (other
 * than classes 4 and 5), 1 node update GEMM (other than class 4), 2 panel
 * block-MGS kernel, 3 other, 4 the FP16 hi/lo products on pre-split operands
 * with 256 x 128 tiles (a single kernel: node updates, the R2 R1
 * composition, large first-pass projections), 5 the int8 four-slice
 * projections.  later_prof_end synchronises and fills 6-entry arrays:
 * launches, total ms, algorithmic flops and algorithmic HBM bytes.
 * Host-side, process-wide state. */
void later_prof_begin(void);
int later_prof_end(int *launches, double *ms, double *flops, double *bytes);

/* Library version string. */
const char *later_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LATER_H_ */
