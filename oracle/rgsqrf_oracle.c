/*
 * rgsqrf_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU restatement, in plain C, of the reference's recursive Gram-Schmidt QR
 * path (later_rgsqrf) and of the input generator and -check metrics built
 * around it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker.  The product
 * library (later_amd/) never links, loads or calls it.
 *
 * What each part follows (paths relative to the reference root):
 *   oracle_rgsqrf ............ QR/later_rgsqrf.cu:25-60 (recursion, n/2 split,
 *                               FP16 staging of Q1, A2 and R12 at :43-44,51,
 *                               GemmEx R12 = Q1^T A2 at :45-49, A2 -= Q1 R12
 *                               at :52-56) with the NMIN=128 leaf at :28-32
 *   panel128 ................. QR/panel.cu:10-63 (32|32 then 64|64 split with
 *                               FP32 Sgemm projections)
 *   caqr32 ................... QR/panel.cu:65-134 (TSQR tree over 256-row
 *                               blocks, stacked R in work, Q-tree apply)
 *   mgs_block ................ QR/panel.cu:246-325 (mgs_kernel2: per-column
 *                               norm, reciprocal scale, projections of the
 *                               later columns; warp butterfly at :241-245)
 *   f32_to_f16 ............... util/util.cu:24-32 (__float2half, RN-even)
 *   orth / backward metrics .. test/test_qr.cu:216-268
 *
 * Generalisations, decided where the reference is undefined (its domain is
 * n = 128*2^k, m >= 256, m % 256 == 0 or m % 256 >= 32):
 *   - any n >= 1 splits at n1 = 32*ceil(n/64); this equals n/2 on the
 *     reference domain, so the reference structure is reproduced exactly;
 *   - a TSQR tail of fewer rows than the panel width is merged into the
 *     previous block (the reference leaves garbage in work there);
 *   - R's strictly-lower triangle is zeroed.
 * The GEMMs accumulate exact products of the (rounded) operands in FP64 and
 * round once to FP32: cuBLAS's accumulation order is not published, so the
 * oracle takes the most accurate form of the same arithmetic.
 *
 * Parity pinning: every function here is checked in tests/test_oracle.py
 * against published known-answer vectors (Philox), numpy's IEEE float16 cast
 * (exhaustively), FP64 LAPACK dgeqrf/dorgqr, brute-force FP64 Gram-Schmidt on
 * small inputs, and closed forms (orthonormal, diagonal, triangular inputs).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                        */
/* ------------------------------------------------------------------------ */
EXPORT void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Element e = i + j*m of matrix `mat` under `seed`: key = (seed, mat),
 * counter = (e low 32 bits, e high 32 bits, 0, 0). */
static void elem_bits(uint64_t e, uint32_t seed, uint32_t mat, uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0u, 0u};
    uint32_t key[2] = {seed, mat};
    oracle_philox4x32_10(ctr, key, out);
}

/* a = 2u - 1, u = (x >> 8) * 2^-24: exact in FP32, range [-1, 1). */
EXPORT void oracle_gen_uniform(int64_t m, int64_t n, float *A, int64_t lda, uint32_t seed, uint32_t mat)
{
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            uint32_t o[4];
            elem_bits((uint64_t)(i + j * m), seed, mat, o);
            float u = (float)(o[0] >> 8) * 0x1p-24f;
            A[i + j * lda] = 2.0f * u - 1.0f;
        }
}

/* Box-Muller in FP64: u1 = ((x0>>8)+1) 2^-24 in (0,1], u2 = (x1>>8) 2^-24. */
EXPORT void oracle_gen_normal(int64_t m, int64_t n, float *A, int64_t lda, uint32_t seed, uint32_t mat)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            uint32_t o[4];
            elem_bits((uint64_t)(i + j * m), seed, mat, o);
            double u1 = (double)((o[0] >> 8) + 1u) * 0x1p-24;
            double u2 = (double)(o[1] >> 8) * 0x1p-24;
            A[i + j * lda] = (float)(sqrt(-2.0 * log(u1)) * cos(two_pi * u2));
        }
}

/* ------------------------------------------------------------------------ */
/* IEEE-754 binary16 / bfloat16 round-to-nearest-even casts.                 */
/* ------------------------------------------------------------------------ */
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

EXPORT uint16_t oracle_f32_to_f16(float x)
{
    uint32_t b = f2u(x);
    uint16_t sign = (uint16_t)((b >> 16) & 0x8000u);
    uint32_t a = b & 0x7fffffffu;
    if (a >= 0x7f800000u) /* inf / nan */
        return (uint16_t)(sign | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (a >= 0x477ff000u) /* >= 65520: rounds to inf */
        return (uint16_t)(sign | 0x7c00u);
    uint32_t E = a >> 23;
    if (E < 113) { /* result is subnormal or zero: |x| < 2^-14 */
        if (a <= 0x33000000u) return sign; /* |x| <= 2^-25 -> 0 (tie to even) */
        uint32_t mant = (a & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - E; /* value in 2^-24 units = mant >> shift */
        uint32_t q = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    uint32_t h = ((E - 112u) << 10) | ((a & 0x7fffffu) >> 13);
    uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

EXPORT float oracle_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, f = h & 0x3ffu;
    if (e == 0) {
        float v = (float)f * 0x1p-24f;
        return sign ? -v : v;
    }
    if (e == 31) return u2f(sign | 0x7f800000u | (f << 13));
    return u2f(sign | ((e + 112u) << 23) | (f << 13));
}

EXPORT uint16_t oracle_f32_to_bf16(float x)
{
    uint32_t b = f2u(x);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

EXPORT float oracle_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

EXPORT void oracle_f32_to_f16_array(const float *x, uint16_t *y, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) y[i] = oracle_f32_to_f16(x[i]);
}

/* Counts inputs start .. start+n-1 (as FP32 bit patterns) whose cast differs
 * from ref[]: bit-exact for non-NaN inputs, any NaN for NaN inputs. */
EXPORT int64_t oracle_f16_mismatches(uint32_t start, int64_t n, const uint16_t *ref)
{
    int64_t bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t b = start + (uint32_t)i;
        uint16_t h = oracle_f32_to_f16(u2f(b));
        if ((b & 0x7fffffffu) > 0x7f800000u)
            bad += ((ref[i] & 0x7c00u) != 0x7c00u) || ((ref[i] & 0x3ffu) == 0) || ((h & 0x3ffu) == 0);
        else
            bad += (h != ref[i]);
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Reference restatement.                                                    */
/* ------------------------------------------------------------------------ */
enum { PREC_F32 = 0, PREC_F16 = 1, PREC_BF16 = 2 };

static inline float rnd(float x, int prec)
{
    if (prec == PREC_F16) return oracle_f16_to_f32(oracle_f32_to_f16(x));
    if (prec == PREC_BF16) return oracle_bf16_to_f32(oracle_f32_to_bf16(x));
    return x;
}

/* Copy an M x N column-major block rounding each value to `prec`. */
static float *stage(int M, int N, const float *X, int ldx, int prec)
{
    float *Y = (float *)malloc(sizeof(float) * (size_t)M * (size_t)(N > 0 ? N : 1));
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) Y[i + (size_t)j * M] = rnd(X[i + (size_t)j * ldx], prec);
    return Y;
}

/* C(MxN) = A^T B, A is K x M, B is K x N (both column-major): the projection
 * shape of QR/later_rgsqrf.cu:45 and QR/panel.cu:21,35,51. */
static void gemm_tn(int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                    float *C, int ldc)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
            const float *a = A + (size_t)i * lda, *b = B + (size_t)j * ldb;
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += (double)a[k] * (double)b[k];
            C[i + (size_t)j * ldc] = (float)s;
        }
}

/* C(MxN) = beta*C + alpha*A B with A M x K, B K x N: the update shape of
 * QR/later_rgsqrf.cu:52 and QR/panel.cu:27,41,57 (alpha=-1, beta=1) and the
 * Q-tree apply of QR/panel.cu:97-130 (alpha=1, beta=0). */
static void gemm_nn(int M, int N, int K, float alpha, const float *A, int lda, const float *B,
                    int ldb, float beta, float *C, int ldc)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        double *acc = (double *)calloc((size_t)M, sizeof(double));
        for (int k = 0; k < K; ++k) {
            double bkj = (double)B[k + (size_t)j * ldb];
            const float *a = A + (size_t)k * lda;
            for (int i = 0; i < M; ++i) acc[i] += (double)a[i] * bkj;
        }
        float *c = C + (size_t)j * ldc;
        for (int i = 0; i < M; ++i)
            c[i] = (float)((beta != 0.0f ? (double)beta * (double)c[i] : 0.0) + (double)alpha * acc[i]);
        free(acc);
    }
}

/* Sum of 32 lane partials by the xor butterfly of QR/panel.cu:241-245
 * (masks 16, 8, 4, 2, 1); every lane ends with the same value. */
static float butterfly32(float p[32])
{
    float t[32];
    for (int mask = 16; mask > 0; mask >>= 1) {
        for (int x = 0; x < 32; ++x) t[x] = p[x] + p[x ^ mask];
        memcpy(p, t, sizeof(t));
    }
    return p[0];
}

/* One block of mgs_kernel2 (QR/panel.cu:246-325): MGS of an mm x w block
 * (mm >= w, w <= 32) in FP32.  Q overwrites A; R (w x w, zeros below the
 * diagonal) goes to R with leading dimension ldr.  Lane x of the reference
 * owns rows x, x+32, x+64, ...; the partial sums follow that order. */
static void mgs_block(int mm, int w, float *A, int lda, float *R, int ldr)
{
    float Rs[32 * 32];
    memset(Rs, 0, sizeof(Rs));
    for (int k = 0; k < w; ++k) {
        float *ak = A + (size_t)k * lda;
        float p[32];
        for (int x = 0; x < 32; ++x) {
            float nu = 0.0f;
            for (int i = x; i < mm; i += 32) nu = fmaf(ak[i], ak[i], nu);
            p[x] = nu;
        }
        float normx = sqrtf(butterfly32(p));
        Rs[k + k * 32] = normx;
        float scale = 1.0f / normx;
        for (int i = 0; i < mm; ++i) ak[i] *= scale;
        for (int y = k + 1; y < w; ++y) {
            float *ay = A + (size_t)y * lda;
            for (int x = 0; x < 32; ++x) {
                float nu = 0.0f;
                for (int i = x; i < mm; i += 32) nu = fmaf(ak[i], ay[i], nu);
                p[x] = nu;
            }
            float r = butterfly32(p);
            for (int i = 0; i < mm; ++i) ay[i] = fmaf(-ak[i], r, ay[i]);
            Rs[k + y * 32] = r;
        }
    }
    for (int y = 0; y < w; ++y)
        for (int x = 0; x < w; ++x) R[x + (size_t)y * ldr] = (x <= y) ? Rs[x + y * 32] : 0.0f;
}

#define BLK 256 /* row block of the reference's panel (QR/panel.cu:87-133) */

/* mgs_caqr_panel_256x32 (QR/panel.cu:65-134) for an m x w panel, w <= 32.
 * work needs the stacked R's of every tree level (sum over levels of
 * nb*w*w floats). */
static void caqr32(int m, int w, float *A, int lda, float *R, int ldr, float *work)
{
    if (m < 2 * BLK && m <= BLK) { /* leaf: QR/panel.cu:74-80 */
        mgs_block(m, w, A, lda, R, ldr);
        return;
    }
    int nb = m / BLK, r = m % BLK;
    /* A tail of >= w rows is its own block (QR/panel.cu:105-130); a shorter
     * tail is merged into the last full block. */
    int ntail = (r >= w) ? 1 : 0;
    int nblk = nb + ntail;
    int ldwork = nblk * w;
    for (int b = 0; b < nblk; ++b) {
        int r0 = b * BLK;
        int mm = (b < nb) ? BLK : r;
        if (b == nb - 1 && !ntail) mm += r;
        mgs_block(mm, w, A + r0, lda, work + (size_t)b * w, ldwork);
    }
    float *next = work + (size_t)ldwork * w;
    caqr32(ldwork, w, work, ldwork, R, ldr, next);
    /* Q-tree apply, A_b <- Q_b * work_b (QR/panel.cu:97-102,117-130). */
    for (int b = 0; b < nblk; ++b) {
        int r0 = b * BLK;
        int mm = (b < nb) ? BLK : r;
        if (b == nb - 1 && !ntail) mm += r;
        float *Qb = stage(mm, w, A + r0, lda, PREC_F32);
        gemm_nn(mm, w, w, 1.0f, Qb, mm, work + (size_t)b * w, ldwork, 0.0f, A + r0, lda);
        free(Qb);
    }
}

/* Split point of the recursion: n/2 on the reference domain. */
static int split_n(int n) { return 32 * ((n + 63) / 64); }

/* The recursion of QR/later_rgsqrf.cu:25-60.  Nodes of width > nmin use
 * `prec` operands (the FP16 tensor-op GemmEx pair); nodes of width <= nmin
 * are inside the 128-panel and use FP32 Sgemm (QR/panel.cu:21-61). */
static void qr_rec(int m, int n, float *A, int lda, float *R, int ldr, float *work, int prec,
                   int nmin)
{
    if (n <= 32) {
        caqr32(m, n, A, lda, R, ldr, work);
        return;
    }
    int n1 = split_n(n), n2 = n - n1;
    int p = (n <= nmin) ? PREC_F32 : prec;
    qr_rec(m, n1, A, lda, R, ldr, work, prec, nmin);
    float *A2 = A + (size_t)n1 * lda;
    float *R12 = R + (size_t)n1 * ldr;
    float *Qh = stage(m, n1, A, lda, p);  /* s2h(Q1): QR/later_rgsqrf.cu:43 */
    float *Bh = stage(m, n2, A2, lda, p); /* s2h(A2): QR/later_rgsqrf.cu:44 */
    gemm_tn(n1, n2, m, Qh, m, Bh, m, R12, ldr);
    free(Bh);
    float *Rh = stage(n1, n2, R12, ldr, p); /* s2h(R12): QR/later_rgsqrf.cu:51 */
    gemm_nn(m, n2, n1, -1.0f, Qh, m, Rh, n1, 1.0f, A2, lda);
    free(Rh);
    free(Qh);
    qr_rec(m, n2, A2, lda, R + n1 + (size_t)n1 * ldr, ldr, work, prec, nmin);
}

/* Floats of panel workspace the restatement needs for an m-row panel. */
EXPORT int64_t oracle_rgsqrf_lwork(int m)
{
    int64_t tot = 0;
    int64_t mm = m;
    while (mm > BLK) {
        int64_t nblk = mm / BLK + 1; /* upper bound for any panel width */
        tot += nblk * 32 * 32;
        mm = nblk * 32;
    }
    return tot + 32 * 32;
}

/* A (m x n, lda) is overwritten by Q; R (n x n, ldr) receives the upper
 * triangle, strictly-lower part zeroed.  prec: 0 FP32, 1 FP16, 2 BF16
 * operands for nodes wider than nmin.  Returns 0, or -1 on bad arguments. */
EXPORT int oracle_rgsqrf(int m, int n, float *A, int lda, float *R, int ldr, int prec, int nmin)
{
    if (m < 1 || n < 1 || m < n || lda < m || ldr < n) return -1;
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) R[i + (size_t)j * ldr] = 0.0f;
    float *work = (float *)calloc((size_t)oracle_rgsqrf_lwork(m), sizeof(float));
    if (!work) return -1;
    qr_rec(m, n, A, lda, R, ldr, work, prec, nmin);
    free(work);
    return 0;
}

/* The 32-column TSQR panel alone (QR/panel.cu:65-134), exposed for the
 * panel-level parity tests (test/test_caqr_panel.cu:83 runs m=4096, n=32). */
EXPORT int oracle_caqr32(int m, int w, float *A, int lda, float *R, int ldr)
{
    if (w < 1 || w > 32 || m < w) return -1;
    float *work = (float *)calloc((size_t)oracle_rgsqrf_lwork(m), sizeof(float));
    caqr32(m, w, A, lda, R, ldr, work);
    free(work);
    return 0;
}

/* One mgs_kernel2 block alone (test/test_mgs_panel.cu:25: m=256, n=32). */
EXPORT int oracle_mgs_block(int mm, int w, float *A, int lda, float *R, int ldr)
{
    if (w < 1 || w > 32 || mm < w) return -1;
    mgs_block(mm, w, A, lda, R, ldr);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* -check metrics (test/test_qr.cu:216-268), accumulated in FP64.            */
/* ------------------------------------------------------------------------ */
/* ||I - Q^T Q||_F (raw; the reference prints this divided by n). */
EXPORT double oracle_orth_fro(int m, int n, const float *Q, int ldq)
{
    double tot = 0.0;
#pragma omp parallel for reduction(+ : tot) schedule(dynamic, 4)
    for (int j = 0; j < n; ++j) {
        const float *qj = Q + (size_t)j * ldq;
        for (int i = 0; i <= j; ++i) {
            const float *qi = Q + (size_t)i * ldq;
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += (double)qi[k] * (double)qj[k];
            double d = s - (i == j ? 1.0 : 0.0);
            tot += (i == j ? 1.0 : 2.0) * d * d;
        }
    }
    return sqrt(tot);
}

/* ||A - Q triu(R)||_F / ||A||_F. */
EXPORT double oracle_backward(int m, int n, const float *A, int lda, const float *Q, int ldq,
                              const float *R, int ldr)
{
    double num = 0.0, den = 0.0;
#pragma omp parallel for reduction(+ : num, den) schedule(static)
    for (int j = 0; j < n; ++j) {
        double *acc = (double *)calloc((size_t)m, sizeof(double));
        for (int k = 0; k <= j; ++k) {
            double rkj = (double)R[k + (size_t)j * ldr];
            const float *qk = Q + (size_t)k * ldq;
            for (int i = 0; i < m; ++i) acc[i] += (double)qk[i] * rkj;
        }
        const float *aj = A + (size_t)j * lda;
        for (int i = 0; i < m; ++i) {
            double d = (double)aj[i] - acc[i];
            num += d * d;
            den += (double)aj[i] * (double)aj[i];
        }
        free(acc);
    }
    return sqrt(num) / sqrt(den);
}
