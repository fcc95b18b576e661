// Internal interface of the node GEMMs (gemm.hip).
#pragma once
#include "common.h"

enum { GEMM_F32 = LATER_GEMM_F32, GEMM_F16 = LATER_GEMM_F16, GEMM_BF16 = LATER_GEMM_BF16,
       GEMM_F16X3 = LATER_GEMM_F16X3, GEMM_F16X6 = LATER_GEMM_F16X6, GEMM_FP8 = LATER_GEMM_FP8,
       GEMM_I8X4 = LATER_GEMM_I8X4 };

// One product of the tiled MFMA kernel (device-side argument block; the
// grouped launches read an array of these).
struct GemmArgs {
    int M, N, K;
    const float *A; int64_t lda;   // TN: K x M (col i of A = row i of op(A)); NN: M x K
    const float *B; int64_t ldb;   // K x N
    float *C; int64_t ldc;         // M x N
    int a_exp;                     // op(A) scaled by 2^a_exp before the cast
    const int *b_exp;              // column j of B scaled by 2^b_exp[j] (nullptr: none)
    int kper;                      // K rows per split (multiple of 64)
    int chunk_rows;                // K rows per FP64 fold, a multiple of 32 (0: no FP64 accumulation)
    void *partial;                 // split-K partials (nsplit x M x N), nullptr: write C
    int update;                    // 0: C = result; 1: C += alpha * result (FP32); 2: the same in FP64
    float alpha;
    int tri;                       // op(A), B upper triangular: skip zero tiles and K ranges
    // Pre-split operands (nullptr: convert A and B in the kernel): 16-bit
    // pieces, K-major (row i of op(A) at Ah + i * ldah, piece c at
    // Ah + c * psa), already scaled by 2^a_exp / 2^b_exp[j].
    const uint16_t *Ah; int64_t ldah, psa;
    const uint16_t *Bh; int64_t ldbh, psb;
    // FP8 operands (GEMM_FP8): Ah / Bh hold bytes (ld in bytes), the E8M0
    // block scales are at byte offset psa / psb with row stride lsa / lsb
    // (scale of row i, K block kb at kb * ls + i).
    // int8 slices (GEMM_I8X4): K, kper, ldah / ldbh and psa / psb count
    // pairs of bytes (so the 16-bit streaming code serves them unchanged);
    // the int32 column exponents of op(A)'s rows and of B's columns are at
    // byte offset lsa / lsb from Ah / Bh.
    int64_t lsa, lsb;
};

// One operand in pre-split form (see gemm_split, gemm_split8).
struct Pieces {
    const uint16_t *p;
    int64_t ld, stride;
    int64_t ls;  // FP8: row stride of the block scales; I8X4: byte offset of the exponents
};

// FP8 staging (GEMM_FP8) of an operand whose K runs down its columns: column
// j of X (K rows) becomes K (padded with zeros to a multiple of 64) e4m3
// bytes at buf + j * fp8_ld(K), scaled by 2^-e for each 128-row block kb, and
// the E8M0 byte 127 - e of that block sits at buf + fp8_scale_off(K, cols) +
// kb * fp8_ls(cols) + j.  e makes the block's largest magnitude land in
// (224, 448], e4m3's top binade.
int64_t fp8_ld(int K);
int64_t fp8_ls(int cols);
int64_t fp8_scale_off(int K, int cols);
int64_t fp8_bytes(int K, int cols);  // whole staging buffer, a multiple of 256
Pieces fp8_pieces(const void *buf, int K, int cols);
int gemm_split8(hipStream_t s, int K, int cols, const float *X, int64_t ldx, void *buf, int *launches);

// int8 staging (GEMM_I8X4) of an operand whose K runs down its columns:
// column j of X gets the exponent e_j with max |X(:,j)| 2^e_j in [32, 64),
// and slice c of its scaled values (rows padded with zeros to a multiple of
// 64) sits at buf + c * i8_stride(K, cols) + j * i8_ld(K), bytes; the int32
// e_j at buf + i8_exp_off(K, cols) + 4 j.
constexpr int kI8MaxK = 131008;  // exact int32 sums: 4 x 64 x 64 x K < 2^31 (2^17 - 64)
int64_t i8_ld(int K);
int64_t i8_stride(int K, int cols);
int64_t i8_exp_off(int K, int cols);
int64_t i8_bytes(int K, int cols);  // whole staging buffer, a multiple of 256
Pieces i8_pieces(const void *buf, int K, int cols);
int gemm_split_i8(hipStream_t s, int K, int cols, const float *X, int64_t ldx, void *buf, int *launches);
// The same slices into an explicit layout: column j at out + j * ld, slice c
// at + c * stride, its exponent at exps[j] (a column range of a wider array).
int gemm_split_i8_into(hipStream_t s, int K, int cols, const float *X, int64_t ldx, void *out, int64_t ld,
                       int64_t stride, int *exps, int *launches);

struct ColexpArgs {
    int m;
    const float *X;
    int64_t ldx;
    int *e;
};

// Grouped launches (64x64 tiles, no split-K): task b = (args index, tile row,
// tile col, split) for the GEMM, (args index, column) for colexp.  proj
// selects the K-major A operand (C = A^T B); otherwise A is M-major (C += alpha A B).
int gemm_grouped(hipStream_t s, int prec, bool proj, bool ch64, const GemmArgs *args, const int4 *tasks,
                 int ntasks);
int colexp_grouped(hipStream_t s, const ColexpArgs *args, const int2 *tasks, int ntasks);

// Floats of workspace the projection needs for its split-K partials.
int64_t gemm_proj_work_floats(int m, int n1, int n2, int chunk);

// 16-bit pieces per value of a precision (1, 2 or 3; 0 for FP32 and FP8).
int gemm_pieces(int prec);

// Splits X (rows x cols FP32, column-major, ldx), scaled by 2^a_exp (col_exp
// null) or 2^col_exp[j], into the MFMA input pieces of `prec`: piece c at
// out + c * stride, column-major with ld `ldo` (trans false: element (i, j)
// at j * ldo + i) or transposed (element (i, j) at i * ldo + j).  out, ldo
// and stride are multiples of 8 elements; the transposed form also writes
// the zero padding of j up to a multiple of 64 (ldo covers it).
int gemm_split(hipStream_t s, int prec, int rows, int cols, const float *X, int64_t ldx, int a_exp,
               const int *col_exp, uint16_t *out, int64_t ldo, int64_t stride, bool trans, int *launches);

// The transposed split writing exactly `cols` columns of `out` (no zero
// padding): a column range of a wider array, e.g. persistent pieces of Q^T.
int gemm_split_t_exact(hipStream_t s, int prec, int rows, int cols, const float *X, int64_t ldx, int a_exp,
                       uint16_t *out, int64_t ldo, int64_t stride, int *launches);

// The transposed split (zero-padded like gemm_split) of X's upper triangle
// (n x n) with `shift` subtracted from the diagonal and zeros below it.
int gemm_split_t_tri(hipStream_t s, int prec, int n, const float *X, int64_t ldx, int a_exp, float shift,
                     uint16_t *out, int64_t ldo, int64_t stride, int *launches);

// gemm_colexp and gemm_split (not transposed) of the same X in one launch:
// col_exp[j] for each column and the pieces scaled by 2^col_exp[j]
// (scaled precisions only).
int gemm_split_exp(hipStream_t s, int prec, int rows, int cols, const float *X, int64_t ldx, int *col_exp,
                   uint16_t *out, int64_t ldo, int64_t stride, int *launches);

// C (n1 x n2) = Q^T B.  a_exp scales Q by 2^a_exp before the cast; b_exp
// (per column of B, may be null) likewise.  chunk > 0 folds FP32 MFMA
// accumulators into FP64 every `chunk` rows of K and reduces split-K in FP64.
// qh, bh (optional): Q and B pre-split (K-major: Q column-major, B
// column-major), for the same bits without the in-kernel conversion.
int gemm_proj(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *B,
              int64_t ldb, float *C, int64_t ldc, int chunk, const int *b_exp, int a_exp, float *work,
              int64_t wfloats, int *launches, const Pieces *qh = nullptr, const Pieces *bh = nullptr);

// B (m x n2) -= Q C with Q m x n1, C n1 x n2.  f64 subtracts in FP64.
// qh (Q transposed), ch (C column-major): optional pre-split operands.
int gemm_update(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *C,
                int64_t ldc, float *B, int64_t ldb, bool f64, const int *c_exp, int a_exp, int *launches,
                const Pieces *qh = nullptr, const Pieces *ch = nullptr);

// C (M x N) += alpha * A B with A M x K (M-contiguous), B K x N.  tri: A and B
// are upper triangular (C's strictly-lower tiles are left untouched).
int gemm_nn(hipStream_t s, int prec, int M, int N, int K, float alpha, const float *A, int64_t lda,
            const float *B, int64_t ldb, float *C, int64_t ldc, bool f64, const int *b_exp, int a_exp, bool tri,
            int *launches, const Pieces *ah = nullptr, const Pieces *bh = nullptr);

// e[j] such that max |X(:,j)| 2^e[j] lies in [2^14, 2^15).
int gemm_colexp(hipStream_t s, int m, int n, const float *X, int64_t ldx, int *e, int *launches);
