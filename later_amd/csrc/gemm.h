// Internal interface of the node GEMMs (gemm.hip).
#pragma once
#include "common.h"

enum { GEMM_F32 = LATER_GEMM_F32, GEMM_F16 = LATER_GEMM_F16, GEMM_BF16 = LATER_GEMM_BF16,
       GEMM_F16X3 = LATER_GEMM_F16X3, GEMM_F16X6 = LATER_GEMM_F16X6 };

// Floats of workspace the projection needs for its split-K partials.
int64_t gemm_proj_work_floats(int m, int n1, int n2, int chunk);

// C (n1 x n2) = Q^T B.  a_exp scales Q by 2^a_exp before the cast; b_exp
// (per column of B, may be null) likewise.  chunk > 0 folds FP32 MFMA
// accumulators into FP64 every `chunk` rows of K and reduces split-K in FP64.
int gemm_proj(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *B,
              int64_t ldb, float *C, int64_t ldc, int chunk, const int *b_exp, int a_exp, float *work,
              int64_t wfloats, int *launches);

// B (m x n2) -= Q C with Q m x n1, C n1 x n2.  f64 subtracts in FP64.
int gemm_update(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *C,
                int64_t ldc, float *B, int64_t ldb, bool f64, const int *c_exp, int a_exp, int *launches);

// C (M x N) += alpha * A B with A M x K (M-contiguous), B K x N.  tri: A and B
// are upper triangular (C's strictly-lower tiles are left untouched).
int gemm_nn(hipStream_t s, int prec, int M, int N, int K, float alpha, const float *A, int64_t lda,
            const float *B, int64_t ldb, float *C, int64_t ldc, bool f64, const int *b_exp, int a_exp, bool tri,
            int *launches);

// e[j] such that max |X(:,j)| 2^e[j] lies in [2^14, 2^15).
int gemm_colexp(hipStream_t s, int m, int n, const float *X, int64_t ldx, int *e, int *launches);
