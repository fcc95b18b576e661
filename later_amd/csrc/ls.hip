// Least-squares solve on top of the QR: X = R^{-1} Q^T B.
//
// Q^T B is the node projection GEMM (gemm.hip).  The triangular solve is a
// blocked back substitution standing in for the reference's recursive TRSM
// (BLAS/later_rtrsm.cu:75-187): one workgroup per right-hand side keeps the
// column in LDS in FP64; each 64-row diagonal block is solved by one wave with
// the block held in registers (lane i owns row i, x_r is broadcast by a
// shuffle), then the rows above are updated by all waves with FP64 sums.
#include "common.h"
#include "gemm.h"

namespace {

constexpr int TB = 64;          // diagonal block
constexpr int NT = 256;
constexpr int kMaxN = 8192;     // FP64 column in LDS: 64 KiB

__global__ __launch_bounds__(NT) void trsm_upper_kernel(int n, const float *R, int64_t ldr, float *B, int64_t ldb,
                                                        int *status)
{
    extern __shared__ double x[];
    float *b = B + (int64_t)blockIdx.x * ldb;
    for (int i = threadIdx.x; i < n; i += NT) x[i] = (double)b[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nblk = (n + TB - 1) / TB;
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = blk * TB;
        const int jn = min(TB, n - j0);
        if (wave == 0) {
            // Rows j0 .. j0+jn-1: lane i owns row j0+i.
            double xi = (lane < jn) ? x[j0 + lane] : 0.0;
            for (int r = jn - 1; r >= 0; --r) {
                const float drr = R[(j0 + r) + (int64_t)(j0 + r) * ldr];
                double xr = __shfl(xi, r, 64);
                if (!(drr != 0.0f) && lane == 0 && status) atomicOr(status, 2);
                xr = (drr != 0.0f) ? xr / (double)drr : 0.0;
                if (lane == r) xi = xr;
                if (lane < r) xi -= (double)R[(j0 + lane) + (int64_t)(j0 + r) * ldr] * xr;
            }
            if (lane < jn) x[j0 + lane] = xi;
        }
        __syncthreads();
        // x[0:j0] -= R[0:j0, j0:j0+jn] x[j0:j0+jn]
        for (int i = threadIdx.x; i < j0; i += NT) {
            double s = 0.0;
            for (int c = 0; c < jn; ++c) s += (double)R[i + (int64_t)(j0 + c) * ldr] * x[j0 + c];
            x[i] -= s;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += NT) b[i] = (float)x[i];
}

__global__ void copy_kernel(int m, int n, const float *X, int64_t ldx, float *Y, int64_t ldy)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    Y[i + (int64_t)j * ldy] = X[i + (int64_t)j * ldx];
}

} // namespace

int trsm_upper(hipStream_t s, int n, int k, const float *R, int64_t ldr, float *B, int64_t ldb, int *status)
{
    if (n < 1 || k < 1 || n > kMaxN) return LATER_ERR_SHAPE;
    hipLaunchKernelGGL(trsm_upper_kernel, dim3(k), dim3(NT), sizeof(double) * n, s, n, R, ldr, B, ldb, status);
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

LATER_API int later_trsm_upper(laterCtxt ctxt, int n, int k, const float *R, int ldr, float *B, int ldb)
{
    if (n < 1 || k < 1 || ldr < n || ldb < n || !R || !B) return LATER_ERR_SHAPE;
    if (n > kMaxN) return LATER_ERR_SHAPE;
    return trsm_upper(ctx_stream(ctxt), n, k, R, ldr, B, ldb, nullptr);
}

LATER_API int later_ls_workspace(int m, int n, int k, const later_opts *opts, int64_t *lwork)
{
    int64_t lw = 0, lh = 0;
    int rc = later_rgsqrf_workspace(m, n, opts, &lw, &lh);
    if (rc) return rc;
    // later_gels reuses `work` after the QR: the status header (64 floats),
    // the exponents of B's columns, then the split-K partials of Q^T B.
    const int64_t proj = 64 + ((int64_t)k + 63) / 64 * 64 + gemm_proj_work_floats(m, n, k, 256);
    if (lwork) *lwork = (lw > proj ? lw : proj);
    return LATER_OK;
}

// Solves min ||A X - B||: A (m x n) <- Q, R (n x n) <- R, X (n x k) <- solution.
// B is read only.  Q^T B uses FP16 hi/lo products with FP64-folded chunks.
LATER_API int later_gels(laterCtxt ctxt, int m, int n, int k, float *A, int lda, float *R, int ldr, const float *B,
                         int ldb, float *X, int ldx, float *work, int64_t lwork, const later_opts *opts,
                         later_info *info)
{
    if (m < 1 || n < 1 || k < 1 || m < n || lda < m || ldr < n || ldb < m || ldx < n || !A || !R || !B || !X)
        return LATER_ERR_SHAPE;
    if (n > kMaxN) return LATER_ERR_SHAPE;
    int64_t need = 0;
    int rc = later_ls_workspace(m, n, k, opts, &need);
    if (rc) return rc;
    if (!work || lwork < need) return LATER_ERR_WORKSPACE;
    rc = later_rgsqrf_ex(ctxt, m, n, A, lda, R, ldr, work, lwork, nullptr, 0, opts, info);
    if (rc) return rc;
    hipStream_t s = ctx_stream(ctxt);
    // X = Q^T B (FP16 hi/lo split products, column scaling of B, FP64 chunks).
    int *bexp = (int *)work + 64;
    rc = gemm_colexp(s, m, k, B, ldb, bexp, nullptr);
    if (rc) return rc;
    const int64_t off = 64 + ((int64_t)k + 63) / 64 * 64;
    rc = gemm_proj(s, GEMM_F16X3, m, n, k, A, lda, B, ldb, X, ldx, 256, bexp, 14, work + off, lwork - off, nullptr);
    if (rc) return rc;
    return trsm_upper(s, n, k, R, ldr, X, ldx, (int *)work);
}
