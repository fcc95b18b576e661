// Tall-skinny panel by Cholesky QR, twice (CholQR2), with FP64 Gram factors.
//
// An alternative to the block-MGS TSQR tree for the leaves of the accurate
// forms.  It has no sequential loop over columns across the grid: for an
// m x w panel (w <= 32)
//   K1  each workgroup forms the Gram matrix of its rows in FP64 (products of
//       FP32 values are exact in FP64); the last workgroup to finish adds the
//       partials in a fixed order, factors G = R1^T R1 (Cholesky, FP64, one
//       wave, columns in registers) and stores R1^{-1};
//   K2  every row is multiplied by R1^{-1} in FP64, Q1 is stored, and the
//       Gram matrix of Q1 is factored the same way: R2, R = R2 R1;
//   K3  Q = Q1 R2^{-1}.
// One pass of Cholesky QR loses orthogonality like kappa^2 u; the second pass
// starts from kappa(Q1) ~ 1 and restores O(u) for kappa(panel) up to ~1e7
// (FP64 Gram).  A non-positive pivot (rank deficiency, NaN/Inf) sets status
// bit 0 and yields a zero column, like the MGS panel.
//
// The "last workgroup" uses an atomic counter in the caller's workspace header
// (zeroed when the factorisation starts, reset by the last workgroup), so
// nothing waits on another workgroup and graph replays see a zero counter.
#include "common.h"
#include "panel.h"

namespace {

constexpr int kW = 32;          // max panel width
constexpr int kRows = 256;      // rows per block = threads per workgroup
constexpr int kMaxGroups = 256; // workgroups (and Gram partials) per launch

struct CholWork {
    double *partial;  // kMaxGroups x kW x kW
    double *rinv;     // kW x kW (current R^{-1}, upper)
    double *r1;       // kW x kW (first factor)
};

__host__ __device__ inline CholWork carve(float *ws)
{
    CholWork c;
    double *d = (double *)ws;
    c.partial = d;
    c.rinv = d + (int64_t)kMaxGroups * kW * kW;
    c.r1 = c.rinv + kW * kW;
    return c;
}

// Partial Gram of the rows this workgroup owns: G += B^T B over blocks
// b = g, g + G, ... of the m x w matrix held (per block) in LDS.  Each thread
// owns a 2 x 2 tile (2bi, 2bj) of G with bi <= bj.
__device__ __forceinline__ void gram_tile(const double (*Bs)[kW + 1], int rows, int w, double acc[4])
{
    const int t = threadIdx.x;
    // Tiles of the upper triangle of a 16 x 16 tile grid: t -> (bi, bj).
    int bi = 0, rem = t;
    while (bi < 16 && rem >= 16 - bi) { rem -= 16 - bi; ++bi; }
    if (bi >= 16) return;
    const int bj = bi + rem;
    const int i0 = 2 * bi, j0 = 2 * bj;
    if (i0 >= w || j0 >= w) return;
    for (int r = 0; r < rows; ++r) {
        const double a0 = Bs[r][i0], a1 = Bs[r][i0 + 1], b0 = Bs[r][j0], b1 = Bs[r][j0 + 1];
        acc[0] = fma(a0, b0, acc[0]);
        acc[1] = fma(a0, b1, acc[1]);
        acc[2] = fma(a1, b0, acc[2]);
        acc[3] = fma(a1, b1, acc[3]);
    }
}

__device__ __forceinline__ void gram_store(double *P, int w, const double acc[4])
{
    const int t = threadIdx.x;
    int bi = 0, rem = t;
    while (bi < 16 && rem >= 16 - bi) { rem -= 16 - bi; ++bi; }
    if (bi >= 16) return;
    const int bj = bi + rem;
    const int i0 = 2 * bi, j0 = 2 * bj;
    const double v[2][2] = {{acc[0], acc[1]}, {acc[2], acc[3]}};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int i = i0 + a, j = j0 + b;
            if (i < kW && j < kW) {
                P[i + j * kW] = v[a][b];
                P[j + i * kW] = v[a][b];
            }
        }
}

// Last workgroup: sum the partials in a fixed order, Cholesky-factor and
// invert.  Writes rinv (upper R^{-1}) and, when rout is given, R (upper).
__device__ void finish_cholesky(const double *partial, int ngroups, int w, double *rinv, double *rout, int *status)
{
    __shared__ double G[kW][kW + 1];
    for (int e = threadIdx.x; e < kW * kW; e += blockDim.x) {
        double s = 0.0;
        for (int g = 0; g < ngroups; ++g) s += partial[(int64_t)g * kW * kW + e];
        G[e % kW][e / kW] = s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        // Right-looking Cholesky G = L L^T by one wave.  Lane j keeps rows
        // k.. of column j of the trailing matrix as a window c[0 .. w-k-1]
        // (c[t] = G[k + t][j]); the window shifts down one row per step, so
        // the pivot row is always c[0] and every register index is static.
        // Lane k writes the finished column k of L to LDS (G[i][k], i >= k).
        double c[kW];
#pragma unroll
        for (int t = 0; t < kW; ++t) c[t] = (lane < w && t < w) ? G[t][lane] : 0.0;
        bool bad = false;
        for (int k = 0; k < w; ++k) {
            const double gkk = __shfl(c[0], k, 64);
            const bool ok = gkk > 0.0 && gkk < INFINITY;
            bad |= !ok;
            const double d = ok ? sqrt(gkk) : 0.0;
            const double inv = ok ? 1.0 / d : 0.0;
            const double ljk = c[0] * inv;  // L[j][k] (trailing matrix symmetric)
#pragma unroll
            for (int t = 0; t < kW; ++t) {
                const double lik = __shfl(c[t], k, 64) * inv;  // L[k + t][k]
                if (lane == k && k + t < w) G[k + t][k] = (t == 0) ? d : lik;
                if (lane == k && t < k) G[t][k] = 0.0;
                c[t] = (lane > k && t > 0) ? c[t] - lik * ljk : c[t];
            }
#pragma unroll
            for (int t = 0; t < kW - 1; ++t) c[t] = c[t + 1];
            c[kW - 1] = 0.0;
        }
        if (bad && lane == 0 && status) atomicOr(status, 1);
        // Columns w..31 of L (unused) are zero.
        if (lane >= w && lane < kW)
#pragma unroll
            for (int i = 0; i < kW; ++i) G[i][lane] = 0.0;
    }
    __syncthreads();
    // G[i][k] = L[i][k] (lower).  R = L^T upper: R[k][j] = L[j][k].
    // Column j of R^{-1} (upper): solve R x = e_j by back substitution, one
    // lane per column.
    if (threadIdx.x < kW) {
        const int j = threadIdx.x;
        // x = column j of R^{-1}: R x = e_j, R[i][k] = L[k][i]; x[i] = 0 for
        // i > j, so the static loops below never read an unset entry.
        double x[kW];
#pragma unroll
        for (int i = kW - 1; i >= 0; --i) {
            double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = i + 1; k < kW; ++k) s -= G[k][i] * x[k];
            const double rii = G[i][i];
            x[i] = (i <= j && j < w && rii != 0.0) ? s / rii : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kW; ++i) rinv[i + j * kW] = x[i];
        if (rout)
#pragma unroll
            for (int i = 0; i < kW; ++i) rout[i + j * kW] = (i <= j && j < w) ? G[j][i] : 0.0;
    }
}

// Loads block b of the m x w panel (rows b*256 ..) into LDS as FP64.
__device__ __forceinline__ int load_block(const float *A, int64_t lda, int m, int w, int b, double (*Bs)[kW + 1])
{
    const int r0 = b * kRows;
    const int rows = min(kRows, m - r0);
    const int t = threadIdx.x;
    const int rc = min(t, rows - 1);
    float v[kW];
#pragma unroll
    for (int k = 0; k < kW; ++k) v[k] = A[r0 + rc + (int64_t)min(k, w - 1) * lda];
#pragma unroll
    for (int k = 0; k < kW; ++k) Bs[t][k] = (t < rows && k < w) ? (double)v[k] : 0.0;
    return rows;
}

// Signals completion of this workgroup's partial; true for the last one.
__device__ __forceinline__ bool arrive_last(int *counter, int ngroups)
{
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const int old = atomicAdd(counter, 1);
        last = (old == ngroups - 1);
        if (last) *counter = 0;  // reset for the next panel (and graph replays)
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

__global__ __launch_bounds__(kRows) void chol_gram_kernel(int m, int w, const float *A, int64_t lda, float *ws,
                                                         int *counter, int *status)
{
    __shared__ double Bs[kRows][kW + 1];
    const CholWork cw = carve(ws);
    const int nblk = (m + kRows - 1) / kRows;
    const int G = gridDim.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = blockIdx.x; b < nblk; b += G) {
        const int rows = load_block(A, lda, m, w, b, Bs);
        __syncthreads();
        gram_tile(Bs, rows, w, acc);
        __syncthreads();
    }
    gram_store(cw.partial + (int64_t)blockIdx.x * kW * kW, w, acc);
    if (arrive_last(counter, G)) finish_cholesky(cw.partial, G, w, cw.rinv, cw.r1, status);
}

// Q1 = A R1^{-1} (stored), then the Gram of Q1 (from the FP64 values) is
// factored: R2; R (ldr) = R2 R1 in FP32.  With gram = false only Q1 is formed.
template <bool GRAM>
__global__ __launch_bounds__(kRows) void chol_apply_kernel(int m, int w, float *A, int64_t lda, float *ws,
                                                          float *R, int64_t ldr, int *counter, int *status)
{
    __shared__ double Bs[GRAM ? kRows : 1][kW + 1];
    __shared__ double Ri[kW][kW + 1];
    const CholWork cw = carve(ws);
    for (int e = threadIdx.x; e < kW * kW; e += kRows) Ri[e % kW][e / kW] = cw.rinv[e];
    __syncthreads();
    const int nblk = (m + kRows - 1) / kRows;
    const int G = gridDim.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = blockIdx.x; b < nblk; b += G) {
        const int r0 = b * kRows;
        const int rows = min(kRows, m - r0);
        const int t = threadIdx.x;
        const int rc = min(t, rows - 1);
        float v[kW];
#pragma unroll
        for (int k = 0; k < kW; ++k) v[k] = A[r0 + rc + (int64_t)min(k, w - 1) * lda];
        double q[kW];
#pragma unroll
        for (int j = 0; j < kW; ++j) q[j] = 0.0;
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            if (k >= w) break;
            const double vk = (double)v[k];
#pragma unroll
            for (int j = 0; j < kW; ++j)
                if (j >= k) q[j] = fma(vk, Ri[k][j], q[j]);
        }
        if (t < rows)
#pragma unroll
            for (int j = 0; j < kW; ++j)
                if (j < w) A[r0 + t + (int64_t)j * lda] = (float)q[j];
        if constexpr (GRAM) {
#pragma unroll
            for (int j = 0; j < kW; ++j) Bs[t][j] = (t < rows && j < w) ? q[j] : 0.0;
            __syncthreads();
            gram_tile(Bs, rows, w, acc);
            __syncthreads();
        }
    }
    if constexpr (GRAM) {
        gram_store(cw.partial + (int64_t)blockIdx.x * kW * kW, w, acc);
        if (arrive_last(counter, G)) {
            __shared__ double R2[kW * kW];
            finish_cholesky(cw.partial, G, w, cw.rinv, R2, status);
            __syncthreads();
            // R = R2 R1 (both upper), FP64 then rounded once.
            for (int e = threadIdx.x; e < w * w; e += kRows) {
                const int i = e % w, j = e / w;
                double s = 0.0;
                for (int k = i; k <= j; ++k) s += R2[i + k * kW] * cw.r1[k + j * kW];
                R[i + (int64_t)j * ldr] = (i <= j) ? (float)s : 0.0f;
            }
        }
    }
}

} // namespace

int64_t panel_chol_work_floats() { return ((int64_t)kMaxGroups * kW * kW + 2 * kW * kW) * 2 + 64; }

int panel_cholqr2(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work, int64_t lwork,
                  int *counters, int *status, int *launches)
{
    if (w < 1 || w > kW || m < w) return LATER_ERR_SHAPE;
    if (lwork < panel_chol_work_floats()) return LATER_ERR_WORKSPACE;
    const int nblk = (m + kRows - 1) / kRows;
    const int G = nblk < kMaxGroups ? nblk : kMaxGroups;
    hipLaunchKernelGGL(chol_gram_kernel, dim3(G), dim3(kRows), 0, s, m, w, (const float *)A, (int64_t)lda, work,
                       counters, status);
    hipLaunchKernelGGL(chol_apply_kernel<true>, dim3(G), dim3(kRows), 0, s, m, w, A, (int64_t)lda, work, R,
                       (int64_t)ldr, counters, status);
    hipLaunchKernelGGL(chol_apply_kernel<false>, dim3(G), dim3(kRows), 0, s, m, w, A, (int64_t)lda, work, R,
                       (int64_t)ldr, counters, status);
    HIP_OK(hipGetLastError());
    if (launches) *launches += 3;
    return LATER_OK;
}
