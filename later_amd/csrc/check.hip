// -check metrics of test/test_qr.cu:216-268 with FP64 accumulation:
//   checkOtho   ||I - Q^T Q||_F            (the reference prints it / n)
//   checkResult ||A - Q triu(R)||_F / ||A||_F
// The reference computes both in FP32 through cuBLAS; here every product of
// two FP32 values is exact in FP64 and the sums are FP64, so the metric is
// not polluted by its own rounding at the 1e-6 level it has to resolve.
// Partial sums are combined in a fixed order (no atomics): results are
// bit-reproducible.
#include "common.h"

namespace {

constexpr int T = 64;   // output tile
constexpr int KT = 32;  // k slab

// Sum of (G - I)^2 over the 64x64 tile (ti, tj), ti <= tj, of G = Q^T Q.
__global__ __launch_bounds__(256) void orth_tile_kernel(int m, int n, const float *Q, int64_t ldq, double *part)
{
    // Decode the upper-triangular tile index.
    const int nt = (n + T - 1) / T;
    int b = blockIdx.x, ti = 0;
    while (b >= nt - ti) { b -= nt - ti; ++ti; }
    const int tj = ti + b;
    __shared__ float Qi[KT][T + 1], Qj[KT][T + 1];
    __shared__ double red[256];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < m; k0 += KT) {
        for (int e = threadIdx.x; e < KT * T; e += 256) {
            const int kk = e % KT, c = e / KT;
            const int k = k0 + kk;
            const int ci = ti * T + c, cj = tj * T + c;
            Qi[kk][c] = (k < m && ci < n) ? Q[k + (int64_t)ci * ldq] : 0.0f;
            Qj[kk][c] = (k < m && cj < n) ? Q[k + (int64_t)cj * ldq] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < KT; ++kk) {
            double a[4], bb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = (double)Qi[kk][tx * 4 + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) bb[c] = (double)Qj[kk][ty * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], bb[c], acc[r][c]);
        }
        __syncthreads();
    }
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = ti * T + tx * 4 + r, j = tj * T + ty * 4 + c;
            if (i < n && j < n) {
                const double d = acc[r][c] - (i == j ? 1.0 : 0.0);
                s += d * d;
            }
        }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = (ti == tj ? 1.0 : 2.0) * red[0];
}

// Tile (ti over rows of A, tj over columns) of A - Q triu(R): writes the sum
// of squared residuals and of squared A entries.
__global__ __launch_bounds__(256) void bwd_tile_kernel(int m, int n, const float *A, int64_t lda, const float *Q,
                                                       int64_t ldq, const float *R, int64_t ldr, double *part)
{
    const int ti = blockIdx.x, tj = blockIdx.y;
    __shared__ float Qs[KT][T + 1], Rs[KT][T + 1];
    __shared__ double red[2][256];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4] = {};
    const int kend = min(n, (tj + 1) * T);  // R upper triangular
    for (int k0 = 0; k0 < kend; k0 += KT) {
        for (int e = threadIdx.x; e < KT * T; e += 256) {
            const int r = e % T, kk = e / T;  // Qs[kk][r]: Q(row, k)
            const int row = ti * T + r, k = k0 + kk;
            Qs[kk][r] = (row < m && k < kend) ? Q[row + (int64_t)k * ldq] : 0.0f;
        }
        for (int e = threadIdx.x; e < KT * T; e += 256) {
            const int kk = e % KT, c = e / KT;  // Rs[kk][c]: R(k, col), zero below the diagonal
            const int k = k0 + kk, col = tj * T + c;
            Rs[kk][c] = (k < kend && col < n && k <= col) ? R[k + (int64_t)col * ldr] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < KT; ++kk) {
            double a[4], bb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = (double)Qs[kk][tx * 4 + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) bb[c] = (double)Rs[kk][ty * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fma(a[r], bb[c], acc[r][c]);
        }
        __syncthreads();
    }
    double s = 0.0, sa = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = ti * T + tx * 4 + r, j = tj * T + ty * 4 + c;
            if (i < m && j < n) {
                const double a = (double)A[i + (int64_t)j * lda];
                const double d = a - acc[r][c];
                s += d * d;
                sa += a * a;
            }
        }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = sa;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    const int64_t id = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    const int64_t nb = (int64_t)gridDim.x * gridDim.y;
    if (threadIdx.x == 0) {
        part[id] = red[0][0];
        part[nb + id] = red[1][0];
    }
}

// out[v] = sum of part[v*n .. v*n+n-1] in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void sum_kernel(int64_t n, int nvec, const double *part, double *out)
{
    __shared__ double red[256];
    for (int v = 0; v < nvec; ++v) {
        double s = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += 256) s += part[v * n + i];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[v] = red[0];
        __syncthreads();
    }
}

} // namespace

LATER_API int later_check_orth(laterCtxt ctxt, int m, int n, const float *Q, int ldq, double *orth)
{
    if (m < 1 || n < 1 || ldq < m || !Q || !orth) return LATER_ERR_SHAPE;
    hipStream_t s = ctx_stream(ctxt);
    const int64_t nt = (n + T - 1) / T;
    const int64_t nb = nt * (nt + 1) / 2;
    double *d = nullptr;
    HIP_OK(hipMalloc(&d, sizeof(double) * (nb + 1)));
    hipLaunchKernelGGL(orth_tile_kernel, dim3((unsigned)nb), dim3(256), 0, s, m, n, Q, (int64_t)ldq, d);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, nb, 1, (const double *)d, d + nb);
    double h = 0.0;
    hipError_t e = hipMemcpyAsync(&h, d + nb, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return LATER_ERR_DEVICE;
    *orth = sqrt(h);
    return LATER_OK;
}

LATER_API int later_check_backward(laterCtxt ctxt, int m, int n, const float *A, int lda, const float *Q,
                                   int ldq, const float *R, int ldr, double *backward)
{
    if (m < 1 || n < 1 || lda < m || ldq < m || ldr < n || !A || !Q || !R || !backward) return LATER_ERR_SHAPE;
    hipStream_t s = ctx_stream(ctxt);
    dim3 grid((unsigned)((m + T - 1) / T), (unsigned)((n + T - 1) / T));
    const int64_t nb = (int64_t)grid.x * grid.y;
    double *d = nullptr;
    HIP_OK(hipMalloc(&d, sizeof(double) * (2 * nb + 2)));
    hipLaunchKernelGGL(bwd_tile_kernel, grid, dim3(256), 0, s, m, n, A, (int64_t)lda, Q, (int64_t)ldq, R,
                       (int64_t)ldr, d);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, nb, 2, (const double *)d, d + 2 * nb);
    double h[2] = {0.0, 0.0};
    hipError_t e = hipMemcpyAsync(h, d + 2 * nb, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return LATER_ERR_DEVICE;
    *backward = (h[1] > 0.0) ? sqrt(h[0]) / sqrt(h[1]) : sqrt(h[0]);
    return LATER_OK;
}
