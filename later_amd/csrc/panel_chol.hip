// Tall-skinny panel by Cholesky QR, twice (CholQR2), with FP64 Gram factors.
//
// An alternative to the block-MGS TSQR tree for the leaves of the accurate
// forms, with no sequential loop over columns across the grid.  For an m x w
// panel (w <= 64; kernels for widths 32 and 64), three launches over the same
// G workgroups (G depends on m only, so every result is bit-reproducible):
//   K1  each workgroup forms the Gram matrix of its rows with FP64 MFMA
//       (products of FP32 values are exact in FP64) and stores the packed
//       upper triangle as its partial;
//   K2  every workgroup adds the G partials in the same fixed order, factors
//       G = R1^T R1 and inverts R1 (FP64, one wave per 32 x 32 diagonal
//       block, columns in registers), forms q = a R1^{-1} for each of its
//       rows in FP64 and stores Q1 in FP32, then forms the partial Gram of
//       the stored Q1;
//   K3  the same for Q1: R2, Q = Q1 R2^{-1}, R = R2 R1.
// All of the small-matrix arithmetic is FP64, so multiplying by the explicit
// inverse adds errors of order u_64 kappa, far below the FP32 rounding of Q.
// Workgroups never wait for each other: the kernel boundaries order the
// partials, so there are no atomics and no device-scope fences (which would
// write back and invalidate the L2 of every XCD).  Tall panels (more than 32
// workgroups, up to 256) add their partials in a small reduction kernel
// before K2 and K3 instead of in every workgroup.
//
// One pass of Cholesky QR loses orthogonality like kappa^2 u_64 (the Gram and
// its factor are FP64) plus the FP32 rounding of Q; the second pass starts
// from kappa(Q1) ~ 1 and restores O(u) for kappa(panel) up to ~1e7.  Panels
// known to be near-orthonormal (a repeated factorisation of a Q) take one
// pass (`once`).
// A pivot that is not positive and finite, or that cancels to below
// 64 u_64 of the column's squared norm (numerical rank deficiency in the FP64
// Gram), sets status bit 0 and yields a zero column, like the MGS panel.
#include "chol_dev.h"
#include "common.h"
#include "panel.h"

namespace {

// Sum of more than kDirect partial Grams, in a fixed tree: workgroup b owns
// entries 32b .. 32b+31; thread (q, e) adds groups q*G/8 .. (q+1)*G/8 - 1 in
// order, then the eight part sums are added in order of q.
template <int W> __global__ __launch_bounds__(256) void chol_reduce_kernel(const double *part, int ngroups, double *out)
{
    __shared__ double ps[8][32];
    const int e = blockIdx.x * 32 + (threadIdx.x & 31), q = threadIdx.x >> 5;
    const int per = (ngroups + 7) / 8, g0 = q * per, g1 = min(ngroups, g0 + per);
    double s = 0.0;
    if (e < Shape<W>::P) {
        constexpr int kB = 32;  // one round of loads for up to 256 partials
        for (int g = g0; g < g1; g += kB) {
            double v[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) v[u] = part[(int64_t)min(g + u, g1 - 1) * Shape<W>::P + e];
#pragma unroll
            for (int u = 0; u < kB; ++u)
                if (g + u < g1) s += v[u];
        }
    }
    ps[q][threadIdx.x & 31] = s;
    __syncthreads();
    if (threadIdx.x < 32 && e < Shape<W>::P) {
        double t = ps[0][threadIdx.x];
#pragma unroll
        for (int k = 1; k < 8; ++k) t += ps[k][threadIdx.x];
        out[e] = t;
    }
}

template <int W>
__global__ __launch_bounds__(kThreads) void chol_gram_kernel(int m, int w, const float *A, int64_t lda, float *ws)
{
    constexpr int CT = Shape<W>::CT;
    __shared__ double Cb[4 * Shape<W>::NT * 4 * 64];
    const CholWork cw = carve<W>(ws);
    const int nblk = (m + kRows - 1) / kRows;
    const int G = gridDim.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15;
    int64_t col[CT];
    bool ok[CT];
#pragma unroll
    for (int x = 0; x < CT; ++x) {
        col[x] = (int64_t)min(16 * x + c, w - 1) * lda;
        ok[x] = 16 * x + c < w;
    }
    Gram<W> g;
    gram_zero<W>(g);
    for (int b = blockIdx.x; b < nblk; b += G) {
#pragma unroll
        for (int sg = 0; sg < kTT; ++sg) {
            const int row0 = b * kRows + (wv * kTT + sg) * 16 + 4 * (lane >> 4);
            double a[CT][4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int row = row0 + p, rc = min(row, m - 1);
#pragma unroll
                for (int x = 0; x < CT; ++x) {
                    const float v = A[rc + col[x]];
                    a[x][p] = (row < m && ok[x]) ? (double)v : 0.0;
                }
            }
            gram_step<W>(g, a);
        }
    }
    gram_finish<W>(g, Cb, cw.p1 + (int64_t)blockIdx.x * Shape<W>::P);
}

// One Cholesky QR pass over the rows of this workgroup.
//   PASS 1: Gram of A -> R1 (stored to r1 by workgroup 0); Q1 = A R1^{-1}
//           stored; partials p2 of the stored Q1.
//   PASS 2: Gram of Q1 -> R2; Q = Q1 R2^{-1} stored; R = R2 R1 (ldr, FP32).
//   PASS 0: the only pass (panels with condition number near 1): Gram of A
//           -> R1; Q = A R1^{-1} stored; R = R1.
// The product with R^{-1} runs on FP64 MFMA as D = R^{-T} A^T per 16 x 16
// tile: the A operand (slices of R^{-1}) stays in registers for the whole
// block, the B operand is the block's rows, and D's lanes walk 16
// consecutive rows of one column (64-byte stores).
template <int PASS, int W>
__global__ __launch_bounds__(kThreads) void chol_apply_kernel(int m, int w, float *A, int64_t lda, float *ws, float *R,
                                                          int64_t ldr, int *status)
{
    constexpr int CT = Shape<W>::CT, KS = Shape<W>::KS, LDG = Shape<W>::LDG;
    __shared__ __attribute__((aligned(16))) char smem[ApplyLds<W>::kBytes];
    double *Gs = (double *)smem;      // W x (W + 1)
    double *Rs = Gs + W * LDG;        // W x W
    double *Ris = Rs + W * W;         // W x W
    double *d0 = Ris + W * W;         // W
    float *Qs = (float *)smem;        // W x (kRows + 4), after the factor
    double *Cb = (double *)smem;      // after the last block
    const CholWork cw = carve<W>(ws);
    const int nblk = (m + kRows - 1) / kRows;
    const int G = gridDim.x;
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;

    // The first block's rows are in flight while the factor is formed.
    float av[kTT][KS];
    load_block_mfma<W>(A, lda, m, w, blockIdx.x * kRows, av);

    if (G > kDirect) sum_partials<W>(cw.g, 1, w, Gs);  // reduced by chol_reduce_kernel
    else sum_partials<W>(PASS == 2 ? cw.p2 : cw.p1, G, w, Gs);
    __syncthreads();
    bool first_order = false;
    if constexpr (PASS != 1) {
        // The panel (Q1, or a near-orthonormal input) has Gram I + E.  When
        // max |E| <= 1e-5, R2 = I + U and R2^{-1} = I - U with
        // U = triu(E, 1) + diag(E) / 2 agree with the Cholesky factor and its
        // inverse to O(w |E|^2) <= 6.4e-9, far below the FP32 rounding of Q;
        // otherwise (and for a zero column of Q1) the Cholesky runs.  Every
        // workgroup sees the same E, so all take the same branch.
        __shared__ float emax_w[4];
        float e = 0.0f;
        for (int q = t; q < W * W; q += kThreads) {
            const int i = q % W, j = q / W;
            if (i <= j) e = fmaxf(e, (float)fabs(Gs[i * LDG + j] - (i == j ? 1.0 : 0.0)));
        }
#pragma unroll
        for (int mask = 32; mask > 0; mask >>= 1) e = fmaxf(e, __shfl_xor(e, mask, 64));
        if ((t & 63) == 0) emax_w[t >> 6] = e;
        __syncthreads();
        const float emax = fmaxf(fmaxf(emax_w[0], emax_w[1]), fmaxf(emax_w[2], emax_w[3]));
        first_order = emax <= 1e-5f;  // NaN compares false: Cholesky path
        if (first_order) {
            for (int q = t; q < W * W; q += kThreads) {
                const int i = q / W, j = q % W;
                const double u = (i < j) ? Gs[i * LDG + j] : (i == j ? 0.5 * (Gs[i * LDG + i] - 1.0) : 0.0);
                Rs[i * W + j] = (i == j ? 1.0 : 0.0) + u;
                Ris[i * W + j] = (i == j ? 1.0 : 0.0) - u;
            }
        }
    }
    if (!first_order) {
        const bool bad = factor<W>(Gs, Rs, Ris, d0);
        if (bad && t == 0 && blockIdx.x == 0 && status) atomicOr(status, 1);
    }
    __syncthreads();
    if constexpr (PASS == 1) {
        if (blockIdx.x == 0)
            for (int e = t; e < W * W; e += kThreads) cw.r1[e] = Rs[e];
    } else {
        // R (FP32, ldr): R1 itself (PASS 0) or R2 R1 (PASS 2, both upper,
        // FP64 then rounded once), four columns per round, column j by
        // workgroup (j / 4) mod G, with column j of R1 staged in LDS (over
        // the Gram, no longer needed).
        for (int j0 = blockIdx.x * 4; j0 < w; j0 += G * 4) {
            const int c = t >> 6, j = j0 + c, i = t & 63;
            if constexpr (PASS == 2) {
                if (j < w && i < W) Gs[c * LDG + i] = (i <= j) ? cw.r1[i * W + j] : 0.0;
                __syncthreads();
            }
            if (j < w && i < w) {
                double s = 0.0;
                if constexpr (PASS == 2) {
                    // Full k range: R2's lower and the staged column's
                    // below-j entries are zeros, so the sum is the same.
#pragma unroll
                    for (int k = 0; k < W; ++k) s = fma(Rs[i * W + k], Gs[c * LDG + k], s);
                } else {
                    s = Rs[i * W + j];
                }
                R[i + (int64_t)j * ldr] = (i <= j) ? (float)s : 0.0f;
            }
            if constexpr (PASS == 2) __syncthreads();
        }
    }

    // A-operand slices of R^{-1}: rv[ct][s] = R^{-1}[4 s + l/16][16 ct + l%16]
    // (zero for k > column, so column tile ct needs k-steps 0 .. 4 ct + 3).
    double rv[CT][KS];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int s2 = 0; s2 < 4 * (ct + 1); ++s2) rv[ct][s2] = Ris[(4 * s2 + (lane >> 4)) * W + 16 * ct + (lane & 15)];

    Gram<W> g;
    gram_zero<W>(g);
#pragma unroll 1
    for (int b = blockIdx.x; b < nblk; b += G) {
        __syncthreads();  // the factor, or Qs of the previous block, has been read
        const int r0 = b * kRows;
#pragma unroll
        for (int tt = 0; tt < kTT; ++tt) {
            const int lrow = (wv * kTT + tt) * 16 + (lane & 15);  // row within the block
            const int row = r0 + lrow;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                doublex4 d = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s2 = 0; s2 < 4 * (ct + 1); ++s2)
                    d = __builtin_amdgcn_mfma_f64_16x16x4f64(rv[ct][s2], (double)av[tt][s2], d, 0, 0, 0);
                // d[r] = Q[row][16 ct + l/16 + 4 r]
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * ct + (lane >> 4) + 4 * r;
                    const bool ok = row < m && col < w;
                    const float q = ok ? (float)d[r] : 0.0f;
                    if (ok) A[row + (int64_t)col * lda] = q;
                    if constexpr (PASS == 1) Qs[col * (kRows + 4) + lrow] = q;
                }
            }
        }
        if (b + G < nblk) load_block_mfma<W>(A, lda, m, w, (b + G) * kRows, av);
        if constexpr (PASS == 1) {
            // Gram of the stored (FP32) Q1 rows of this block.
            __syncthreads();
#pragma unroll
            for (int sg = 0; sg < kTT; ++sg) gram_lds<W>(g, Qs, (wv * kTT + sg) * 16);
        }
    }
    if constexpr (PASS == 1) {
        __syncthreads();
        gram_finish<W>(g, Cb, cw.p2 + (int64_t)blockIdx.x * Shape<W>::P);
    }
}

template <int W>
void launch_cholqr(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work, int *status,
                   int G, bool once)
{
    const CholWork cw = carve<W>(work);
    const dim3 rgrid((Shape<W>::P + 31) / 32);
    hipLaunchKernelGGL(chol_gram_kernel<W>, dim3(G), dim3(kThreads), 0, s, m, w, (const float *)A, (int64_t)lda, work);
    if (G > kDirect) hipLaunchKernelGGL(chol_reduce_kernel<W>, rgrid, dim3(256), 0, s, cw.p1, G, cw.g);
    if (once) {
        hipLaunchKernelGGL((chol_apply_kernel<0, W>), dim3(G), dim3(kThreads), 0, s, m, w, A, (int64_t)lda, work, R,
                           (int64_t)ldr, status);
        return;
    }
    hipLaunchKernelGGL((chol_apply_kernel<1, W>), dim3(G), dim3(kThreads), 0, s, m, w, A, (int64_t)lda, work, R,
                       (int64_t)ldr, status);
    if (G > kDirect) hipLaunchKernelGGL(chol_reduce_kernel<W>, rgrid, dim3(256), 0, s, cw.p2, G, cw.g);
    hipLaunchKernelGGL((chol_apply_kernel<2, W>), dim3(G), dim3(kThreads), 0, s, m, w, A, (int64_t)lda, work, R,
                       (int64_t)ldr, status);
}

} // namespace

int64_t panel_chol_work_floats(int w) { return w <= 32 ? work_floats<32>() : work_floats<64>(); }

int panel_cholqr2(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work, int64_t lwork,
                  int *status, int *launches, bool once)
{
    if (w < 1 || w > kWMax || m < w) return LATER_ERR_SHAPE;
    if (lwork < panel_chol_work_floats(w)) return LATER_ERR_WORKSPACE;
    const int nblk = (m + kRows - 1) / kRows;
    const int G = nblk < kMaxGroups ? nblk : kMaxGroups;
    if (w <= 32) launch_cholqr<32>(s, m, w, A, lda, R, ldr, work, status, G, once);
    else launch_cholqr<64>(s, m, w, A, lda, R, ldr, work, status, G, once);
    HIP_OK(hipGetLastError());
    if (launches) *launches += (once ? 2 : 3) + (G > kDirect ? (once ? 1 : 2) : 0);
    return LATER_OK;
}
