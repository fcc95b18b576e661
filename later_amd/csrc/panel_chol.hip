// Tall-skinny panel by Cholesky QR, twice (CholQR2), with FP64 Gram factors.
//
// An alternative to the block-MGS TSQR tree for the leaves of the accurate
// forms, with no sequential loop over columns across the grid.  For an m x w
// panel (w <= 64; kernels for widths 32 and 64), per pass (the partition of
// the rows depends on m only, so every result is bit-reproducible):
//   gram    G workgroups each form the Gram matrix of their 64-row blocks
//           with FP64 MFMA (products of FP32 values are exact in FP64) from
//           coalesced loads staged in LDS, and store the packed upper
//           triangle as a partial (a fixed-tree reduction kernel adds them
//           when there are more than 32);
//   factor  one workgroup adds the partials in a fixed order, factors
//           G = R^T R and inverts R (FP64, one wave per 32 x 32 diagonal
//           block, columns in registers), and stores R^{-1} (and R);
//   apply   up to 1024 workgroups form q = a R^{-1} for their rows on FP64
//           MFMA and store Q in FP32; in the first of two passes they also
//           form the partial Grams of the stored Q1 for the second, which
//           repeats the factor and apply on Q1 and sets R = R2 R1.
// All of the small-matrix arithmetic is FP64, so multiplying by the explicit
// inverse adds errors of order u_64 kappa, far below the FP32 rounding of Q.
// Workgroups never wait for each other: the kernel boundaries order the
// partials and the factor, so there are no atomics and no device-scope
// fences (which would write back and invalidate the L2 of every XCD).  The
// factor runs once per pass instead of in every apply workgroup, so the
// apply kernels hold little LDS and several of them share a CU.
//
// One pass of Cholesky QR loses orthogonality like kappa^2 u_64 (the Gram and
// its factor are FP64) plus the FP32 rounding of Q; the second pass starts
// from kappa(Q1) ~ 1 and restores O(u) for kappa(panel) up to ~1e7.  Panels
// known to be near-orthonormal (a repeated factorisation of a Q) take one
// pass (`once`).
// A pivot that is not positive and finite, or that cancels to below
// 64 u_64 of the column's squared norm (numerical rank deficiency in the FP64
// Gram), sets status bit 0 and yields a zero column, like the MGS panel.
#include <type_traits>

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

// Rows r0 .. r0 + 63 of the panel's (up to W) columns, loaded with 16-byte
// loads (16 lanes cover the 64 rows of one column) into registers, then
// stored column-major into LDS (stride kRows + 4) for the Gram MFMA
// operands; zeros past m and past w.
template <int W> struct BlockRegs {
    static constexpr int NV = W / 16;  // float4 per thread
    floatx4 v[NV];
    __device__ __forceinline__ void load(const float *A, int64_t lda, int m, int w, int r0)
    {
        const int t = threadIdx.x, row = r0 + 4 * (t & 15);
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int c = (t >> 4) + 16 * j;
            const float *src = A + (int64_t)min(c, w - 1) * lda + min(row, m - 1);
            if (c < w && row + 4 <= m && ((((uintptr_t)src) & 15) == 0)) {
                v[j] = *(const floatx4 *)src;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    v[j][e] = (c < w && row + e < m) ? A[(int64_t)min(c, w - 1) * lda + min(row + e, m - 1)] : 0.0f;
            }
        }
    }
    __device__ __forceinline__ void store_ld(float *Qs, int ld) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int j = 0; j < NV; ++j)
            *(floatx4 *)&Qs[((t >> 4) + 16 * j) * ld + 4 * (t & 15)] = v[j];
    }
    __device__ __forceinline__ void store(float *Qs) const { store_ld(Qs, kRows + 4); }
};

// Partial Gram of the rows of this workgroup (64-row blocks b, b + G, ...)
// into p1[blockIdx.x]: coalesced loads two blocks ahead, the block staged in
// LDS, FP64 MFMA per 16-row group of each wave.
template <int W>
__global__ __launch_bounds__(kThreads) void chol_gram_kernel(int m, int w, const float *A, int64_t lda, float *ws)
{
    constexpr int kQs = W * (kRows + 4) * 4, kCb = 2 * Shape<W>::NT * 4 * 64 * 8;
    __shared__ __attribute__((aligned(16))) char smem[kQs > kCb ? kQs : kCb];
    float *Qs = (float *)smem;
    const CholWork cw = carve<W>(ws);
    const int nblk = (m + kRows - 1) / kRows;
    const int G = gridDim.x;
    const int wv = threadIdx.x >> 6;
    Gram<W> g;
    gram_zero<W>(g);
    BlockRegs<W> r[2];
    int b = blockIdx.x;
    if (b < nblk) r[0].load(A, lda, m, w, b * kRows);
    if (b + G < nblk) r[1].load(A, lda, m, w, (b + G) * kRows);
    for (int it = 0; b < nblk; b += G, ++it) {
        __syncthreads();  // the previous block has been read
        if (it & 1) r[1].store(Qs); else r[0].store(Qs);
        __syncthreads();
        if (b + 2 * G < nblk) {
            if (it & 1) r[1].load(A, lda, m, w, (b + 2 * G) * kRows);
            else r[0].load(A, lda, m, w, (b + 2 * G) * kRows);
        }
#pragma unroll
        for (int sg = 0; sg < kTT; ++sg) gram_lds<W>(g, Qs, (wv * kTT + sg) * 16);
    }
    __syncthreads();
    gram_finish<W>(g, (double *)smem, cw.p1 + (int64_t)blockIdx.x * Shape<W>::P);
}

// The factor of one Cholesky QR pass, by one workgroup, from the G partial
// Grams (or their reduced sum when G > kDirect):
//   PASS 1: R1 -> r1; PASS 0: R = R1 (ldr, FP32); PASS 2: R = R2 R1.
// R^{-1} goes to ri for the apply kernel.  A panel known to be
// near-orthonormal (PASS 0 or 2) whose Gram deviation E = G - I has
// max |E| <= 1e-5 uses R = I + U, R^{-1} = I - U with U = triu(E, 1) +
// diag(E) / 2, which agree with the Cholesky factor and its inverse to
// O(w |E|^2) <= 6.4e-9, far below the FP32 rounding of Q.
template <int PASS, int W>
__global__ __launch_bounds__(kThreads) void chol_factor_kernel(int w, float *ws, int G, float *R, int64_t ldr,
                                                            int *status)
{
    constexpr int LDG = Shape<W>::LDG;
    __shared__ double Gs[W * LDG], Rs[W * W], Ris[W * W], d0[W];
    __shared__ float emax_w[4];
    const CholWork cw = carve<W>(ws);
    const int t = threadIdx.x;
    if (G > kDirect) sum_partials<W>(cw.g, 1, w, Gs);
    else sum_partials<W>(PASS == 2 ? cw.p2 : cw.p1, G, w, Gs);
    __syncthreads();
    bool first_order = false;
    if constexpr (PASS != 1) {
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
        if (bad && t == 0 && status) atomicOr(status, 1);
    }
    __syncthreads();
    for (int e = t; e < W * W; e += kThreads) cw.ri[e] = Ris[e];
    if constexpr (PASS == 1) {
        for (int e = t; e < W * W; e += kThreads) cw.r1[e] = Rs[e];
    } else {
        // R (FP32, ldr): R1 itself (PASS 0) or R2 R1 (PASS 2, both upper, FP64
        // then rounded once).
        for (int e = t; e < w * w; e += kThreads) {
            const int i = e % w, j = e / w;
            double s = 0.0;
            if constexpr (PASS == 2) {
                for (int k = i; k <= j; ++k) s = fma(Rs[i * W + k], cw.r1[k * W + j], s);
            } else {
                s = Rs[i * W + j];
            }
            R[i + (int64_t)j * ldr] = (i <= j) ? (float)s : 0.0f;
        }
    }
}

// Q = A R^{-1} over the rows of this workgroup (64-row blocks b, b + G, ...),
// with R^{-1} from ri.  Each block is loaded with coalesced 16-byte row
// loads two blocks ahead (as in the Gram kernel) and staged column-major in
// LDS with a row stride of kRows + 16 floats, so that a wave's MFMA operand
// reads (16 rows x 4 columns) cover 64 distinct banks.  The product runs on
// FP64 MFMA as D = R^{-T} A^T per 16 x 16 tile: the A operand (slices of
// R^{-1}) stays in registers, the B operand is the block's rows, and the
// four column tiles' accumulation chains are interleaved k-step by k-step
// (up to four independent MFMAs in flight); D's lanes walk 16 consecutive
// rows of one column (64-byte stores).  PASS 1 also forms the partial Gram
// of the stored Q1 rows into p2[blockIdx.x] (then the grid is the Gram
// kernel's).
constexpr int kApplyLd = kRows + 16;

template <int PASS, int W>
__global__ __launch_bounds__(kThreads) void chol_apply_kernel(int m, int w, float *A, int64_t lda, float *ws)
{
    constexpr int CT = Shape<W>::CT, KS = Shape<W>::KS;
    constexpr int kAs = W * kApplyLd * 4;                       // the block, for the product
    constexpr int kQs = (PASS == 1) ? W * (kRows + 4) * 4 : 0;  // the stored Q1 rows, for their Gram
    constexpr int kCb = (PASS == 1) ? 2 * Shape<W>::NT * 4 * 64 * 8 : 0;
    constexpr int kBytes = (kAs + kQs) > kCb ? (kAs + kQs) : kCb;
    __shared__ __attribute__((aligned(16))) char smem[kBytes];
    float *As = (float *)smem;
    float *Qs = (float *)(smem + kAs);
    const CholWork cw = carve<W>(ws);
    const int nblk = (m + kRows - 1) / kRows;
    const int G = gridDim.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

    BlockRegs<W> r[2];
    int b = blockIdx.x;
    if (b < nblk) r[0].load(A, lda, m, w, b * kRows);
    if (b + G < nblk) r[1].load(A, lda, m, w, (b + G) * kRows);
    // A-operand slices of R^{-1}: rv[ct][s] = R^{-1}[4 s + l/16][16 ct + l%16]
    // (zero for k > column, so column tile ct needs k-steps 0 .. 4 ct + 3).
    double rv[CT][KS];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int s2 = 0; s2 < 4 * (ct + 1); ++s2)
            rv[ct][s2] = cw.ri[(4 * s2 + (lane >> 4)) * W + 16 * ct + (lane & 15)];

    Gram<W> g;
    if constexpr (PASS == 1) gram_zero<W>(g);
#pragma unroll 1
    for (int it = 0; b < nblk; b += G, ++it) {
        __syncthreads();  // the previous block's operands have been read
        if (it & 1) r[1].store_ld(As, kApplyLd); else r[0].store_ld(As, kApplyLd);
        __syncthreads();
        if (b + 2 * G < nblk) {
            if (it & 1) r[1].load(A, lda, m, w, (b + 2 * G) * kRows);
            else r[0].load(A, lda, m, w, (b + 2 * G) * kRows);
        }
        const int lrow = wv * 16 + (lane & 15);  // row within the block
        const int row = b * kRows + lrow;
        float av[KS];
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2) av[s2] = As[(4 * s2 + (lane >> 4)) * kApplyLd + lrow];
        doublex4 d[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) d[ct] = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                if (s2 < 4 * (ct + 1)) d[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(rv[ct][s2], (double)av[s2], d[ct], 0, 0, 0);
        // d[ct][r] = Q[row][16 ct + l/16 + 4 r]
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = 16 * ct + (lane >> 4) + 4 * q;
                const bool ok = row < m && col < w;
                const float v = ok ? (float)d[ct][q] : 0.0f;
                if (ok) A[row + (int64_t)col * lda] = v;
                if constexpr (PASS == 1) Qs[col * (kRows + 4) + lrow] = v;
            }
        if constexpr (PASS == 1) {
            // Gram of the stored (FP32) Q1 rows of this block.
            __syncthreads();
#pragma unroll
            for (int sg = 0; sg < kTT; ++sg) gram_lds<W>(g, Qs, (wv * kTT + sg) * 16);
        }
    }
    if constexpr (PASS == 1) {
        __syncthreads();
        gram_finish<W>(g, (double *)smem, cw.p2 + (int64_t)blockIdx.x * Shape<W>::P);
    }
}

// One Cholesky QR pass over the rows of this workgroup, the factor formed
// by every workgroup (panels of at most kFusedBlocks 64-row blocks: one
// launch less than the separate factor kernel).
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
__global__ __launch_bounds__(kThreads) void chol_apply_fused_kernel(int m, int w, float *A, int64_t lda, float *ws, float *R,
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

// Workgroups of the apply kernels without a partial Gram: several per CU, so
// tall panels keep enough loads in flight.
constexpr int kMaxApplyGroups = 1024;
constexpr int kFusedBlocks = 256;

template <int W>
int launch_cholqr(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work, int *status,
                  int G, bool once)
{
    const CholWork cw = carve<W>(work);
    const int nblk = (m + kRows - 1) / kRows;
    const int Ga = nblk < kMaxApplyGroups ? nblk : kMaxApplyGroups;
    // Up to kFusedBlocks blocks every apply workgroup forms the factor itself
    // (the latency of a separate factor launch outweighs the redundancy);
    // taller panels factor once and apply with more workgroups.
    const bool fused = nblk <= kFusedBlocks;
    const dim3 rgrid((Shape<W>::P + 31) / 32);
    int n = 0;
    hipLaunchKernelGGL(chol_gram_kernel<W>, dim3(G), dim3(kThreads), 0, s, m, w, (const float *)A, (int64_t)lda, work);
    ++n;
    if (G > kDirect) {
        hipLaunchKernelGGL(chol_reduce_kernel<W>, rgrid, dim3(256), 0, s, cw.p1, G, cw.g);
        ++n;
    }
    auto pass = [&](auto P) {
        constexpr int PASS = decltype(P)::value;
        if (fused) {
            hipLaunchKernelGGL((chol_apply_fused_kernel<PASS, W>), dim3(G), dim3(kThreads), 0, s, m, w, A,
                               (int64_t)lda, work, R, (int64_t)ldr, status);
            ++n;
        } else {
            hipLaunchKernelGGL((chol_factor_kernel<PASS, W>), dim3(1), dim3(kThreads), 0, s, w, work, G, R,
                               (int64_t)ldr, status);
            hipLaunchKernelGGL((chol_apply_kernel<PASS, W>), dim3(PASS == 1 ? G : Ga), dim3(kThreads), 0, s, m, w, A,
                               (int64_t)lda, work);
            n += 2;
        }
    };
    if (once) {
        pass(std::integral_constant<int, 0>());
        return n;
    }
    pass(std::integral_constant<int, 1>());
    if (G > kDirect) {
        hipLaunchKernelGGL(chol_reduce_kernel<W>, rgrid, dim3(256), 0, s, cw.p2, G, cw.g);
        ++n;
    }
    pass(std::integral_constant<int, 2>());
    return n;
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
    const int n = (w <= 32) ? launch_cholqr<32>(s, m, w, A, lda, R, ldr, work, status, G, once)
                            : launch_cholqr<64>(s, m, w, A, lda, R, ldr, work, status, G, once);
    HIP_OK(hipGetLastError());
    if (launches) *launches += n;
    return LATER_OK;
}
