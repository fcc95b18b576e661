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
#include "common.h"
#include "panel.h"

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kWMax = 64;                 // max panel width
constexpr int kThreads = 256;             // threads per workgroup (four waves)
constexpr int kTT = 1;                    // 16-row tiles per wave in a block
constexpr int kRows = 64 * kTT;           // rows per block (workgroup b takes blocks b, b + G, ...)
constexpr int kMaxGroups = 256;           // workgroups (and Gram partials) per launch
constexpr int kDirect = 32;               // up to this many partials, every workgroup adds them itself
constexpr double kPivotRel = 64.0 * 2.220446049250313e-16;

// Sizes of the width-W kernels (W = 32 or 64).
template <int W> struct Shape {
    static constexpr int CT = W / 16;             // 16-column tiles
    static constexpr int NT = CT * (CT + 1) / 2;  // distinct 16 x 16 tiles of the symmetric Gram
    static constexpr int P = W * (W + 1) / 2;     // packed upper triangle
    static constexpr int LDG = W + 1;             // row stride of the Gram in LDS
    static constexpr int KS = W / 4;              // MFMA k-steps over the columns
};

// Index of Gram tile (a, b), a <= b < CT, in row-major upper-triangular order.
template <int CT> __host__ __device__ constexpr int tile_index(int a, int b)
{
    return a * CT - a * (a - 1) / 2 + (b - a);
}

// Workspace of the width-W kernels (doubles; the partials have stride P).
struct CholWork {
    double *p1;  // kMaxGroups x P: partial Grams of A
    double *p2;  // kMaxGroups x P: partial Grams of Q1
    double *r1;  // W x W, row-major: R1 (upper)
    double *g;   // P: a reduced Gram (more than kDirect partials)
};

template <int W> __host__ __device__ inline CholWork carve(float *ws)
{
    constexpr int P = Shape<W>::P;
    CholWork c;
    c.p1 = (double *)ws;
    c.p2 = c.p1 + (int64_t)kMaxGroups * P;
    c.r1 = c.p2 + (int64_t)kMaxGroups * P;
    c.g = c.r1 + W * W;
    return c;
}

template <int W> constexpr int64_t work_floats()
{
    return ((int64_t)2 * kMaxGroups * Shape<W>::P + W * W + Shape<W>::P) * 2 + 64;
}

// Packed index e of the upper triangle (column-major, i <= j) -> (i, j).
__device__ __forceinline__ void unpack(int e, int &i, int &j)
{
    int c = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
    while (c * (c + 1) / 2 > e) --c;
    while ((c + 1) * (c + 2) / 2 <= e) ++c;
    j = c;
    i = e - c * (c + 1) / 2;
}

// The distinct 16 x 16 tiles (a, b), a <= b, of a W x W Gram.
template <int W> struct Gram {
    doublex4 t[Shape<W>::NT];
};

template <int W> __device__ __forceinline__ void gram_zero(Gram<W> &g)
{
#pragma unroll
    for (int q = 0; q < Shape<W>::NT; ++q) g.t[q] = doublex4{0.0, 0.0, 0.0, 0.0};
}

// v_mfma_f64_16x16x4f64: A (16 x 4) lane l holds A[l%16][l/16], B (4 x 16)
// lane l holds B[l/16][l%16], D (16 x 16) lane l item r holds
// D[l/16 + 4r][l%16] (rows interleaved, unlike the FP32 16 x 16 forms).  For
// the Gram, the k index is a row of the panel: the lane's value of column
// c = 16 t + l%16 at row k serves as both A[c][k] and B[k][c].  a[t][p] holds
// column tile t at four rows (one per MFMA).
template <int W> __device__ __forceinline__ void gram_step(Gram<W> &g, const double (*a)[4])
{
    constexpr int CT = Shape<W>::CT;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int x = 0; x < CT; ++x)
#pragma unroll
            for (int y = x; y < CT; ++y)
                g.t[tile_index<CT>(x, y)] =
                    __builtin_amdgcn_mfma_f64_16x16x4f64(a[x][p], a[y][p], g.t[tile_index<CT>(x, y)], 0, 0, 0);
}

// Adds the four waves' Grams in wave order and stores this workgroup's
// packed partial.  Cb: 4 x NT x 4 x 64 doubles of LDS.
template <int W> __device__ __forceinline__ void gram_finish(const Gram<W> &g, double *Cb, double *part)
{
    constexpr int NT = Shape<W>::NT, CT = Shape<W>::CT;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) Cb[((wv * NT + t) * 4 + r) * 64 + lane] = g.t[t][r];
    __syncthreads();
    for (int e = threadIdx.x; e < Shape<W>::P; e += kThreads) {
        int i, j;
        unpack(e, i, j);
        const int t = tile_index<CT>(i >> 4, j >> 4);
        const int ii = i & 15, jj = j & 15;
        const int ln = jj + 16 * (ii & 3), r = ii >> 2;
        double s = Cb[((0 * NT + t) * 4 + r) * 64 + ln];
        s += Cb[((1 * NT + t) * 4 + r) * 64 + ln];
        s += Cb[((2 * NT + t) * 4 + r) * 64 + ln];
        s += Cb[((3 * NT + t) * 4 + r) * 64 + ln];
        part[e] = s;
    }
}

// G (LDS, W x (W + 1)) = sum over g < ngroups of the packed partials, in order
// g = 0, 1, ...; entries outside the w x w block are the identity's.
template <int W>
__device__ __forceinline__ void sum_partials(const double *part, int ngroups, int w, double *Gs)
{
    constexpr int P = Shape<W>::P, LDG = Shape<W>::LDG;
    constexpr int kPer = (P + kThreads - 1) / kThreads;  // 3 or 9
    int ei[kPer];
    double s[kPer];
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        ei[x] = min((int)threadIdx.x + x * kThreads, P - 1);
        s[x] = 0.0;
    }
    // Many loads in flight per thread: the partials were written by other
    // XCDs and come from memory, so each round costs a full memory latency.
    constexpr int kBatch = W == 32 ? 16 : 4;
    for (int g0 = 0; g0 < ngroups; g0 += kBatch) {
        double v[kPer][kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int g = min(g0 + u, ngroups - 1);
#pragma unroll
            for (int x = 0; x < kPer; ++x) v[x][u] = part[(int64_t)g * P + ei[x]];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
#pragma unroll
            for (int x = 0; x < kPer; ++x)
                if (g0 + u < ngroups) s[x] += v[x][u];
    }
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        const int e = threadIdx.x + x * kThreads;
        if (e < P) {
            int i, j;
            unpack(e, i, j);
            double val = s[x];
            if (i >= w || j >= w) val = (i == j) ? 1.0 : 0.0;
            Gs[i * LDG + j] = val;
            Gs[j * LDG + i] = val;
        }
    }
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Cholesky S = R^T R of a 32 x 32 block in LDS by one wave, and R^{-1}, in
// one right-looking elimination, fully unrolled.  Lane j < 32 keeps column j
// of the trailing matrix in registers; lane 32 + c keeps column c of the
// identity, to which the same row operations are applied (Gauss-Jordan on
// [S | I]): it ends as column c of the unit-lower M with M S = D L^T, and
// R^{-1}[c][k] = M[k][c] / sqrt(d_k).  Column k's multipliers S[i][k] are read
// from lane k with v_readlane and serve both halves.
// S: row stride LDS_; d0[k]: the squared norm of column k that the pivot is
// compared with; R, Ri (row stride LDR) get R[k][j] (upper) and R^{-1}[c][k]
// (upper).  A bad pivot k zeroes row k of R and column k of R^{-1} (so column k
// of Q), and drops row k from the later columns' elimination.  Returns true
// if a pivot was bad.
template <int LDS_, int LDR>
__device__ __forceinline__ bool cholesky_wave(const double *S, const double *d0, double *R, double *Ri)
{
    constexpr int kW = 32;
    const int lane = threadIdx.x & 63;
    const bool aug = lane >= kW;
    const int j = lane & (kW - 1);
    double g[kW];
#pragma unroll
    for (int i = 0; i < kW; ++i) g[i] = aug ? (i == j ? 1.0 : 0.0) : S[i * LDS_ + j];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        const double p = readlane_f64(g[k], k);
        const double p0 = d0[k];
        const bool ok = p0 > 0.0 && p > kPivotRel * p0 && p <= 1.7976931348623157e308;
        bad |= !ok;
        // 1/sqrt(p): hardware estimate and two Newton steps.
        const double pp = ok ? p : 1.0;
        double y = __builtin_amdgcn_rsq(pp);
        y = y * fma(-0.5 * pp * y, y, 1.5);
        y = y * fma(-0.5 * pp * y, y, 1.5);
        const double inv = ok ? y : 0.0;
        const double d = ok ? p * y : 0.0;
        const double val = g[k] * inv;
        if (aug)
            Ri[j * LDR + k] = val;
        else
            R[k * LDR + j] = (j < k) ? 0.0 : (j == k ? d : val);
        const double a = val * inv;  // row k of the trailing matrix (or of M) / p
#pragma unroll
        for (int i = k + 1; i < kW; ++i) g[i] = fma(-readlane_f64(g[i], k), a, g[i]);
    }
    return bad;
}

// G = R^T R and R^{-1} for the W x W Gram in LDS (Gs, row stride W + 1; R and
// Ri row stride W).  W = 64 runs the blocked form with 32 x 32 blocks
// G = [A B; B^T C]: A = R11^T R11, R12 = R11^{-T} B, C - R12^T R12 = R22^T R22,
// R^{-1} = [R11^{-1}, -R11^{-1} R12 R22^{-1}; 0, R22^{-1}] (all FP64); a pivot
// of the second block is compared with the diagonal of C.  Destroys Gs.
// Called by every thread of the workgroup; returns (on thread 0) whether a
// pivot was bad.
template <int W> __device__ __forceinline__ bool factor(double *Gs, double *Rs, double *Ris, double *d0)
{
    constexpr int LDG = Shape<W>::LDG;
    const int t = threadIdx.x;
    if (t < W) d0[t] = Gs[t * LDG + t];
    __syncthreads();
    bool bad = false;
    if (t < 64) bad = cholesky_wave<LDG, W>(Gs, d0, Rs, Ris);
    if constexpr (W == 64) {
        __syncthreads();
        // The four 32 x 32 products below: thread t makes column j = t % 32
        // of rows i = t / 32 + 8 u, u < 4, over the full k range (the
        // triangular factors' zero halves are stored, so the sums are the
        // same), with the shared operand loaded once per k and every load
        // independent of the running sums.
        const int j = t & 31, i0 = t >> 5;
        double s[4];
        // R12 = R11^{-T} B -> Rs[i][32 + j]; zero lower-left blocks.
#pragma unroll
        for (int u = 0; u < 4; ++u) s[u] = 0.0;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const double b = Gs[k * LDG + 32 + j];
#pragma unroll
            for (int u = 0; u < 4; ++u) s[u] = fma(Ris[k * W + i0 + 8 * u], b, s[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + 8 * u;
            Rs[i * W + 32 + j] = s[u];
            Rs[(32 + i) * W + j] = 0.0;
            Ris[(32 + i) * W + j] = 0.0;
        }
        __syncthreads();
        // S = C - R12^T R12, in place over C.
#pragma unroll
        for (int u = 0; u < 4; ++u) s[u] = Gs[(32 + i0 + 8 * u) * LDG + 32 + j];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const double b = Rs[k * W + 32 + j];
#pragma unroll
            for (int u = 0; u < 4; ++u) s[u] = fma(-Rs[k * W + 32 + i0 + 8 * u], b, s[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) Gs[(32 + i0 + 8 * u) * LDG + 32 + j] = s[u];
        __syncthreads();
        if (t < 64) bad |= cholesky_wave<LDG, W>(Gs + 32 * LDG + 32, d0 + 32, Rs + 32 * W + 32, Ris + 32 * W + 32);
        __syncthreads();
        // T = R12 R22^{-1} -> over A (Gs[i][j]).
#pragma unroll
        for (int u = 0; u < 4; ++u) s[u] = 0.0;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const double b = Ris[(32 + k) * W + 32 + j];
#pragma unroll
            for (int u = 0; u < 4; ++u) s[u] = fma(Rs[(i0 + 8 * u) * W + 32 + k], b, s[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) Gs[(i0 + 8 * u) * LDG + j] = s[u];
        __syncthreads();
        // R^{-1}[i][32 + j] = -(R11^{-1} T)[i][j].  A bad pivot k leaves, as
        // in one block, row k of R and column k of R^{-1} zero.
#pragma unroll
        for (int u = 0; u < 4; ++u) s[u] = 0.0;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const double b = Gs[k * LDG + j];
#pragma unroll
            for (int u = 0; u < 4; ++u) s[u] = fma(Ris[(i0 + 8 * u) * W + k], b, s[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) Ris[(i0 + 8 * u) * W + 32 + j] = -s[u];
    }
    return bad;
}

// Gram MFMA operands of one 16-row group at local row r0 + 4 (l/16) of a
// column-major LDS block Qs[col][kRows + 4].
template <int W> __device__ __forceinline__ void gram_lds(Gram<W> &g, const float *Qs, int r0)
{
    constexpr int CT = Shape<W>::CT;
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, r = r0 + 4 * (lane >> 4);
    double a[CT][4];
#pragma unroll
    for (int x = 0; x < CT; ++x) {
        const floatx4 v = *(const floatx4 *)&Qs[(16 * x + c) * (kRows + 4) + r];
#pragma unroll
        for (int p = 0; p < 4; ++p) a[x][p] = (double)v[p];
    }
    gram_step<W>(g, a);
}

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

// Rows of one block in the operand layout of the apply MFMA: wave wv owns
// kTT 16-row tiles; lane l holds, for tile t and k-step s,
// A[r0 + 16 (kTT wv + t) + l%16][4 s + l/16] (zero past m or w).
template <int W>
__device__ __forceinline__ void load_block_mfma(const float *A, int64_t lda, int m, int w, int r0, float (*av)[W / 4])
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kk = lane >> 4;
#pragma unroll
    for (int t = 0; t < kTT; ++t) {
        const int row = r0 + (wv * kTT + t) * 16 + (lane & 15);
        const int rc = min(row, m - 1);
#pragma unroll
        for (int s2 = 0; s2 < W / 4; ++s2) {
            const int col = 4 * s2 + kk;
            const float x = A[rc + (int64_t)min(col, w - 1) * lda];
            av[t][s2] = (row < m && col < w) ? x : 0.0f;
        }
    }
}

// LDS of the apply kernel: the factor's matrices while it is formed, then
// (PASS 1) the block of stored Q1 rows, then the Gram reduction buffer.
template <int W> struct ApplyLds {
    static constexpr int kFactor = (W * Shape<W>::LDG + 2 * W * W + W) * 8;
    static constexpr int kQs = W * (kRows + 4) * 4;
    static constexpr int kCb = 4 * Shape<W>::NT * 4 * 64 * 8;
    static constexpr int kBytes = kFactor > kQs ? (kFactor > kCb ? kFactor : kCb) : (kQs > kCb ? kQs : kCb);
};

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
