// Device helpers of the Cholesky QR panels (FP64 MFMA Gram, FP64
// Cholesky with its inverse), shared by the grid-wide panel kernels
// (panel_chol.hip) and the one-workgroup-per-matrix batched kernel
// (batch_wg.hip).
#pragma once
#include "common.h"

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kWMax = 64;                 // max panel width
constexpr int kThreads = 256;             // threads per workgroup (four waves)
constexpr int kTT = 1;                    // 16-row tiles per wave in a block
constexpr int kRows = 64 * kTT;           // rows per block (workgroup b takes blocks b, b + G, ...)
constexpr int kMaxGroups = 512;           // Gram workgroups (and partials) per launch: two per CU
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
    double *ri;  // W x W, row-major: the inverse factor the apply kernel multiplies by
};

template <int W> __host__ __device__ inline CholWork carve(float *ws)
{
    constexpr int P = Shape<W>::P;
    CholWork c;
    c.p1 = (double *)ws;
    c.p2 = c.p1 + (int64_t)kMaxGroups * P;
    c.r1 = c.p2 + (int64_t)kMaxGroups * P;
    c.g = c.r1 + W * W;
    c.ri = c.g + P;
    return c;
}

template <int W> constexpr int64_t work_floats()
{
    return ((int64_t)2 * kMaxGroups * Shape<W>::P + 2 * W * W + Shape<W>::P) * 2 + 64;
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

// Adds the four waves' Grams, (w0 + w2) + (w1 + w3), in two lane-aligned
// LDS rounds (every wave reads back exactly the slots it wrote: no bank
// conflicts), and wave 0 stores this workgroup's packed partial from its
// registers.  Cb: 2 x NT x 4 x 64 doubles of LDS.
template <int W> __device__ __forceinline__ void gram_finish(const Gram<W> &g, double *Cb, double *part)
{
    constexpr int NT = Shape<W>::NT, CT = Shape<W>::CT;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Gram<W> s = g;
    if (wv >= 2) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) Cb[(((wv - 2) * NT + t) * 4 + r) * 64 + lane] = g.t[t][r];
    }
    __syncthreads();
    if (wv < 2) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) s.t[t][r] += Cb[((wv * NT + t) * 4 + r) * 64 + lane];
    }
    __syncthreads();
    if (wv == 1) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) Cb[(t * 4 + r) * 64 + lane] = s.t[t][r];
    }
    __syncthreads();
    if (wv == 0) {
        // Lane l, item r of tile (a, b) holds entry (16 a + l/16 + 4 r, 16 b + l%16).
#pragma unroll
        for (int a = 0; a < CT; ++a)
#pragma unroll
            for (int b = a; b < CT; ++b) {
                const int t = tile_index<CT>(a, b);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 16 * a + (lane >> 4) + 4 * r, j = 16 * b + (lane & 15);
                    if (i <= j) part[j * (j + 1) / 2 + i] = s.t[t][r] + Cb[(t * 4 + r) * 64 + lane];
                }
            }
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
    // The reference norms in registers before the elimination: read inside
    // the loop, each would put an LDS latency on the serial chain (the loop's
    // stores to R and Ri keep the compiler from hoisting them).
    double dk[kW];
#pragma unroll
    for (int k = 0; k < kW; ++k) dk[k] = d0[k];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        const double p = readlane_f64(g[k], k);
        const double p0 = dk[k];
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
        // The four 32 x 32 x 32 products below run on FP64 MFMA: wave w makes
        // the 16 x 16 tile (w / 2, w % 2) of each, over the full k range (the
        // triangular factors' zero halves are stored, so the sums are the
        // same).  Result layout (v_mfma_f64_16x16x4f64): lane l, entry r holds
        // tile element (l / 16 + 4 r, l % 16).
        const int lane = t & 63, wv = t >> 6, ti = 16 * (wv >> 1), tj = 16 * (wv & 1);
        const int li = lane & 15, lk = lane >> 4;
        auto product = [&](auto a, auto b, doublex4 c) {
#pragma unroll
            for (int s = 0; s < 8; ++s)
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(a(ti + li, 4 * s + lk), b(4 * s + lk, tj + li), c, 0, 0, 0);
            return c;
        };
        const doublex4 zero = doublex4{0.0, 0.0, 0.0, 0.0};
        // R12 = R11^{-T} B -> Rs[i][32 + j]; zero lower-left blocks.
        {
            const doublex4 c = product([&](int i, int k) { return Ris[k * W + i]; },
                                       [&](int k, int j) { return Gs[k * LDG + 32 + j]; }, zero);
#pragma unroll
            for (int r = 0; r < 4; ++r) Rs[(ti + lk + 4 * r) * W + 32 + tj + li] = c[r];
            const int j = t & 31, i0 = t >> 5;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + 8 * u;
                Rs[(32 + i) * W + j] = 0.0;
                Ris[(32 + i) * W + j] = 0.0;
            }
        }
        __syncthreads();
        // S = C - R12^T R12, in place over C.
        {
            doublex4 c;
#pragma unroll
            for (int r = 0; r < 4; ++r) c[r] = Gs[(32 + ti + lk + 4 * r) * LDG + 32 + tj + li];
            c = product([&](int i, int k) { return -Rs[k * W + 32 + i]; },
                        [&](int k, int j) { return Rs[k * W + 32 + j]; }, c);
            __syncthreads();  // every wave has read C before any writes S over it
#pragma unroll
            for (int r = 0; r < 4; ++r) Gs[(32 + ti + lk + 4 * r) * LDG + 32 + tj + li] = c[r];
        }
        __syncthreads();
        if (t < 64) bad |= cholesky_wave<LDG, W>(Gs + 32 * LDG + 32, d0 + 32, Rs + 32 * W + 32, Ris + 32 * W + 32);
        __syncthreads();
        // T = R12 R22^{-1} -> over A (Gs[i][j]).
        {
            const doublex4 c = product([&](int i, int k) { return Rs[i * W + 32 + k]; },
                                       [&](int k, int j) { return Ris[(32 + k) * W + 32 + j]; }, zero);
#pragma unroll
            for (int r = 0; r < 4; ++r) Gs[(ti + lk + 4 * r) * LDG + tj + li] = c[r];
        }
        __syncthreads();
        // R^{-1}[i][32 + j] = -(R11^{-1} T)[i][j].  A bad pivot k leaves, as
        // in one block, row k of R and column k of R^{-1} zero.
        {
            const doublex4 c = product([&](int i, int k) { return -Ris[i * W + k]; },
                                       [&](int k, int j) { return Gs[k * LDG + j]; }, zero);
#pragma unroll
            for (int r = 0; r < 4; ++r) Ris[(ti + lk + 4 * r) * W + 32 + tj + li] = c[r];
        }
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
    static constexpr int kCb = 2 * Shape<W>::NT * 4 * 64 * 8;
    static constexpr int kBytes = kFactor > kQs ? (kFactor > kCb ? kFactor : kCb) : (kQs > kCb ? kQs : kCb);
};

} // namespace
