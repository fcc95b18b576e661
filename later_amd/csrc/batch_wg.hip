// Batched QR in one persistent launch: every workgroup takes whole matrices
// from a device work queue and factors each one by itself.
//
// The queue is the task table, ordered by size bucket and, inside a bucket,
// by decreasing cost (largest matrices first, so the long tasks start
// early); a workgroup claims the next task with one atomic add on the queue
// head and leaves when the head has passed the last task -- every workgroup
// reaches that exit, and none waits for another.  The last workgroup to
// leave resets the head for the next launch of the same plan (stream order
// makes the reset visible to it).
//
// Per matrix (m x n, column-major, in place: A <- Q, R n x n), left-looking
// block Gram-Schmidt with reorthogonalisation (BCGS2):
//   for each panel P of BW columns:
//     twice: for each block T of NTB earlier (final) columns of Q:
//              S' = P^T T, R_T,P += S'^T, P -= T S'^T
//     Cholesky QR twice of P with FP64 Gram matrices and factors (the
//       FP64-MFMA helpers of the Cholesky QR leaves): P <- Q_P, R_PP.
// Projecting each panel twice against all of Q keeps the loss of
// orthogonality at the level of the arithmetic, independent of the
// condition number (a right-looking order, which projects the later columns
// once per panel, loses it like kappa u).  The products run on MFMA (FP32
// 32x32x2, or BF16 / FP16 32x32x16 with the operands rounded on the way
// from LDS), staged through LDS in chunks of 64 rows whose successors are
// loaded into registers while the current chunk is used; the update walks
// P's rows with the MFMA computing the transposed tile, so each lane holds
// one row (coalesced loads and stores).  BW and NTB come from the
// per-bucket tuning table (the variant of each task).
#include <algorithm>
#include <vector>

#include <type_traits>

#include "chol_dev.h"
#include "common.h"

namespace {

constexpr int kKC = 64;            // rows per staged chunk
constexpr int kLDC = kKC + 1;      // stride of a column-major LDS chunk (odd: conflict-free)
constexpr int kNTBMax = 128;       // earlier Q columns per block (max)

struct WgTask {
    int m, n, variant, pad;
    float *A;
    int64_t lda;
    float *R;
    int64_t ldr;
    int *status;
};

struct WgQueue {
    unsigned int head, done;
};

// Variant v: panel width 16 << (v & 1), blocks of 64 << (v >> 1) earlier Q columns.
__host__ __device__ constexpr int variant_bw(int v) { return 16 << (v & 1); }
__host__ __device__ constexpr int variant_ntb(int v) { return 64 << ((v >> 1) & 1); }

// LDS: the chunk of Q_P (column-major, 32 x kLDC floats) stays apart; the
// Cholesky factor's matrices and the Q block's chunk and S' share one
// region (they are live at different times).
struct WgLds {
    static constexpr int kPc = 32 * kLDC * 4;
    static constexpr int kTc = kNTBMax * kLDC * 4;
    static constexpr int kSs = 128 * 33 * 4 + 2 * 16 * 64 * 4;  // -S'^T (k x y) and the k-half partials
    static constexpr int kChol = (32 * 33 + 3 * 32 * 32 + 32 + 528) * 8 + 4 * 3 * 4 * 64 * 8;
    static constexpr int kShared = (kTc + kSs) > kChol ? (kTc + kSs) : kChol;
    static constexpr int kBytes = kPc + kShared;
};

// MFMA step over two k of FP32 operands, or 16 k of 16-bit operands
// rounded from FP32 (lane l supplies row l % 32 and k 16 s + 8 (l / 32) + e).
template <int PREC> struct WgMfma;
template <> struct WgMfma<LATER_GEMM_F32> {
    static constexpr int KSTEP = 2;
    // value for lane l: element (row, k = k0 + l / 32) of a column-major (row, k) array with stride ld
    __device__ static float frag(const float *base, int ld_k, int ld_r, int row, int k0, int lane)
    {
        return base[(k0 + (lane >> 5)) * ld_k + row * ld_r];
    }
    __device__ static floatx16 mma(float a, float b, floatx16 c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }
};
template <int PREC16> struct WgMfma16 {
    static constexpr int KSTEP = 16;
    typedef typename std::conditional<PREC16 == LATER_GEMM_BF16, bf16x8_t, half8_t>::type V;
    typedef typename std::conditional<PREC16 == LATER_GEMM_BF16, __bf16, _Float16>::type T;
    __device__ static V frag(const float *base, int ld_k, int ld_r, int row, int k0, int lane)
    {
        V v;
        const float *p = base + (k0 + 8 * (lane >> 5)) * ld_k + row * ld_r;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (T)p[e * ld_k];
        return v;
    }
    __device__ static floatx16 mma(V a, V b, floatx16 c)
    {
        if constexpr (PREC16 == LATER_GEMM_BF16) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
        else return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct WgMfma<LATER_GEMM_BF16> : WgMfma16<LATER_GEMM_BF16> {};
template <> struct WgMfma<LATER_GEMM_F16> : WgMfma16<LATER_GEMM_F16> {};

// A chunk of rows r0 .. r0 + 63 of columns 0 .. NC - 1 (zeros past m and
// past nc), held in registers between its global loads and its LDS store so
// that the next chunks' loads are in flight while the current one is used:
// thread t holds row t % 64 of columns t / 64 + 4 e, so each wave load reads
// 64 consecutive rows of one column (two whole 128-byte lines) and each LDS
// store writes 64 consecutive words.
template <int NC> struct Chunk {
    static constexpr int PER = NC / 4;  // 8 (32 columns), 16 (64), 32 (128)
    static_assert(kKC == 64 && kThreads == 256, "one row per lane, four column sets");
    float v[PER];
    unsigned ok;  // bit e: the value of column set e is inside the matrix
    // Every load reads an in-bounds address (rows and columns clamped); the
    // values outside the matrix are replaced by zeros when stored, so no load
    // sits behind a branch.
    __device__ __forceinline__ void load(const float *A, int64_t lda, int m, int r0, int nc)
    {
        const auto *G = gmem(A);
        const int t = threadIdx.x, row = r0 + (t & 63), rc = min(row, m - 1);
        ok = 0u;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int c = (t >> 6) + 4 * e;
            v[e] = G[(int64_t)min(c, nc - 1) * lda + rc];
            ok |= (c < nc && row < m) ? (1u << e) : 0u;
        }
    }
    // Column-major LDS chunk (stride kLDC).
    __device__ __forceinline__ void store(float *C) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int e = 0; e < PER; ++e) C[((t >> 6) + 4 * e) * kLDC + (t & 63)] = ((ok >> e) & 1u) ? v[e] : 0.0f;
    }
};

// One Cholesky QR pass of the m x w panel P (w <= 32) by this workgroup:
// FP64-MFMA Gram of all rows, FP64 Cholesky with its inverse (one wave),
// P <- P R^{-1} on FP64 MFMA.  R (row-major 32 x 32 doubles) is left in Rs.
// Returns true (uniform) when a pivot was bad.
__device__ __noinline__ bool wg_cholqr(int m, int w, float *P, int64_t lda, char *lds)
{
    constexpr int W = 32, CT = 2, KS = 8, LDG = 33;
    static_assert(kTT == 1, "one 16-row tile per wave and block");
    double *Gs = (double *)lds;
    double *Rs = Gs + W * LDG;
    double *Ris = Rs + W * W;
    double *d0 = Ris + W * W;
    double *part = d0 + W + W * W;  // after the saved R1 (Rs of the first pass is copied to d0 + W)
    double *Cb = part + 528;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15;
    const int nblk = (m + kRows - 1) / kRows;
    auto *G = gmem(P);
    // Gram
    Gram<W> g;
    gram_zero<W>(g);
    int64_t col[CT];
    bool okc[CT];
#pragma unroll
    for (int x = 0; x < CT; ++x) {
        col[x] = (int64_t)min(16 * x + c, w - 1) * lda;
        okc[x] = 16 * x + c < w;
    }
    // Blocks are read in groups of kGB, the next group's loads in flight
    // while the current group's products run.
    constexpr int kGB = 4;
    float an[kGB][CT][4];
    auto load_group = [&](int b0) {
#pragma unroll
        for (int q = 0; q < kGB; ++q) {
            const int row0 = (b0 + q) * kRows + wv * 16 + 4 * (lane >> 4);
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int row = row0 + p, rc = min(row, m - 1);
#pragma unroll
                for (int x = 0; x < CT; ++x) an[q][x][p] = G[rc + col[x]];
            }
        }
    };
    load_group(0);
    for (int b0 = 0; b0 < nblk; b0 += kGB) {
        float ac[kGB][CT][4];
#pragma unroll
        for (int q = 0; q < kGB; ++q)
#pragma unroll
            for (int x = 0; x < CT; ++x)
#pragma unroll
                for (int p = 0; p < 4; ++p) ac[q][x][p] = an[q][x][p];
        if (b0 + kGB < nblk) load_group(b0 + kGB);
#pragma unroll
        for (int q = 0; q < kGB; ++q) {
            const int row0 = (b0 + q) * kRows + wv * 16 + 4 * (lane >> 4);
            double a[CT][4];
#pragma unroll
            for (int x = 0; x < CT; ++x)
#pragma unroll
                for (int p = 0; p < 4; ++p) a[x][p] = (row0 + p < m && okc[x]) ? (double)ac[q][x][p] : 0.0;
            gram_step<W>(g, a);
        }
    }
    gram_finish<W>(g, Cb, part);
    __syncthreads();
    sum_partials<W>(part, 1, w, Gs);
    __syncthreads();
    __shared__ int bad_s;
    bool bad = factor<W>(Gs, Rs, Ris, d0);
    if (threadIdx.x == 0) bad_s = bad ? 1 : 0;
    __syncthreads();
    // P <- P R^{-1}
    double rv[CT][KS];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2)
            rv[ct][s2] = (s2 < 4 * (ct + 1)) ? Ris[(4 * s2 + (lane >> 4)) * W + 16 * ct + (lane & 15)] : 0.0;
    // Groups of kGB blocks with the next group's loads in flight, as above
    // (the loads read rows the stores of the current group do not touch).
    // Lane l of wave v reads row 16 v + l % 16 of the block and columns
    // 4 s2 + l / 16 (the MFMA B layout); addresses clamped, zeros at use.
    float avn[kGB][W / 4];
    auto load_apply = [&](int b0) {
#pragma unroll
        for (int q = 0; q < kGB; ++q) {
            const int rc = min((b0 + q) * kRows + wv * 16 + (lane & 15), m - 1);
#pragma unroll
            for (int s2 = 0; s2 < W / 4; ++s2) avn[q][s2] = G[rc + (int64_t)min(4 * s2 + (lane >> 4), w - 1) * lda];
        }
    };
    load_apply(0);
    for (int b0 = 0; b0 < nblk; b0 += kGB) {
        float avc[kGB][W / 4];
#pragma unroll
        for (int q = 0; q < kGB; ++q)
#pragma unroll
            for (int s2 = 0; s2 < W / 4; ++s2) avc[q][s2] = avn[q][s2];
        if (b0 + kGB < nblk) load_apply(b0 + kGB);
#pragma unroll
        for (int q = 0; q < kGB; ++q) {
            const int r0 = (b0 + q) * kRows;
            const int row = r0 + wv * 16 + (lane & 15);
            float av[W / 4];
#pragma unroll
            for (int s2 = 0; s2 < W / 4; ++s2) av[s2] = (row < m && 4 * s2 + (lane >> 4) < w) ? avc[q][s2] : 0.0f;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                doublex4 d = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s2 = 0; s2 < 4 * (ct + 1); ++s2)
                    d = __builtin_amdgcn_mfma_f64_16x16x4f64(rv[ct][s2], (double)av[s2], d, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int cl = 16 * ct + (lane >> 4) + 4 * r;
                    if (row < m && cl < w) G[row + (int64_t)cl * lda] = (float)d[r];
                }
            }
        }
    }
    __syncthreads();
    return bad_s != 0;
}

// S' (w x nb) = P^T T over all m rows, P the panel (w <= 32 columns), T a
// block of nb <= NTB earlier columns; wave v holds columns 32 (v % (NTB / 32))
// .. + 31 of S' (MFMA result layout: column l % 32, rows 8 (r / 4) + 4 (l / 32)
// + r % 4).  With NTB = 64 the waves 2 and 3 take the second k half of every
// chunk and their sums are added through LDS (Xs) at the end.  Two chunks of
// loads are in flight (register sets pc0/tc0 and pc1/tc1 alternate): one
// workgroup streams a whole matrix alone, so its rate is the bytes it keeps in
// flight over the memory latency.
template <int PREC, int NTB>
__device__ __noinline__ floatx16 wg_project(int m, int w, const float *P, const float *T, int64_t lda, int nb, float *Pc,
                               float *Tc, float *Xs)
{
    typedef WgMfma<PREC> M;
    constexpr int KW = (NTB == 64) ? kKC / 2 : kKC;  // k per wave and chunk
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wc = (NTB == 64) ? (wv & 1) : wv, kq = (NTB == 64) ? (wv >> 1) * KW : 0;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    Chunk<32> pc0, pc1;
    Chunk<NTB> tc0, tc1;
    pc0.load(P, lda, m, 0, w);
    tc0.load(T, lda, m, 0, nb);
    if (kKC < m) {
        pc1.load(P, lda, m, kKC, w);
        tc1.load(T, lda, m, kKC, nb);
    }
    auto step = [&](Chunk<32> &pc, Chunk<NTB> &tc, int r0) {
        __syncthreads();  // the previous chunk has been read
        pc.store(Pc);
        tc.store(Tc);
        __syncthreads();
        if (r0 + 2 * kKC < m) {
            pc.load(P, lda, m, r0 + 2 * kKC, w);
            tc.load(T, lda, m, r0 + 2 * kKC, nb);
        }
        if (32 * wc < NTB) {
#pragma unroll 4
            for (int k0 = kq; k0 < kq + KW; k0 += M::KSTEP) {
                // A = P^T: A[i][k] = Pc[i][k]; B = T: B[k][j] = Tc[j][k]
                const auto a = M::frag(Pc, 1, kLDC, lane & 31, k0, lane);
                const auto b = M::frag(Tc, 1, kLDC, 32 * wc + (lane & 31), k0, lane);
                acc = M::mma(a, b, acc);
            }
        }
    };
    for (int r0 = 0; r0 < m; r0 += 2 * kKC) {
        step(pc0, tc0, r0);
        if (r0 + kKC < m) step(pc1, tc1, r0 + kKC);
    }
    if constexpr (NTB == 64) {
        if (wv >= 2) {
#pragma unroll
            for (int r = 0; r < 16; ++r) Xs[((wv - 2) * 16 + r) * 64 + lane] = acc[r];
        }
        __syncthreads();
        if (wv < 2) {
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] += Xs[(wv * 16 + r) * 64 + lane];
        }
    }
    return acc;
}

// P (m x w) -= T S'^T with T a block of nb <= NTB earlier columns and -S'^T
// in LDS as St[k][y] (k < NTB rows of stride kLDT).  The MFMA computes the
// transposed tile (-S') T^T, so lane l owns row l % 32 of P and columns
// 8 (r / 4) + 4 (l / 32) + r % 4.  Waves 2 v' + h: rows 32 h .. of each 64-row
// chunk and k half v' of the block; the second k half is added through LDS.
constexpr int kLDT = 33;
template <int PREC, int NTB>
__device__ __noinline__ void wg_update_left(int m, int w, float *P, const float *T, int64_t lda, int nb, float *Tc, const float *St,
                               float *Xs)
{
    typedef WgMfma<PREC> M;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int h = wv & 1, kh = wv >> 1;
    constexpr int KH = NTB / 2;
    // Two chunks in flight, as in wg_project: (tc0, pv0) and (tc1, pv1).
    Chunk<NTB> tc0, tc1;
    float pv0[16], pv1[16];
    // P values at clamped addresses; the zeros outside the matrix (and for
    // the second k half, which adds into the first) are applied at use.
    auto *G = gmem(P);
    auto load_p = [&](float (&pv)[16], int r0) {
        const int rc = min(r0 + 32 * h + (lane & 31), m - 1);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int y = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            pv[r] = G[rc + (int64_t)min(y, w - 1) * lda];
        }
    };
    tc0.load(T, lda, m, 0, nb);
    load_p(pv0, 0);
    if (kKC < m) {
        tc1.load(T, lda, m, kKC, nb);
        load_p(pv1, kKC);
    }
    auto step = [&](Chunk<NTB> &tc, float (&pv)[16], int r0) {
        __syncthreads();
        tc.store(Tc);
        __syncthreads();
        floatx16 acc;
        {
            const int row = r0 + 32 * h + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int y = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                acc[r] = (kh == 0 && row < m && y < w) ? pv[r] : 0.0f;
            }
        }
        if (r0 + 2 * kKC < m) {
            tc.load(T, lda, m, r0 + 2 * kKC, nb);
            load_p(pv, r0 + 2 * kKC);
        }
        const int x0 = 32 * h;
#pragma unroll 4
        for (int k0 = kh * KH; k0 < kh * KH + KH; k0 += M::KSTEP) {
            // A = -S': A[y][k] = St[k][y]; B = T^T: B[k][x] = Tc[k][x]
            const auto a = M::frag(St, kLDT, 1, lane & 31, k0, lane);
            const auto b = M::frag(Tc + x0, kLDC, 1, lane & 31, k0, lane);
            acc = M::mma(a, b, acc);
        }
        if (kh == 1) {
#pragma unroll
            for (int r = 0; r < 16; ++r) Xs[(h * 16 + r) * 64 + lane] = acc[r];
        }
        __syncthreads();
        if (kh == 0) {
            const int row = r0 + x0 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int y = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (row < m && y < w) G[row + (int64_t)y * lda] = acc[r] + Xs[(h * 16 + r) * 64 + lane];
            }
        }
    };
    for (int r0 = 0; r0 < m; r0 += 2 * kKC) {
        step(tc0, pv0, r0);
        if (r0 + kKC < m) step(tc1, pv1, r0 + kKC);
    }
}

template <int PREC, int BW, int NTB>
__device__ void wg_factor(const WgTask &tk, char *lds, bool reorth)
{
    const int m = tk.m, n = tk.n;
    float *A = tk.A, *R = tk.R;
    const int64_t lda = tk.lda, ldr = tk.ldr;
    float *Pc = (float *)lds;
    char *sh = lds + WgLds::kPc;
    float *Tc = (float *)sh;
    float *St = (float *)(sh + WgLds::kTc);
    float *Xs = St + kNTBMax * kLDT;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // R's strictly-lower triangle is zero (the blocks above are all written).
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kThreads) {
        const int i = (int)(e % n), j = (int)(e / n);
        if (i > j) gmem(R)[i + (int64_t)j * ldr] = 0.0f;
    }
    bool bad = false;
    for (int c0 = 0; c0 < n; c0 += BW) {
        const int w = min(BW, n - c0);
        float *P = A + (int64_t)c0 * lda;
        // Left-looking block Gram-Schmidt: P is projected out of every
        // earlier block of Q, block by block, twice.
        for (int pass = 0; pass < (reorth ? 2 : 1); ++pass) {
            for (int q0 = 0; q0 < c0; q0 += NTB) {
                const int nb = min(NTB, c0 - q0);
                const float *T = A + (int64_t)q0 * lda;
                const floatx16 acc = wg_project<PREC, NTB>(m, w, P, T, lda, nb, Pc, Tc, Xs);
                __syncthreads();  // Tc and Pc have been read
                if (32 * wv < NTB) {
                    const int j = 32 * wv + (lane & 31);  // column of the block
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);  // column of the panel
                        St[j * kLDT + i] = -acc[r];
                        if (i < w && j < nb) {
                            auto *d = gmem(R) + q0 + j + (int64_t)(c0 + i) * ldr;
                            *d = pass ? *d + acc[r] : acc[r];
                        }
                    }
                }
                wg_update_left<PREC, NTB>(m, w, P, T, lda, nb, Tc, St, Xs);
                __syncthreads();
            }
        }
        // Cholesky QR twice: P = Q_P R_PP, R_PP = R2 R1.
        bad |= wg_cholqr(m, w, P, lda, sh);
        {
            double *Rs = (double *)sh + 32 * 33;
            double *R1 = Rs + 2 * 32 * 32 + 32;
            for (int e = threadIdx.x; e < 32 * 32; e += kThreads) R1[e] = Rs[e];
            __syncthreads();
        }
        bad |= wg_cholqr(m, w, P, lda, sh);
        {
            double *Rs = (double *)sh + 32 * 33;
            double *R1 = Rs + 2 * 32 * 32 + 32;
            for (int e = threadIdx.x; e < w * w; e += kThreads) {
                const int i = e % w, j = e / w;
                double s = 0.0;
                for (int k = i; k <= j; ++k) s = fma(Rs[i * 32 + k], R1[k * 32 + j], s);
                gmem(R)[c0 + i + (int64_t)(c0 + j) * ldr] = (i <= j) ? (float)s : 0.0f;
            }
            __syncthreads();
        }
    }
    if (bad && threadIdx.x == 0 && tk.status) atomicOr(tk.status, 1);
}

template <int PREC>
__global__ __launch_bounds__(kThreads, 2) void batch_wg_kernel(const WgTask *tasks, int ntasks, WgQueue *q, int reorth)
{
    __shared__ __attribute__((aligned(16))) char lds[WgLds::kBytes];
    __shared__ unsigned int cur;
    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(&q->head, 1u);
        __syncthreads();
        const unsigned int t = cur;
        __syncthreads();
        if (t >= (unsigned int)ntasks) break;
        const WgTask tk = tasks[t];
        if (threadIdx.x == 0 && tk.status) *tk.status = 0;
        __syncthreads();
        switch (tk.variant) {
        case 0: wg_factor<PREC, 16, 64>(tk, lds, reorth != 0); break;
        case 1: wg_factor<PREC, 32, 64>(tk, lds, reorth != 0); break;
        case 2: wg_factor<PREC, 16, 128>(tk, lds, reorth != 0); break;
        default: wg_factor<PREC, 32, 128>(tk, lds, reorth != 0); break;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned int d = atomicAdd(&q->done, 1u);
        if (d == gridDim.x - 1) {
            atomicExch(&q->head, 0u);
            atomicExch(&q->done, 0u);
        }
    }
}

} // namespace

namespace {

// Size buckets of the tuning table: n class min(ceil(n / 32), 8) - 1 and
// m class clamp(floor(log2 m) - 6, 0, 6); entry 7 n_class + m_class.
constexpr int kMClasses = 7, kNClasses = 8, kBuckets = kMClasses * kNClasses;

int bucket_of(int m, int n)
{
    int nc = (n + 31) / 32;
    nc = (nc > kNClasses ? kNClasses : nc) - 1;
    int lg = 0;
    while ((2 << lg) <= m) ++lg;  // floor(log2 m)
    int mc = lg - 6;
    mc = mc < 0 ? 0 : (mc >= kMClasses ? kMClasses - 1 : mc);
    return kMClasses * nc + mc;
}

#include "batch_tuning.h"  // kDefaultTuning[kBuckets]: variant per bucket from the autotuner

int wg_prec(const later_opts &o)
{
    switch (o.gemm) {
    case LATER_GEMM_F16: return LATER_GEMM_F16;
    case LATER_GEMM_BF16: return LATER_GEMM_BF16;
    default: return LATER_GEMM_F32;  // F32 and the split forms: FP32 products
    }
}

struct WgPlan {
    void *dev = nullptr;
    WgTask *tasks = nullptr;
    WgQueue *q = nullptr;
    int ntasks = 0, grid = 0, prec = 0, reorth = 0;
};

} // namespace

LATER_API int later_batch_wg_create(int count, const int *m, const int *n, float *const *A, const int *lda,
                                    float *const *R, const int *ldr, int *status, const later_opts *opts,
                                    const int *tuning, void **plan)
{
    if (!plan) return LATER_ERR_SHAPE;
    *plan = nullptr;
    later_opts o;
    if (opts) o = *opts; else later_opts_default(&o);
    if (count < 1 || !m || !n || !A || !lda || !R || !ldr) return LATER_ERR_SHAPE;
    if (o.gemm < LATER_GEMM_F32 || o.gemm > LATER_GEMM_I8X4 || o.reorth < 0 || o.reorth > 3) return LATER_ERR_OPTION;
    std::vector<WgTask> tasks(count);
    std::vector<std::pair<double, int>> order(count);
    for (int i = 0; i < count; ++i) {
        if (m[i] < 1 || n[i] < 1 || m[i] < n[i] || lda[i] < m[i] || ldr[i] < n[i] || !A[i] || !R[i])
            return LATER_ERR_SHAPE;
        const int b = bucket_of(m[i], n[i]);
        int v = tuning ? tuning[b] : kDefaultTuning[b];
        if (v < 0 || v > 3) v = kDefaultTuning[b];
        tasks[i] = WgTask{m[i], n[i], v, 0, A[i], (int64_t)lda[i], R[i], (int64_t)ldr[i], status ? status + i : nullptr};
        // Cost: the products (4 m n^2 with the second projection) plus the
        // sequential panel factorisations (~ n / BW Cholesky factors).
        const double cost = 4.0 * m[i] * (double)n[i] * n[i] + 2.0e6 * ((n[i] + variant_bw(v) - 1) / variant_bw(v));
        order[i] = {-cost, i};
    }
    std::stable_sort(order.begin(), order.end());
    std::vector<WgTask> sorted(count);
    for (int i = 0; i < count; ++i) sorted[i] = tasks[order[i].second];
    WgPlan *P = new WgPlan();
    const size_t tbytes = ((size_t)count * sizeof(WgTask) + 255) / 256 * 256;
    if (hipMalloc(&P->dev, tbytes + 256) != hipSuccess) { delete P; return LATER_ERR_DEVICE; }
    P->tasks = (WgTask *)P->dev;
    P->q = (WgQueue *)((char *)P->dev + tbytes);
    if (hipMemcpy(P->tasks, sorted.data(), (size_t)count * sizeof(WgTask), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(P->q, 0, sizeof(WgQueue)) != hipSuccess) {
        hipFree(P->dev);
        delete P;
        return LATER_ERR_DEVICE;
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    // Two workgroups per CU (the kernel's LDS allows two).
    P->ntasks = count;
    P->grid = count < 2 * cus ? count : 2 * cus;
    P->prec = wg_prec(o);
    P->reorth = o.reorth != 0 ? 1 : 0;
    *plan = P;
    return LATER_OK;
}

LATER_API int later_batch_wg_execute(void *plan, void *stream)
{
    WgPlan *P = (WgPlan *)plan;
    if (!P || !P->dev) return LATER_ERR_SHAPE;
    hipStream_t s = (hipStream_t)stream;
    switch (P->prec) {
    case LATER_GEMM_F16:
        hipLaunchKernelGGL(batch_wg_kernel<LATER_GEMM_F16>, dim3(P->grid), dim3(kThreads), 0, s, P->tasks, P->ntasks,
                           P->q, P->reorth);
        break;
    case LATER_GEMM_BF16:
        hipLaunchKernelGGL(batch_wg_kernel<LATER_GEMM_BF16>, dim3(P->grid), dim3(kThreads), 0, s, P->tasks,
                           P->ntasks, P->q, P->reorth);
        break;
    default:
        hipLaunchKernelGGL(batch_wg_kernel<LATER_GEMM_F32>, dim3(P->grid), dim3(kThreads), 0, s, P->tasks, P->ntasks,
                           P->q, P->reorth);
        break;
    }
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

LATER_API int later_batch_wg_destroy(void *plan)
{
    WgPlan *P = (WgPlan *)plan;
    if (!P) return LATER_ERR_SHAPE;
    hipFree(P->dev);
    delete P;
    return LATER_OK;
}

LATER_API int later_batch_wg_bucket(int m, int n)
{
    if (m < 1 || n < 1) return -1;
    return bucket_of(m, n);
}
