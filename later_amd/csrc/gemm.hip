// MFMA GEMMs of the recursion nodes (gfx950).
//
// The reference's two tensor-op GemmEx calls per node (QR/later_rgsqrf.cu:45-56),
// the FP32 Sgemm projections of its 128-column panel (QR/panel.cu:21-61) and
// the FP32->FP16 staging kernel s2h (util/util.cu:24-32) become one tiled MFMA
// kernel with the cast fused into the global->LDS path:
//   projection  C (n1 x n2)  = Q^T B     Q: m x n1, B: m x n2   (K = m, split-K)
//   update      B (m x n2)  -= Q C       C: n1 x n2             (K = n1)
// Operands are read as FP32 from HBM and converted in registers to the MFMA
// input type: FP32 (v_mfma_f32_32x32x2f32), FP16 / BF16 (v_mfma_f32_32x32x16_*)
// or an FP16 hi/lo pair (hi = fl16(x), lo = fl16((x - hi) 2^11)) that gives
// three products hi*hi + 2^-11 (hi*lo + lo*hi) accumulated separately.
// With `chunk` set, the FP32 MFMA accumulators are folded into FP64 registers
// every `chunk` rows of K, and split-K partials are summed in FP64 in a fixed
// order, so results are bit-reproducible and independent of the launch.
#include "common.h"
#include "gemm.h"
#include "prof.h"

namespace {

constexpr int NT = 256;  // 4 waves in a 2 x 2 grid
constexpr int kMaxBK = 64;

// K depth of one LDS slab.  64 for the split-FP16 forms without the FP64
// fold (their three or six products per slab then cover the next slab's
// global-load latency; measured +20% on the node updates); 32 elsewhere: for
// FP16/BF16 the deeper slab halves the workgroups per CU, for FP32 its LDS
// tiles are twice as big, and with the FP64 fold the extra staging
// registers spill.
template <int PREC, bool CH64> constexpr int bk()
{
    return ((PREC == GEMM_F16X3 || PREC == GEMM_F16X6) && !CH64) ? 64 : 32;
}

template <int PREC> struct Lds;
template <> struct Lds<GEMM_F32> { typedef float T; };
template <> struct Lds<GEMM_F16> { typedef _Float16 T; };
template <> struct Lds<GEMM_BF16> { typedef __bf16 T; };
template <> struct Lds<GEMM_F16X3> { typedef _Float16 T; };
template <> struct Lds<GEMM_F16X6> { typedef _Float16 T; };

// FP16 pieces per operand value: x = p0 + p1 2^-11 + p2 2^-22, with
// p0 = fl16(x), p1 = fl16((x - p0) 2^11), p2 = fl16((x - p0 - p1 2^-11) 2^22).
// Products are summed per power of 2^-11 in separate FP32 accumulators:
// group g holds sum_{i+j=g} p_i q_j.
template <int PREC> constexpr int pieces() { return PREC == GEMM_F16X6 ? 3 : PREC == GEMM_F16X3 ? 2 : 1; }



// Loads of one BK-deep slab of an R-row operand tile into registers.
//   KM (rows contiguous along K): thread owns SEG consecutive k of one row.
//   !KM (rows contiguous along M): thread owns 4 consecutive rows x KG k.
template <int R, bool KM, int BK> struct Slab {
    static constexpr int NV = R * BK / NT;
    static constexpr int SEG = NV;              // KM
    static constexpr int TPR = BK / SEG;        // KM: threads per row
    static constexpr int KG = NV / 4;           // !KM: k per thread
    float v[NV];

    // Loads with clamped in-bounds addresses and a select afterwards, so that
    // no load sits behind a branch (all of a thread's loads issue before the
    // first wait).  rmax, kmax >= 1.
    __device__ __forceinline__ void load(const float *P, int64_t ld, int r0, int rmax, int k0, int kmax,
                                         bool vec)
    {
        const int t = threadIdx.x;
        if (KM) {
            const int row = r0 + t / TPR;
            const int k = k0 + (t % TPR) * SEG;
            const bool rin = row < rmax;
            const float *p = P + (int64_t)min(row, rmax - 1) * ld;
            if (vec && rin && k + SEG <= kmax) {
#pragma unroll
                for (int e = 0; e < SEG; e += 4) {
                    floatx4 x = *(const floatx4 *)(p + k + e);
                    v[e] = x[0]; v[e + 1] = x[1]; v[e + 2] = x[2]; v[e + 3] = x[3];
                }
            } else {
#pragma unroll
                for (int e = 0; e < SEG; ++e) v[e] = p[min(k + e, kmax - 1)];
#pragma unroll
                for (int e = 0; e < SEG; ++e)
                    if (!(rin && k + e < kmax)) v[e] = 0.0f;
            }
        } else {
            const int rg = r0 + (t % (R / 4)) * 4;
            const int kb = k0 + (t / (R / 4)) * KG;
            if (vec && rg + 4 <= rmax && kb + KG <= kmax) {
#pragma unroll
                for (int kk = 0; kk < KG; ++kk) {
                    floatx4 x = *(const floatx4 *)(P + (int64_t)(kb + kk) * ld + rg);
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) v[kk * 4 + rr] = x[rr];
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < KG; ++kk)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        v[kk * 4 + rr] = P[(int64_t)min(kb + kk, kmax - 1) * ld + min(rg + rr, rmax - 1)];
#pragma unroll
                for (int kk = 0; kk < KG; ++kk)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        if (!(kb + kk < kmax && rg + rr < rmax)) v[kk * 4 + rr] = 0.0f;
            }
        }
    }

    // Row and k (within the tile) of value e.
    __device__ __forceinline__ int row_of(int e) const
    {
        const int t = threadIdx.x;
        return KM ? t / TPR : (t % (R / 4)) * 4 + (e & 3);
    }
    __device__ __forceinline__ int k_of(int e) const
    {
        const int t = threadIdx.x;
        return KM ? (t % TPR) * SEG + e : (t / (R / 4)) * KG + (e >> 2);
    }
};

// LDS tile layouts.  16-bit types: [row][BK+8] (k contiguous, 16-byte MFMA
// fragment reads).  FP32: [k][R+4] (row contiguous, 4-byte fragment reads).
template <int PREC, int R, int BK> struct Tile {
    typedef typename Lds<PREC>::T T;
    static constexpr bool F32 = (PREC == GEMM_F32);
    static constexpr int LD = F32 ? (R + 4) : (BK + 8);
    static constexpr int SZ = F32 ? BK * LD : R * LD;
    alignas(16) T p[pieces<PREC>()][SZ];

    __device__ __forceinline__ int idx(int row, int k) const { return F32 ? k * LD + row : row * LD + k; }
};

template <typename T, int N> struct Vec {
    typedef T type __attribute__((ext_vector_type(N)));
};

// Writes a thread's slab to LDS, converted to the MFMA input type, with the
// widest stores the layout allows: 16-bit types keep k contiguous per row, so
// a K-major slab (SEG consecutive k of one row) goes out as 16-byte vectors
// and an M-major slab (KG consecutive k for each of 4 rows) as KG-wide ones;
// FP32 keeps rows contiguous per k, so an M-major slab writes 4-row vectors.
template <int PREC, int R, bool KM, int BK>
__device__ __forceinline__ void store_slab(Tile<PREC, R, BK> &t, const Slab<R, KM, BK> &s, float scale)
{
    typedef typename Lds<PREC>::T T;
    typedef Slab<R, KM, BK> S;
    if constexpr (PREC == GEMM_F32) {
        if constexpr (KM) {
#pragma unroll
            for (int e = 0; e < S::NV; ++e) t.p[0][t.idx(s.row_of(e), s.k_of(e))] = s.v[e] * scale;
        } else {
#pragma unroll
            for (int kk = 0; kk < S::KG; ++kk) {
                floatx4 x;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) x[rr] = s.v[kk * 4 + rr] * scale;
                *(floatx4 *)&t.p[0][t.idx(s.row_of(kk * 4), s.k_of(kk * 4))] = x;
            }
        }
    } else {
        constexpr int W = KM ? 8 : S::KG;      // elements per store
        constexpr int NST = S::NV / W;          // stores per array
        constexpr int NP = pieces<PREC>();
        typedef typename Vec<T, W>::type VT;
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            VT pv[NP];
            int row = 0, k = 0;
#pragma unroll
            for (int e = 0; e < W; ++e) {
                // KM: values q*W .. q*W+W-1 are consecutive k of one row.
                // M-major: value kk*4+rr is row rr, k kk; store q = row rr.
                const int idx = KM ? q * W + e : e * 4 + q;
                float x = s.v[idx] * scale;
                if constexpr (NP == 1) {
                    pv[0][e] = (T)x;
                } else {
                    const _Float16 h = (_Float16)x;
                    const float r = x - (float)h;  // exact
                    const _Float16 m = (_Float16)(r * 2048.0f);
                    pv[0][e] = h;
                    pv[1][e] = m;
                    if constexpr (NP == 3) pv[2][e] = (_Float16)((r - (float)m * 0x1p-11f) * 0x1p22f);
                }
                if (e == 0) { row = s.row_of(idx); k = s.k_of(idx); }
            }
#pragma unroll
            for (int c = 0; c < NP; ++c) *(VT *)&t.p[c][t.idx(row, k)] = pv[c];
        }
    }
}

template <int PREC> struct Frag;
template <> struct Frag<GEMM_F32> { typedef float T; };
template <> struct Frag<GEMM_F16> { typedef half8_t T; };
template <> struct Frag<GEMM_BF16> { typedef bf16x8_t T; };
template <> struct Frag<GEMM_F16X3> { typedef half8_t T; };
template <> struct Frag<GEMM_F16X6> { typedef half8_t T; };

template <int PREC>
__device__ __forceinline__ floatx16 mfma(typename Frag<PREC>::T a, typename Frag<PREC>::T b, floatx16 c)
{
    if constexpr (PREC == GEMM_F32) return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    else if constexpr (PREC == GEMM_BF16) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// Fragment of 32 rows starting at `row0` for k-step `ks`.
// 32x32x16 (16-bit): lane l supplies row l%32, k = 16 ks + 8 (l/32) + 0..7.
// 32x32x2 (FP32):    lane l supplies row l%32, k = 2 ks + l/32.
template <int PREC, int R, int BK>
__device__ __forceinline__ typename Frag<PREC>::T frag(const typename Lds<PREC>::T *base, int row0, int ks)
{
    const int lane = threadIdx.x & 63;
    typedef Tile<PREC, R, BK> TT;
    if constexpr (PREC == GEMM_F32) {
        return base[(2 * ks + (lane >> 5)) * TT::LD + row0 + (lane & 31)];
    } else {
        return *(const typename Frag<PREC>::T *)(base + (row0 + (lane & 31)) * TT::LD + 16 * ks + 8 * (lane >> 5));
    }
}

template <int PREC, int BM, int BN, bool AKM, bool CH64>
__device__ __forceinline__ void gemm_body(const GemmArgs &g, const int tm, const int tn, const int split)
{
    constexpr int WM = BM / 2, WN = BN / 2;     // wave tile
    constexpr int FM = WM / 32, FN = WN / 32;   // 32x32 MFMA blocks per wave
    constexpr int NP = pieces<PREC>();          // FP16 pieces per value
    constexpr int NG = NP;                       // accumulator groups
    constexpr int BK = bk<PREC, CH64>();
    constexpr int KR = (PREC == GEMM_F32) ? 2 : 16;  // K rows per MFMA step
    constexpr int KSTEPS = BK / KR;

    // Double-buffered operand tiles: one barrier per BK slab.
    __shared__ Tile<PREC, BM, BK> As[2];
    __shared__ Tile<PREC, BN, BK> Bs[2];

    const int m0 = tm * BM, n0 = tn * BN;
    int kbeg = split * g.kper;
    int kend = min(g.K, kbeg + g.kper);
    if (g.tri) {
        // Upper-triangular product: rows >= cols of the tile contribute nothing
        // below the diagonal, and only k in [m0, n0 + BN) is non-zero.
        if (m0 > n0 + BN - 1) return;
        kbeg = max(kbeg, m0);
        kend = min(kend, n0 + BN);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave >> 1, wn = wave & 1;

    const bool avec = ((g.lda & 3) == 0) && ((((uintptr_t)g.A) & 15) == 0);
    const bool bvec = ((g.ldb & 3) == 0) && ((((uintptr_t)g.B) & 15) == 0);

    // Operand scales (powers of two, so exact).  A has one global scale; a
    // thread's B values all lie in one column of B (B rows are K-contiguous).
    Slab<BM, AKM, BK> sa;
    Slab<BN, true, BK> sb;
    const float ascale = ldexpf(1.0f, g.a_exp);
    const int bcol = n0 + sb.row_of(0);
    const float bscale = (g.b_exp && bcol < g.N) ? ldexpf(1.0f, g.b_exp[bcol]) : 1.0f;

    // The MFMA computes the transposed tile (B operand rows x A operand rows),
    // so that lane l%32 walks the rows of C: stores of column-major C are
    // then 128-byte contiguous per half-wave.
    floatx16 acc[NG][FM][FN];
    double acc64[CH64 ? FM : 1][CH64 ? FN : 1][CH64 ? 16 : 1];
#pragma unroll
    for (int q = 0; q < NG; ++q)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[q][i][j][r] = 0.0f;
    if constexpr (CH64) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc64[i][j][r] = 0.0;
    }

    // Value of accumulator element (i, j, r): groups combined in FP64.
    auto combined = [&](int i, int j, int r) -> double {
        double v = (double)acc[0][i][j][r];
        if constexpr (NG > 1) v += (double)acc[1][i][j][r] * 0x1p-11;
        if constexpr (NG > 2) v += (double)acc[2][i][j][r] * 0x1p-22;
        return v;
    };
    auto fold = [&]() {
        if constexpr (CH64) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc64[i][j][r] += combined(i, j, r);
#pragma unroll
                        for (int q = 0; q < NG; ++q) acc[q][i][j][r] = 0.0f;
                    }
        }
    };

    const int ntiles = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
    if (ntiles > 0) {
        sa.load(g.A, g.lda, m0, g.M, kbeg, kend, avec);
        sb.load(g.B, g.ldb, n0, g.N, kbeg, kend, bvec);
        store_slab<PREC, BM, AKM, BK>(As[0], sa, ascale);
        store_slab<PREC, BN, true, BK>(Bs[0], sb, bscale);
        __syncthreads();
    }
    for (int it = 0; it < ntiles; ++it) {
        const int cur = it & 1;
        const bool more = it + 1 < ntiles;
        if (more) {
            const int k0 = kbeg + (it + 1) * BK;
            sa.load(g.A, g.lda, m0, g.M, k0, kend, avec);
            sb.load(g.B, g.ldb, n0, g.N, k0, kend, bvec);
        }
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            typename Frag<PREC>::T fa[NP][FM], fb[NP][FN];
#pragma unroll
            for (int c = 0; c < NP; ++c) {
#pragma unroll
                for (int i = 0; i < FM; ++i) fa[c][i] = frag<PREC, BM, BK>(As[cur].p[c], wm * WM + 32 * i, ks);
#pragma unroll
                for (int j = 0; j < FN; ++j) fb[c][j] = frag<PREC, BN, BK>(Bs[cur].p[c], wn * WN + 32 * j, ks);
            }
            // Group g gets products p_a q_b with a + b = g; the loops run over
            // the blocks innermost so that updates of one accumulator are
            // separated by independent MFMAs.
#pragma unroll
            for (int gq = 0; gq < NG; ++gq)
#pragma unroll
                for (int a = 0; a <= gq; ++a)
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j)
                            acc[gq][i][j] = mfma<PREC>(fb[gq - a][j], fa[a][i], acc[gq][i][j]);
            // FP64 fold every chunk_rows rows of K (a multiple of 32).
            if (CH64 && ks + 1 < KSTEPS && (it * BK + (ks + 1) * KR) % g.chunk_rows == 0) fold();
        }
        if (CH64 && (((it + 1) * BK) % g.chunk_rows == 0 || it + 1 == ntiles)) fold();
        if (more) {
            store_slab<PREC, BM, AKM, BK>(As[cur ^ 1], sa, ascale);
            store_slab<PREC, BN, true, BK>(Bs[cur ^ 1], sb, bscale);
        }
        __syncthreads();
    }

    // Epilogue.  Transposed 32x32 accumulator layout: lane l holds row l%32
    // of the block and columns 8 (r/4) + 4 (l/32) + r%4, r = 0..15.
    const float oscale_a = ldexpf(1.0f, -g.a_exp);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int row = m0 + wm * WM + 32 * i + (lane & 31);
        if (row >= g.M) continue;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int col = n0 + wn * WN + 32 * j + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (col >= g.N) continue;
                const float cs = oscale_a * ((g.b_exp) ? ldexpf(1.0f, -g.b_exp[col]) : 1.0f);
                double v;
                if constexpr (CH64) v = acc64[i][j][r];
                else v = combined(i, j, r);
                v *= (double)cs;
                if (g.partial) {
                    const int64_t off = ((int64_t)split * g.N + col) * g.M + row;
                    if constexpr (CH64) ((double *)g.partial)[off] = v;
                    else ((float *)g.partial)[off] = (float)v;
                } else if (g.update) {
                    float *c = g.C + (int64_t)col * g.ldc + row;
                    if constexpr (CH64) *c = (float)((double)*c + (double)g.alpha * v);
                    else *c = fmaf(g.alpha, (float)v, *c);
                } else {
                    g.C[(int64_t)col * g.ldc + row] = (float)v;
                }
            }
        }
    }
}

// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2: the linear id is renumbered so that every XCD works
// through a contiguous range, and each range walks the tiles in groups of 8
// tile-rows (column by column inside a group), so the workgroups resident on
// one XCD share their A and B tiles in its L2.
template <int PREC, int BM, int BN, bool AKM, bool CH64>
__global__ __launch_bounds__(NT) void gemm_kernel(GemmArgs g)
{
    constexpr int kXcd = 8, kGroup = 8;
    const int nx = gridDim.x, ny = gridDim.y;
    const int total = nx * ny * gridDim.z;
    int L = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const int q = total / kXcd;
    if (L < kXcd * q) L = (L % kXcd) * q + L / kXcd;
    const int tiles = nx * ny;
    const int split = L / tiles;
    const int r = L % tiles;
    const int first = (r / (kGroup * ny)) * kGroup;
    const int gsz = min(nx - first, kGroup);
    const int rr = r - first * ny;
    gemm_body<PREC, BM, BN, AKM, CH64>(g, first + rr % gsz, rr / gsz, split);
}

// Grouped launch: workgroup b computes tile (t.y, t.z) of split t.w of the
// product described by args[t.x], t = tasks[b].
template <int PREC, int BM, int BN, bool AKM, bool CH64>
__global__ __launch_bounds__(NT) void gemm_grouped_kernel(const GemmArgs *args, const int4 *tasks)
{
    const int4 t = tasks[blockIdx.x];
    const GemmArgs g = args[t.x];
    gemm_body<PREC, BM, BN, AKM, CH64>(g, t.y, t.z, t.w);
}

// C = sum over splits (fixed order, FP64) of the partials.
template <bool F64>
__global__ void splitk_reduce_kernel(int M, int N, int nsplit, const void *partial, float *C, int64_t ldc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    const int row = (int)(idx % M), col = (int)(idx / M);
    double s = 0.0;
    for (int p = 0; p < nsplit; ++p) {
        const int64_t off = (int64_t)p * M * N + idx;
        s += F64 ? ((const double *)partial)[off] : (double)((const float *)partial)[off];
    }
    C[(int64_t)col * ldc + row] = (float)s;
}

// Per-column exponents e_j with max_i |X_ij| 2^e_j in [2^14, 2^15) (0 for a
// zero column).  One wave per column.
__global__ void colexp_kernel(int m, int n, const float *X, int64_t ldx, int *e)
{
    const int j = blockIdx.x;
    const float *x = X + (int64_t)j * ldx;
    float mx = 0.0f;
    for (int i = threadIdx.x; i < m; i += 64) mx = fmaxf(mx, fabsf(x[i]));
#pragma unroll
    for (int mask = 32; mask > 0; mask >>= 1) mx = fmaxf(mx, __shfl_xor(mx, mask, 64));
    if (threadIdx.x == 0) {
        int ex = 0;
        if (mx > 0.0f && mx < INFINITY) {
            int fe;
            frexpf(mx, &fe);  // mx in [2^(fe-1), 2^fe)
            ex = 15 - fe;
        }
        e[j] = ex;
    }
}

// Grouped colexp: workgroup b handles column t.y of matrix args[t.x].
__global__ void colexp_grouped_kernel(const ColexpArgs *args, const int2 *tasks)
{
    const int2 t = tasks[blockIdx.x];
    const ColexpArgs a = args[t.x];
    const float *x = a.X + (int64_t)t.y * a.ldx;
    float mx = 0.0f;
    for (int i = threadIdx.x; i < a.m; i += 64) mx = fmaxf(mx, fabsf(x[i]));
#pragma unroll
    for (int mask = 32; mask > 0; mask >>= 1) mx = fmaxf(mx, __shfl_xor(mx, mask, 64));
    if (threadIdx.x == 0) {
        int ex = 0;
        if (mx > 0.0f && mx < INFINITY) {
            int fe;
            frexpf(mx, &fe);
            ex = 15 - fe;
        }
        a.e[t.y] = ex;
    }
}

template <int PREC, int BM, int BN, bool AKM, bool CH64>
void launch_one(hipStream_t s, const GemmArgs &g, dim3 grid)
{
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, AKM, CH64>), grid, dim3(NT), 0, s, g);
}

template <int PREC, bool AKM, bool CH64>
void launch_tile(hipStream_t s, const GemmArgs &g, int nsplit, bool small)
{
    const dim3 g64((unsigned)ceil_div64(g.M, 64), (unsigned)ceil_div64(g.N, 64), (unsigned)nsplit);
    if constexpr (PREC == GEMM_F16X6) {
        // Three accumulator groups: a 128x128 tile would spill, so the
        // three-piece form always runs 64x64 tiles.
        launch_one<PREC, 64, 64, AKM, CH64>(s, g, g64);
    } else if (small) {
        launch_one<PREC, 64, 64, AKM, CH64>(s, g, g64);
    } else {
        dim3 grid((unsigned)ceil_div64(g.M, 128), (unsigned)ceil_div64(g.N, 128), (unsigned)nsplit);
        launch_one<PREC, 128, 128, AKM, CH64>(s, g, grid);
    }
}

template <bool AKM, bool CH64>
int dispatch(hipStream_t s, int prec, const GemmArgs &g, int nsplit, bool small)
{
    switch (prec) {
    case GEMM_F32: launch_tile<GEMM_F32, AKM, CH64>(s, g, nsplit, small); break;
    case GEMM_F16: launch_tile<GEMM_F16, AKM, CH64>(s, g, nsplit, small); break;
    case GEMM_BF16: launch_tile<GEMM_BF16, AKM, CH64>(s, g, nsplit, small); break;
    case GEMM_F16X3: launch_tile<GEMM_F16X3, AKM, CH64>(s, g, nsplit, small); break;
    case GEMM_F16X6: launch_tile<GEMM_F16X6, AKM, CH64>(s, g, nsplit, small); break;
    default: return LATER_ERR_OPTION;
    }
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

template <int PREC, bool AKM, bool CH64>
void launch_grouped_t(hipStream_t s, const GemmArgs *args, const int4 *tasks, int ntasks)
{
    hipLaunchKernelGGL((gemm_grouped_kernel<PREC, 64, 64, AKM, CH64>), dim3(ntasks), dim3(NT), 0, s, args, tasks);
}

template <bool AKM, bool CH64>
int launch_grouped_p(hipStream_t s, int prec, const GemmArgs *args, const int4 *tasks, int ntasks)
{
    switch (prec) {
    case GEMM_F32: launch_grouped_t<GEMM_F32, AKM, CH64>(s, args, tasks, ntasks); break;
    case GEMM_F16: launch_grouped_t<GEMM_F16, AKM, CH64>(s, args, tasks, ntasks); break;
    case GEMM_BF16: launch_grouped_t<GEMM_BF16, AKM, CH64>(s, args, tasks, ntasks); break;
    case GEMM_F16X3: launch_grouped_t<GEMM_F16X3, AKM, CH64>(s, args, tasks, ntasks); break;
    case GEMM_F16X6: launch_grouped_t<GEMM_F16X6, AKM, CH64>(s, args, tasks, ntasks); break;
    default: return LATER_ERR_OPTION;
    }
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

} // namespace

int gemm_grouped(hipStream_t s, int prec, bool proj, bool ch64, const GemmArgs *args, const int4 *tasks,
                 int ntasks)
{
    if (ntasks <= 0) return LATER_OK;
    if (proj) return ch64 ? launch_grouped_p<true, true>(s, prec, args, tasks, ntasks)
                          : launch_grouped_p<true, false>(s, prec, args, tasks, ntasks);
    return ch64 ? launch_grouped_p<false, true>(s, prec, args, tasks, ntasks)
                : launch_grouped_p<false, false>(s, prec, args, tasks, ntasks);
}

int colexp_grouped(hipStream_t s, const ColexpArgs *args, const int2 *tasks, int ntasks)
{
    if (ntasks <= 0) return LATER_OK;
    hipLaunchKernelGGL(colexp_grouped_kernel, dim3(ntasks), dim3(64), 0, s, args, tasks);
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

// Split plan of the projection, a function of (m, n1, n2, chunk) only, so the
// summation order (and the result bits) never depends on the caller's
// workspace: enough workgroups to cover the CUs twice, at least 256 rows of K
// per split, and a partial buffer of at most 1/8 of the operands' size.
static void proj_plan(int m, int n1, int n2, int chunk, bool *small, int *nsplit, int *kper)
{
    const int64_t tiles128 = ceil_div64(n1, 128) * ceil_div64(n2, 128);
    *small = (n1 <= 64 || n2 <= 64 || tiles128 < 128);
    const int bm = *small ? 64 : 128;
    const int64_t tiles = ceil_div64(n1, bm) * ceil_div64(n2, bm);
    int64_t ns = ceil_div64(512, tiles);
    const int64_t maxs = m / 256 > 1 ? m / 256 : 1;
    if (ns > maxs) ns = maxs;
    const int64_t per = (int64_t)n1 * n2 * (chunk > 0 ? 2 : 1);
    const int64_t budget = (int64_t)m * (n1 + n2) / 8 / per;
    if (ns > budget) ns = budget;
    if (ns < 1) ns = 1;
    int64_t kp = ceil_div64(m, ns);
    const int gran = (chunk > kMaxBK) ? chunk : kMaxBK;
    kp = ceil_div64(kp, gran) * gran;
    *nsplit = (int)ceil_div64(m, kp);
    *kper = (int)kp;
}

int64_t gemm_proj_work_floats(int m, int n1, int n2, int chunk)
{
    bool small;
    int ns, kp;
    proj_plan(m, n1, n2, chunk, &small, &ns, &kp);
    return ns > 1 ? (int64_t)ns * n1 * n2 * (chunk > 0 ? 2 : 1) : 0;
}

int gemm_proj(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *B,
              int64_t ldb, float *C, int64_t ldc, int chunk, const int *b_exp, int a_exp, float *work,
              int64_t wfloats, int *launches)
{
    const bool ch64 = chunk > 0;
    if (ch64 && (chunk % 32 || (chunk < kMaxBK && kMaxBK % chunk))) return LATER_ERR_OPTION;
    bool small;
    int ns, kp;
    proj_plan(m, n1, n2, chunk, &small, &ns, &kp);
    if (ns > 1 && wfloats < (int64_t)ns * n1 * n2 * (ch64 ? 2 : 1)) return LATER_ERR_WORKSPACE;
    GemmArgs g;
    g.M = n1; g.N = n2; g.K = m;
    g.A = Q; g.lda = ldq; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    g.a_exp = a_exp; g.b_exp = b_exp;
    g.kper = kp;
    g.chunk_rows = ch64 ? chunk : 0;
    g.partial = (ns > 1) ? (void *)work : nullptr;
    g.update = 0;
    g.alpha = 1.0f;
    g.tri = 0;
    const int tok = prof_start(s);
    int rc = ch64 ? dispatch<true, true>(s, prec, g, ns, small) : dispatch<true, false>(s, prec, g, ns, small);
    prof_stop(tok, s, PROF_GEMM_PROJ, 2.0 * m * n1 * n2, 4.0 * ((double)m * (n1 + n2) + (double)n1 * n2));
    if (rc) return rc;
    if (launches) ++*launches;
    if (ns > 1) {
        const int64_t tot = (int64_t)n1 * n2;
        dim3 grid((unsigned)ceil_div64(tot, 256));
        if (ch64) hipLaunchKernelGGL(splitk_reduce_kernel<true>, grid, dim3(256), 0, s, n1, n2, ns, work, C, ldc);
        else hipLaunchKernelGGL(splitk_reduce_kernel<false>, grid, dim3(256), 0, s, n1, n2, ns, work, C, ldc);
        HIP_OK(hipGetLastError());
        if (launches) ++*launches;
    }
    return LATER_OK;
}

int gemm_update(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *Cm,
                int64_t ldc, float *B, int64_t ldb, bool f64, const int *c_exp, int a_exp, int *launches)
{
    return gemm_nn(s, prec, m, n2, n1, -1.0f, Q, ldq, Cm, ldc, B, ldb, f64, c_exp, a_exp, false, launches);
}

int gemm_nn(hipStream_t s, int prec, int M, int N, int K, float alpha, const float *A, int64_t lda,
            const float *B, int64_t ldb, float *C, int64_t ldc, bool f64, const int *b_exp, int a_exp, bool tri,
            int *launches)
{
    GemmArgs g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    g.a_exp = a_exp; g.b_exp = b_exp;
    g.kper = K;
    g.chunk_rows = f64 ? kFoldAtEnd : 0;
    g.partial = nullptr;
    g.update = 1;
    g.alpha = alpha;
    g.tri = tri ? 1 : 0;
    const bool small = (N <= 64);
    const int tok = prof_start(s);
    int rc = f64 ? dispatch<false, true>(s, prec, g, 1, small) : dispatch<false, false>(s, prec, g, 1, small);
    prof_stop(tok, s, PROF_GEMM_UPDATE, tri ? (double)M * N * K / 3.0 : 2.0 * M * N * K,
              4.0 * ((double)M * K + 2.0 * M * N + (double)K * N));
    if (rc) return rc;
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_colexp(hipStream_t s, int m, int n, const float *X, int64_t ldx, int *e, int *launches)
{
    if (n <= 0) return LATER_OK;
    hipLaunchKernelGGL(colexp_kernel, dim3(n), dim3(64), 0, s, m, n, X, ldx, e);
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}
