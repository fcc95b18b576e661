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
#include <type_traits>

#include "common.h"
#include "gemm.h"
#include "prof.h"

namespace {

constexpr int NT = 256;  // 4 waves in a 2 x 2 grid
constexpr int kMaxBK = 64;
// FP64 folds of the split forms (chunk_rows > 0): every chunk_rows rows of K
// and after the last k, the leading accumulator group (the hi x hi products)
// is added into the FP64 sums and cleared; every kLeadFolds-th time, and
// after the last k, the other groups follow.  The lower groups hold products 2^-11 and 2^-22 smaller, so their FP32
// sums may run kLeadFolds times longer for the same error relative to the
// result.  Both GEMM kernels fold at the same k, so their bits agree.
constexpr int kLeadFolds = 8;

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

// Slabs of global loads in flight (see the main loop of gemm_body).
template <int PREC, bool CH64, int BM> constexpr int stages()
{
    if (PREC == GEMM_F32) return 2;
    if (BM == 64) return 4;
    if (PREC == GEMM_F16X3 && CH64) return 1;  // the FP64 fold's registers leave no room for more
    return bk<PREC, CH64>() == 64 ? 2 : 3;      // 64-deep slabs of the hi/lo pair: 3 would spill
}

template <int PREC> struct Lds;
template <> struct Lds<GEMM_F32> { typedef float T; };
template <> struct Lds<GEMM_F16> { typedef _Float16 T; };
template <> struct Lds<GEMM_BF16> { typedef __bf16 T; };
template <> struct Lds<GEMM_F16X3> { typedef _Float16 T; };
template <> struct Lds<GEMM_F16X6> { typedef _Float16 T; };
template <> struct Lds<GEMM_I8X4> { typedef int8_t T; };

// FP16 pieces per operand value: x = p0 + p1 2^-11 + p2 2^-22, with
// p0 = fl16(x), p1 = fl16((x - p0) 2^11), p2 = fl16((x - p0 - p1 2^-11) 2^22).
// Products are summed per power of 2^-11 in separate FP32 accumulators:
// group g holds sum_{i+j=g} p_i q_j.
template <int PREC> constexpr int pieces()
{
    return PREC == GEMM_I8X4 ? 4 : PREC == GEMM_F16X6 ? 3 : PREC == GEMM_F16X3 ? 2 : 1;
}

// The MFMA input pieces of one (already scaled) FP32 value.  The fused
// conversion of the slab stores and the split kernels both use this, so
// pre-split operands give the MFMAs the same bits.
template <int PREC>
__device__ __forceinline__ void split_value(float x, typename Lds<PREC>::T *out)
{
    typedef typename Lds<PREC>::T T;
    if constexpr (pieces<PREC>() == 1) {
        out[0] = (T)x;
    } else {
        const _Float16 h = (_Float16)x;
        const float r = x - (float)h;  // exact
        const _Float16 m = (_Float16)(r * 2048.0f);
        out[0] = h;
        out[1] = m;
        if constexpr (pieces<PREC>() == 3) out[2] = (_Float16)((r - (float)m * 0x1p-11f) * 0x1p22f);
    }
}



// Loads of one BK-deep slab of an R-row operand tile into registers.
//   KM (rows contiguous along K): the row's BK values are CPL 16-byte chunks;
//     consecutive lanes take consecutive chunks of one row, so every wave
//     load covers whole 128-byte lines (8 rows of 32 k, or 4 rows of 64 k).
//     A thread owns 4 consecutive k of NPASS rows, RPI rows apart.
//   !KM (rows contiguous along M): thread owns 4 consecutive rows x KG k.
template <int R, bool KM, int BK> struct Slab {
    static constexpr int NV = R * BK / NT;
    static constexpr int CPL = BK / 4;          // KM: chunks per row
    static constexpr int RPI = NT / CPL;        // KM: rows per pass of the workgroup
    static constexpr int NPASS = NV / 4;        // KM: rows per thread
    static constexpr int KG = NV / 4;           // !KM: k per thread
    float v[NV];

    // Loads with clamped in-bounds addresses and a select afterwards, so that
    // no load sits behind a branch (all of a thread's loads issue before the
    // first wait).  rmax, kmax >= 1.
    __device__ __forceinline__ void load(const float *P, int64_t ld, int r0, int rmax, int k0, int kmax,
                                         bool vec)
    {
        // Global view: the grouped kernels read P from an argument table.
        const auto *G = gmem(P);
        typedef const __attribute__((address_space(1))) floatx4 GV4;
        const int t = threadIdx.x;
        if (KM) {
            const int k = k0 + 4 * (t % CPL);
#pragma unroll
            for (int p = 0; p < NPASS; ++p) {
                const int row = r0 + p * RPI + t / CPL;
                const bool rin = row < rmax;
                const auto *q = G + (int64_t)min(row, rmax - 1) * ld;
                if (vec && rin && k + 4 <= kmax) {
                    floatx4 x = *(GV4 *)(q + k);
                    v[4 * p] = x[0]; v[4 * p + 1] = x[1]; v[4 * p + 2] = x[2]; v[4 * p + 3] = x[3];
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[4 * p + e] = q[min(k + e, kmax - 1)];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (!(rin && k + e < kmax)) v[4 * p + e] = 0.0f;
                }
            }
        } else {
            const int rg = r0 + (t % (R / 4)) * 4;
            const int kb = k0 + (t / (R / 4)) * KG;
            if (vec && rg + 4 <= rmax && kb + KG <= kmax) {
#pragma unroll
                for (int kk = 0; kk < KG; ++kk) {
                    floatx4 x = *(GV4 *)(G + (int64_t)(kb + kk) * ld + rg);
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) v[kk * 4 + rr] = x[rr];
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < KG; ++kk)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        v[kk * 4 + rr] = G[(int64_t)min(kb + kk, kmax - 1) * ld + min(rg + rr, rmax - 1)];
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
        return KM ? (e >> 2) * RPI + t / CPL : (t % (R / 4)) * 4 + (e & 3);
    }
    __device__ __forceinline__ int k_of(int e) const
    {
        const int t = threadIdx.x;
        return KM ? 4 * (t % CPL) + (e & 3) : (t / (R / 4)) * KG + (e >> 2);
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

// Writes a thread's slab to LDS, converted to the MFMA input type.  A K-major
// slab holds 4 consecutive k of each of its rows (one 8-byte store per row
// and piece for the 16-bit types); an M-major slab holds KG consecutive k of
// each of 4 rows (KG-wide stores for the 16-bit types, 4-row vectors for
// FP32, whose LDS rows are contiguous along M).  sc: the scale of each row
// (K-major: one per pass) or of the whole slab (M-major).
template <int PREC, int R, bool KM, int BK>
__device__ __forceinline__ void store_slab(Tile<PREC, R, BK> &t, const Slab<R, KM, BK> &s, const float *sc)
{
    typedef typename Lds<PREC>::T T;
    typedef Slab<R, KM, BK> S;
    constexpr int NP = pieces<PREC>();
    if constexpr (PREC == GEMM_F32) {
        if constexpr (KM) {
#pragma unroll
            for (int e = 0; e < S::NV; ++e) t.p[0][t.idx(s.row_of(e), s.k_of(e))] = s.v[e] * sc[e >> 2];
        } else {
#pragma unroll
            for (int kk = 0; kk < S::KG; ++kk) {
                floatx4 x;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) x[rr] = s.v[kk * 4 + rr] * sc[0];
                *(floatx4 *)&t.p[0][t.idx(s.row_of(kk * 4), s.k_of(kk * 4))] = x;
            }
        }
    } else {
        constexpr int W = KM ? 4 : S::KG;      // elements per store
        constexpr int NST = S::NV / W;          // stores per piece
        typedef typename Vec<T, W>::type VT;
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            VT pv[NP];
            int row = 0, k = 0;
#pragma unroll
            for (int e = 0; e < W; ++e) {
                // K-major: values 4q .. 4q+3 are consecutive k of row q.
                // M-major: value kk*4+rr is row rr, k kk; store q = row rr.
                const int idx = KM ? q * W + e : e * 4 + q;
                T pc[NP];
                split_value<PREC>(s.v[idx] * (KM ? sc[q] : sc[0]), pc);
#pragma unroll
                for (int c = 0; c < NP; ++c) pv[c][e] = pc[c];
                if (e == 0) { row = s.row_of(idx); k = s.k_of(idx); }
            }
#pragma unroll
            for (int c = 0; c < NP; ++c) *(VT *)&t.p[c][t.idx(row, k)] = pv[c];
        }
    }
}

typedef unsigned short ushort8 __attribute__((ext_vector_type(8)));

// Loads of one BK-deep slab of pre-split 16-bit pieces (K-major, piece c at
// P + c * piece): consecutive lanes take consecutive 16-byte chunks (8 k) of
// one row.  A thread owns one chunk of NPASS rows, RPI rows apart.
template <int R, int BK, int NP> struct Slab16 {
    static constexpr int CPL = BK / 8;      // chunks per row
    static constexpr int RPI = NT / CPL;    // rows per pass of the workgroup
    static constexpr int NPASS = R / RPI;
    ushort8 v[NP][NPASS];

    __device__ __forceinline__ void load(const uint16_t *P, int64_t ld, int64_t piece, int r0, int rmax, int k0,
                                         int kmax, bool vec)
    {
        const auto *G = gmem(P);
        const int t = threadIdx.x;
        const int k = k0 + 8 * (t % CPL);
#pragma unroll
        for (int p = 0; p < NPASS; ++p) {
            const int row = r0 + p * RPI + t / CPL;
            const bool rin = row < rmax;
#pragma unroll
            for (int c = 0; c < NP; ++c) {
                const auto *q = G + c * piece + (int64_t)min(row, rmax - 1) * ld;
                if (vec && rin && k + 8 <= kmax) {
                    v[c][p] = *(const __attribute__((address_space(1))) ushort8 *)(q + k);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const unsigned short u = q[min(k + e, kmax - 1)];
                        v[c][p][e] = (rin && k + e < kmax) ? u : (unsigned short)0;
                    }
                }
            }
        }
    }
    __device__ __forceinline__ int row_of(int p) const { return p * RPI + (int)threadIdx.x / CPL; }
    __device__ __forceinline__ int k_of() const { return 8 * ((int)threadIdx.x % CPL); }
};

template <int PREC, int R, int BK, int NP>
__device__ __forceinline__ void store_slab16(Tile<PREC, R, BK> &t, const Slab16<R, BK, NP> &s)
{
#pragma unroll
    for (int c = 0; c < NP; ++c)
#pragma unroll
        for (int p = 0; p < Slab16<R, BK, NP>::NPASS; ++p)
            *(ushort8 *)&t.p[c][t.idx(s.row_of(p), s.k_of())] = s.v[c][p];
}

template <int PREC> struct Frag;
template <> struct Frag<GEMM_F32> { typedef float T; };
template <> struct Frag<GEMM_F16> { typedef half8_t T; };
template <> struct Frag<GEMM_BF16> { typedef bf16x8_t T; };
template <> struct Frag<GEMM_F16X3> { typedef half8_t T; };
template <> struct Frag<GEMM_F16X6> { typedef half8_t T; };
template <> struct Frag<GEMM_I8X4> { typedef intx4 T; };  // 16 int8 k per lane

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

// Epilogue of one wave's FM x FN accumulator blocks (value(i, j, r), FP64,
// still scaled).  Transposed 32x32 layout: lane l holds row l%32 of the block
// and columns 8 (r/4) + 4 (l/32) + r%4, r = 0..15.
// ROWEXP: rows of op(A) carry their own exponents (int8 slices, at byte
// offset lsa from Ah); the result is scaled by 2^-(a_exp + row's + column's).
template <bool CH64, int FM, int FN, typename V, bool ROWEXP = false>
__device__ __forceinline__ void store_tile(const GemmArgs &g, int split, int row0, int col0, V value)
{
    const int lane = threadIdx.x & 63;
    const bool rmw = !g.partial && g.update;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int row = row0 + 32 * i + (lane & 31);
        if (row >= g.M) continue;
        // An update reads all of its C values first, so that the loads are in
        // flight together (interleaved with the stores they would wait one
        // memory latency each: the compiler cannot move a load of C above a
        // store to C).
        const int rexp = ROWEXP ? gmem((const int *)((const char *)g.Ah + g.lsa))[row] : 0;
        float cv[FN][16];
        if (rmw) {
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int col = col0 + 32 * j + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    cv[j][r] = (col < g.N) ? gmem(g.C)[(int64_t)col * g.ldc + row] : 0.0f;
                }
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int col = col0 + 32 * j + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (col >= g.N) continue;
                const int esum = g.a_exp + rexp + ((g.b_exp) ? gmem(g.b_exp)[col] : 0);
                const double v = value(i, j, r) * ldexp(1.0, -esum);
                if (g.partial) {
                    const int64_t off = ((int64_t)split * g.N + col) * g.M + row;
                    if constexpr (CH64) gmem((double *)g.partial)[off] = v;
                    else gmem((float *)g.partial)[off] = (float)v;
                } else if (g.update) {
                    auto *c = gmem(g.C) + (int64_t)col * g.ldc + row;
                    if (CH64 || g.update == 2) *c = (float)((double)cv[j][r] + (double)g.alpha * v);
                    else *c = fmaf(g.alpha, (float)v, cv[j][r]);
                } else {
                    gmem(g.C)[(int64_t)col * g.ldc + row] = (float)v;
                }
            }
        }
    }
}

// PRE: both operands come pre-split (GemmArgs.Ah / Bh, K-major 16-bit
// pieces); otherwise they are read as FP32 and converted on the way to LDS.
template <int PREC, int BM, int BN, bool AKM, bool CH64, bool PRE>
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
    constexpr int S = stages<PREC, CH64, BM>();
    typedef typename std::conditional<PRE, Slab16<BM, BK, NP>, Slab<BM, AKM, BK>>::type SA;
    typedef typename std::conditional<PRE, Slab16<BN, BK, NP>, Slab<BN, true, BK>>::type SB;
    SA sa[S];
    SB sb[S];
    // Scales of the rows a thread converts: A has one; B's rows are columns
    // of B with one exponent each.  (Pre-split pieces are already scaled.)
    constexpr int NSA = PRE ? 1 : (AKM ? SA::NPASS : 1), NSB = PRE ? 1 : SB::NPASS;
    float ascale[NSA], bscale[NSB];
#pragma unroll
    for (int p = 0; p < NSA; ++p) ascale[p] = ldexpf(1.0f, g.a_exp);
#pragma unroll
    for (int p = 0; p < NSB; ++p) {
        int bcol = 0;
        if constexpr (!PRE) bcol = n0 + sb[0].row_of(4 * p);
        bscale[p] = (g.b_exp && bcol < g.N) ? ldexpf(1.0f, g.b_exp[bcol]) : 1.0f;
    }
    const bool avec16 = PRE && ((g.ldah & 7) == 0) && ((((uintptr_t)g.Ah) & 15) == 0) && ((g.psa & 7) == 0);
    const bool bvec16 = PRE && ((g.ldbh & 7) == 0) && ((((uintptr_t)g.Bh) & 15) == 0) && ((g.psb & 7) == 0);
    auto store = [&](int buf, const SA &a, const SB &b) {
        if constexpr (PRE) {
            store_slab16<PREC, BM, BK, NP>(As[buf], a);
            store_slab16<PREC, BN, BK, NP>(Bs[buf], b);
        } else {
            store_slab<PREC, BM, AKM, BK>(As[buf], a, ascale);
            store_slab<PREC, BN, true, BK>(Bs[buf], b, bscale);
        }
    };

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
    auto fold_lead = [&]() {
        if constexpr (CH64) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc64[i][j][r] += (double)acc[0][i][j][r];
                        acc[0][i][j][r] = 0.0f;
                    }
        }
    };
    auto fold_rest = [&]() {
        if constexpr (CH64 && NG > 1) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        double v = (double)acc[1][i][j][r] * 0x1p-11;
                        if constexpr (NG > 2) v += (double)acc[2][i][j][r] * 0x1p-22;
                        acc64[i][j][r] += v;
#pragma unroll
                        for (int q = 1; q < NG; ++q) acc[q][i][j][r] = 0.0f;
                    }
        }
    };

    const int ntiles = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;

    const int per_ks = CH64 ? g.chunk_rows / KR : 1;  // k-steps per FP64 fold
    int ksince = 0, nlead = 0;
    // One slab of MFMAs from LDS buffer `cur` (slab `it` of this split).
    auto compute = [&](int cur, int it) {
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
            // FP64 folds (kLeadFolds), counted rather than divided.
            if constexpr (CH64) {
                const bool last = ks + 1 == KSTEPS && it + 1 == ntiles;
                if (++ksince == per_ks || last) {
                    ksince = 0;
                    fold_lead();
                    if (++nlead == kLeadFolds || last) {
                        nlead = 0;
                        fold_rest();
                    }
                }
            }
        }
    };
    auto load = [&](int slot_it, SA &a, SB &b) {
        const int k0 = kbeg + slot_it * BK;
        if constexpr (PRE) {
            a.load(g.Ah, g.ldah, g.psa, m0, g.M, k0, kend, avec16);
            b.load(g.Bh, g.ldbh, g.psb, n0, g.N, k0, kend, bvec16);
        } else {
            a.load(g.A, g.lda, m0, g.M, k0, kend, avec);
            b.load(g.B, g.ldb, n0, g.N, k0, kend, bvec);
        }
    };

    // Global loads run S slabs ahead of the MFMAs through a ring of register
    // slabs: slab j lives in slot j % S from its load until it is written to
    // LDS, S iterations later, so S slabs of compute cover its latency.
    // LDS itself is double-buffered (one barrier per slab).
    if (ntiles > 0) {
#pragma unroll
        for (int u = 0; u < S; ++u)
            if (u < ntiles) load(u, sa[u], sb[u]);
        store(0, sa[0], sb[0]);
        __syncthreads();
        if (S < ntiles) load(S, sa[0], sb[0]);
    }
    for (int it0 = 0; it0 < ntiles; it0 += S) {
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int it = it0 + u;
            if (it >= ntiles) break;
            const int cur = it & 1;
            compute(cur, it);
            const int nxt = (u + 1) % S;  // static once the loop over u is unrolled
            if (it + 1 < ntiles) store(cur ^ 1, sa[nxt], sb[nxt]);
            __syncthreads();
            if (it + 1 + S < ntiles) load(it + 1 + S, sa[nxt], sb[nxt]);
        }
    }

    store_tile<CH64, FM, FN>(g, split, m0 + wm * WM, n0 + wn * WN, [&](int i, int j, int r) -> double {
        if constexpr (CH64) return acc64[i][j][r];
        else return combined(i, j, r);
    });
}

// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2: the linear id is renumbered so that every XCD works
// through a contiguous range, and each range walks the tiles in groups of
// kGroup tile-rows (column by column inside a group), so the workgroups
// resident on one XCD share their A and B tiles in its L2.  With 32
// resident tiles per XCD the group is kGroup rows x 32 / kGroup columns:
// 8 x 4 for square tiles, 4 x 8 for the 256 x 128 tiles (half the A strips
// in flight, whose slabs are twice the B slabs).
template <int kGroup = 8>
__device__ __forceinline__ void tile_of(int &tm, int &tn, int &split)
{
    constexpr int kXcd = 8;
    const int nx = gridDim.x, ny = gridDim.y;
    const int total = nx * ny * gridDim.z;
    int L = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const int q = total / kXcd;
    if (L < kXcd * q) L = (L % kXcd) * q + L / kXcd;
    const int tiles = nx * ny;
    split = L / tiles;
    const int r = L % tiles;
    const int first = (r / (kGroup * ny)) * kGroup;
    const int gsz = min(nx - first, kGroup);
    const int rr = r - first * ny;
    tm = first + rr % gsz;
    tn = rr / gsz;
}

template <int PREC, int BM, int BN, bool AKM, bool CH64, bool PRE>
__global__ __launch_bounds__(NT) void gemm_kernel(GemmArgs g)
{
    int tm, tn, split;
    tile_of(tm, tn, split);
    gemm_body<PREC, BM, BN, AKM, CH64, PRE>(g, tm, tn, split);
}

// s_waitcnt with vmcnt <= N and the other counters left alone (gfx9 layout:
// vmcnt[3:0] and [15:14], expcnt[6:4], lgkmcnt[11:8]).
template <int N> __device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Pre-split operands streamed straight into LDS (global_load_lds_dwordx4: each
// lane's 16 bytes land at the wave's LDS base + 16 * lane, no registers).
// BM x 128 tiles, slabs of 32 k, waves in a WGM x 2 grid of (BM / WGM) x 64
// wave tiles.  In a stage, each piece of each operand tile is a run of rows
// of 64 bytes whose four 16-byte chunks are XOR-swizzled by (row / 4) % 4; a
// fragment read (16 rows x one chunk per 16 lanes) then covers all 64 banks
// once.  NS stages are in flight: a wave waits on its own loads with
// s_waitcnt, the barrier publishes everyone's, and the freed stage is refilled
// right after it.  The piece arrays must be zero-padded in K to a multiple of
// 32 (rows of A or B past M or N are clamped: they only feed outputs that are
// not stored).
template <int PREC, bool CH64, int BM, int WGM>
__device__ __forceinline__ void gemm_lds_body(const GemmArgs &g, const int tm, const int tn, const int split)
{
    // BM 128, WGM 2: four waves of 64 x 64; BM 128, WGM 4: eight waves of
    // 32 x 64 (half the accumulators per wave, for the three-piece form);
    // BM 256, WGM 4: eight waves of 64 x 64 (two per SIMD, and a third less
    // operand traffic per product than the 128 x 128 tile).
    constexpr int NW = 2 * WGM, BN = 128;
    constexpr int BK = 32, WM = BM / WGM, WN = 64, FM = WM / 32, FN = 2, KSTEPS = 2;
    constexpr int BPA = BM / 16 / NW, BPB = BN / 16 / NW;  // 1 KB blocks per wave per piece tile
    static_assert(BPA >= 1 && BPB >= 1 && FM >= 1, "tile shape");
    constexpr bool I8 = PREC == GEMM_I8X4;  // int8 slices: 64 k per slab, exact int32 sums
    // Accumulator groups: slice pairs of weight 2^-11g (FP16 pieces, g below
    // the piece count) or 2^-7g (int8 slices, g <= 4: the pairs of weight
    // 2^-28 are as large as a column's typical product times 2^-21, so they
    // stay; the first pairs dropped weigh 2^-35).
    constexpr int NP = pieces<PREC>(), NG = I8 ? 5 : NP;
    static_assert(!(I8 && CH64), "int8 products need no FP64 folds");
    constexpr int PBA = BM * BK * 2, PBB = BN * BK * 2;  // bytes of one piece of one operand tile
    constexpr int SB = NP * (PBA + PBB);                 // bytes of one stage
    constexpr int NS = (163840 / SB) < 4 ? (163840 / SB) : 4;
    static_assert(NS >= 2, "stages");
    constexpr int LPS = NP * (BPA + BPB);                // loads per wave per stage
    typedef typename Frag<PREC>::T FT;
    __shared__ __attribute__((aligned(16))) char lds[NS * SB];

    const int m0 = tm * BM, n0 = tn * BN;
    int kbeg = split * g.kper;
    int kend = min(g.K, kbeg + g.kper);
    if (g.tri) {
        // Upper-triangular operands: tiles below the diagonal get nothing and
        // only k in [m0, n0 + BN) contributes (both multiples of the slab).
        if (m0 > n0 + BN - 1) return;
        kbeg = max(kbeg, m0);
        kend = min(kend, n0 + BN);
    }
    const int ntiles = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave >> 1, wn = wave & 1;

    auto issue = [&](int slab, int stage) {
        const int k0 = kbeg + slab * BK;
        char *sbase = lds + stage * SB;
#pragma unroll
        for (int op = 0; op < 2; ++op) {
            const uint16_t *P = op ? g.Bh : g.Ah;
            const int64_t ld = op ? g.ldbh : g.ldah, ps = op ? g.psb : g.psa;
            const int r0 = op ? n0 : m0, rmax = op ? g.N : g.M;
            const int bpw = op ? BPB : BPA, pb = op ? PBB : PBA;
            char *obase = sbase + (op ? NP * PBA : 0);
#pragma unroll
            for (int i = 0; i < (op ? BPB : BPA); ++i) {
                const int blk = bpw * wave + i;  // 1 KB block of the piece tile
                const int q = blk * 64 + lane;
                const int r = q >> 2;
                const int cg = (q & 3) ^ ((r >> 2) & 3);
                const uint16_t *src = P + (int64_t)min(r0 + r, rmax - 1) * ld + k0 + 8 * cg;
#pragma unroll
                for (int c = 0; c < NP; ++c)
                    __builtin_amdgcn_global_load_lds(
                        (const void *)(src + c * ps),
                        (__attribute__((address_space(3))) void *)(obase + c * pb + blk * 1024), 16, 0, 0);
            }
        }
    };
    auto frag = [&](const char *pbase, int row, int ks) -> FT {
        const int ch = (2 * ks + (lane >> 5)) ^ ((row >> 2) & 3);
        return *(const FT *)(pbase + row * 64 + ch * 16);
    };

    typedef typename std::conditional<I8, intx16, floatx16>::type AccT;
    AccT acc[NG][FM][FN];
    double acc64[CH64 ? FM : 1][CH64 ? FN : 1][CH64 ? 16 : 1];
#pragma unroll
    for (int q = 0; q < NG; ++q)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[q][i][j][r] = 0;
    if constexpr (CH64) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc64[i][j][r] = 0.0;
    }
    auto combined = [&](int i, int j, int r) -> double {
        if constexpr (I8) {
            // Slice-pair sums of weight 2^-7g; each is exact, the FP64 sum
            // rounds once per term.
            return ((((double)acc[0][i][j][r] + (double)acc[1][i][j][r] * 0x1p-7) +
                     (double)acc[2][i][j][r] * 0x1p-14) + (double)acc[3][i][j][r] * 0x1p-21) +
                   (double)acc[4][i][j][r] * 0x1p-28;
        } else {
            double v = (double)acc[0][i][j][r];
            if constexpr (NG > 1) v += (double)acc[1][i][j][r] * 0x1p-11;
            if constexpr (NG > 2) v += (double)acc[2][i][j][r] * 0x1p-22;
            return v;
        }
    };
    auto fold_lead = [&]() {
        if constexpr (CH64) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc64[i][j][r] += (double)acc[0][i][j][r];
                        acc[0][i][j][r] = 0.0f;
                    }
        }
    };
    auto fold_rest = [&]() {
        if constexpr (CH64 && NG > 1) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        double v = (double)acc[1][i][j][r] * 0x1p-11;
                        if constexpr (NG > 2) v += (double)acc[2][i][j][r] * 0x1p-22;
                        acc64[i][j][r] += v;
#pragma unroll
                        for (int q = 1; q < NG; ++q) acc[q][i][j][r] = 0.0f;
                    }
        }
    };

    const int per = CH64 ? g.chunk_rows / BK : 1;  // slabs per FP64 fold
    int since = 0, nlead = 0;
#pragma unroll
    for (int st = 0; st < NS - 1; ++st)
        if (st < ntiles) issue(st, st);
    for (int it = 0; it < ntiles; ++it) {
        if (it + NS - 2 < ntiles) wait_vmcnt<(NS - 2) * LPS>();
        else wait_vmcnt<0>();
        // A bare barrier: this wave's loads of stage `it` are complete (the
        // wait above), so after it every wave's are; __syncthreads() would
        // also wait for the later stages still in flight.  The memory clobber
        // keeps LDS reads and the refill below on their side of it.
        asm volatile("s_barrier" ::: "memory");
        if (it + NS - 1 < ntiles) issue(it + NS - 1, (it + NS - 1) % NS);
        const char *sb = lds + (it % NS) * SB;
        // Fragments double-buffered across the slab's k-steps: the reads of
        // step 1 are in flight while step 0's MFMAs run.
        FT fa[KSTEPS][NP][FM], fb[KSTEPS][NP][FN];
        auto load_frags = [&](int ks) {
#pragma unroll
            for (int c = 0; c < NP; ++c) {
#pragma unroll
                for (int i = 0; i < FM; ++i) fa[ks][c][i] = frag(sb + c * PBA, wm * WM + 32 * i + (lane & 31), ks);
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    fb[ks][c][j] = frag(sb + NP * PBA + c * PBB, wn * WN + 32 * j + (lane & 31), ks);
            }
        };
        load_frags(0);
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ++ks) {
            if (ks + 1 < KSTEPS) load_frags(ks + 1);
#pragma unroll
            for (int gq = 0; gq < NG; ++gq)
#pragma unroll
                for (int a = (gq < NP ? 0 : gq - NP + 1); a <= (gq < NP ? gq : NP - 1); ++a)
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j) {
                            if constexpr (I8)
                                acc[gq][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[ks][gq - a][j], fa[ks][a][i],
                                                                                      acc[gq][i][j], 0, 0, 0);
                            else
                                acc[gq][i][j] = mfma<PREC>(fb[ks][gq - a][j], fa[ks][a][i], acc[gq][i][j]);
                        }
        }
        // FP64 folds (kLeadFolds), counted rather than divided.
        if constexpr (CH64) {
            const bool last = it + 1 == ntiles;
            if (++since == per || last) {
                since = 0;
                fold_lead();
                if (++nlead == kLeadFolds || last) {
                    nlead = 0;
                    fold_rest();
                }
            }
        }
    }

    auto value = [&](int i, int j, int r) -> double {
        if constexpr (CH64) return acc64[i][j][r];
        else return combined(i, j, r);
    };
    store_tile<CH64, FM, FN, decltype(value), I8>(g, split, m0 + wm * WM, n0 + wn * WN, value);
}

// Tile order of a triangular product (one split).  Tile row tm has work
// proportional to the K range [m0, n0 + BN), so the rows are far from equal;
// tile_of's contiguous range per XCD would give XCD 0 the heaviest eight rows.
// Here XCD x (workgroups L with L % 8 == x) takes rows 16 b + x and
// 16 b + 15 - x of every block of 16 rows, heavier blocks first.  Returns
// false (use tile_of) unless the tile rows are a multiple of 16.
__device__ __forceinline__ bool tile_of_tri(int &tm, int &tn, int &split)
{
    const int nx = gridDim.x, ny = gridDim.y;
    if ((nx & 15) || gridDim.z != 1) return false;
    const int L = blockIdx.x + nx * blockIdx.y;
    const int x = L & 7, r = L >> 3;
    const int b = r / (2 * ny), w = r % (2 * ny);
    tn = w % ny;
    tm = 16 * b + (w < ny ? x : 15 - x);
    split = 0;
    return true;
}

template <int PREC, bool CH64, int BM, int WGM>
__global__ __launch_bounds__(128 * WGM) void gemm_lds_kernel(GemmArgs g)
{
    int tm, tn, split;
    if (!(g.tri && tile_of_tri(tm, tn, split))) tile_of<BM == 256 ? 4 : 8>(tm, tn, split);
    gemm_lds_body<PREC, CH64, BM, WGM>(g, tm, tn, split);
}

// FP8 (e4m3) products on staged operands: the streaming kernel's LDS ring
// with 64 bytes per row and slab (64 k instead of 32), one
// v_mfma_scale_f32_32x32x64_f8f6f4 per 32 x 32 block and slab, and the E8M0
// scales of the slab's 128-row K block streamed beside the data (one dword
// load per lane by waves 0 and 1: the tile's row scales are contiguous).
// Lane l of a fragment supplies row l % 32 and k = 32 (l / 32) .. + 31 (two
// 16-byte chunks), with the scale of that row; both operands use the same k
// order, so the contraction is over matching k.  BM x 128 tiles, eight waves
// of (BM / 4) x 64, FP32 accumulation, split-K partials in FP32.
template <int BM>
__device__ __forceinline__ void gemm8_body(const GemmArgs &g, const int tm, const int tn, const int split)
{
    constexpr int NW = 8, WGM = 4, BN = 128;
    constexpr int BK = 64, WM = BM / WGM, WN = 64, FM = WM / 32, FN = 2;
    constexpr int BPA = BM / 16 / NW, BPB = BN / 16 / NW;  // 1 KB blocks (16 rows x 64 B) per wave
    static_assert(BPA >= 1 && BPB >= 1 && FM >= 1, "tile shape");
    constexpr int PBA = BM * 64, PBB = BN * 64, SCB = 256;  // bytes per stage: A, B, each scale row
    constexpr int SB = PBA + PBB + 2 * SCB;
    constexpr int NS = 4;
    static_assert(NS * SB <= 163840, "LDS");
    constexpr int LPS = BPA + BPB;  // data loads per wave per stage (+1 scale load on waves 0 and 1)
    typedef int v8i __attribute__((ext_vector_type(8)));
    typedef int v4i __attribute__((ext_vector_type(4)));
    __shared__ __attribute__((aligned(16))) char lds[NS * SB];

    const int m0 = tm * BM, n0 = tn * BN;
    const int kbeg = split * g.kper;
    const int kend = min(g.K, kbeg + g.kper);
    const int ntiles = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const uint8_t *Pa = (const uint8_t *)g.Ah, *Pb = (const uint8_t *)g.Bh;

    auto issue = [&](int slab, int stage) {
        const int k0 = kbeg + slab * BK;
        char *sbase = lds + stage * SB;
#pragma unroll
        for (int op = 0; op < 2; ++op) {
            const uint8_t *P = op ? Pb : Pa;
            const int64_t ld = op ? g.ldbh : g.ldah;
            const int r0 = op ? n0 : m0, rmax = op ? g.N : g.M;
            char *obase = sbase + (op ? PBA : 0);
#pragma unroll
            for (int i = 0; i < (op ? BPB : BPA); ++i) {
                const int blk = (op ? BPB : BPA) * wave + i;
                const int q = blk * 64 + lane;
                const int r = q >> 2;
                const int cg = (q & 3) ^ ((r >> 2) & 3);
                const uint8_t *src = P + (int64_t)min(r0 + r, rmax - 1) * ld + k0 + 16 * cg;
                __builtin_amdgcn_global_load_lds((const void *)src,
                                                 (__attribute__((address_space(3))) void *)(obase + blk * 1024), 16,
                                                 0, 0);
            }
        }
        if (wave < 2) {
            // Scales of rows r0 .. r0 + 255 of K block k0 / 128 (the scale
            // rows are padded to a multiple of 256, so every dword is in
            // bounds; rows past M or N only feed outputs that are not stored).
            const int kb = k0 >> 7;
            const uint8_t *S = (wave ? Pb + g.psb + (int64_t)kb * g.lsb + n0 : Pa + g.psa + (int64_t)kb * g.lsa + m0);
            __builtin_amdgcn_global_load_lds((const void *)(S + 4 * lane),
                                             (__attribute__((address_space(3))) void *)(sbase + PBA + PBB +
                                                                                        wave * SCB),
                                             4, 0, 0);
        }
    };
    auto frag = [&](const char *pbase, int row) -> v8i {
        const int sw = (row >> 2) & 3, c0 = 2 * (lane >> 5);
        const v4i x = *(const v4i *)(pbase + row * 64 + ((c0 ^ sw) * 16));
        const v4i y = *(const v4i *)(pbase + row * 64 + (((c0 + 1) ^ sw) * 16));
        return v8i{x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3]};
    };

    floatx16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

#pragma unroll
    for (int st = 0; st < NS - 1; ++st)
        if (st < ntiles) issue(st, st);
    for (int it = 0; it < ntiles; ++it) {
        if (it + NS - 2 < ntiles) {
            if (wave < 2) wait_vmcnt<(NS - 2) * (LPS + 1)>();
            else wait_vmcnt<(NS - 2) * LPS>();
        } else {
            wait_vmcnt<0>();
        }
        asm volatile("s_barrier" ::: "memory");
        if (it + NS - 1 < ntiles) issue(it + NS - 1, (it + NS - 1) % NS);
        const char *sb = lds + (it % NS) * SB;
        const uint8_t *sca = (const uint8_t *)(sb + PBA + PBB), *scb = sca + SCB;
        v8i fa[FM], fb[FN];
        int sa[FM], sbv[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int row = wm * WM + 32 * i + (lane & 31);
            fa[i] = frag(sb, row);
            sa[i] = sca[row];
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int row = wn * WN + 32 * j + (lane & 31);
            fb[j] = frag(sb + PBA, row);
            sbv[j] = scb[row];
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fb[j], fa[i], acc[i][j], 0, 0, 0, sbv[j],
                                                                            0, sa[i]);
    }

    store_tile<false, FM, FN>(g, split, m0 + wm * WM, n0 + wn * WN,
                              [&](int i, int j, int r) -> double { return (double)acc[i][j][r]; });
}

template <int BM>
__global__ __launch_bounds__(512) void gemm8_kernel(GemmArgs g)
{
    int tm, tn, split;
    tile_of(tm, tn, split);
    gemm8_body<BM>(g, tm, tn, split);
}

// FP32 -> e4m3 staging (see fp8_ld): 32 lanes per 128-row block of one
// column, four rows per lane; the block's largest magnitude (over the 32
// lanes) sets the power of two e, and the bytes are v_cvt_pk_fp8_f32's
// round-to-nearest-even of x 2^e.  Rows K .. roundup(K, 64) - 1 are zeros.
__global__ __launch_bounds__(256) void split8_kernel(int K, const float *X, int64_t ldx, uint8_t *out, int64_t ldo,
                                                     uint8_t *sc, int64_t ls)
{
    const int j = blockIdx.y;
    const int k = (blockIdx.x * 256 + threadIdx.x) * 4;
    const int prow = (K + 63) / 64 * 64;
    if (blockIdx.x * 1024 >= prow) return;  // uniform over the workgroup
    const float *x = X + (int64_t)j * ldx;
    float v[4];
    if (k + 4 <= K && ((((uintptr_t)(x + k)) & 15) == 0)) {
        const floatx4 a = *(const floatx4 *)(x + k);
        v[0] = a[0]; v[1] = a[1]; v[2] = a[2]; v[3] = a[3];
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (k + e < K) ? x[k + e] : 0.0f;
    }
    float mx = fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3])));
#pragma unroll
    for (int mask = 1; mask < 32; mask <<= 1) mx = fmaxf(mx, __shfl_xor(mx, mask, 32));
    int e = 0;
    if (mx > 0.0f && mx < INFINITY) {
        int fe;
        const float f = frexpf(mx, &fe);  // mx = f 2^fe, f in [0.5, 1); 448 = 0.875 2^9
        e = (f <= 0.875f) ? 9 - fe : 8 - fe;
        e = e < -126 ? -126 : (e > 126 ? 126 : e);
    }
    const float s = ldexpf(1.0f, e);
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * s, v[1] * s, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * s, v[3] * s, w, true);
    if (k < prow) *(int *)(out + (int64_t)j * ldo + k) = w;
    if ((threadIdx.x & 31) == 0 && k < K) sc[(int64_t)(k >> 7) * ls + j] = (uint8_t)(127 - e);
}

// FP32 -> int8 slices (see i8_ld), one workgroup per column: the column's
// largest magnitude mx = f 2^E (f in [0.5, 1)) gives e = 6 - E, so |x 2^e| < 64;
// then s0 = rint(x 2^e) and s_c = rint(128 r_{c-1}) with r_c the remainder
// after slice c (each remainder is exact in FP32 and |r| <= 1/2, so every
// slice lies in [-64, 64]).  Rows K .. roundup(K, 64) - 1 are zeros.
// Thread t owns rows 4 t + 1024 i .. + 3: columns of up to kI8RegRows rows
// are read once into registers (16-byte loads when the column is aligned),
// longer ones are read twice (the maximum, then the slices).
constexpr int kI8RegRows = 16384;

__device__ __forceinline__ void i8_words(floatx4 v, float sc, int (&w)[4])
{
#pragma unroll
    for (int c = 0; c < 4; ++c) w[c] = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float r = v[q] * sc;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float sl = rintf(r);
            r = (r - sl) * 128.0f;
            w[c] |= ((int)sl & 0xff) << (8 * q);
        }
    }
}

__device__ __forceinline__ floatx4 i8_load4(const __attribute__((address_space(1))) float *x, int k, int K, bool vec)
{
    if (vec && k + 4 <= K) return *(const __attribute__((address_space(1))) floatx4 *)(x + k);
    floatx4 v;
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (k + q < K) ? x[k + q] : 0.0f;
    return v;
}

__global__ __launch_bounds__(256) void split_i8_kernel(int K, const float *X, int64_t ldx, int8_t *out, int64_t ld,
                                                       int64_t stride, int *exps)
{
    __shared__ float red[4];
    const int j = blockIdx.x, t = threadIdx.x;
    const auto *x = gmem(X + (int64_t)j * ldx);
    const bool vec = (((uintptr_t)(X + (int64_t)j * ldx)) & 15) == 0;
    const int prow = (K + 63) / 64 * 64;
    constexpr int NI = kI8RegRows / 1024;
    floatx4 v[NI];
    float mx = 0.0f;
    const bool inreg = prow <= kI8RegRows;
    if (inreg) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int k = 4 * t + 1024 * i;
            v[i] = (k < K) ? i8_load4(x, k, K, vec) : floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int i = 0; i < NI; ++i)
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[i][0]), fabsf(v[i][1])), fmaxf(fabsf(v[i][2]), fabsf(v[i][3]))));
    } else {
        for (int k = 4 * t; k < K; k += 1024) {
            const floatx4 a = i8_load4(x, k, K, vec);
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(a[0]), fabsf(a[1])), fmaxf(fabsf(a[2]), fabsf(a[3]))));
        }
    }
#pragma unroll
    for (int mask = 1; mask < 64; mask <<= 1) mx = fmaxf(mx, __shfl_xor(mx, mask, 64));
    if ((t & 63) == 0) red[t >> 6] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    int e = 0;
    if (mx > 0.0f && mx < INFINITY) {
        int fe;
        frexpf(mx, &fe);
        e = 6 - fe;
        e = e < -126 ? -126 : (e > 126 ? 126 : e);
    }
    if (t == 0) gmem(exps)[j] = e;
    const float sc = ldexpf(1.0f, e);
    auto *o = gmem(out + (int64_t)j * ld);
    auto put = [&](int k, floatx4 a) {
        int w[4];
        i8_words(a, sc, w);
#pragma unroll
        for (int c = 0; c < 4; ++c) *(__attribute__((address_space(1))) int *)(o + c * stride + k) = w[c];
    };
    if (inreg) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int k = 4 * t + 1024 * i;
            if (k < prow) put(k, v[i]);
        }
    } else {
        for (int k = 4 * t; k < prow; k += 1024) put(k, (k < K) ? i8_load4(x, k, K, vec) : floatx4{0.0f, 0.0f, 0.0f, 0.0f});
    }
}

// Grouped launch: workgroup b computes tile (t.y, t.z) of split t.w of the
// product described by args[t.x], t = tasks[b].
template <int PREC, int BM, int BN, bool AKM, bool CH64>
__global__ __launch_bounds__(NT) void gemm_grouped_kernel(const GemmArgs *args, const int4 *tasks)
{
    const int4 t = tasks[blockIdx.x];
    const GemmArgs g = args[t.x];
    gemm_body<PREC, BM, BN, AKM, CH64, false>(g, t.y, t.z, t.w);
}

// C = sum over splits (fixed order, FP64) of the partials; 16 loads in
// flight per thread (the sum itself stays sequential in p).
template <bool F64>
__global__ void splitk_reduce_kernel(int M, int N, int nsplit, const void *partial, float *C, int64_t ldc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)M * N;
    if (idx >= tot) return;
    const int row = (int)(idx % M), col = (int)(idx / M);
    constexpr int kB = 16;
    double s = 0.0;
    for (int p0 = 0; p0 < nsplit; p0 += kB) {
        double v[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int64_t off = (int64_t)min(p0 + u, nsplit - 1) * tot + idx;
            v[u] = F64 ? ((const double *)partial)[off] : (double)((const float *)partial)[off];
        }
#pragma unroll
        for (int u = 0; u < kB; ++u)
            if (p0 + u < nsplit) s += v[u];
    }
    C[(int64_t)col * ldc + row] = (float)s;
}

// Per-column exponents e_j with max_i |X_ij| 2^e_j in [2^14, 2^15) (0 for a
// zero column).  One 256-thread workgroup per column, 16-byte loads where the
// column is aligned (the maximum does not depend on the order).
// Exponent of column x (m values) by the whole 256-thread workgroup:
// e with max |x| 2^e in [2^14, 2^15) (0 for a zero or non-finite maximum).
__device__ __forceinline__ int column_exp(int m, const float *x)
{
    __shared__ float wmax[4];
    float mx = 0.0f;
    const int head = (int)min((int64_t)m, (int64_t)((16 - ((uintptr_t)x & 15)) & 15) / 4);
    for (int i = threadIdx.x; i < head; i += 256) mx = fmaxf(mx, fabsf(x[i]));
    const int nv = (m - head) / 4;
    const floatx4 *xv = (const floatx4 *)(x + head);
    for (int v = threadIdx.x; v < nv; v += 256) {
        const floatx4 q = xv[v];
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(q[0]), fabsf(q[1])), fmaxf(fabsf(q[2]), fabsf(q[3]))));
    }
    for (int i = head + 4 * nv + threadIdx.x; i < m; i += 256) mx = fmaxf(mx, fabsf(x[i]));
#pragma unroll
    for (int mask = 32; mask > 0; mask >>= 1) mx = fmaxf(mx, __shfl_xor(mx, mask, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    int ex = 0;
    if (mx > 0.0f && mx < INFINITY) {
        int fe;
        frexpf(mx, &fe);  // mx in [2^(fe-1), 2^fe)
        ex = 15 - fe;
    }
    return ex;
}

__global__ __launch_bounds__(256) void colexp_kernel(int m, int n, const float *X, int64_t ldx, int *e)
{
    const int j = blockIdx.x;
    const int ex = column_exp(m, X + (int64_t)j * ldx);
    if (threadIdx.x == 0) e[j] = ex;
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

// Pieces of X scaled by 2^e (e = a_exp or col_exp[j]), same layout: each
// thread converts 4 consecutive rows of one column.
// Rows rows .. roundup(rows, 64) - 1 (the K padding the streaming GEMM
// reads) are written as zeros.
// Rows i0 .. i0 + 7 of column x (zeros past `rows`): two 16-byte loads where
// the column is aligned.
__device__ __forceinline__ void load_rows8(const float *x, int rows, int i0, float (&v)[8])
{
    if (i0 + 8 <= rows && ((((uintptr_t)(x + i0)) & 15) == 0)) {
        const floatx4 a = *(const floatx4 *)(x + i0), b = *(const floatx4 *)(x + i0 + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) { v[e] = a[e]; v[4 + e] = b[e]; }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (i0 + e < rows) ? x[i0 + e] : 0.0f;
    }
}

// Pieces of rows i0 .. i0 + 7 (values v, zeros past `rows`) at o + c * stride.
template <int PREC>
__device__ __forceinline__ void store_rows8(const float (&v)[8], uint16_t *o, int rows, int prow, int64_t stride,
                                            int i0, float sc)
{
    typedef typename Lds<PREC>::T T;
    constexpr int NP = pieces<PREC>();
    const bool vec = i0 + 8 <= rows && ((((uintptr_t)(o + i0)) & 15) == 0) && (stride & 7) == 0;
    ushort8 pv[NP];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        T pc[NP];
        split_value<PREC>(v[e] * sc, pc);
#pragma unroll
        for (int c = 0; c < NP; ++c) pv[c][e] = (i0 + e < rows) ? __builtin_bit_cast(uint16_t, pc[c]) : (uint16_t)0;
    }
    if (vec) {
#pragma unroll
        for (int c = 0; c < NP; ++c) *(ushort8 *)(o + c * stride + i0) = pv[c];
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e)
            if (i0 + e < prow)
#pragma unroll
                for (int c = 0; c < NP; ++c) o[c * stride + i0 + e] = pv[c][e];
    }
}

// Eight rows i0 .. i0 + 7 of column x into its pieces o + c * stride: two
// 16-byte loads and one 16-byte store per piece where the column and its
// piece rows are aligned; rows rows .. prow - 1 are written as zeros.
template <int PREC>
__device__ __forceinline__ void split_rows8(const float *x, uint16_t *o, int rows, int prow, int64_t stride, int i0,
                                            float sc)
{
    float v[8];
    load_rows8(x, rows, i0, v);
    store_rows8<PREC>(v, o, rows, prow, stride, i0, sc);
}

// Pieces of X scaled by 2^e (e = a_exp or col_exp[j]), same layout: each
// thread converts 8 consecutive rows of one column.
// Rows rows .. roundup(rows, 64) - 1 (the K padding the streaming GEMM
// reads) are written as zeros.
template <int PREC>
__global__ void split_kernel(int rows, int cols, const float *X, int64_t ldx, int a_exp, const int *col_exp,
                             uint16_t *out, int64_t ldo, int64_t stride)
{
    const int j = blockIdx.y;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 8;
    const int prow = (rows + 63) / 64 * 64;
    if (i0 >= prow) return;
    const float sc = ldexpf(1.0f, col_exp ? col_exp[j] : a_exp);
    split_rows8<PREC>(X + (int64_t)j * ldx, out + (int64_t)j * ldo, rows, prow, stride, i0, sc);
}

// The column exponents (as colexp_kernel) and the pieces scaled by them (as
// split_kernel) in one pass per column: workgroup j finds column j's maximum,
// writes e[j], then splits the column, whose second read comes from cache.
// Columns of up to kRegRows rows are read once, into registers (thread t
// owns rows 8 t + 2048 i .. + 7), for both the maximum and the pieces.
constexpr int kRegRows = 16384;

template <int PREC>
__global__ __launch_bounds__(256) void split_exp_kernel(int rows, const float *X, int64_t ldx, int *col_exp,
                                                        uint16_t *out, int64_t ldo, int64_t stride)
{
    const int j = blockIdx.x;
    const float *x = X + (int64_t)j * ldx;
    const int prow_r = (rows + 63) / 64 * 64;
    if (prow_r <= kRegRows) {
        constexpr int NI = kRegRows / 2048;
        __shared__ float wmax[4];
        float v[NI][8];
        float mx = 0.0f;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int i0 = 8 * (int)threadIdx.x + 2048 * i;
            if (i0 < rows) load_rows8(x, rows, i0, v[i]);
            else
#pragma unroll
                for (int e = 0; e < 8; ++e) v[i][e] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int e = 0; e < 8; ++e) mx = fmaxf(mx, fabsf(v[i][e]));
#pragma unroll
        for (int mask = 32; mask > 0; mask >>= 1) mx = fmaxf(mx, __shfl_xor(mx, mask, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        int ex = 0;
        if (mx > 0.0f && mx < INFINITY) {
            int fe;
            frexpf(mx, &fe);
            ex = 15 - fe;
        }
        if (threadIdx.x == 0) col_exp[j] = ex;
        const float sc = ldexpf(1.0f, ex);
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int i0 = 8 * (int)threadIdx.x + 2048 * i;
            if (i0 < prow_r) store_rows8<PREC>(v[i], out + (int64_t)j * ldo, rows, prow_r, stride, i0, sc);
        }
        return;
    }
    const int ex = column_exp(rows, x);
    if (threadIdx.x == 0) col_exp[j] = ex;
    const float sc = ldexpf(1.0f, ex);
    const int prow = (rows + 63) / 64 * 64;
    for (int i0 = threadIdx.x * 8; i0 < prow; i0 += 256 * 8)
        split_rows8<PREC>(x, out + (int64_t)j * ldo, rows, prow, stride, i0, sc);
}

// Transposed pieces: element (i, j) of X goes to out + i * ldo + j.  64 x 64
// tiles through LDS, so both the FP32 reads (along i) and the 16-bit writes
// (along j) are contiguous.
// jw: the columns written, cols rounded up to 64 (zero K padding) or cols
// exactly (a column range of a wider persistent array: nothing past it is
// touched).  tri: X is read as its upper triangle minus `shift` on the
// diagonal (zeros below), e.g. R2 - I without a copy of it.
template <int PREC>
__global__ __launch_bounds__(256) void split_t_kernel(int rows, int cols, const float *X, int64_t ldx, int a_exp,
                                                     const int *col_exp, uint16_t *out, int64_t ldo, int64_t stride,
                                                     int jw, int tri, float shift)
{
    typedef typename Lds<PREC>::T T;
    constexpr int NP = pieces<PREC>();
    __shared__ float tile[64][65];
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // All sixteen loads (and exponents) from clamped in-bounds addresses
    // first, so they are in flight together; the values outside the matrix
    // (or below the diagonal) are replaced by zeros afterwards.
    const int i = i0 + tx, ic = min(i, rows - 1);
    float xv[16];
    int ev[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int jc = min(j0 + ty + 4 * q, cols - 1);
        xv[q] = X[ic + (int64_t)jc * ldx];
        ev[q] = col_exp ? col_exp[jc] : a_exp;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int jj = ty + 4 * q, j = j0 + jj;
        float x = xv[q];
        if (tri && i == j) x -= shift;
        const float v = x * ldexpf(1.0f, ev[q]);
        tile[jj][tx] = (i < rows && j < cols && !(tri && i > j)) ? v : 0.0f;
    }
    __syncthreads();
    // Thread (ii, jq): 16 consecutive j of row i0 + ii, two 16-byte stores per
    // piece (columns cols .. roundup(cols, 64) - 1 are the zero K padding).
    const int ii = threadIdx.x >> 2, jq = (threadIdx.x & 3) * 16, io = i0 + ii;
    if (io < rows) {
        ushort8 pv[NP][2];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            T pc[NP];
            split_value<PREC>(tile[jq + u][ii], pc);
#pragma unroll
            for (int c = 0; c < NP; ++c) pv[c][u >> 3][u & 7] = __builtin_bit_cast(uint16_t, pc[c]);
        }
        if (j0 + jq + 16 <= jw) {
#pragma unroll
            for (int c = 0; c < NP; ++c) {
                ushort8 *dst = (ushort8 *)(out + c * stride + (int64_t)io * ldo + j0 + jq);
                dst[0] = pv[c][0];
                dst[1] = pv[c][1];
            }
        } else {
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (j0 + jq + u < jw)
#pragma unroll
                    for (int c = 0; c < NP; ++c) out[c * stride + (int64_t)io * ldo + j0 + jq + u] = pv[c][u >> 3][u & 7];
        }
    }
}

template <int PREC>
void launch_split(hipStream_t s, int rows, int cols, const float *X, int64_t ldx, int a_exp, const int *col_exp,
                  uint16_t *out, int64_t ldo, int64_t stride, bool trans, bool pad = true, bool tri = false,
                  float shift = 0.0f)
{
    if (trans) {
        dim3 grid((unsigned)ceil_div64(rows, 64), (unsigned)ceil_div64(cols, 64));
        const int jw = pad ? (int)(ceil_div64(cols, 64) * 64) : cols;
        hipLaunchKernelGGL(split_t_kernel<PREC>, grid, dim3(256), 0, s, rows, cols, X, ldx, a_exp, col_exp, out, ldo,
                           stride, jw, tri ? 1 : 0, shift);
    } else {
        dim3 grid((unsigned)ceil_div64((rows + 63) / 64 * 64, 2048), (unsigned)cols);
        hipLaunchKernelGGL(split_kernel<PREC>, grid, dim3(256), 0, s, rows, cols, X, ldx, a_exp, col_exp, out, ldo,
                           stride);
    }
}

template <int PREC, int BM, int BN, bool AKM, bool CH64, bool PRE>
void launch_one(hipStream_t s, const GemmArgs &g, dim3 grid)
{
    hipLaunchKernelGGL((gemm_kernel<PREC, BM, BN, AKM, CH64, PRE>), grid, dim3(NT), 0, s, g);
}

template <int PREC, bool AKM, bool CH64, bool PRE>
void launch_tile(hipStream_t s, const GemmArgs &g, int nsplit, bool small)
{
    if constexpr (PRE) {
        // Pre-split operands stream through LDS on 128- or 256-row tiles;
        // thin products (a side of 64 or less) keep the register-staged
        // kernel on 64 x 64 tiles.  The tile size does not change any
        // element's summation order, so the bits are the same.
        if (small && (g.M <= 64 || g.N <= 64)) {
            const dim3 g64((unsigned)ceil_div64(g.M, 64), (unsigned)ceil_div64(g.N, 64), (unsigned)nsplit);
            launch_one<PREC, 64, 64, true, CH64, true>(s, g, g64);
        } else {
            // The split forms run eight waves (two per SIMD): 256-row tiles
            // for the two-piece form without FP64 folds (unless that leaves
            // fewer tiles than CUs: small), 128-row tiles where the folds'
            // FP64 accumulators take the registers.
            constexpr int WGM = (PREC == GEMM_F16X6 || PREC == GEMM_F16X3) ? 4 : 2;
            bool big = false;
            if constexpr (PREC == GEMM_F16X3 && !CH64) {
                big = !small;
                if (big) {
                    const dim3 grid((unsigned)ceil_div64(g.M, 256), (unsigned)ceil_div64(g.N, 128),
                                    (unsigned)nsplit);
                    hipLaunchKernelGGL((gemm_lds_kernel<PREC, CH64, 256, WGM>), grid, dim3(128 * WGM), 0, s, g);
                }
            }
            if (!big) {
                const dim3 grid((unsigned)ceil_div64(g.M, 128), (unsigned)ceil_div64(g.N, 128), (unsigned)nsplit);
                hipLaunchKernelGGL((gemm_lds_kernel<PREC, CH64, 128, WGM>), grid, dim3(128 * WGM), 0, s, g);
            }
        }
    } else {
        const dim3 g64((unsigned)ceil_div64(g.M, 64), (unsigned)ceil_div64(g.N, 64), (unsigned)nsplit);
        if constexpr (PREC == GEMM_F16X6) {
            // Three accumulator groups: a 128x128 tile would spill, so the
            // fused three-piece form always runs 64x64 tiles.
            launch_one<PREC, 64, 64, AKM, CH64, false>(s, g, g64);
        } else if (small) {
            launch_one<PREC, 64, 64, AKM, CH64, false>(s, g, g64);
        } else {
            dim3 grid((unsigned)ceil_div64(g.M, 128), (unsigned)ceil_div64(g.N, 128), (unsigned)nsplit);
            launch_one<PREC, 128, 128, AKM, CH64, false>(s, g, grid);
        }
    }
}

// Pre-split operands are K-major on both sides, so those kernels exist in
// the AKM = true form only.
template <bool AKM, bool CH64>
int dispatch(hipStream_t s, int prec, const GemmArgs &g, int nsplit, bool small)
{
    const bool pre = g.Ah && g.Bh;
    if (pre && prec == GEMM_F32) return LATER_ERR_OPTION;
    switch (prec) {
    case GEMM_F32: launch_tile<GEMM_F32, AKM, CH64, false>(s, g, nsplit, small); break;
#define LATER_DISPATCH_16(P)                                                            \
    case P:                                                                             \
        if (pre) launch_tile<P, true, CH64, true>(s, g, nsplit, small);                 \
        else launch_tile<P, AKM, CH64, false>(s, g, nsplit, small);                     \
        break;
    LATER_DISPATCH_16(GEMM_F16)
    LATER_DISPATCH_16(GEMM_BF16)
    LATER_DISPATCH_16(GEMM_F16X3)
    LATER_DISPATCH_16(GEMM_F16X6)
#undef LATER_DISPATCH_16
    case GEMM_FP8: {
        // Staged e4m3 operands only; 256-row tiles when they still give every
        // CU a tile.
        if (!pre || !AKM) return LATER_ERR_WORKSPACE;
        const int64_t t256 = ceil_div64(g.M, 256) * ceil_div64(g.N, 128) * nsplit;
        if (t256 >= 256) {
            const dim3 grid((unsigned)ceil_div64(g.M, 256), (unsigned)ceil_div64(g.N, 128), (unsigned)nsplit);
            hipLaunchKernelGGL(gemm8_kernel<256>, grid, dim3(512), 0, s, g);
        } else {
            const dim3 grid((unsigned)ceil_div64(g.M, 128), (unsigned)ceil_div64(g.N, 128), (unsigned)nsplit);
            hipLaunchKernelGGL(gemm8_kernel<128>, grid, dim3(512), 0, s, g);
        }
        break;
    }
    case GEMM_I8X4: {
        // Staged int8 slices only, 128 x 128 tiles on eight waves (four
        // int32 accumulator groups per 32 x 64 wave tile), every shape.
        if (!pre || !AKM || CH64) return LATER_ERR_WORKSPACE;
        const dim3 grid((unsigned)ceil_div64(g.M, 128), (unsigned)ceil_div64(g.N, 128), (unsigned)nsplit);
        hipLaunchKernelGGL((gemm_lds_kernel<GEMM_I8X4, false, 128, 4>), grid, dim3(512), 0, s, g);
        break;
    }
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
    (void)ch64;  // updates accumulate in FP32 and subtract in FP64 when GemmArgs.update == 2
    return launch_grouped_p<false, false>(s, prec, args, tasks, ntasks);
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
// Split-K plan of a projection: 64 x 64 tiles for thin or few-tile products
// (small), and enough K splits for ~512 tiles x splits, within m / 256
// splits and a partials budget of m (n1 + n2) / 2 floats (m (n1 + n2) / 8
// for thin products).  Tiles are counted at 128 x 128 unless a side is 64 or
// less, so the plan (and every element's summation order) is the same
// whichever kernel runs it.
static void proj_plan(int m, int n1, int n2, int chunk, bool *small, int *nsplit, int *kper)
{
    const int64_t tiles128 = ceil_div64(n1, 128) * ceil_div64(n2, 128);
    const bool thin = n1 <= 64 || n2 <= 64;
    *small = thin || tiles128 < 128;
    const int64_t tiles = thin ? ceil_div64(n1, 64) * ceil_div64(n2, 64) : tiles128;
    int64_t ns = ceil_div64(512, tiles);
    const int64_t maxs = m / 256 > 1 ? m / 256 : 1;
    if (ns > maxs) ns = maxs;
    const int64_t per = (int64_t)n1 * n2 * (chunk > 0 ? 2 : 1);
    // The 64 Ki floats held back from the wide budget keep the partials, with
    // the rest of the factorisation's FP32 workspace, inside m n / 2 floats:
    // the reference's callers size the 16-bit arena at m n halves
    // (test/test_qr.cu:76, EVD/later_qdwh_polar.cu:75), which then holds
    // everything.  Thin products stay within m n / 8 without it.
    const int64_t budget = thin ? (int64_t)m * (n1 + n2) / 8 / per : ((int64_t)m * (n1 + n2) / 2 - 65536) / per;
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

static void set_pieces(GemmArgs &g, const Pieces *ah, const Pieces *bh)
{
    const bool pre = ah && bh && ah->p && bh->p;
    g.Ah = pre ? ah->p : nullptr;
    g.ldah = pre ? ah->ld : 0;
    g.psa = pre ? ah->stride : 0;
    g.lsa = pre ? ah->ls : 0;
    g.Bh = pre ? bh->p : nullptr;
    g.ldbh = pre ? bh->ld : 0;
    g.psb = pre ? bh->stride : 0;
    g.lsb = pre ? bh->ls : 0;
}

int gemm_proj(hipStream_t s, int prec, int m, int n1, int n2, const float *Q, int64_t ldq, const float *B,
              int64_t ldb, float *C, int64_t ldc, int chunk, const int *b_exp, int a_exp, float *work,
              int64_t wfloats, int *launches, const Pieces *qh, const Pieces *bh)
{
    const bool i8 = prec == GEMM_I8X4;
    if (prec == GEMM_FP8 || i8) chunk = 0;  // exact e4m3 / int8 products: no FP64 folds
    const bool ch64 = chunk > 0;
    if (ch64 && (chunk % 32 || (chunk < kMaxBK && kMaxBK % chunk))) return LATER_ERR_OPTION;
    bool small;
    int ns, kp;
    proj_plan(m, n1, n2, chunk, &small, &ns, &kp);
    if (i8 && (kp > kI8MaxK || !qh || !bh)) return kp > kI8MaxK ? LATER_ERR_SHAPE : LATER_ERR_WORKSPACE;
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
    set_pieces(g, qh, bh);
    if (i8) {
        // Byte operands counted in pairs of bytes (K padded to 64 rows);
        // the exponents come with the staged slices.
        g.K = (m + 63) / 64 * 32;
        g.kper = kp / 2;
        g.ldah /= 2; g.psa /= 2; g.ldbh /= 2; g.psb /= 2;
        g.a_exp = 0;
        g.b_exp = (const int *)((const char *)bh->p + bh->ls);
    }
    const int tok = prof_start(s);
    int rc = ch64 ? dispatch<true, true>(s, prec, g, ns, small) : dispatch<true, false>(s, prec, g, ns, small);
    // Unfolded hi/lo projections on pre-split pieces run the updates' 256-row
    // streaming kernel: they are timed in its class.
    const bool x3 = prec == GEMM_F16X3 && !ch64 && !small && g.Ah && g.Bh;
    prof_stop(tok, s, i8 ? PROF_PROJ_I8 : (x3 ? PROF_UPDATE_X3 : PROF_GEMM_PROJ), 2.0 * m * n1 * n2,
              4.0 * ((double)m * (n1 + n2) + (double)n1 * n2));
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
                int64_t ldc, float *B, int64_t ldb, bool f64, const int *c_exp, int a_exp, int *launches,
                const Pieces *qh, const Pieces *ch)
{
    return gemm_nn(s, prec, m, n2, n1, -1.0f, Q, ldq, Cm, ldc, B, ldb, f64, c_exp, a_exp, false, launches, qh, ch);
}

int gemm_nn(hipStream_t s, int prec, int M, int N, int K, float alpha, const float *A, int64_t lda,
            const float *B, int64_t ldb, float *C, int64_t ldc, bool f64, const int *b_exp, int a_exp, bool tri,
            int *launches, const Pieces *ah, const Pieces *bh)
{
    GemmArgs g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    g.a_exp = a_exp; g.b_exp = b_exp;
    g.kper = K;
    g.chunk_rows = 0;
    g.partial = nullptr;
    g.update = f64 ? 2 : 1;
    g.alpha = alpha;
    g.tri = tri ? 1 : 0;
    set_pieces(g, ah, bh);
    // Thin products, and products with fewer 128 x 128 tiles than the chip's
    // 256 CUs, run the smaller tiles.
    const bool small = (N <= 64) || ceil_div64(M, 128) * ceil_div64(N, 128) < 256;
    const int tok = prof_start(s);
    int rc = dispatch<false, false>(s, prec, g, 1, small);
    const bool x3 = prec == GEMM_F16X3 && !small && g.Ah && g.Bh;
    prof_stop(tok, s, x3 ? PROF_UPDATE_X3 : PROF_GEMM_UPDATE, tri ? (double)M * N * K / 3.0 : 2.0 * M * N * K,
              4.0 * ((double)M * K + 2.0 * M * N + (double)K * N));
    if (rc) return rc;
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_colexp(hipStream_t s, int m, int n, const float *X, int64_t ldx, int *e, int *launches)
{
    if (n <= 0) return LATER_OK;
    hipLaunchKernelGGL(colexp_kernel, dim3(n), dim3(256), 0, s, m, n, X, ldx, e);
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_pieces(int prec)
{
    switch (prec) {
    case GEMM_F16: case GEMM_BF16: return 1;
    case GEMM_F16X3: return 2;
    case GEMM_F16X6: return 3;
    default: return 0;
    }
}

int gemm_split_exp(hipStream_t s, int prec, int rows, int cols, const float *X, int64_t ldx, int *col_exp,
                   uint16_t *out, int64_t ldo, int64_t stride, int *launches)
{
    if (rows <= 0 || cols <= 0) return LATER_OK;
    switch (prec) {
    case GEMM_F16X3:
        hipLaunchKernelGGL(split_exp_kernel<GEMM_F16X3>, dim3(cols), dim3(256), 0, s, rows, X, ldx, col_exp, out, ldo,
                           stride);
        break;
    case GEMM_F16X6:
        hipLaunchKernelGGL(split_exp_kernel<GEMM_F16X6>, dim3(cols), dim3(256), 0, s, rows, X, ldx, col_exp, out, ldo,
                           stride);
        break;
    default: return LATER_ERR_OPTION;
    }
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_split(hipStream_t s, int prec, int rows, int cols, const float *X, int64_t ldx, int a_exp,
               const int *col_exp, uint16_t *out, int64_t ldo, int64_t stride, bool trans, int *launches)
{
    if (rows <= 0 || cols <= 0) return LATER_OK;
    switch (prec) {
    case GEMM_F16: launch_split<GEMM_F16>(s, rows, cols, X, ldx, a_exp, col_exp, out, ldo, stride, trans); break;
    case GEMM_BF16: launch_split<GEMM_BF16>(s, rows, cols, X, ldx, a_exp, col_exp, out, ldo, stride, trans); break;
    case GEMM_F16X3: launch_split<GEMM_F16X3>(s, rows, cols, X, ldx, a_exp, col_exp, out, ldo, stride, trans); break;
    case GEMM_F16X6: launch_split<GEMM_F16X6>(s, rows, cols, X, ldx, a_exp, col_exp, out, ldo, stride, trans); break;
    default: return LATER_ERR_OPTION;
    }
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_split_t_exact(hipStream_t s, int prec, int rows, int cols, const float *X, int64_t ldx, int a_exp,
                       uint16_t *out, int64_t ldo, int64_t stride, int *launches)
{
    if (rows <= 0 || cols <= 0) return LATER_OK;
    switch (prec) {
    case GEMM_F16: launch_split<GEMM_F16>(s, rows, cols, X, ldx, a_exp, nullptr, out, ldo, stride, true, false); break;
    case GEMM_BF16: launch_split<GEMM_BF16>(s, rows, cols, X, ldx, a_exp, nullptr, out, ldo, stride, true, false); break;
    case GEMM_F16X3: launch_split<GEMM_F16X3>(s, rows, cols, X, ldx, a_exp, nullptr, out, ldo, stride, true, false); break;
    case GEMM_F16X6: launch_split<GEMM_F16X6>(s, rows, cols, X, ldx, a_exp, nullptr, out, ldo, stride, true, false); break;
    default: return LATER_ERR_OPTION;
    }
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_split_t_tri(hipStream_t s, int prec, int n, const float *X, int64_t ldx, int a_exp, float shift,
                     uint16_t *out, int64_t ldo, int64_t stride, int *launches)
{
    if (n <= 0) return LATER_OK;
    switch (prec) {
    case GEMM_F16: launch_split<GEMM_F16>(s, n, n, X, ldx, a_exp, nullptr, out, ldo, stride, true, true, true, shift); break;
    case GEMM_BF16: launch_split<GEMM_BF16>(s, n, n, X, ldx, a_exp, nullptr, out, ldo, stride, true, true, true, shift); break;
    case GEMM_F16X3: launch_split<GEMM_F16X3>(s, n, n, X, ldx, a_exp, nullptr, out, ldo, stride, true, true, true, shift); break;
    case GEMM_F16X6: launch_split<GEMM_F16X6>(s, n, n, X, ldx, a_exp, nullptr, out, ldo, stride, true, true, true, shift); break;
    default: return LATER_ERR_OPTION;
    }
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int64_t fp8_ld(int K) { return ((int64_t)K + 63) / 64 * 64 + 64; }
int64_t fp8_ls(int cols) { return ((int64_t)cols + 255) / 256 * 256 + 256; }
int64_t fp8_scale_off(int K, int cols) { return (fp8_ld(K) * cols + 255) / 256 * 256; }
int64_t fp8_bytes(int K, int cols)
{
    return fp8_scale_off(K, cols) + ((int64_t)K + 127) / 128 * fp8_ls(cols);
}

Pieces fp8_pieces(const void *buf, int K, int cols)
{
    Pieces p;
    p.p = (const uint16_t *)buf;
    p.ld = fp8_ld(K);
    p.stride = fp8_scale_off(K, cols);
    p.ls = fp8_ls(cols);
    return p;
}

int gemm_split8(hipStream_t s, int K, int cols, const float *X, int64_t ldx, void *buf, int *launches)
{
    if (K <= 0 || cols <= 0) return LATER_OK;
    if (cols > 65535) return LATER_ERR_SHAPE;
    uint8_t *b = (uint8_t *)buf;
    const int prow = (K + 63) / 64 * 64;
    const dim3 grid((unsigned)ceil_div64(prow, 1024), (unsigned)cols);
    hipLaunchKernelGGL(split8_kernel, grid, dim3(256), 0, s, K, X, ldx, b, fp8_ld(K), b + fp8_scale_off(K, cols),
                       fp8_ls(cols));
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int64_t i8_ld(int K) { return ((int64_t)K + 63) / 64 * 64 + 64; }
int64_t i8_stride(int K, int cols) { return (i8_ld(K) * cols + 255) / 256 * 256; }
int64_t i8_exp_off(int K, int cols) { return 4 * i8_stride(K, cols); }
int64_t i8_bytes(int K, int cols) { return i8_exp_off(K, cols) + (4 * (int64_t)cols + 255) / 256 * 256; }

Pieces i8_pieces(const void *buf, int K, int cols)
{
    Pieces p;
    p.p = (const uint16_t *)buf;
    p.ld = i8_ld(K);
    p.stride = i8_stride(K, cols);
    p.ls = i8_exp_off(K, cols);
    return p;
}

int gemm_split_i8_into(hipStream_t s, int K, int cols, const float *X, int64_t ldx, void *out, int64_t ld,
                       int64_t stride, int *exps, int *launches)
{
    if (K <= 0 || cols <= 0) return LATER_OK;
    if (ld % 64 || ld < (K + 63) / 64 * 64 || stride % 64) return LATER_ERR_SHAPE;
    hipLaunchKernelGGL(split_i8_kernel, dim3((unsigned)cols), dim3(256), 0, s, K, X, ldx, (int8_t *)out, ld, stride,
                       exps);
    HIP_OK(hipGetLastError());
    if (launches) ++*launches;
    return LATER_OK;
}

int gemm_split_i8(hipStream_t s, int K, int cols, const float *X, int64_t ldx, void *buf, int *launches)
{
    int8_t *b = (int8_t *)buf;
    return gemm_split_i8_into(s, K, cols, X, ldx, b, i8_ld(K), i8_stride(K, cols), (int *)(b + i8_exp_off(K, cols)),
                              launches);
}
