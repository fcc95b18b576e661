// Tall-skinny panel: TSQR tree of block modified Gram-Schmidt factorisations.
//
// Follows the structure of mgs_caqr_panel_256x32 (QR/panel.cu:65-134): every
// 256-row block is factored by MGS, the blocks' R factors are stacked and the
// stack is factored recursively, then each block's Q is multiplied by its
// rows of the stack's Q.  The block kernel is laid out for 64-lane
// wavefronts instead of the reference's 32x32 thread block of 32-lane warps
// (QR/panel.cu:246-325): 8 waves, wave v owns columns v, v+8, v+16, v+24 and
// lane l holds rows l, l+64, ... of them in registers, so the column norm and
// the projections are single-wave reductions and one barrier per column
// suffices (the q vector is double-buffered in LDS).
#include "common.h"
#include "panel.h"
#include "prof.h"

namespace {

constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr int kCols = 32;                 // max panel width
constexpr int kColsPerWave = kCols / kWaves;
constexpr int kRowsPerLane = 5;           // up to 320 rows per block
constexpr int kMaxRows = 64 * kRowsPerLane;

template <typename Acc>
__device__ __forceinline__ void mgs_block_body(int m, int w, float *A, int64_t lda, float *Rout, int64_t ldr,
                                               int bh, int nblk, int *status, const int b)
{
    __shared__ float qs[2][kMaxRows];
    __shared__ float Rs[kCols][kCols + 1];

    const int r0 = b * bh;
    // The last block takes every remaining row (a tail shorter than w is merged).
    const int mm = (b == nblk - 1) ? (m - r0) : bh;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    auto *Ab = gmem(A) + r0;  // global view (the grouped kernel reads A from a table)

    // Loads from clamped in-bounds addresses, selected afterwards, so that all
    // of them issue before the first wait.
    float a[kColsPerWave][kRowsPerLane];
#pragma unroll
    for (int c = 0; c < kColsPerWave; ++c) {
        const int j = wave + kWaves * c;
        const int jc = min(j, w - 1);
#pragma unroll
        for (int t = 0; t < kRowsPerLane; ++t) {
            const int row = lane + 64 * t;
            a[c][t] = Ab[min(row, mm - 1) + (int64_t)jc * lda];
        }
    }
#pragma unroll
    for (int c = 0; c < kColsPerWave; ++c)
#pragma unroll
        for (int t = 0; t < kRowsPerLane; ++t)
            if (!(wave + kWaves * c < w && lane + 64 * t < mm)) a[c][t] = 0.0f;
    for (int i = threadIdx.x; i < kCols * (kCols + 1); i += kThreads) (&Rs[0][0])[i] = 0.0f;
    __syncthreads();

    for (int k = 0; k < w; ++k) {
        const int buf = k & 1;
        if (wave == (k % kWaves)) {
            const int slot = k / kWaves;
            Acc s = 0;
#pragma unroll
            for (int t = 0; t < kRowsPerLane; ++t) s += (Acc)a[slot][t] * (Acc)a[slot][t];
            s = wave_sum(s);
            Acc nrm = sqrt(s);
            bool ok = (nrm > (Acc)0) && (nrm < (Acc)INFINITY);
            if (!ok && lane == 0 && status) atomicOr(status, 1);
            Acc inv = ok ? (Acc)1 / nrm : (Acc)0;
#pragma unroll
            for (int t = 0; t < kRowsPerLane; ++t) {
                a[slot][t] = (float)((Acc)a[slot][t] * inv);
                const int row = lane + 64 * t;
                if (row < kMaxRows) qs[buf][row] = a[slot][t];
            }
            if (lane == 0) Rs[k][k] = ok ? (float)nrm : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kColsPerWave; ++c) {
            const int j = wave + kWaves * c;
            if (j > k && j < w) {
                Acc d = 0;
#pragma unroll
                for (int t = 0; t < kRowsPerLane; ++t) d += (Acc)qs[buf][lane + 64 * t] * (Acc)a[c][t];
                d = wave_sum(d);
#pragma unroll
                for (int t = 0; t < kRowsPerLane; ++t)
                    a[c][t] = (float)((Acc)a[c][t] - (Acc)qs[buf][lane + 64 * t] * d);
                if (lane == 0) Rs[k][j] = (float)d;
            }
        }
    }
    __syncthreads();

#pragma unroll
    for (int c = 0; c < kColsPerWave; ++c) {
        const int j = wave + kWaves * c;
        if (j < w) {
#pragma unroll
            for (int t = 0; t < kRowsPerLane; ++t) {
                const int row = lane + 64 * t;
                if (row < mm) Ab[row + (int64_t)j * lda] = a[c][t];
            }
        }
    }
    // R block (w x w, zeros below the diagonal) at rows b*w of Rout.
    auto *Rb = gmem(Rout) + (int64_t)b * w;
    for (int i = threadIdx.x; i < w * w; i += kThreads) {
        const int x = i % w, y = i / w;
        Rb[x + (int64_t)y * ldr] = (x <= y) ? Rs[x][y] : 0.0f;
    }
}

template <typename Acc>
__global__ __launch_bounds__(kThreads) void mgs_block_kernel(int m, int w, float *A, int lda, float *Rout,
                                                             int ldr, int bh, int nblk, int *status)
{
    mgs_block_body<Acc>(m, w, A, lda, Rout, ldr, bh, nblk, status, blockIdx.x);
}

template <typename Acc>
__global__ __launch_bounds__(kThreads) void mgs_block_grouped_kernel(const PanelArgs *args, const int2 *tasks)
{
    const int2 t = tasks[blockIdx.x];
    const PanelArgs a = args[t.x];
    mgs_block_body<Acc>(a.m, a.w, a.A, a.lda, a.W, a.ldw, kPanelBlockRows, a.nblk, a.status, t.y);
}

// A_b <- A_b * W_b for every row block b, W_b = rows b*w .. b*w+w-1 of the
// stacked factor (ld ldw).  Each thread owns whole rows, so the in-place
// update never reads a value another thread writes (the reference relies on
// an aliased batched Sgemm here, QR/panel.cu:97-102).
template <typename Acc>
__device__ __forceinline__ void tsqr_apply_body(int m, int w, float *A, int64_t lda, const float *W, int64_t ldw,
                                                int bh, int nblk, const int b)
{
    __shared__ float Ws[kCols][kCols];
    const int r0 = b * bh;
    const int mm = (b == nblk - 1) ? (m - r0) : bh;
    const auto *Wb = gmem(W) + (int64_t)b * w;
    for (int i = threadIdx.x; i < w * w; i += 256) {
        const int x = i % w, y = i / w;
        Ws[x][y] = Wb[x + (int64_t)y * ldw];
    }
    __syncthreads();
    auto *Ab = gmem(A) + r0;
    for (int row = threadIdx.x; row < mm; row += 256) {
        // Unconditional loads (clamped column) so all of them are in flight at once.
        float v[kCols];
#pragma unroll
        for (int k = 0; k < kCols; ++k) v[k] = Ab[row + (int64_t)min(k, w - 1) * lda];
#pragma unroll
        for (int k = 0; k < kCols; ++k)
            if (k >= w) v[k] = 0.0f;
#pragma unroll 4
        for (int j = 0; j < w; ++j) {
            Acc s = 0;
#pragma unroll
            for (int k = 0; k < kCols; ++k)
                if (k < w) s += (Acc)v[k] * (Acc)Ws[k][j];
            Ab[row + (int64_t)j * lda] = (float)s;
        }
    }
}

template <typename Acc>
__global__ __launch_bounds__(256) void tsqr_apply_grouped_kernel(const PanelArgs *args, const int2 *tasks)
{
    const int2 t = tasks[blockIdx.x];
    const PanelArgs a = args[t.x];
    tsqr_apply_body<Acc>(a.m, a.w, a.A, a.lda, a.W, a.ldw, kPanelBlockRows, a.nblk, t.y);
}

// The same product over 64-row tiles of the panel (four waves per tile: the
// tile is staged in LDS, wave q forms output columns q, q+4, ...), so that a
// 16384-row panel spreads over 256 workgroups instead of 64.  Every output
// keeps the row-per-thread form's order of summation over k, so the two give
// the same bits.  A tile never straddles two 256-row blocks; the tiles of a
// merged tail belong to the last block.
constexpr int kApplyRows = 64;

template <typename Acc>
__global__ __launch_bounds__(256) void tsqr_apply_tile_kernel(int m, int w, float *A, int lda, const float *W,
                                                              int ldw, int bh, int nblk)
{
    __shared__ float Ws[kCols][kCols];
    __shared__ float Vs[kCols][kApplyRows];
    const int row0 = blockIdx.x * kApplyRows;
    const int b = min(row0 / bh, nblk - 1);
    const int rows = min(kApplyRows, m - row0);
    const auto *Wb = gmem(W) + (int64_t)b * w;
    for (int i = threadIdx.x; i < w * w; i += 256) {
        const int x = i % w, y = i / w;
        Ws[x][y] = Wb[x + (int64_t)y * ldw];
    }
    const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
    float *At = A + row0;
    for (int c = q; c < w; c += 4) Vs[c][r] = (r < rows) ? At[r + (int64_t)c * lda] : 0.0f;
    __syncthreads();
    float v[kCols];
#pragma unroll
    for (int k = 0; k < kCols; ++k) v[k] = (k < w) ? Vs[k][r] : 0.0f;
    for (int j = q; j < w; j += 4) {
        Acc s = 0;
#pragma unroll
        for (int k = 0; k < kCols; ++k)
            if (k < w) s += (Acc)v[k] * (Acc)Ws[k][j];
        if (r < rows) At[r + (int64_t)j * lda] = (float)s;
    }
}

struct Blocks {
    int nblk;
};

Blocks plan_blocks(int m, int w)
{
    const int bh = kPanelBlockRows;
    int nb = m / bh, r = m % bh;
    if (nb == 0) return {1};
    return {nb + ((r >= w) ? 1 : 0)};
}

} // namespace

int panel_blocks(int m, int w) { return plan_blocks(m, w).nblk; }

int mgs_grouped(hipStream_t s, bool fp64, const PanelArgs *args, const int2 *tasks, int ntasks)
{
    if (ntasks <= 0) return LATER_OK;
    if (fp64) hipLaunchKernelGGL(mgs_block_grouped_kernel<double>, dim3(ntasks), dim3(kThreads), 0, s, args, tasks);
    else hipLaunchKernelGGL(mgs_block_grouped_kernel<float>, dim3(ntasks), dim3(kThreads), 0, s, args, tasks);
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

int apply_grouped(hipStream_t s, bool fp64, const PanelArgs *args, const int2 *tasks, int ntasks)
{
    if (ntasks <= 0) return LATER_OK;
    if (fp64) hipLaunchKernelGGL(tsqr_apply_grouped_kernel<double>, dim3(ntasks), dim3(256), 0, s, args, tasks);
    else hipLaunchKernelGGL(tsqr_apply_grouped_kernel<float>, dim3(ntasks), dim3(256), 0, s, args, tasks);
    HIP_OK(hipGetLastError());
    return LATER_OK;
}

int64_t panel_work_floats_any(int m)
{
    // A ragged leaf can be narrower than 32 columns; its tree can need a
    // stack where the 32-wide tree has none (a short tail is merged only when
    // it has fewer rows than the panel has columns).
    int64_t mx = 0;
    for (int w = 1; w <= kCols; ++w) {
        const int64_t f = panel_work_floats(m, w);
        if (f > mx) mx = f;
    }
    return mx;
}

int64_t panel_work_floats(int m, int w)
{
    if (w < 1) w = 1;
    int64_t tot = 0;
    int64_t mm = m;
    for (;;) {
        Blocks p = plan_blocks((int)mm, w);
        if (p.nblk <= 1) break;
        tot += (int64_t)p.nblk * w * w;
        mm = (int64_t)p.nblk * w;
    }
    return tot;
}

int panel_tsqr(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work,
               int64_t lwork, bool fp64, int *status, int *launches)
{
    if (w < 1 || w > kCols || m < w) return LATER_ERR_SHAPE;
    Blocks p = plan_blocks(m, w);
    const int bh = kPanelBlockRows;
    if (p.nblk <= 1) {
        if (m > kMaxRows) return LATER_ERR_SHAPE;
        if (fp64)
            hipLaunchKernelGGL(mgs_block_kernel<double>, dim3(1), dim3(kThreads), 0, s, m, w, A, lda, R, ldr, bh,
                               1, status);
        else
            hipLaunchKernelGGL(mgs_block_kernel<float>, dim3(1), dim3(kThreads), 0, s, m, w, A, lda, R, ldr, bh,
                               1, status);
        if (launches) ++*launches;
        HIP_OK(hipGetLastError());
        return LATER_OK;
    }
    const int ldst = p.nblk * w;
    const int64_t need = (int64_t)ldst * w;
    if (lwork < need) return LATER_ERR_WORKSPACE;
    float *stack = work;
    const int tok = prof_start(s);
    if (fp64)
        hipLaunchKernelGGL(mgs_block_kernel<double>, dim3(p.nblk), dim3(kThreads), 0, s, m, w, A, lda, stack,
                           ldst, bh, p.nblk, status);
    else
        hipLaunchKernelGGL(mgs_block_kernel<float>, dim3(p.nblk), dim3(kThreads), 0, s, m, w, A, lda, stack,
                           ldst, bh, p.nblk, status);
    prof_stop(tok, s, PROF_PANEL, 4.0 * m * w * w, 8.0 * m * w + 4.0 * ldst * w);
    if (launches) ++*launches;
    HIP_OK(hipGetLastError());
    int rc = panel_tsqr(s, ldst, w, stack, ldst, R, ldr, work + need, lwork - need, fp64, status, launches);
    if (rc) return rc;
    const int tiles = (m + kApplyRows - 1) / kApplyRows;
    if (fp64)
        hipLaunchKernelGGL(tsqr_apply_tile_kernel<double>, dim3(tiles), dim3(256), 0, s, m, w, A, lda, stack, ldst,
                           bh, p.nblk);
    else
        hipLaunchKernelGGL(tsqr_apply_tile_kernel<float>, dim3(tiles), dim3(256), 0, s, m, w, A, lda, stack, ldst,
                           bh, p.nblk);
    if (launches) ++*launches;
    HIP_OK(hipGetLastError());
    return LATER_OK;
}
