// Out-of-core recursive Gram-Schmidt QR: the role of later_oc_qr_rec
// (include/LATER.h:49, QR/later_oc_qr.cu:29-94) for matrices whose columns do
// not fit on the device at once.
//
// A (m x n, column-major, lda) and R (n x n, ldr) live in host memory.
// Memory from hipHostMalloc lets the copies overlap the products; pageable
// memory also works, with copies that block the host.  The device holds at
// most `device_bytes`, allocated once per call and carved anew at every
// step (the recursion runs one step at a time).
//
// The recursion splits the columns as the in-core factorisation does
// (n1 = 32 ceil(n/64)) until a column block, its R and its workspace fit the
// budget; such a block is copied in, factored by later_rgsqrf_ex with the
// caller's options and copied back.  A node too wide to fit is a block
// Gram-Schmidt step out of core: row blocks of Q1 (the left half, already
// factored) and A2 (the right half) stream through the device on two
// streams, so that one row block's copies overlap the other's products:
//   sweep 1: R12 = sum_b Q1_b^T A2_b, each block's product in the node
//            precision, the blocks summed in FP64 per stream and the two
//            stream sums added in a fixed order (bit-reproducible);
//   sweep 2: A2_b -= Q1_b R12, copied back.
// With re-orthogonalisation requested (opts.reorth != NONE) every
// out-of-core node is block CGS2: sweep 2 also forms Q1_b^T A2_b of the
// block it has just updated (still on the device), and a third sweep
// subtracts that second projection in the second-pass precision; R12 is
// the sum of the two.
//
// Node precisions: projections opts.gemm (FP8 runs as FP16 hi/lo here: its
// staged e4m3 operands are not kept out of core), updates as the in-core
// nodes pair them.  The reference's OC_gemm clears its device C tile and
// never copies C in (BLAS/later_gemm.cu, the float overload), so its update
// call (beta = 1) stores -Q1 R12 in place of A2 - Q1 R12; this driver reads A2.
//
// Before anything is copied, a dry run of the recursion checks that every
// leaf and every node fits the budget (LATER_ERR_WORKSPACE otherwise), so a
// rejected call leaves A and R untouched.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"

namespace {

constexpr int64_t kAlign = 256;    // bytes: every carved buffer starts on a 256-byte boundary
constexpr int kMinRows = 256;      // smallest row block of a node sweep

int64_t up(int64_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }
int split_cols(int n) { return 32 * ((n + 63) / 64); }

int upd_prec(const later_opts &o, int p)
{
    if (o.gemm_update >= 0) return o.gemm_update;
    return (p == LATER_GEMM_F16X6 || p == LATER_GEMM_FP8 || p == LATER_GEMM_I8X4) ? LATER_GEMM_F16X3 : p;
}
// The sweeps' block products convert in the kernel: int8 slices (which need
// a staged, column-scaled operand) run as the three-piece FP16 split.
int node_prec(int p)
{
    return p == LATER_GEMM_FP8 ? LATER_GEMM_F16X3 : (p == LATER_GEMM_I8X4 ? LATER_GEMM_F16X6 : p);
}

// FP64 running sum of one stream's block projections.
__global__ void acc_kernel(int n1, int n2, const float *C, double *acc, int first)
{
    const int64_t total = (int64_t)n1 * n2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        acc[i] = first ? (double)C[i] : acc[i] + (double)C[i];
}

// R12 = acc0 + acc1 (acc1 only when the second stream ran a block).
__global__ void fold_kernel(int64_t total, const double *acc0, const double *acc1, int two, float *out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)(two ? acc0[i] + acc1[i] : acc0[i]);
}

// out += x (the second CGS2 projection added to the first).
__global__ void add_kernel(int64_t total, const float *x, float *out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        out[i] += x[i];
}

dim3 grid_of(int64_t total) { return dim3((unsigned)std::min<int64_t>((total + 255) / 256, 4096)); }

struct Oc {
    later_opts o;
    bool cgs2;
    hipStream_t s[2];
    char *arena;
    int64_t bytes;
    later_info acc;  // status, kernels, tree_levels summed over the in-core calls
    bool dry;        // plan only: check the budget, copy and launch nothing
};

// Device bytes of an in-core leaf of n columns (m rows), with or without the
// 16-bit operand arena.
int leaf_bytes(const Oc &c, int m, int n, int64_t *with_pieces, int64_t *plain, int64_t *lw, int64_t *lh)
{
    int rc = later_rgsqrf_workspace(m, n, &c.o, lw, lh);
    if (rc) return rc;
    *plain = up((int64_t)m * n * 4) + up((int64_t)n * n * 4) + up(*lw * 4);
    *with_pieces = *plain + up(*lh * 2);
    return LATER_OK;
}

// Device bytes of a node sweep with row blocks of mb rows.
struct NodeLayout {
    int mb;
    int64_t q, a, cb, wk, acc, r12, r12b, total;
    int64_t lw;  // FP32 workspace floats of the block products
};

NodeLayout node_layout(const Oc &c, int m, int n1, int n2, int mb)
{
    NodeLayout L{};
    L.mb = mb;
    const int p1 = node_prec(c.o.gemm), p2 = node_prec(c.o.gemm_pass2);
    int64_t lw = 0, t = 0;
    for (int p : {p1, c.cgs2 ? p2 : p1}) {
        const int ch = (p == p2 && c.cgs2 && c.o.chunk2 > 0) ? c.o.chunk2 : c.o.chunk;
        if (later_gemm_workspace(mb, n1, n2, p, ch, 0, &t, nullptr) == LATER_OK) lw = std::max(lw, t);
        if (later_gemm_workspace(mb, n1, n2, upd_prec(c.o, p), 0, 1, &t, nullptr) == LATER_OK) lw = std::max(lw, t);
    }
    L.lw = std::max<int64_t>(lw, 1);
    L.q = up((int64_t)mb * n1 * 4);
    L.a = up((int64_t)mb * n2 * 4);
    L.cb = up((int64_t)n1 * n2 * 4);
    L.wk = up(L.lw * 4);
    L.acc = up((int64_t)n1 * n2 * 8);
    L.r12 = up((int64_t)n1 * n2 * 4);
    L.r12b = c.cgs2 ? L.r12 : 0;
    L.total = 2 * (L.q + L.a + L.cb + L.wk + L.acc) + L.r12 + L.r12b;
    (void)m;
    return L;
}

// Largest row block (a multiple of 256, m rounded up at most) whose sweep fits.
int choose_rows(const Oc &c, int m, int n1, int n2, NodeLayout *out)
{
    for (int parts = 1;; parts *= 2) {
        int64_t mb = ((int64_t)m + parts - 1) / parts;
        mb = (mb + kMinRows - 1) / kMinRows * kMinRows;
        if (mb < kMinRows) return LATER_ERR_WORKSPACE;
        NodeLayout L = node_layout(c, m, n1, n2, (int)mb);
        if (L.total <= c.bytes) {
            *out = L;
            return LATER_OK;
        }
        if (mb == kMinRows) return LATER_ERR_WORKSPACE;
    }
}

// Host <-> device copies of a column-major block (rows x cols).
int h2d(hipStream_t s, float *dst, int64_t ldd, const float *src, int64_t lds, int rows, int cols)
{
    HIP_OK(hipMemcpy2DAsync(dst, ldd * 4, src, lds * 4, (size_t)rows * 4, cols, hipMemcpyHostToDevice, s));
    return LATER_OK;
}
int d2h(hipStream_t s, float *dst, int64_t ldd, const float *src, int64_t lds, int rows, int cols)
{
    HIP_OK(hipMemcpy2DAsync(dst, ldd * 4, src, lds * 4, (size_t)rows * 4, cols, hipMemcpyDeviceToHost, s));
    return LATER_OK;
}

int leaf(Oc &c, int m, int n, float *A, int64_t lda, float *R, int64_t ldr, int depth)
{
    int64_t wp = 0, pl = 0, lw = 0, lh = 0;
    int rc = leaf_bytes(c, m, n, &wp, &pl, &lw, &lh);
    if (rc) return rc;
    if (pl > c.bytes) return LATER_ERR_WORKSPACE;
    const bool pieces = wp <= c.bytes && lh > 0;
    if (c.dry) {
        // FP8 projections need the operand arena.
        const bool fp8 = n > c.o.nmin && (c.o.gemm == LATER_GEMM_FP8 || (c.o.reorth & LATER_REORTH_PASS &&
                                                                         c.o.gemm_pass2 == LATER_GEMM_FP8));
        return (fp8 && !pieces) ? LATER_ERR_WORKSPACE : LATER_OK;
    }
    char *p = c.arena;
    float *dA = (float *)p; p += up((int64_t)m * n * 4);
    float *dR = (float *)p; p += up((int64_t)n * n * 4);
    float *work = (float *)p; p += up(lw * 4);
    uint16_t *hwork = pieces ? (uint16_t *)p : nullptr;
    hipStream_t s = c.s[0];
    rc = h2d(s, dA, m, A, lda, m, n);
    if (rc) return rc;
    laterCtxt ctx{(void *)s, nullptr};
    later_info info{};
    rc = later_rgsqrf_ex(ctx, m, n, dA, m, dR, n, work, lw, hwork, pieces ? lh : 0, &c.o, &info);
    if (rc) return rc;
    int st = 0;
    rc = later_read_status(ctx, work, &st);  // synchronises the stream
    if (rc) return rc;
    c.acc.status |= st;
    c.acc.kernels += info.kernels;
    c.acc.tree_levels = std::max(c.acc.tree_levels, depth + info.tree_levels);
    if ((rc = d2h(s, A, lda, dA, m, m, n))) return rc;
    if ((rc = d2h(s, R, ldr, dR, n, n, n))) return rc;
    HIP_OK(hipStreamSynchronize(s));
    return LATER_OK;
}

// One sweep over the row blocks.  kind 0: C = Q1^T A2 (into the stream's
// accumulator); kind 1: A2 -= Q1 R12, copied back, and, with `proj`, the
// updated block's Q1^T A2 accumulated as in kind 0.
struct Bufs {
    float *q[2], *a[2], *cb[2], *wk[2];
    double *acc[2];
    float *r12, *r12b;
};

int sweep(Oc &c, const NodeLayout &L, const Bufs &b, int m, int n1, int n2, float *Q1, float *A2, int64_t lda,
          int kind, const float *rmat, int pp, int pu, int chunk, bool proj, int *blocks_on)
{
    const int nb = (m + L.mb - 1) / L.mb;
    blocks_on[0] = blocks_on[1] = 0;
    for (int i = 0; i < nb; ++i) {
        const int st = i & 1;
        hipStream_t s = c.s[st];
        laterCtxt ctx{(void *)s, nullptr};
        const int r0 = i * L.mb, rows = std::min(L.mb, m - r0);
        int rc = h2d(s, b.q[st], L.mb, Q1 + r0, lda, rows, n1);
        if (!rc) rc = h2d(s, b.a[st], L.mb, A2 + r0, lda, rows, n2);
        if (rc) return rc;
        if (kind == 1) {
            rc = later_gemm_update(ctx, rows, n1, n2, b.q[st], L.mb, rmat, n1, b.a[st], L.mb, pu, b.wk[st], L.lw,
                                   nullptr, 0);
            if (!rc) rc = d2h(s, A2 + r0, lda, b.a[st], L.mb, rows, n2);
            if (rc) return rc;
            c.acc.kernels += 2;
        }
        if (kind == 0 || proj) {
            rc = later_gemm_proj(ctx, rows, n1, n2, b.q[st], L.mb, b.a[st], L.mb, b.cb[st], n1, pp, chunk, b.wk[st],
                                 L.lw, nullptr, 0);
            if (rc) return rc;
            hipLaunchKernelGGL(acc_kernel, grid_of((int64_t)n1 * n2), dim3(256), 0, s, n1, n2, b.cb[st], b.acc[st],
                               blocks_on[st] == 0 ? 1 : 0);
            HIP_OK(hipGetLastError());
            c.acc.kernels += 3;
        }
        ++blocks_on[st];
    }
    return LATER_OK;
}

// R12 (into `out`) from the two streams' accumulators, on stream 0 after
// stream 1's work; stream 1 then waits for it.
int fold(Oc &c, const Bufs &b, int n1, int n2, const int *blocks_on, float *out)
{
    hipEvent_t e1, e0;
    HIP_OK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
    if (hipEventCreateWithFlags(&e0, hipEventDisableTiming) != hipSuccess) {
        hipEventDestroy(e1);
        return LATER_ERR_DEVICE;
    }
    int rc = LATER_OK;
    if (hipEventRecord(e1, c.s[1]) != hipSuccess || hipStreamWaitEvent(c.s[0], e1, 0) != hipSuccess) rc = LATER_ERR_DEVICE;
    if (!rc) {
        const int64_t total = (int64_t)n1 * n2;
        hipLaunchKernelGGL(fold_kernel, grid_of(total), dim3(256), 0, c.s[0], total, b.acc[0], b.acc[1],
                           blocks_on[1] > 0 ? 1 : 0, out);
        ++c.acc.kernels;
        if (hipGetLastError() != hipSuccess || hipEventRecord(e0, c.s[0]) != hipSuccess ||
            hipStreamWaitEvent(c.s[1], e0, 0) != hipSuccess)
            rc = LATER_ERR_DEVICE;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int node(Oc &c, int m, int n1, int n2, float *Q1, float *A2, int64_t lda, float *R12, int64_t ldr)
{
    NodeLayout L;
    int rc = choose_rows(c, m, n1, n2, &L);
    if (rc || c.dry) return rc;
    Bufs b{};
    char *p = c.arena;
    for (int st = 0; st < 2; ++st) {
        b.q[st] = (float *)p; p += L.q;
        b.a[st] = (float *)p; p += L.a;
        b.cb[st] = (float *)p; p += L.cb;
        b.wk[st] = (float *)p; p += L.wk;
        b.acc[st] = (double *)p; p += L.acc;
    }
    b.r12 = (float *)p; p += L.r12;
    b.r12b = c.cgs2 ? (float *)p : nullptr;
    const int p1 = node_prec(c.o.gemm), p2 = node_prec(c.o.gemm_pass2);
    const int ch2 = c.o.chunk2 > 0 ? c.o.chunk2 : c.o.chunk;
    int on[2];
    // Both streams start after everything before this node.
    HIP_OK(hipStreamSynchronize(c.s[0]));
    HIP_OK(hipStreamSynchronize(c.s[1]));
    rc = sweep(c, L, b, m, n1, n2, Q1, A2, lda, 0, nullptr, p1, upd_prec(c.o, p1), c.o.chunk, false, on);
    if (!rc) rc = fold(c, b, n1, n2, on, b.r12);
    if (!rc) rc = sweep(c, L, b, m, n1, n2, Q1, A2, lda, 1, b.r12, p2, upd_prec(c.o, p1), ch2, c.cgs2, on);
    if (!rc && c.cgs2) {
        rc = fold(c, b, n1, n2, on, b.r12b);
        if (!rc) rc = sweep(c, L, b, m, n1, n2, Q1, A2, lda, 1, b.r12b, p2, upd_prec(c.o, p2), ch2, false, on);
        if (!rc) {
            // Stream 0 adds the second projection once stream 1's last update has read it.
            hipEvent_t e;
            HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            if (hipEventRecord(e, c.s[1]) != hipSuccess || hipStreamWaitEvent(c.s[0], e, 0) != hipSuccess)
                rc = LATER_ERR_DEVICE;
            hipEventDestroy(e);
            if (!rc) {
                const int64_t total = (int64_t)n1 * n2;
                hipLaunchKernelGGL(add_kernel, grid_of(total), dim3(256), 0, c.s[0], total, b.r12b, b.r12);
                ++c.acc.kernels;
                if (hipGetLastError() != hipSuccess) rc = LATER_ERR_DEVICE;
            }
        }
    }
    if (!rc) rc = d2h(c.s[0], R12, ldr, b.r12, n1, n1, n2);
    if (hipStreamSynchronize(c.s[0]) != hipSuccess || hipStreamSynchronize(c.s[1]) != hipSuccess)
        rc = rc ? rc : LATER_ERR_DEVICE;
    return rc;
}

int rec(Oc &c, int m, int n, float *A, int64_t lda, float *R, int64_t ldr, int depth)
{
    int64_t wp = 0, pl = 0, lw = 0, lh = 0;
    int rc = leaf_bytes(c, m, n, &wp, &pl, &lw, &lh);
    if (rc) return rc;
    if (pl <= c.bytes) return leaf(c, m, n, A, lda, R, ldr, depth);
    if (n <= 32) return LATER_ERR_WORKSPACE;
    const int n1 = split_cols(n), n2 = n - n1;
    if ((rc = rec(c, m, n1, A, lda, R, ldr, depth + 1))) return rc;
    if ((rc = node(c, m, n1, n2, A, A + (int64_t)n1 * lda, lda, R + (int64_t)n1 * ldr, ldr))) return rc;
    if (!c.dry)
        for (int j = 0; j < n1; ++j) memset(R + n1 + (int64_t)j * ldr, 0, (size_t)n2 * 4);
    return rec(c, m, n2, A + (int64_t)n1 * lda, lda, R + n1 + (int64_t)n1 * ldr, ldr, depth + 1);
}

} // namespace

LATER_API int later_oc_qr_rec(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr,
                              int64_t device_bytes, const later_opts *opts, later_info *info)
{
    if (info) memset(info, 0, sizeof(*info));
    Oc c{};
    if (opts) c.o = *opts; else later_opts_default(&c.o);
    c.cgs2 = c.o.reorth != LATER_REORTH_NONE;
    c.bytes = device_bytes / kAlign * kAlign;
    int rc = LATER_OK;
    if (m < 1 || n < 1 || m < n || lda < m || ldr < n || !A || !R) rc = LATER_ERR_SHAPE;
    else if (c.o.gemm < LATER_GEMM_F32 || c.o.gemm > LATER_GEMM_I8X4 || c.o.gemm_pass2 < LATER_GEMM_F32 ||
             c.o.gemm_pass2 > LATER_GEMM_I8X4 || c.o.gemm_update < -1 || c.o.gemm_update > LATER_GEMM_F16X3 ||
             c.o.panel < 0 || c.o.panel > 3 || c.o.reorth < 0 || c.o.reorth > 3 || c.o.nmin < 0 || c.o.nmin2 < 0 ||
             c.o.chunk < 0 || c.o.chunk % 32 || c.o.chunk2 < 0 || c.o.chunk2 % 32)
        rc = LATER_ERR_OPTION;
    if (!rc) {
        // Dry run: every leaf and node must fit before anything is touched.
        c.dry = true;
        rc = rec(c, m, n, A, lda, R, ldr, 0);
        c.dry = false;
    }
    hipStream_t own = nullptr;
    if (!rc) {
        c.s[0] = ctx_stream(ctxt);
        if (hipStreamCreateWithFlags(&own, hipStreamNonBlocking) != hipSuccess) rc = LATER_ERR_DEVICE;
        c.s[1] = own;
    }
    if (!rc && hipMalloc((void **)&c.arena, (size_t)c.bytes) != hipSuccess) rc = LATER_ERR_DEVICE;
    if (!rc) {
        HIP_OK(hipStreamSynchronize(c.s[0]));
        rc = rec(c, m, n, A, lda, R, ldr, 0);
    }
    if (c.s[0]) hipStreamSynchronize(c.s[0]);
    if (own) {
        hipStreamSynchronize(own);
        hipStreamDestroy(own);
    }
    if (c.arena) hipFree(c.arena);
    if (!rc && c.acc.status) rc = LATER_ERR_BREAKDOWN;
    if (info) {
        info->status = rc;
        info->kernels = c.acc.kernels;
        info->tree_levels = c.acc.tree_levels;
        info->flops_ref = 2.0 * (double)n * n * ((double)m - n / 3.0);
        info->flops_mgs = 2.0 * (double)m * n * n;
    }
    return rc;
}
