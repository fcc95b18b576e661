// Recursive Gram-Schmidt QR: host recursion and the C-ABI of include/later.h.
//
// The recursion is the reference's (QR/later_rgsqrf.cu:25-60):
//     [A1 | A2] = [Q1 | Q2] [R11 R12; 0 R22]
//     Q1 R11 = qr(A1);  R12 = Q1^T A2;  A2 -= Q1 R12;  Q2 R22 = qr(A2)
// split at n1 = 32 ceil(n/64) (n/2 on the reference domain n = 128*2^k), with
// nodes of width <= nmin computed with FP32 products as in the reference's
// 128-column panel (QR/panel.cu:10-63) and leaves of width <= 32 factored by
// the TSQR panel (QR/panel.cu:65-134).  Everything is enqueued on the caller's
// stream; nothing synchronises or allocates.
//
// Accuracy recovery (absent from the reference):
//   LATER_REORTH_NODE: a second projection S = Q1^T A2, A2 -= Q1 S, R12 += S
//                      at every node (block CGS2);
//   LATER_REORTH_PASS: the whole factorisation is repeated on Q (whose
//                      condition number is ~1 after the first pass), and
//                      R := R2 R1 = R1 + (R2 - I) R1.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "gemm.h"
#include "panel.h"

namespace {

constexpr int kHeaderFloats = 64;  // [0] device status word
constexpr int kQExp = 14;          // |Q| <= 1: Q scaled by 2^14 before 16-bit casts

int split_n(int n) { return 32 * ((n + 63) / 64); }

bool prec_ok(int p) { return p >= LATER_GEMM_F32 && p <= LATER_GEMM_I8X4; }
bool chol_panel(int p) { return p == LATER_PANEL_CHOLQR2 || p == LATER_PANEL_CHOLQR1; }
bool scaled(int p) { return p == LATER_GEMM_F16X3 || p == LATER_GEMM_F16X6; }
// F16X6, FP8 and I8X4 are projection precisions; their updates use the hi/lo pair.
int update_prec(int p)
{
    return (p == LATER_GEMM_F16X6 || p == LATER_GEMM_FP8 || p == LATER_GEMM_I8X4) ? LATER_GEMM_F16X3 : p;
}
// Byte-staged projection operands (each node splits its Q1 and A2 itself).
bool staged8(int p) { return p == LATER_GEMM_FP8 || p == LATER_GEMM_I8X4; }
int64_t staged_bytes(int p, int m, int n) { return p == LATER_GEMM_FP8 ? fp8_bytes(m, n) : i8_bytes(m, n); }
// The projection precision a node of height m runs: int8 slices need their
// staging arena (and no block CGS2 at the nodes, whose re-projection
// converts in the kernel) and at most kI8MaxK rows (exact int32 sums);
// otherwise the three-piece FP16 split, the next most accurate form.
int proj_eff(int p, int m, bool staged)
{
    return (p == LATER_GEMM_I8X4 && (!staged || m > kI8MaxK)) ? LATER_GEMM_F16X6 : p;
}
// The update precision of the nodes whose projections run `p`.
int update_of(const later_opts &o, int p) { return o.gemm_update >= 0 ? o.gemm_update : update_prec(p); }
bool update_ok(int p) { return p == -1 || (p >= LATER_GEMM_F32 && p <= LATER_GEMM_F16X3); }
// FP32-product widths of a pass: the first pass runs nmin for both products;
// the second runs nmin2 (0: nmin) for its projections, which set the final
// orthogonality, and the wider of the two for its updates (FP32 products are
// more accurate than the hi/lo pair and faster on these memory-bound shapes).
int pass_nmin_p(const later_opts &o, bool second) { return (second && o.nmin2 > 0) ? o.nmin2 : o.nmin; }
int pass_nmin_u(const later_opts &o, bool second) { return second ? std::max(o.nmin, pass_nmin_p(o, true)) : o.nmin; }

struct Layout {
    int64_t header, exps, panel, gemm, snode, r2, tcopy, total;
    int64_t hpieces;  // 16-bit elements of `hwork` for the pre-split operand pieces
    int64_t hrow;     // offset of Q's transposed pieces in hwork
    int64_t hnode;    // offset of the per-node operand pieces in hwork
};

// Nodes at least this wide stage their 16-bit GEMM operands as pre-split
// pieces (below it the four extra launches cost more than they save).
constexpr int kPieceMinN = 512;

// Leading dimension of a piece array: a multiple of 8 (16-byte rows), kept
// off powers of two so that the rows of one tile spread over the memory
// channels.
int64_t piece_ld(int64_t x) { return (x + 7) / 8 * 8 + 64; }

// Pieces of the projection operands.
int proj_pieces(int prec) { return gemm_pieces(prec); }

// Regions of hwork are multiples of 128 elements (256 bytes).
int64_t up128(int64_t x) { return (x + 127) / 128 * 128; }

// Does a pass whose nodes run projections `p` and updates `pu` stage its
// operands as 16-bit pieces (at nodes of kPieceMinN columns and more)?
bool pass_pieces(int n, int nmin_p, int p, int pu)
{
    return n >= kPieceMinN && n > nmin_p && p != LATER_GEMM_F32 && gemm_pieces(pu) > 0;
}

// 16-bit elements the per-node staged operands of one node need: the e4m3
// bytes of Q1 and A2 for FP8 projections, the int8 slices of A2 for I8X4
// (whose Q1 slices persist like Q's pieces), at any width above nmin; for nodes
// of kPieceMinN columns and more, A2 as 16-bit pieces for the projection and
// R12 for the update (Q's pieces persist across the nodes of a pass).
int64_t node_scratch(int m, int n, int p, int pu)
{
    const int n1 = split_n(n), n2 = n - n1;
    const int64_t ldm = piece_ld(m), ldn1 = piece_ld(n1);
    int64_t tot = 0;
    if (p == LATER_GEMM_FP8) tot += (fp8_bytes(m, n1) + fp8_bytes(m, n2)) / 2;
    if (p == LATER_GEMM_I8X4) tot += i8_bytes(m, n2) / 2;
    if (n >= kPieceMinN && gemm_pieces(pu) > 0) {
        if (!staged8(p)) tot += proj_pieces(p) * ldm * n2;
        tot += gemm_pieces(pu) * ldn1 * n2;
    }
    return tot;
}

// Largest split-K partial buffer over the nodes of the recursion.
void walk_gemm(int m, int n, int nmin, int chunk_big, int64_t *mx)
{
    if (n <= 32) return;
    const int n1 = split_n(n), n2 = n - n1;
    const int chunk = (n <= nmin) ? 0 : chunk_big;
    int64_t w = gemm_proj_work_floats(m, n1, n2, chunk);
    if (w > *mx) *mx = w;
    walk_gemm(m, n1, nmin, chunk_big, mx);
    walk_gemm(m, n2, nmin, chunk_big, mx);
}

Layout layout(int m, int n, const later_opts &o)
{
    Layout L;
    L.header = kHeaderFloats;
    L.exps = ((int64_t)n + 63) / 64 * 64;
    L.panel = chol_panel(o.panel) ? panel_chol_work_floats(n) : panel_work_floats_any(m);
    int64_t g = 0;
    walk_gemm(m, n, pass_nmin_p(o, false), o.chunk, &g);
    if (o.reorth & LATER_REORTH_PASS) walk_gemm(m, n, pass_nmin_p(o, true), o.chunk2 > 0 ? o.chunk2 : o.chunk, &g);
    L.gemm = g;
    const int n1 = split_n(n);
    L.snode = ((o.reorth & LATER_REORTH_NODE) && n > 32) ? (int64_t)n1 * (n - n1) : 0;
    L.r2 = (o.reorth & LATER_REORTH_PASS) ? (int64_t)n * n : 0;
    L.tcopy = L.r2;
    L.total = L.header + L.exps + L.panel + L.gemm + L.snode + L.r2 + L.tcopy;
    // hwork: Q's persistent pieces (column-major for the projections,
    // transposed for the updates; written once per column and pass), then
    // the largest node's own staged operands (the top node's).
    int64_t hcol = 0, hrow = 0, hnode = 0;
    for (int pass = 0; pass < ((o.reorth & LATER_REORTH_PASS) ? 2 : 1); ++pass) {
        const int p = proj_eff(pass ? o.gemm_pass2 : o.gemm, m, !(o.reorth & LATER_REORTH_NODE)),
                  pu = update_of(o, p);
        const int nmin_p = pass_nmin_p(o, pass == 1);
        if (p == LATER_GEMM_F32 || n <= nmin_p) continue;
        hnode = std::max(hnode, node_scratch(m, n, p, pu));
        if (p == LATER_GEMM_I8X4) hcol = std::max(hcol, i8_bytes(m, n) / 2);  // Q's slices, any width
        if (pass_pieces(n, nmin_p, p, pu)) {
            if (!staged8(p)) hcol = std::max(hcol, proj_pieces(p) * piece_ld(m) * n);
            hrow = std::max(hrow, gemm_pieces(pu) * piece_ld(n) * m);
        }
    }
    L.hrow = up128(hcol);
    L.hnode = L.hrow + up128(hrow);
    int64_t ph = hnode > 0 ? L.hnode + hnode : 0;
    // The final R = R1 + (R2 - I) R1 stages R2 (transposed) and R1.
    if ((o.reorth & LATER_REORTH_PASS) && o.gemm_pass2 != LATER_GEMM_F32 && n >= kPieceMinN) {
        const int64_t pt = 2 * gemm_pieces(update_of(o, o.gemm_pass2)) * piece_ld(n) * n;
        if (pt > ph) ph = pt;
    }
    L.hpieces = ph;
    return L;
}

struct Run {
    hipStream_t s;
    later_opts o;
    int *status;
    int *exps;
    float *panel; int64_t panel_n;
    float *gemm; int64_t gemm_n;
    float *snode;
    uint16_t *pcs;     // pre-split operand pieces (nullptr: fused conversion)
    uint16_t *qcol;    // Q's pieces, column-major (ld ldqc, piece stride qcol_ps)
    uint16_t *qrow;    // Q's pieces, transposed (ld ldqr, piece stride qrow_ps)
    uint16_t *npcs;    // the current node's own staged operands
    int64_t ldqc, ldqr, qcol_ps, qrow_ps;
    int64_t q8ld, q8ps, q8exp;  // int8 passes: Q's slices at qcol (bytes: ld, slice stride, exponents)
    std::vector<unsigned char> col_done, row_done;  // columns of Q split in this pass
    int launches;
    int depth;
    bool node_reorth;  // block CGS2 at the nodes of this pass
    int chunk;         // FP64 fold chunk of this pass
    bool chol_once;    // one Cholesky QR pass per leaf (the input is a Q)
    int nmin_p;        // projections of nodes this wide or narrower run FP32 products
    int nmin_u;        // the same for the updates
};


// Column-wise helpers on a (ceil(rows / 256), min(cols, 65535)) grid: thread
// row i of columns j = blockIdx.y, blockIdx.y + gridDim.y, ...
dim3 grid_cols(int rows, int cols) { return dim3((unsigned)ceil_div64(rows, 256), (unsigned)(cols < 65535 ? cols : 65535)); }

__global__ void zero_lower_kernel(int n, float *R, int64_t ldr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = blockIdx.y; j < n && j < i; j += gridDim.y) R[i + (int64_t)j * ldr] = 0.0f;
}

__global__ void fill_kernel(int64_t n, float *x, float v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void status_clear_kernel(int *st) { st[threadIdx.x] = 0; }

// C (m x n, ldc) += S (m x n, lds)
__global__ void add_kernel(int m, int n, const float *S, int64_t lds, float *C, int64_t ldc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    C[i + (int64_t)j * ldc] += S[i + (int64_t)j * lds];
}

// Y (n x n, ld n) = X (upper triangle, ldx); diag(Y) -= shift.
__global__ void copy_upper_kernel(int n, const float *X, int64_t ldx, float *Y, float shift)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        float v = (i <= j) ? X[i + (int64_t)j * ldx] : 0.0f;
        if (i == j) v -= shift;
        Y[i + (int64_t)j * n] = v;
    }
}

dim3 grid1(int64_t n) { return dim3((unsigned)ceil_div64(n, 256)); }

// Q's columns [c0, c0 + cols) (the node's Q1, FP32 at A) in the persistent
// pieces of this pass: the column ranges not split yet are split now.  A
// column of Q is final once its leaf has run, so each is split at most once
// per pass, by the first node whose Q1 holds it (the pieces are the bits the
// per-node split would give: Q is scaled by the fixed 2^aexp).
template <bool ROW>
int ensure_q(Run &r, int prec, int m, int c0, int cols, const float *A, int64_t lda, int aexp)
{
    std::vector<unsigned char> &done = ROW ? r.row_done : r.col_done;
    int rc = LATER_OK;
    for (int a = c0; a < c0 + cols && !rc;) {
        if (done[a]) { ++a; continue; }
        int b = a;
        while (b < c0 + cols && !done[b]) done[b++] = 1;
        const float *X = A + (int64_t)(a - c0) * lda;
        if (!ROW && prec == LATER_GEMM_I8X4)
            rc = gemm_split_i8_into(r.s, m, b - a, X, lda, (char *)r.qcol + a * r.q8ld, r.q8ld, r.q8ps,
                                    (int *)((char *)r.qcol + r.q8exp) + a, &r.launches);
        else if (ROW)
            rc = gemm_split_t_exact(r.s, prec, m, b - a, X, lda, aexp, r.qrow + a, r.ldqr, r.qrow_ps, &r.launches);
        else
            rc = gemm_split(r.s, prec, m, b - a, X, lda, aexp, nullptr, r.qcol + (int64_t)a * r.ldqc, r.ldqc,
                            r.qcol_ps, false, &r.launches);
        a = b;
    }
    return rc;
}

int qr_rec(Run &r, int m, int n, float *A, int64_t lda, float *R, int64_t ldr, int prec, int depth, int c0)
{
    if (depth > r.depth) r.depth = depth;
    // Cholesky QR leaves take up to 64 columns (one Gram and factor for what
    // would otherwise be two leaves and a node); the MGS panels take 32.
    if (n <= (chol_panel(r.o.panel) ? 64 : 32)) {
        if (chol_panel(r.o.panel))
            return panel_cholqr2(r.s, m, n, A, (int)lda, R, (int)ldr, r.panel, r.panel_n, r.status, &r.launches,
                                 r.chol_once || r.o.panel == LATER_PANEL_CHOLQR1);
        return panel_tsqr(r.s, m, n, A, (int)lda, R, (int)ldr, r.panel, r.panel_n,
                          r.o.panel == LATER_PANEL_MGS64, r.status, &r.launches);
    }
    const int n1 = split_n(n), n2 = n - n1;
    const bool inner = (n <= r.nmin_p), inner_u = (n <= r.nmin_u);
    // projection precision
    const int p = inner ? LATER_GEMM_F32 : proj_eff(prec, m, r.pcs && !(r.o.reorth & LATER_REORTH_NODE));
    const int pu = inner_u ? LATER_GEMM_F32 : update_of(r.o, prec);  // update precision
    const int chunk = inner ? 0 : r.chunk;
    const bool f64 = r.chunk > 0 && !inner_u;  // split updates subtract in FP64 when the pass folds
    int rc = qr_rec(r, m, n1, A, lda, R, ldr, prec, depth + 1, c0);
    if (rc) return rc;
    float *A2 = A + (int64_t)n1 * lda;
    float *R12 = R + (int64_t)n1 * ldr;
    const int aexp = scaled(p) ? kQExp : 0, aexp_u = scaled(pu) ? kQExp : 0;
    const int *bexp = scaled(p) ? r.exps : nullptr, *bexp_u = scaled(pu) ? r.exps : nullptr;
    // Pre-split operands (the reference's s2h staging, util/util.cu:24-32,
    // as FP16 pieces): the same MFMA inputs as the fused conversion, made
    // once per node instead of once per tile.  FP8 projections are always
    // staged (their block scales need a pass over each 128-row block).
    const bool fp8 = staged8(p);  // FP8 or int8: staged per node, always
    const bool pre = r.pcs && !inner_u && n >= kPieceMinN && gemm_pieces(pu) > 0;
    const bool pre_proj = fp8 ? (r.pcs && !inner) : (pre && proj_pieces(p) > 0);
    if (fp8 && !pre_proj) return LATER_ERR_WORKSPACE;
    Pieces q1h{}, a2h{}, q1t{}, r12h{};
    {
        uint16_t *b = r.npcs;
        if (p == LATER_GEMM_FP8) {
            q1h = fp8_pieces(b, m, n1);
            b += fp8_bytes(m, n1) / 2;
            a2h = fp8_pieces(b, m, n2);
            b += fp8_bytes(m, n2) / 2;
        } else if (p == LATER_GEMM_I8X4) {
            // Q1: the pass's persistent slices from column c0 (its exponents
            // at a fixed offset from the array base); A2: this node's own.
            q1h.p = (const uint16_t *)((const char *)r.qcol + c0 * r.q8ld);
            q1h.ld = r.q8ld;
            q1h.stride = r.q8ps;
            q1h.ls = r.q8exp + 4 * (int64_t)c0 - c0 * r.q8ld;
            a2h = i8_pieces(b, m, n2);
            b += i8_bytes(m, n2) / 2;
        } else if (pre_proj) {
            const int64_t ldm = piece_ld(m);
            q1h = {r.qcol + (int64_t)c0 * r.ldqc, r.ldqc, r.qcol_ps};
            a2h = {b, ldm, ldm * n2}; b += proj_pieces(p) * ldm * n2;
        }
        if (pre) {
            const int64_t ldn1 = piece_ld(n1);
            q1t = {r.qrow + c0, r.ldqr, r.qrow_ps};
            r12h = {b, ldn1, ldn1 * n2};
        }
    }
    // R12 = Q1^T A2   (QR/later_rgsqrf.cu:43-49); A2's column exponents and
    // pieces in one launch when it is staged.
    if (p == LATER_GEMM_FP8) {
        if ((rc = gemm_split8(r.s, m, n1, A, lda, (void *)q1h.p, &r.launches))) return rc;
        if ((rc = gemm_split8(r.s, m, n2, A2, lda, (void *)a2h.p, &r.launches))) return rc;
    } else if (p == LATER_GEMM_I8X4) {
        if ((rc = ensure_q<false>(r, p, m, c0, n1, A, lda, 0))) return rc;
        if ((rc = gemm_split_i8(r.s, m, n2, A2, lda, (void *)a2h.p, &r.launches))) return rc;
    } else {
        if (pre_proj && (rc = ensure_q<false>(r, p, m, c0, n1, A, lda, aexp))) return rc;
        if (pre_proj && scaled(p)) {
            if ((rc = gemm_split_exp(r.s, p, m, n2, A2, lda, r.exps, (uint16_t *)a2h.p, a2h.ld, a2h.stride,
                                     &r.launches)))
                return rc;
        } else {
            if (scaled(p) && (rc = gemm_colexp(r.s, m, n2, A2, lda, r.exps, &r.launches))) return rc;
            if (pre_proj && (rc = gemm_split(r.s, p, m, n2, A2, lda, aexp, bexp, (uint16_t *)a2h.p, a2h.ld,
                                             a2h.stride, false, &r.launches)))
                return rc;
        }
    }
    rc = gemm_proj(r.s, p, m, n1, n2, A, lda, A2, lda, R12, ldr, chunk, bexp, aexp, r.gemm, r.gemm_n,
                   &r.launches, pre_proj ? &q1h : nullptr, pre_proj ? &a2h : nullptr);
    if (rc) return rc;
    // A2 -= Q1 R12    (QR/later_rgsqrf.cu:51-56); R12's column exponents
    // replace A2's.
    if (pre) {
        if ((rc = ensure_q<true>(r, pu, m, c0, n1, A, lda, aexp_u))) return rc;
        if (scaled(pu))
            rc = gemm_split_exp(r.s, pu, n1, n2, R12, ldr, r.exps, (uint16_t *)r12h.p, r12h.ld, r12h.stride,
                                &r.launches);
        else
            rc = gemm_split(r.s, pu, n1, n2, R12, ldr, 0, nullptr, (uint16_t *)r12h.p, r12h.ld, r12h.stride, false,
                            &r.launches);
        if (rc) return rc;
    } else if (scaled(pu)) {
        if ((rc = gemm_colexp(r.s, n1, n2, R12, ldr, r.exps, &r.launches))) return rc;
    }
    rc = gemm_update(r.s, pu, m, n1, n2, A, lda, R12, ldr, A2, lda, f64, bexp_u, aexp_u, &r.launches,
                     pre ? &q1t : nullptr, pre ? &r12h : nullptr);
    if (rc) return rc;
    if (r.node_reorth) {
        // S = Q1^T A2; A2 -= Q1 S; R12 += S
        float *S = r.snode;
        if (scaled(p)) {
            if ((rc = gemm_colexp(r.s, m, n2, A2, lda, r.exps, &r.launches))) return rc;
        }
        rc = gemm_proj(r.s, p, m, n1, n2, A, lda, A2, lda, S, n1, chunk, bexp, aexp, r.gemm, r.gemm_n,
                       &r.launches);
        if (rc) return rc;
        if (scaled(pu)) {
            if ((rc = gemm_colexp(r.s, n1, n2, S, n1, r.exps, &r.launches))) return rc;
        }
        rc = gemm_update(r.s, pu, m, n1, n2, A, lda, S, n1, A2, lda, f64, bexp_u, aexp_u, &r.launches);
        if (rc) return rc;
        hipLaunchKernelGGL(add_kernel, grid1((int64_t)n1 * n2), dim3(256), 0, r.s, n1, n2, S, (int64_t)n1, R12,
                           ldr);
        HIP_OK(hipGetLastError());
        ++r.launches;
    }
    return qr_rec(r, m, n2, A2, lda, R + n1 + (int64_t)n1 * ldr, ldr, prec, depth + 1, c0 + n1);
}

bool overlap(const void *a, int64_t abytes, const void *b, int64_t bbytes)
{
    if (!a || !b || abytes <= 0 || bbytes <= 0) return false;
    const char *x = (const char *)a, *y = (const char *)b;
    return x < y + bbytes && y < x + abytes;
}

} // namespace

LATER_API void later_opts_default(later_opts *o)
{
    o->gemm = LATER_GEMM_F16;
    o->gemm_pass2 = LATER_GEMM_F16;
    o->panel = LATER_PANEL_MGS;
    o->reorth = LATER_REORTH_NONE;
    o->nmin = 128;
    o->chunk = 0;
    o->chunk2 = 0;
    o->gemm_update = -1;
    o->nmin2 = 0;
}

LATER_API void later_opts_fp32(later_opts *o)
{
    later_opts_default(o);
    o->gemm = LATER_GEMM_F32;
    o->gemm_pass2 = LATER_GEMM_F32;
}

// Mixed form: FP16 hi/lo products accumulated in FP32 over the whole depth
// in the first pass (its projections only have to leave a Q that the second
// pass can repair: FP64 folds every 256 rows change the final orthogonality
// from 8.66e-6 to 8.60e-6 at 16384^2 and cost 2.5 ms); int8 four-slice
// products summed exactly in int32 for the second pass's projections, which
// set the orthogonality floor (the three-piece FP16 form with FP64 folds
// every 32 rows where the slices cannot be staged or a node is taller than
// 131008 rows); FP32 products at nodes of up to 256 columns (128 for
// the second pass's projections), which are memory-bound: the FP32 MFMA
// runs them faster than the split forms and rounds less (262144 x 256:
// 2.18 vs 2.85 ms); Cholesky QR leaves of up to 64 columns, one pass each
// (the second factorisation repairs what a second Cholesky QR pass of a
// first-pass leaf would).  At 16384^2 (N(0,1)) the form gives
// ||I - Q^T Q||_F ~ 6.6e-6 and ||A - QR||_F / ||A||_F = 3.4e-7 (graded
// singular values 1 .. 1e-4: 7.6e-6 and 5.8e-7); with three-piece
// second-pass projections 8.7e-6 (graded 9.6e-6); single FP16 first-pass
// projections give 1.02e-5 (graded 4.1e-5), two-piece second-pass
// projections 1.3e-5, int8 slices without the weight-2^-28 pairs 1.4e-5,
// single FP16 second-pass updates 2.2e-4.
LATER_API void later_opts_mixed(later_opts *o)
{
    later_opts_default(o);
    o->gemm = LATER_GEMM_F16X3;
    o->gemm_pass2 = LATER_GEMM_I8X4;
    o->panel = LATER_PANEL_CHOLQR1;
    o->reorth = LATER_REORTH_PASS;
    o->nmin = 256;
    o->nmin2 = 128;
    o->chunk = 0;
    o->chunk2 = 32;
}

LATER_API int later_rgsqrf_workspace(int m, int n, const later_opts *opts, int64_t *lwork, int64_t *lhwork)
{
    later_opts o;
    if (opts) o = *opts; else later_opts_default(&o);
    if (m < 1 || n < 1 || m < n) return LATER_ERR_SHAPE;
    Layout L = layout(m, n, o);
    if (lwork) *lwork = L.total;
    if (lhwork) *lhwork = L.hpieces;
    return LATER_OK;
}

LATER_API int later_rgsqrf_ex(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr, float *work,
                              int64_t lwork, uint16_t *hwork, int64_t lhwork, const later_opts *opts,
                              later_info *info)
{
    later_opts o;
    if (opts) o = *opts; else later_opts_default(&o);
    if (info) memset(info, 0, sizeof(*info));
    int rc = LATER_OK;
    if (m < 1 || n < 1 || m < n || lda < m || ldr < n || !A || !R) rc = LATER_ERR_SHAPE;
    else if (!prec_ok(o.gemm) || !prec_ok(o.gemm_pass2) || !update_ok(o.gemm_update) || o.panel < 0 ||
             o.panel > 3 || o.reorth < 0 || o.reorth > 3 || o.nmin < 0 || o.nmin2 < 0 || o.chunk < 0 ||
             (o.chunk % 32) != 0 ||
             o.chunk2 < 0 || (o.chunk2 % 32) != 0 ||
             ((o.reorth & LATER_REORTH_NODE) && (o.gemm == LATER_GEMM_FP8 || o.gemm_pass2 == LATER_GEMM_FP8)))
        rc = LATER_ERR_OPTION;
    Layout L = layout(m > 0 ? m : 1, n > 0 ? n : 1, o);
    // Everything lives in `work` when it is large enough.  Otherwise the
    // 16-bit arena `hwork` (which the reference's caller sizes at m*n halves,
    // test/test_qr.cu:76, next to lwork = m/256*32*n floats, :73) takes the
    // split-K partials, or, when `work` cannot even hold the rest, the whole
    // FP32 layout.  The status word stays at work[0] whenever `work` has room
    // for the header (lwork is 0 for m < 256 there: then it is hwork's).
    const bool h_aligned = hwork && (((uintptr_t)hwork) & 255) == 0;
    const bool work_fits = work && lwork >= L.total;
    const bool hdr_in_work = work && lwork >= L.header;
    const bool partial_in_h = !work_fits && work && lwork >= L.total - L.gemm && h_aligned && lhwork / 2 >= L.gemm;
    const bool all_in_h = !work_fits && !partial_in_h && h_aligned &&
                          lhwork / 2 >= L.total - (hdr_in_work ? L.header : 0);
    if (!rc && !work_fits && !partial_in_h && !all_in_h) rc = LATER_ERR_WORKSPACE;
    // The 16-bit arena holds the pre-split GEMM operands when it is large
    // enough (and not used for FP32 workspace); otherwise the GEMMs convert
    // their operands in the kernel, with the same result bits.
    const bool with_pieces = work_fits && hwork && L.hpieces > 0 && lhwork >= L.hpieces;
    // FP8 projections need their staged operands.
    const bool fp8 = n > o.nmin && (o.gemm == LATER_GEMM_FP8 ||
                                    ((o.reorth & LATER_REORTH_PASS) && o.gemm_pass2 == LATER_GEMM_FP8));
    if (!rc && fp8 && !with_pieces) rc = LATER_ERR_WORKSPACE;
    if (!rc) {
        const int64_t abytes = ((int64_t)(n - 1) * lda + m) * 4, rbytes = ((int64_t)(n - 1) * ldr + n) * 4;
        const int64_t wbytes = lwork * 4, hbytes = lhwork * 2;
        if (overlap(A, abytes, R, rbytes) || overlap(A, abytes, work, wbytes) || overlap(R, rbytes, work, wbytes) ||
            overlap(hwork, hbytes, work, wbytes) || overlap(hwork, hbytes, R, rbytes) ||
            overlap(hwork, hbytes, A, abytes))
            rc = LATER_ERR_ALIAS;
    }
    if (rc) {
        if (info) info->status = rc;
        return rc;
    }
    Run r;
    r.s = ctx_stream(ctxt);
    r.o = o;
    float *p = all_in_h ? (float *)hwork : work;
    if (all_in_h && hdr_in_work) {
        r.status = (int *)work;
    } else {
        r.status = (int *)p;
        p += L.header;
    }
    r.exps = (int *)p; p += L.exps;
    r.panel = p; r.panel_n = L.panel; p += L.panel;
    if (partial_in_h) {
        r.gemm = (float *)hwork;
    } else {
        r.gemm = p;
        p += L.gemm;
    }
    r.gemm_n = L.gemm;
    r.snode = p; p += L.snode;
    float *R2 = p; p += L.r2;
    float *T = p; p += L.tcopy;
    r.pcs = with_pieces ? hwork : nullptr;
    r.qcol = r.pcs;
    r.qrow = r.pcs ? r.pcs + L.hrow : nullptr;
    r.npcs = r.pcs ? r.pcs + L.hnode : nullptr;
    r.ldqc = piece_ld(m);
    r.ldqr = piece_ld(n);
    r.qcol_ps = r.ldqc * n;
    r.qrow_ps = r.ldqr * m;
    r.q8ld = i8_ld(m);
    r.q8ps = i8_stride(m, n);
    r.q8exp = i8_exp_off(m, n);
    r.col_done.assign(n, 0);
    r.row_done.assign(n, 0);
    r.launches = 0;
    r.depth = 0;
    r.node_reorth = (o.reorth & LATER_REORTH_NODE) != 0;
    r.chunk = o.chunk;
    r.chol_once = false;
    r.nmin_p = pass_nmin_p(o, false);
    r.nmin_u = pass_nmin_u(o, false);

    hipLaunchKernelGGL(status_clear_kernel, dim3(1), dim3(64), 0, r.s, r.status);
    hipLaunchKernelGGL(zero_lower_kernel, grid_cols(n, n), dim3(256), 0, r.s, n, R, (int64_t)ldr);
    r.launches += 2;
    rc = hipGetLastError() == hipSuccess ? LATER_OK : LATER_ERR_DEVICE;
    if (!rc) rc = qr_rec(r, m, n, A, lda, R, ldr, o.gemm, 0, 0);
    if (!rc && (o.reorth & LATER_REORTH_PASS)) {
        // Second pass on Q, then R = R1 + (R2 - I) R1 (upper-triangular
        // product).  R2's blocks below the diagonal are never written: its
        // products read the upper triangle only.
        r.node_reorth = false;  // Q has condition number ~1: one projection per node suffices
        r.chol_once = true;     // and one Cholesky QR pass per leaf
        r.nmin_p = pass_nmin_p(o, true);
        r.nmin_u = pass_nmin_u(o, true);
        r.chunk = o.chunk2 > 0 ? o.chunk2 : o.chunk;
        r.col_done.assign(n, 0);  // the second pass writes a new Q
        r.row_done.assign(n, 0);
        rc = qr_rec(r, m, n, A, lda, R2, n, o.gemm_pass2, 0, 0);
        if (!rc) {
            const int pu = update_of(o, o.gemm_pass2);
            const int *bexp = scaled(pu) ? r.exps : nullptr;
            const int aexp = scaled(pu) ? kQExp : 0;
            const bool pre = r.pcs && n >= kPieceMinN && gemm_pieces(pu) > 0;
            Pieces rt{}, tt{};
            if (pre) {
                // Staged operands straight from R2 and R (R1, zero below its
                // diagonal): R2 - I is formed in the split, and the product
                // reads only the pieces, so C = R is safe.
                const int64_t ld = piece_ld(n);
                rt = {r.pcs, ld, ld * n};
                tt = {r.pcs + gemm_pieces(pu) * ld * n, ld, ld * n};
                rc = gemm_split_t_tri(r.s, pu, n, R2, n, aexp, 1.0f, (uint16_t *)rt.p, ld, ld * n, &r.launches);
                if (!rc && scaled(pu))
                    rc = gemm_split_exp(r.s, pu, n, n, R, ldr, r.exps, (uint16_t *)tt.p, ld, ld * n, &r.launches);
                else if (!rc)
                    rc = gemm_split(r.s, pu, n, n, R, ldr, 0, bexp, (uint16_t *)tt.p, ld, ld * n, false, &r.launches);
            } else {
                hipLaunchKernelGGL(copy_upper_kernel, grid_cols(n, n), dim3(256), 0, r.s, n, R, (int64_t)ldr, T,
                                   0.0f);
                hipLaunchKernelGGL(copy_upper_kernel, grid_cols(n, n), dim3(256), 0, r.s, n, R2, (int64_t)n, R2,
                                   1.0f);
                r.launches += 2;
                if (scaled(pu)) rc = gemm_colexp(r.s, n, n, T, n, r.exps, &r.launches);
            }
            if (!rc)
                rc = gemm_nn(r.s, pu, n, n, n, 1.0f, R2, n, T, n, R, ldr, o.chunk > 0, bexp, aexp, true,
                             &r.launches, pre ? &rt : nullptr, pre ? &tt : nullptr);
        }
    }
    if (!rc && hipGetLastError() != hipSuccess) rc = LATER_ERR_DEVICE;
    if (info) {
        info->status = rc;
        info->kernels = r.launches;
        info->tree_levels = r.depth;
        info->flops_ref = 2.0 * (double)n * n * ((double)m - n / 3.0);
        info->flops_mgs = 2.0 * (double)m * n * n;
    }
    return rc;
}

LATER_API void later_rgsqrf(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr, float *work,
                            int lwork, uint16_t *hwork, int lhwork)
{
    const int rc = later_rgsqrf_ex(ctxt, m, n, A, lda, R, ldr, work, lwork, hwork, lhwork, nullptr, nullptr);
    // The reference's polar-decomposition caller passes R == work and reads
    // only Q afterwards (EVD/later_qdwh_polar.cu:79-84).  The status form
    // rejects that aliasing; here the factorisation then keeps its whole
    // FP32 workspace in the 16-bit arena (the caller's 2n*n halves hold it),
    // so that Q and R are both written.
    if (rc == LATER_ERR_ALIAS && hwork && R && work && lwork >= 0) {
        const int64_t rbytes = ((int64_t)(n - 1) * ldr + n) * 4;
        if (overlap(R, rbytes, work, (int64_t)lwork * 4 + rbytes))
            later_rgsqrf_ex(ctxt, m, n, A, lda, R, ldr, nullptr, 0, hwork, lhwork, nullptr, nullptr);
    }
}

LATER_API int later_read_status(laterCtxt ctxt, const float *work, int *device_status)
{
    if (!work || !device_status) return LATER_ERR_SHAPE;
    hipStream_t s = ctx_stream(ctxt);
    HIP_OK(hipMemcpyAsync(device_status, work, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return LATER_OK;
}

LATER_API int later_caqr_panel_workspace(int m, int w, int fp64_sums, int64_t *lwork)
{
    if (w < 1 || w > 32 || m < w) return LATER_ERR_SHAPE;
    if (fp64_sums < 0 || fp64_sums > 2) return LATER_ERR_OPTION;
    if (lwork) *lwork = kHeaderFloats + (fp64_sums == 2 ? panel_chol_work_floats(w) : panel_work_floats(m, w));
    return LATER_OK;
}

LATER_API int later_caqr_panel(laterCtxt ctxt, int m, int w, float *A, int lda, float *R, int ldr, float *work,
                               int64_t lwork, int fp64_sums)
{
    if (w < 1 || w > 32 || m < w || lda < m || ldr < w || !A || !R) return LATER_ERR_SHAPE;
    int64_t need = 0;
    const int rc = later_caqr_panel_workspace(m, w, fp64_sums, &need);
    if (rc) return rc;
    if (!work || lwork < need) return LATER_ERR_WORKSPACE;
    hipStream_t s = ctx_stream(ctxt);
    hipLaunchKernelGGL(status_clear_kernel, dim3(1), dim3(64), 0, s, (int *)work);
    if (fp64_sums == 2)
        return panel_cholqr2(s, m, w, A, lda, R, ldr, work + kHeaderFloats, need - kHeaderFloats, (int *)work,
                             nullptr);
    return panel_tsqr(s, m, w, A, lda, R, ldr, work + kHeaderFloats, need - kHeaderFloats, fp64_sums != 0,
                      (int *)work, nullptr);
}

// 16-bit elements of the pre-split operand pieces of one component GEMM, in
// the layout the recursion's nodes use: the projection stages Q and B
// column-major (ld piece_ld(m)); the update stages Q transposed (ld
// piece_ld(n1), m rows) and C column-major (ld piece_ld(n1)).
static int64_t component_pieces(int m, int n1, int n2, int prec, bool update)
{
    if (staged8(prec)) return update ? 0 : (staged_bytes(prec, m, n1) + staged_bytes(prec, m, n2)) / 2;
    const int np = gemm_pieces(prec);
    if (np == 0) return 0;
    if (!update) return (int64_t)np * piece_ld(m) * ((int64_t)n1 + n2);
    return (int64_t)np * piece_ld(n1) * ((int64_t)m + n2);
}

LATER_API int later_gemm_workspace(int m, int n1, int n2, int prec, int chunk, int update, int64_t *lwork,
                                   int64_t *lhwork)
{
    if (m < 1 || n1 < 1 || n2 < 1) return LATER_ERR_SHAPE;
    if (!prec_ok(prec) || chunk < 0 || chunk % 32 || (update && staged8(prec))) return LATER_ERR_OPTION;
    if (staged8(prec)) chunk = 0;
    const int64_t exps = ((int64_t)n2 + 63) / 64 * 64;
    if (lwork) *lwork = update ? exps : exps + gemm_proj_work_floats(m, n1, n2, chunk);
    if (lhwork) *lhwork = component_pieces(m, n1, n2, prec, update != 0);
    return LATER_OK;
}

LATER_API int later_gemm_proj(laterCtxt ctxt, int m, int n1, int n2, const float *Q, int ldq, const float *B,
                              int ldb, float *C, int ldc, int prec, int chunk, float *work, int64_t lwork,
                              uint16_t *hwork, int64_t lhwork)
{
    if (m < 1 || n1 < 1 || n2 < 1 || ldq < m || ldb < m || ldc < n1 || !Q || !B || !C) return LATER_ERR_SHAPE;
    if (!prec_ok(prec) || chunk < 0 || chunk % 32) return LATER_ERR_OPTION;
    const bool fp8 = staged8(prec);
    if (fp8) chunk = 0;
    const int64_t exps = ((int64_t)n2 + 63) / 64 * 64;
    const int64_t need = exps + gemm_proj_work_floats(m, n1, n2, chunk);
    if (lwork < need || (!work && need > 0)) return LATER_ERR_WORKSPACE;
    const int64_t hneed = component_pieces(m, n1, n2, prec, false);
    const bool pre = hwork && hneed > 0 && lhwork >= hneed && (((uintptr_t)hwork) & 255) == 0;
    hipStream_t s = ctx_stream(ctxt);
    if (fp8) {
        // e4m3 operands with per-128-row block scales, or int8 slices with
        // column exponents, staged in hwork.
        if (!pre) return LATER_ERR_WORKSPACE;
        const bool i8 = prec == LATER_GEMM_I8X4;
        if (i8 && m > kI8MaxK) return LATER_ERR_SHAPE;
        const Pieces qh = i8 ? i8_pieces(hwork, m, n1) : fp8_pieces(hwork, m, n1);
        const uint16_t *hb = hwork + staged_bytes(prec, m, n1) / 2;
        const Pieces bh = i8 ? i8_pieces(hb, m, n2) : fp8_pieces(hb, m, n2);
        int rc = i8 ? gemm_split_i8(s, m, n1, Q, ldq, (void *)qh.p, nullptr)
                    : gemm_split8(s, m, n1, Q, ldq, (void *)qh.p, nullptr);
        if (!rc) rc = i8 ? gemm_split_i8(s, m, n2, B, ldb, (void *)bh.p, nullptr)
                         : gemm_split8(s, m, n2, B, ldb, (void *)bh.p, nullptr);
        if (rc) return rc;
        return gemm_proj(s, prec, m, n1, n2, Q, ldq, B, ldb, C, ldc, 0, nullptr, 0, work + exps, lwork - exps,
                         nullptr, &qh, &bh);
    }
    const int aexp = scaled(prec) ? kQExp : 0;
    const int *bexp = scaled(prec) ? (const int *)work : nullptr;
    Pieces qh{}, bh{};
    int rc = LATER_OK;
    if (pre) {
        const int64_t ld = piece_ld(m), np = gemm_pieces(prec);
        qh = {hwork, ld, ld * n1};
        bh = {hwork + np * ld * n1, ld, ld * n2};
        rc = gemm_split(s, prec, m, n1, Q, ldq, aexp, nullptr, hwork, qh.ld, qh.stride, false, nullptr);
        if (!rc && scaled(prec))
            rc = gemm_split_exp(s, prec, m, n2, B, ldb, (int *)work, (uint16_t *)bh.p, bh.ld, bh.stride, nullptr);
        else if (!rc)
            rc = gemm_split(s, prec, m, n2, B, ldb, 0, nullptr, (uint16_t *)bh.p, bh.ld, bh.stride, false, nullptr);
    } else if (scaled(prec)) {
        rc = gemm_colexp(s, m, n2, B, ldb, (int *)work, nullptr);
    }
    if (rc) return rc;
    return gemm_proj(s, prec, m, n1, n2, Q, ldq, B, ldb, C, ldc, chunk, bexp, aexp, work + exps, lwork - exps,
                     nullptr, pre ? &qh : nullptr, pre ? &bh : nullptr);
}

LATER_API int later_gemm_update(laterCtxt ctxt, int m, int n1, int n2, const float *Q, int ldq, const float *C,
                                int ldc, float *B, int ldb, int prec, float *work, int64_t lwork, uint16_t *hwork,
                                int64_t lhwork)
{
    if (m < 1 || n1 < 1 || n2 < 1 || ldq < m || ldb < m || ldc < n1 || !Q || !B || !C) return LATER_ERR_SHAPE;
    if (!prec_ok(prec) || staged8(prec)) return LATER_ERR_OPTION;
    const int64_t exps = ((int64_t)n2 + 63) / 64 * 64;
    if (scaled(prec) && (lwork < exps || !work)) return LATER_ERR_WORKSPACE;
    const int64_t hneed = component_pieces(m, n1, n2, prec, true);
    const bool pre = hwork && hneed > 0 && lhwork >= hneed && (((uintptr_t)hwork) & 15) == 0;
    hipStream_t s = ctx_stream(ctxt);
    const int aexp = scaled(prec) ? kQExp : 0;
    const int *bexp = scaled(prec) ? (const int *)work : nullptr;
    Pieces qt{}, ch{};
    int rc = LATER_OK;
    if (pre) {
        const int64_t ld = piece_ld(n1), np = gemm_pieces(prec);
        qt = {hwork, ld, ld * m};
        ch = {hwork + np * ld * m, ld, ld * n2};
        rc = gemm_split(s, prec, m, n1, Q, ldq, aexp, nullptr, hwork, qt.ld, qt.stride, true, nullptr);
        if (!rc && scaled(prec))
            rc = gemm_split_exp(s, prec, n1, n2, C, ldc, (int *)work, (uint16_t *)ch.p, ch.ld, ch.stride, nullptr);
        else if (!rc)
            rc = gemm_split(s, prec, n1, n2, C, ldc, 0, nullptr, (uint16_t *)ch.p, ch.ld, ch.stride, false, nullptr);
    } else if (scaled(prec)) {
        rc = gemm_colexp(s, n1, n2, C, ldc, (int *)work, nullptr);
    }
    if (rc) return rc;
    return gemm_update(s, prec, m, n1, n2, Q, ldq, C, ldc, B, ldb, false, bexp, aexp, nullptr, pre ? &qt : nullptr,
                       pre ? &ch : nullptr);
}

LATER_API const char *later_version(void) { return "later_amd 0.1 (gfx950)"; }
