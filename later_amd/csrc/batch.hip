// Batched QR of many independent matrices of uneven shape.
//
// Matrices are bucketed by k = ceil(n/32).  Within a bucket every matrix has
// the same recursion tree in units of 32 columns (the split n1 = 32 ceil(n/64)
// depends on n only through k; only the rightmost leaf's width differs), so
// each step of the tree -- a node's projection, its update, a TSQR level of a
// leaf -- is one grouped launch over all matrices of the bucket: every
// workgroup reads its (matrix, tile) from a task table.  Buckets are
// independent and are recorded on separate streams, forked and joined inside
// one captured HIP graph, so their launches overlap on the device.  The plan
// (task tables, argument blocks and the graph) is built once; executing it is
// one graph launch.  Nothing splits K: the per-matrix products are small
// enough that the tiles of all matrices fill the GPU.
#include <algorithm>
#include <map>
#include <vector>

#include "common.h"
#include "gemm.h"
#include "panel.h"

namespace {

int split_n(int n) { return 32 * ((n + 63) / 64); }
bool scaled(int p) { return p == LATER_GEMM_F16X3 || p == LATER_GEMM_F16X6; }
int update_prec(int p) { return p == LATER_GEMM_F16X6 ? LATER_GEMM_F16X3 : p; }
constexpr int kQExp = 14;

// Per-matrix elementwise n x n operations, one workgroup per matrix.
struct ElemArgs {
    int n;
    const float *src;
    int64_t lds;
    float *dst;
    int64_t ldd;
    float shift;
    int mode;  // 0: zero the strictly-lower triangle of dst; 1: dst = 0; 2: dst = triu(src) - shift I;
               // 3: clear the status word at dst
};

__global__ void elem_grouped_kernel(const ElemArgs *args)
{
    const ElemArgs a = args[blockIdx.x];
    if (a.mode == 3) {
        if (threadIdx.x == 0) *(int *)a.dst = 0;
        return;
    }
    const int64_t tot = (int64_t)a.n * a.n;
    for (int64_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const int i = (int)(idx % a.n), j = (int)(idx / a.n);
        float *d = a.dst + i + (int64_t)j * a.ldd;
        if (a.mode == 0) {
            if (i > j) *d = 0.0f;
        } else if (a.mode == 1) {
            *d = 0.0f;
        } else {
            float v = (i <= j) ? a.src[i + (int64_t)j * a.lds] : 0.0f;
            if (i == j) v -= a.shift;
            *d = v;
        }
    }
}

enum LaunchType { L_GEMM_PROJ, L_GEMM_UPD, L_COLEXP, L_MGS, L_APPLY, L_ELEM, L_STATUS };

struct Launch {
    int type, prec;
    bool ch64;
    int off, count;  // into the task (or args) table of its type
};

struct Mat {
    int m, n;
    float *A, *R, *R2, *T;
    int64_t lda, ldr;
    int *exps;
    int *status;
    float *stack;
};

struct Builder {
    std::vector<GemmArgs> gargs;
    std::vector<int4> gtasks;
    std::vector<ColexpArgs> cargs;
    std::vector<int2> ctasks;
    std::vector<PanelArgs> pargs;
    std::vector<int2> ptasks;
    std::vector<ElemArgs> eargs;
    std::vector<std::vector<Launch>> streams;  // one launch list per bucket
    later_opts o;
    int *status_all = nullptr;
    int count = 0;

    std::vector<Launch> *cur = nullptr;

    void colexp(const std::vector<Mat *> &ms, const std::vector<ColexpArgs> &a)
    {
        Launch L{L_COLEXP, 0, false, (int)ctasks.size(), 0};
        for (size_t q = 0; q < a.size(); ++q) {
            const int ai = (int)cargs.size();
            cargs.push_back(a[q]);
            (void)ms;
            for (int j = 0; j < ncols_[q]; ++j) ctasks.push_back(make_int2(ai, j));
        }
        L.count = (int)ctasks.size() - L.off;
        if (L.count) cur->push_back(L);
    }
    std::vector<int> ncols_;

    void gemm(int prec, bool proj, bool ch64, const std::vector<GemmArgs> &a)
    {
        Launch L{proj ? L_GEMM_PROJ : L_GEMM_UPD, prec, ch64, (int)gtasks.size(), 0};
        for (const GemmArgs &g : a) {
            const int ai = (int)gargs.size();
            gargs.push_back(g);
            const int tm = (g.M + 63) / 64, tn = (g.N + 63) / 64;
            for (int x = 0; x < tm; ++x)
                for (int y = 0; y < tn; ++y) {
                    if (g.tri && x * 64 > y * 64 + 63) continue;
                    gtasks.push_back(make_int4(ai, x, y, 0));
                }
        }
        L.count = (int)gtasks.size() - L.off;
        if (L.count) cur->push_back(L);
    }

    void elem(const std::vector<ElemArgs> &a)
    {
        if (a.empty()) return;
        Launch L{L_ELEM, 0, false, (int)eargs.size(), (int)a.size()};
        eargs.insert(eargs.end(), a.begin(), a.end());
        cur->push_back(L);
    }

    // One TSQR level across matrices: MGS (or apply) of the listed panels.
    // Batched leaves always use the block-MGS tree; the accurate panels
    // (LATER_PANEL_MGS64 and LATER_PANEL_CHOLQR2) run it with FP64 sums.
    void panel(bool apply, const std::vector<PanelArgs> &a)
    {
        Launch L{apply ? L_APPLY : L_MGS, 0, o.panel != LATER_PANEL_MGS, (int)ptasks.size(), 0};
        for (const PanelArgs &p : a) {
            const int ai = (int)pargs.size();
            pargs.push_back(p);
            for (int b = 0; b < p.nblk; ++b) ptasks.push_back(make_int2(ai, b));
        }
        L.count = (int)ptasks.size() - L.off;
        if (L.count) cur->push_back(L);
    }

    // Leaf: TSQR of the columns [c0, c0 + w_i) of every matrix.
    void leaf(const std::vector<Mat *> &ms, int c0, bool pass2)
    {
        // levels[l] = panel args of level l for every matrix that reaches it.
        std::vector<std::vector<PanelArgs>> levels;
        for (Mat *M : ms) {
            const int w = std::min(32, M->n - c0);
            float *Rt = pass2 ? M->R2 : M->R;
            const int64_t ldr = pass2 ? M->n : M->ldr;
            float *A = M->A + (int64_t)c0 * M->lda;
            int64_t lda = M->lda;
            int m = M->m;
            float *stack = M->stack;
            for (int l = 0;; ++l) {
                const int nb = panel_blocks(m, w);
                PanelArgs p;
                p.m = m; p.w = w; p.nblk = nb; p.A = A; p.lda = lda; p.status = M->status;
                if (nb <= 1) {
                    p.W = Rt + c0 + (int64_t)c0 * ldr;
                    p.ldw = ldr;
                } else {
                    p.W = stack;
                    p.ldw = (int64_t)nb * w;
                }
                if ((int)levels.size() <= l) levels.resize(l + 1);
                levels[l].push_back(p);
                if (nb <= 1) break;
                A = stack;
                lda = (int64_t)nb * w;
                m = nb * w;
                stack += (int64_t)nb * w * w;
            }
        }
        for (auto &lv : levels) panel(false, lv);
        // Apply from the deepest level up, for panels that have a tree below.
        for (int l = (int)levels.size() - 1; l >= 0; --l) {
            std::vector<PanelArgs> ap;
            for (const PanelArgs &p : levels[l])
                if (p.nblk > 1) ap.push_back(p);
            panel(true, ap);
        }
    }

    void node(const std::vector<Mat *> &ms, int c0, int n1, int w2, int prec, int pu, int chunk, bool pass2)
    {
        const bool sc = scaled(prec), ch64 = chunk > 0, scu = scaled(pu);
        const int aexp = sc ? kQExp : 0, aexp_u = scu ? kQExp : 0;
        // Projection R12 = Q1^T A2.
        if (sc) {
            std::vector<ColexpArgs> ca;
            ncols_.clear();
            for (Mat *M : ms) {
                const int n2 = std::min(w2, M->n - c0 - n1);
                ca.push_back({M->m, M->A + (int64_t)(c0 + n1) * M->lda, M->lda, M->exps});
                ncols_.push_back(n2);
            }
            colexp(ms, ca);
        }
        std::vector<GemmArgs> pa, ua;
        for (Mat *M : ms) {
            const int n2 = std::min(w2, M->n - c0 - n1);
            float *Rt = pass2 ? M->R2 : M->R;
            const int64_t ldr = pass2 ? M->n : M->ldr;
            GemmArgs g{};
            g.M = n1; g.N = n2; g.K = M->m;
            g.A = M->A + (int64_t)c0 * M->lda; g.lda = M->lda;
            g.B = M->A + (int64_t)(c0 + n1) * M->lda; g.ldb = M->lda;
            g.C = Rt + c0 + (int64_t)(c0 + n1) * ldr; g.ldc = ldr;
            g.a_exp = aexp; g.b_exp = sc ? M->exps : nullptr;
            g.kper = (M->m + 31) / 32 * 32;
            g.chunk_rows = ch64 ? chunk : 0;
            g.partial = nullptr; g.update = 0; g.alpha = 1.0f; g.tri = 0;
            pa.push_back(g);
        }
        gemm(prec, true, ch64, pa);
        // Update A2 -= Q1 R12.
        if (scu) {
            std::vector<ColexpArgs> ca;
            ncols_.clear();
            for (Mat *M : ms) {
                const int n2 = std::min(w2, M->n - c0 - n1);
                float *Rt = pass2 ? M->R2 : M->R;
                const int64_t ldr = pass2 ? M->n : M->ldr;
                ca.push_back({n1, Rt + c0 + (int64_t)(c0 + n1) * ldr, ldr, M->exps});
                ncols_.push_back(n2);
            }
            colexp(ms, ca);
        }
        for (size_t q = 0; q < ms.size(); ++q) {
            Mat *M = ms[q];
            const GemmArgs &p = pa[q];
            GemmArgs g{};
            g.M = M->m; g.N = p.N; g.K = n1;
            g.A = p.A; g.lda = M->lda;
            g.B = p.C; g.ldb = p.ldc;
            g.C = (float *)p.B; g.ldc = M->lda;
            g.a_exp = aexp_u; g.b_exp = scu ? M->exps : nullptr;
            g.kper = (n1 + 31) / 32 * 32;
            g.chunk_rows = 0;
            g.partial = nullptr; g.update = (ch64 && pu != LATER_GEMM_F32) ? 2 : 1; g.alpha = -1.0f; g.tri = 0;
            ua.push_back(g);
        }
        gemm(pu, false, false, ua);
    }

    void tree(const std::vector<Mat *> &ms, int c0, int wn, int prec, int chunk, bool pass2)
    {
        if (wn <= 32) {
            leaf(ms, c0, pass2);
            return;
        }
        const int n1 = split_n(wn);
        // FP32-product widths per pass, as the single-matrix recursion runs them.
        const int nmin_p = (pass2 && o.nmin2 > 0) ? o.nmin2 : o.nmin;
        const int nmin_u = pass2 ? std::max(o.nmin, nmin_p) : o.nmin;
        const bool inner = wn <= nmin_p;
        const int p = inner ? LATER_GEMM_F32 : prec;
        const int pu = wn <= nmin_u ? LATER_GEMM_F32 : update_prec(prec);
        const int ch = inner ? 0 : chunk;
        tree(ms, c0, n1, prec, chunk, pass2);
        node(ms, c0, n1, wn - n1, p, pu, ch, pass2);
        tree(ms, c0 + n1, wn - n1, prec, chunk, pass2);
    }

    void bucket(const std::vector<Mat *> &ms, int k)
    {
        const int wn = 32 * k;
        std::vector<ElemArgs> z;
        for (Mat *M : ms) z.push_back({M->n, nullptr, 0, M->R, M->ldr, 0.0f, 0});
        for (Mat *M : ms) z.push_back({1, nullptr, 0, (float *)M->status, 1, 0.0f, 3});
        elem(z);
        tree(ms, 0, wn, o.gemm, o.chunk, false);
        if (o.reorth & LATER_REORTH_PASS) {
            std::vector<ElemArgs> f;
            for (Mat *M : ms) f.push_back({M->n, nullptr, 0, M->R2, M->n, 0.0f, 1});
            elem(f);
            tree(ms, 0, wn, o.gemm_pass2, o.chunk2 > 0 ? o.chunk2 : o.chunk, true);
            // R = R1 + (R2 - I) R1
            std::vector<ElemArgs> c;
            for (Mat *M : ms) c.push_back({M->n, M->R, M->ldr, M->T, M->n, 0.0f, 2});
            for (Mat *M : ms) c.push_back({M->n, M->R2, M->n, M->R2, M->n, 1.0f, 2});
            elem(c);
            const int p2 = update_prec(o.gemm_pass2);
            if (scaled(p2)) {
                std::vector<ColexpArgs> ca;
                ncols_.clear();
                for (Mat *M : ms) {
                    ca.push_back({M->n, M->T, M->n, M->exps});
                    ncols_.push_back(M->n);
                }
                colexp(ms, ca);
            }
            std::vector<GemmArgs> ta;
            for (Mat *M : ms) {
                GemmArgs g{};
                g.M = M->n; g.N = M->n; g.K = M->n;
                g.A = M->R2; g.lda = M->n; g.B = M->T; g.ldb = M->n; g.C = M->R; g.ldc = M->ldr;
                g.a_exp = scaled(p2) ? kQExp : 0; g.b_exp = scaled(p2) ? M->exps : nullptr;
                g.kper = (M->n + 31) / 32 * 32;
                g.chunk_rows = 0;
                g.partial = nullptr; g.update = o.chunk > 0 ? 2 : 1; g.alpha = 1.0f; g.tri = 1;
                ta.push_back(g);
            }
            gemm(p2, false, false, ta);
        }
    }
};

int64_t mat_work(int m, int n, const later_opts &o)
{
    int64_t f = ((int64_t)n + 63) / 64 * 64 + 64;  // exps + status
    f += panel_work_floats_any(m);
    if (o.reorth & LATER_REORTH_PASS) f += 2LL * n * n;
    return (f + 63) / 64 * 64;
}

struct BatchPlan {
    void *dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int launches = 0;
};

template <typename T>
T *upload(char *&p, const std::vector<T> &v)
{
    T *d = (T *)p;
    p += (v.size() * sizeof(T) + 255) / 256 * 256;
    return d;
}

} // namespace

LATER_API int later_batch_workspace(int count, const int *m, const int *n, const later_opts *opts, int64_t *lwork)
{
    later_opts o;
    if (opts) o = *opts; else later_opts_default(&o);
    if (count < 1 || !m || !n || !lwork) return LATER_ERR_SHAPE;
    int64_t tot = 0;
    for (int i = 0; i < count; ++i) {
        if (m[i] < 1 || n[i] < 1 || m[i] < n[i]) return LATER_ERR_SHAPE;
        tot += mat_work(m[i], n[i], o);
    }
    *lwork = tot;
    return LATER_OK;
}

LATER_API int later_batch_plan_create(int count, const int *m, const int *n, float *const *A, const int *lda,
                                      float *const *R, const int *ldr, float *work, int64_t lwork,
                                      const later_opts *opts, void **plan)
{
    if (!plan) return LATER_ERR_SHAPE;
    *plan = nullptr;
    later_opts o;
    if (opts) o = *opts; else later_opts_default(&o);
    if (o.reorth & LATER_REORTH_NODE) return LATER_ERR_OPTION;  // not offered for batches
    // The grouped kernels convert their operands in the kernel; int8 slices
    // need a staged, column-scaled operand, so those nodes run the
    // three-piece FP16 split (as later_rgsqrf_ex does without its arena).
    if (o.gemm == LATER_GEMM_I8X4) o.gemm = LATER_GEMM_F16X6;
    if (o.gemm_pass2 == LATER_GEMM_I8X4) o.gemm_pass2 = LATER_GEMM_F16X6;
    // The same option checks as later_rgsqrf_ex (a chunk below 32 would fold
    // the FP64 sums only after the last k).
    auto prec_ok = [](int p) { return p >= LATER_GEMM_F32 && p <= LATER_GEMM_F16X6; };
    if (!prec_ok(o.gemm) || !prec_ok(o.gemm_pass2) || o.gemm_update != -1 || o.panel < 0 || o.panel > 3 ||
        o.reorth < 0 || o.reorth > 3 || o.nmin < 0 || o.nmin2 < 0 || o.chunk < 0 || (o.chunk % 32) != 0 || o.chunk2 < 0 ||
        (o.chunk2 % 32) != 0)
        return LATER_ERR_OPTION;
    int64_t need = 0;
    int rc = later_batch_workspace(count, m, n, &o, &need);
    if (rc) return rc;
    if (!work || lwork < need) return LATER_ERR_WORKSPACE;
    for (int i = 0; i < count; ++i)
        if (!A[i] || !R[i] || lda[i] < m[i] || ldr[i] < n[i]) return LATER_ERR_SHAPE;

    std::vector<Mat> mats(count);
    std::map<int, std::vector<Mat *>> buckets;
    float *p = work;
    for (int i = 0; i < count; ++i) {
        Mat &M = mats[i];
        M.m = m[i]; M.n = n[i]; M.A = A[i]; M.R = R[i]; M.lda = lda[i]; M.ldr = ldr[i];
        float *q = p;
        M.status = (int *)q; q += 64;
        M.exps = (int *)q; q += ((int64_t)n[i] + 63) / 64 * 64;
        M.stack = q; q += panel_work_floats_any(m[i]);
        if (o.reorth & LATER_REORTH_PASS) {
            M.R2 = q; q += (int64_t)n[i] * n[i];
            M.T = q; q += (int64_t)n[i] * n[i];
        } else {
            M.R2 = M.T = nullptr;
        }
        p += mat_work(m[i], n[i], o);
        buckets[(n[i] + 31) / 32].push_back(&M);
    }

    Builder B;
    B.o = o;
    B.count = count;
    for (auto &kv : buckets) {
        // Largest matrices first inside a bucket, so the long tiles start early.
        std::sort(kv.second.begin(), kv.second.end(), [](Mat *a, Mat *b) { return a->m > b->m; });
        B.streams.emplace_back();
        B.cur = &B.streams.back();
        B.bucket(kv.second, kv.first);
    }

    // Upload the tables.
    const int64_t bytes = 256 * 8 + (B.gargs.size() * sizeof(GemmArgs) + B.gtasks.size() * sizeof(int4) +
                                     B.cargs.size() * sizeof(ColexpArgs) + B.ctasks.size() * sizeof(int2) +
                                     B.pargs.size() * sizeof(PanelArgs) + B.ptasks.size() * sizeof(int2) +
                                     B.eargs.size() * sizeof(ElemArgs));
    BatchPlan *P = new BatchPlan();
    if (hipMalloc(&P->dev, bytes) != hipSuccess) { delete P; return LATER_ERR_DEVICE; }
    char *d = (char *)P->dev;
    GemmArgs *d_gargs = upload(d, B.gargs);
    int4 *d_gtasks = upload(d, B.gtasks);
    ColexpArgs *d_cargs = upload(d, B.cargs);
    int2 *d_ctasks = upload(d, B.ctasks);
    PanelArgs *d_pargs = upload(d, B.pargs);
    int2 *d_ptasks = upload(d, B.ptasks);
    ElemArgs *d_eargs = upload(d, B.eargs);
    auto put = [&](void *dst, const void *src, size_t n) {
        return n == 0 || hipMemcpy(dst, src, n, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = put(d_gargs, B.gargs.data(), B.gargs.size() * sizeof(GemmArgs)) &&
              put(d_gtasks, B.gtasks.data(), B.gtasks.size() * sizeof(int4)) &&
              put(d_cargs, B.cargs.data(), B.cargs.size() * sizeof(ColexpArgs)) &&
              put(d_ctasks, B.ctasks.data(), B.ctasks.size() * sizeof(int2)) &&
              put(d_pargs, B.pargs.data(), B.pargs.size() * sizeof(PanelArgs)) &&
              put(d_ptasks, B.ptasks.data(), B.ptasks.size() * sizeof(int2)) &&
              put(d_eargs, B.eargs.data(), B.eargs.size() * sizeof(ElemArgs));
    if (!ok) { hipFree(P->dev); delete P; return LATER_ERR_DEVICE; }

    // Record: status clear on the main stream, each bucket on its own stream
    // (fork / join through events), all captured into one graph.
    const int nb = (int)B.streams.size();
    hipStream_t main_s;
    std::vector<hipStream_t> ss(nb);
    std::vector<hipEvent_t> ev(nb + 1);
    hipStreamCreateWithFlags(&main_s, hipStreamNonBlocking);
    for (int i = 0; i < nb; ++i) hipStreamCreateWithFlags(&ss[i], hipStreamNonBlocking);
    for (int i = 0; i <= nb; ++i) hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
    rc = hipStreamBeginCapture(main_s, hipStreamCaptureModeThreadLocal) == hipSuccess ? LATER_OK : LATER_ERR_DEVICE;
    int launches = 0;
    if (!rc) {
        hipEventRecord(ev[nb], main_s);
        for (int i = 0; i < nb && !rc; ++i) {
            hipStreamWaitEvent(ss[i], ev[nb], 0);
            for (const Launch &L : B.streams[i]) {
                switch (L.type) {
                case L_GEMM_PROJ:
                case L_GEMM_UPD:
                    rc = gemm_grouped(ss[i], L.prec, L.type == L_GEMM_PROJ, L.ch64, d_gargs, d_gtasks + L.off, L.count);
                    break;
                case L_COLEXP: rc = colexp_grouped(ss[i], d_cargs, d_ctasks + L.off, L.count); break;
                case L_MGS: rc = mgs_grouped(ss[i], L.ch64, d_pargs, d_ptasks + L.off, L.count); break;
                case L_APPLY: rc = apply_grouped(ss[i], L.ch64, d_pargs, d_ptasks + L.off, L.count); break;
                case L_ELEM:
                    hipLaunchKernelGGL(elem_grouped_kernel, dim3(L.count), dim3(256), 0, ss[i], d_eargs + L.off);
                    rc = hipGetLastError() == hipSuccess ? LATER_OK : LATER_ERR_DEVICE;
                    break;
                default: break;
                }
                ++launches;
                if (rc) break;
            }
            hipEventRecord(ev[i], ss[i]);
            hipStreamWaitEvent(main_s, ev[i], 0);
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(main_s, &g);
        if (!rc && (e != hipSuccess || !g)) rc = LATER_ERR_DEVICE;
        P->graph = g;
    }
    if (!rc && hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0) != hipSuccess) rc = LATER_ERR_DEVICE;
    for (int i = 0; i < nb; ++i) hipStreamDestroy(ss[i]);
    for (int i = 0; i <= nb; ++i) hipEventDestroy(ev[i]);
    hipStreamDestroy(main_s);
    if (rc) {
        if (P->exec) hipGraphExecDestroy(P->exec);
        if (P->graph) hipGraphDestroy(P->graph);
        hipFree(P->dev);
        delete P;
        return rc;
    }
    P->launches = launches;
    *plan = P;
    return LATER_OK;
}

LATER_API int later_batch_plan_execute(void *plan, void *stream)
{
    BatchPlan *P = (BatchPlan *)plan;
    if (!P || !P->exec) return LATER_ERR_SHAPE;
    HIP_OK(hipGraphLaunch(P->exec, (hipStream_t)stream));
    return LATER_OK;
}

LATER_API int later_batch_plan_launches(void *plan)
{
    BatchPlan *P = (BatchPlan *)plan;
    return P ? P->launches : -1;
}

LATER_API int later_batch_plan_destroy(void *plan)
{
    BatchPlan *P = (BatchPlan *)plan;
    if (!P) return LATER_ERR_SHAPE;
    if (P->exec) hipGraphExecDestroy(P->exec);
    if (P->graph) hipGraphDestroy(P->graph);
    hipFree(P->dev);
    delete P;
    return LATER_OK;
}
