// Least-squares serving runtime: continuous batching of QR solves on one GPU.
//
// Requests (A: m x n, B: m x k, device pointers owned by the caller) arrive
// at any time.  Each is admitted when a slot (a HIP stream) is free and its
// working set fits in the server's HBM pool, which is allocated once at
// creation and carved by a first-fit allocator, so the solve path itself
// never calls hipMalloc.  Admitted requests run concurrently on their own
// streams and progress through a small state machine driven by
// later_server_poll():
//   COPY+QR   A and B are copied into the pool; A is factored in place
//             (FP16 hi/lo products, FP64 panel sums, one pass); a condition
//             estimate max|r_ii| / min|r_ii| and an orthogonality sketch
//             eta = ||Q^T Q g - g|| / ||g|| are written to host memory;
//   QR2       only when eta exceeds 1e-5 (one pass of block Gram-Schmidt
//             loses orthogonality like kappa^2): A is refactored from the
//             kept copy with CGS2 at every node plus a second pass, and the
//             sketch is taken again;
//   SOLVE     X = R^{-1} Q^T B, then `iters` rounds of refinement
//             X += R^{-1} Q^T (B - A X) with the residual summed in FP64 on
//             the original A; iters is chosen from the condition estimate;
//   DONE      Q, R and X stay resident in the pool until released.
// Every kernel a request launches is a function of its own (m, n, k), never
// of the load, so a request's result is bitwise the same alone and in a mix.
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "gemm.h"

int trsm_upper(hipStream_t s, int n, int k, const float *R, int64_t ldr, float *B, int64_t ldb, int *status);

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// cond estimate from R's diagonal (a lower bound of kappa_2(R)).
__global__ void diag_ratio_kernel(int n, const float *R, int64_t ldr, double *out)
{
    __shared__ float mx[256], mn[256];
    float a = 0.0f, b = INFINITY;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float d = fabsf(R[i + (int64_t)i * ldr]);
        a = fmaxf(a, d);
        b = fminf(b, d);
    }
    mx[threadIdx.x] = a;
    mn[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            mx[threadIdx.x] = fmaxf(mx[threadIdx.x], mx[threadIdx.x + w]);
            mn[threadIdx.x] = fminf(mn[threadIdx.x], mn[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (mn[0] > 0.0f) ? (double)mx[0] / (double)mn[0] : INFINITY;
}

// Rs (m x k) = B - A X with FP64 sums.
__global__ void residual_kernel(int m, int n, int k, const float *A, int64_t lda, const float *X, int64_t ldx,
                                const float *B, int64_t ldb, float *Rs, int64_t ldrs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= m) return;
    double s = (double)B[i + (int64_t)j * ldb];
    const float *x = X + (int64_t)j * ldx;
    for (int c = 0; c < n; ++c) s -= (double)A[i + (int64_t)c * lda] * (double)x[c];
    Rs[i + (int64_t)j * ldrs] = (float)s;
}

__global__ void axpy_kernel(int n, int k, const float *D, int64_t ldd, float *X, int64_t ldx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) X[i + (int64_t)j * ldx] += D[i + (int64_t)j * ldd];
}

__global__ void copy2d_kernel(int m, int n, const float *X, int64_t ldx, float *Y, int64_t ldy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m) Y[i + (int64_t)j * ldy] = X[i + (int64_t)j * ldx];
}

// Orthogonality sketch eta = ||Q^T (Q g) - g|| / ||g|| for a fixed
// pseudo-random g in [-1, 1)^n, FP64 sums: two matrix-vector products.
__device__ __forceinline__ double sketch_g(int j)
{
    const uint32_t h = (uint32_t)j * 2654435761u + 0x9E3779B9u;
    return (double)((h ^ (h >> 15)) >> 8) * 0x1p-23 - 1.0;
}

__global__ void sketch_qg_kernel(int m, int n, const float *Q, int64_t ldq, double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += (double)Q[i + (int64_t)j * ldq] * sketch_g(j);
    y[i] = s;
}

__global__ __launch_bounds__(256) void sketch_qty_kernel(int m, int n, const float *Q, int64_t ldq, const double *y,
                                                         double *d2)
{
    __shared__ double red[256];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) s += (double)Q[i + (int64_t)j * ldq] * y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double e = red[0] - sketch_g(j);
        d2[j] = e * e;
    }
}

__global__ __launch_bounds__(256) void sketch_sum_kernel(int n, const double *d2, double *eta)
{
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        a += d2[j];
        const double g = sketch_g(j);
        b += g * g;
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *eta = sqrt(red[0][0] / red[1][0]);
}

enum State { WAITING = 0, RUNNING_QR = 1, RUNNING_SOLVE = 2, DONE = 3, FAILED = 4, RUNNING_QR2 = 5,
             CANCELLED = 6 };

struct Req {
    uint64_t id;
    int m, n, k;
    const float *srcA, *srcB;
    int64_t lda, ldb;
    int64_t off = -1, bytes = 0;
    float *A0 = nullptr, *A = nullptr, *R = nullptr, *B = nullptr, *X = nullptr, *D = nullptr, *Rs = nullptr;
    float *work = nullptr;
    double *sk = nullptr;         // sketch scratch: m + n doubles
    int64_t lwork = 0;
    int slot = -1;
    bool robust = false;          // refactored with the robust form
    double eta = -1.0;
    hipEvent_t ev = nullptr;
    double *cond_host = nullptr;  // the slot's pinned word, valid while the request holds the slot
    double cond = -1.0;
    int iters = 0;
    int state = WAITING;
    int status = LATER_OK;
    bool cancel = false;          // cancellation requested while running
    double t_submit = 0, t_admit = 0, t_done = 0;
};

struct Server {
    std::mutex mu;
    char *pool = nullptr;
    int64_t pool_bytes = 0;
    std::map<int64_t, int64_t> free_blocks;  // offset -> size
    std::vector<hipStream_t> streams;
    std::vector<Req *> slot_req;
    std::deque<Req *> waiting;
    std::unordered_map<uint64_t, Req *> reqs;
    double *cond_pinned = nullptr;           // two FP64 words per slot: cond, eta
    uint64_t next_id = 1;
    later_opts qr_opts;    // first attempt
    later_opts qr_robust;  // after an orthogonality sketch above the threshold
    double eta_max = 1e-5;
    int64_t peak_used = 0, used = 0;
    int completed = 0, rejected = 0;
};

int64_t need_bytes(int m, int n, int k, const later_opts &o, int64_t *lwork)
{
    int64_t lw = 0;
    if (later_ls_workspace(m, n, k, &o, &lw)) return -1;
    int64_t wfl = std::max<int64_t>(lw, ((int64_t)k + 63) / 64 * 64 + 64 + gemm_proj_work_floats(m, n, k, 256));
    *lwork = wfl;
    const int64_t f = 4;
    return align256(f * m * n) * 2 + align256(f * n * n) + align256(f * m * k) * 2 + align256(f * n * k) * 2 +
           align256(f * wfl) + align256(8LL * (m + n));
}

int64_t pool_alloc(Server &S, int64_t bytes)
{
    for (auto it = S.free_blocks.begin(); it != S.free_blocks.end(); ++it) {
        if (it->second >= bytes) {
            const int64_t off = it->first, sz = it->second;
            S.free_blocks.erase(it);
            if (sz > bytes) S.free_blocks[off + bytes] = sz - bytes;
            S.used += bytes;
            S.peak_used = std::max(S.peak_used, S.used);
            return off;
        }
    }
    return -1;
}

void pool_free(Server &S, int64_t off, int64_t bytes)
{
    S.used -= bytes;
    auto it = S.free_blocks.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != S.free_blocks.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        S.free_blocks.erase(nx);
    }
    if (it != S.free_blocks.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            S.free_blocks.erase(it);
        }
    }
}

dim3 g2(int rows, int cols) { return dim3((unsigned)ceil_div64(rows, 256), (unsigned)cols); }

// QR of r->A with `o`, then the condition estimate and the orthogonality
// sketch into the slot's pinned words.
int launch_factor(Server &S, Req *r, const later_opts &o)
{
    hipStream_t s = S.streams[r->slot];
    laterCtxt c{(void *)s, nullptr};
    int rc = later_rgsqrf_ex(c, r->m, r->n, r->A, r->m, r->R, r->n, r->work, r->lwork, nullptr, 0, &o, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(diag_ratio_kernel, dim3(1), dim3(256), 0, s, r->n, r->R, (int64_t)r->n, r->cond_host);
    hipLaunchKernelGGL(sketch_qg_kernel, dim3((unsigned)ceil_div64(r->m, 256)), dim3(256), 0, s, r->m, r->n, r->A,
                       (int64_t)r->m, r->sk);
    hipLaunchKernelGGL(sketch_qty_kernel, dim3(r->n), dim3(256), 0, s, r->m, r->n, r->A, (int64_t)r->m, r->sk,
                       r->sk + r->m);
    hipLaunchKernelGGL(sketch_sum_kernel, dim3(1), dim3(256), 0, s, r->n, r->sk + r->m, r->cond_host + 1);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(r->ev, s));
    return LATER_OK;
}

// Refactor from the kept copy of A with the robust form.
int launch_refactor(Server &S, Req *r)
{
    hipStream_t s = S.streams[r->slot];
    hipLaunchKernelGGL(copy2d_kernel, g2(r->m, r->n), dim3(256), 0, s, r->m, r->n, r->A0, (int64_t)r->m, r->A,
                       (int64_t)r->m);
    r->robust = true;
    return launch_factor(S, r, S.qr_robust);
}

// Enqueue copy + QR + condition estimate for an admitted request.
int launch_qr(Server &S, Req *r)
{
    hipStream_t s = S.streams[r->slot];
    hipLaunchKernelGGL(copy2d_kernel, g2(r->m, r->n), dim3(256), 0, s, r->m, r->n, r->srcA, r->lda, r->A0,
                       (int64_t)r->m);
    hipLaunchKernelGGL(copy2d_kernel, g2(r->m, r->n), dim3(256), 0, s, r->m, r->n, r->srcA, r->lda, r->A,
                       (int64_t)r->m);
    hipLaunchKernelGGL(copy2d_kernel, g2(r->m, r->k), dim3(256), 0, s, r->m, r->k, r->srcB, r->ldb, r->B,
                       (int64_t)r->m);
    return launch_factor(S, r, S.qr_opts);
}

// X = R^{-1} Q^T B' for B' (m x k) already in `src`.
int qtb_solve(hipStream_t s, Req *r, const float *src, float *dst)
{
    int *bexp = (int *)r->work + 64;
    int rc = gemm_colexp(s, r->m, r->k, src, r->m, bexp, nullptr);
    if (rc) return rc;
    const int64_t off = 64 + ((int64_t)r->k + 63) / 64 * 64;
    rc = gemm_proj(s, GEMM_F16X3, r->m, r->n, r->k, r->A, r->m, src, r->m, dst, r->n, 256, bexp, 14,
                   r->work + off, r->lwork - off, nullptr);
    if (rc) return rc;
    return trsm_upper(s, r->n, r->k, r->R, r->n, dst, r->n, (int *)r->work);
}

int launch_solve(Server &S, Req *r)
{
    hipStream_t s = S.streams[r->slot];
    const double cond = *r->cond_host;
    r->cond = cond;
    r->eta = r->cond_host[1];
    r->iters = (cond < 1e2) ? 0 : (cond < 1e4) ? 1 : (cond < 1e6) ? 2 : 3;
    int rc = qtb_solve(s, r, r->B, r->X);
    for (int it = 0; !rc && it < r->iters; ++it) {
        hipLaunchKernelGGL(residual_kernel, g2(r->m, r->k), dim3(256), 0, s, r->m, r->n, r->k, r->A0,
                           (int64_t)r->m, r->X, (int64_t)r->n, r->B, (int64_t)r->m, r->Rs, (int64_t)r->m);
        rc = qtb_solve(s, r, r->Rs, r->D);
        if (!rc)
            hipLaunchKernelGGL(axpy_kernel, g2(r->n, r->k), dim3(256), 0, s, r->n, r->k, r->D, (int64_t)r->n, r->X,
                               (int64_t)r->n);
    }
    if (rc) return rc;
    HIP_OK(hipEventRecord(r->ev, s));
    return LATER_OK;
}

void retire(Server &S, Req *r, int state, int status)
{
    r->state = state;
    r->status = status;
    r->t_done = now_s();
    if (r->slot >= 0) S.slot_req[r->slot] = nullptr;
    r->slot = -1;
    if (state == DONE) ++S.completed;
}

// Admit waiting requests in arrival order while a slot and memory are free.
void admit(Server &S)
{
    while (!S.waiting.empty()) {
        int slot = -1;
        for (size_t i = 0; i < S.slot_req.size(); ++i)
            if (!S.slot_req[i]) { slot = (int)i; break; }
        if (slot < 0) return;
        Req *r = S.waiting.front();
        const int64_t off = pool_alloc(S, r->bytes);
        if (off < 0) return;  // wait for memory (head of line keeps its place)
        S.waiting.pop_front();
        r->off = off;
        char *p = S.pool + off;
        auto take = [&](int64_t bytes) { float *q = (float *)p; p += align256(bytes); return q; };
        r->A0 = take(4LL * r->m * r->n);
        r->A = take(4LL * r->m * r->n);
        r->R = take(4LL * r->n * r->n);
        r->B = take(4LL * r->m * r->k);
        r->Rs = take(4LL * r->m * r->k);
        r->X = take(4LL * r->n * r->k);
        r->D = take(4LL * r->n * r->k);
        r->work = take(4LL * r->lwork);
        r->sk = (double *)take(8LL * (r->m + r->n));
        r->slot = slot;
        r->cond_host = S.cond_pinned + 2 * slot;
        S.slot_req[slot] = r;
        r->t_admit = now_s();
        r->state = RUNNING_QR;
        int rc = launch_qr(S, r);
        if (rc) retire(S, r, FAILED, rc);
    }
}

void progress(Server &S)
{
    for (size_t i = 0; i < S.slot_req.size(); ++i) {
        Req *r = S.slot_req[i];
        if (!r) continue;
        hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) { retire(S, r, FAILED, LATER_ERR_DEVICE); continue; }
        if (r->cancel) {
            // The enqueued stage has finished; nothing further is launched.
            retire(S, r, CANCELLED, LATER_ERR_CANCELLED);
        } else if (r->state == RUNNING_QR && !(r->cond_host[1] <= S.eta_max)) {
            // Q is not orthogonal enough for the solve: refactor robustly.
            r->state = RUNNING_QR2;
            int rc = launch_refactor(S, r);
            if (rc) retire(S, r, FAILED, rc);
        } else if (r->state == RUNNING_QR || r->state == RUNNING_QR2) {
            r->state = RUNNING_SOLVE;
            int rc = launch_solve(S, r);
            if (rc) retire(S, r, FAILED, rc);
        } else if (r->state == RUNNING_SOLVE) {
            retire(S, r, DONE, LATER_OK);
        }
    }
    admit(S);
}

} // namespace

LATER_API int later_server_create(int slots, int64_t hbm_budget_bytes, const later_opts *qr_opts, void **out)
{
    if (!out || slots < 1 || slots > 64 || hbm_budget_bytes < (1 << 20)) return LATER_ERR_SHAPE;
    Server *S = new Server();
    if (qr_opts) S->qr_opts = *qr_opts;
    else {
        later_opts_default(&S->qr_opts);
        S->qr_opts.gemm = LATER_GEMM_F16X3;
        S->qr_opts.panel = LATER_PANEL_MGS64;
        S->qr_opts.chunk = 256;
    }
    later_opts_mixed(&S->qr_robust);
    // Block CGS2 at every node re-projects with operands converted in the
    // kernel, so the second pass's projections use the three-piece FP16
    // split rather than staged int8 slices.
    S->qr_robust.gemm_pass2 = LATER_GEMM_F16X6;
    S->qr_robust.reorth = LATER_REORTH_NODE | LATER_REORTH_PASS;
    S->qr_robust.chunk = 256;  // ill-conditioned requests: FP64 folds in both passes
    S->pool_bytes = hbm_budget_bytes / 256 * 256;
    if (hipMalloc((void **)&S->pool, S->pool_bytes) != hipSuccess) { delete S; return LATER_ERR_DEVICE; }
    if (hipHostMalloc((void **)&S->cond_pinned, 2 * sizeof(double) * slots) != hipSuccess) {
        hipFree(S->pool);
        delete S;
        return LATER_ERR_DEVICE;
    }
    S->free_blocks[0] = S->pool_bytes;
    S->streams.resize(slots);
    S->slot_req.assign(slots, nullptr);
    for (int i = 0; i < slots; ++i) hipStreamCreateWithFlags(&S->streams[i], hipStreamNonBlocking);
    *out = S;
    return LATER_OK;
}

LATER_API int later_server_submit(void *srv, int m, int n, int k, const float *A, int lda, const float *B, int ldb,
                                  uint64_t *id)
{
    Server *S = (Server *)srv;
    if (!S || !id || m < 1 || n < 1 || k < 1 || m < n || lda < m || ldb < m || !A || !B) return LATER_ERR_SHAPE;
    std::lock_guard<std::mutex> lk(S->mu);
    Req *r = new Req();
    r->id = S->next_id++;
    r->m = m; r->n = n; r->k = k;
    r->srcA = A; r->lda = lda; r->srcB = B; r->ldb = ldb;
    r->t_submit = now_s();
    int64_t lw_fast = 0;
    const int64_t b_fast = need_bytes(m, n, k, S->qr_opts, &lw_fast);
    r->bytes = need_bytes(m, n, k, S->qr_robust, &r->lwork);
    if (b_fast < 0) r->bytes = -1;
    if (lw_fast > r->lwork) r->lwork = lw_fast;
    *id = r->id;
    S->reqs[r->id] = r;
    if (r->bytes < 0 || r->bytes > S->pool_bytes) {
        r->state = FAILED;
        r->status = LATER_ERR_BUDGET;
        r->t_done = r->t_submit;
        ++S->rejected;
        return LATER_ERR_BUDGET;
    }
    if (hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess) {
        // Recorded as failed (never admitted), so it can still be released.
        r->ev = nullptr;
        r->state = FAILED;
        r->status = LATER_ERR_DEVICE;
        r->t_done = r->t_submit;
        return LATER_ERR_DEVICE;
    }
    S->waiting.push_back(r);
    progress(*S);
    return LATER_OK;
}

// Advances every in-flight request and admits waiting ones; never blocks.
// *active receives the number of requests waiting or running.
LATER_API int later_server_poll(void *srv, int *active)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    std::lock_guard<std::mutex> lk(S->mu);
    progress(*S);
    if (active) {
        int a = (int)S->waiting.size();
        for (Req *r : S->slot_req) a += (r != nullptr);
        *active = a;
    }
    return LATER_OK;
}

// Result of a finished request: device pointers into the pool (valid until
// later_server_release), the state (3 done, 4 failed, else still in flight),
// its status, latency from submission to observed completion, queueing time,
// the condition estimate and the number of refinement rounds.
LATER_API int later_server_result(void *srv, uint64_t id, int *state, int *status, const float **X, const float **Q,
                                  const float **R, double *latency_s, double *queue_s, double *cond, int *iters)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return LATER_ERR_SHAPE;
    Req *r = it->second;
    if (state) *state = r->state;
    if (status) *status = r->status;
    const bool fin = r->state == DONE;
    if (X) *X = fin ? r->X : nullptr;
    if (Q) *Q = fin ? r->A : nullptr;
    if (R) *R = fin ? r->R : nullptr;
    const bool ended = (r->state == DONE || r->state == FAILED || r->state == CANCELLED);
    if (latency_s) *latency_s = ended ? r->t_done - r->t_submit : -1.0;
    if (queue_s) *queue_s = (r->t_admit > 0) ? r->t_admit - r->t_submit : -1.0;
    if (cond) *cond = r->cond;
    if (iters) *iters = r->iters;
    return LATER_OK;
}

// Copies X (n x k), R (n x n) and Q (m x n) of a finished request into caller
// buffers (any may be null); synchronous.
LATER_API int later_server_fetch(void *srv, uint64_t id, float *X, int ldx, float *R, int ldr, float *Q, int ldq)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    // Held through the copies: a concurrent later_server_release cannot free
    // the request (and its pool memory) under them.
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || it->second->state != DONE) return LATER_ERR_SHAPE;
    const Req *r = it->second;
    if (X && ldx < r->n) return LATER_ERR_SHAPE;
    if (R && ldr < r->n) return LATER_ERR_SHAPE;
    if (Q && ldq < r->m) return LATER_ERR_SHAPE;
    if (X) HIP_OK(hipMemcpy2D(X, 4 * (size_t)ldx, r->X, 4 * (size_t)r->n, 4 * (size_t)r->n, r->k, hipMemcpyDeviceToDevice));
    if (R) HIP_OK(hipMemcpy2D(R, 4 * (size_t)ldr, r->R, 4 * (size_t)r->n, 4 * (size_t)r->n, r->n, hipMemcpyDeviceToDevice));
    if (Q) HIP_OK(hipMemcpy2D(Q, 4 * (size_t)ldq, r->A, 4 * (size_t)r->m, 4 * (size_t)r->m, r->n, hipMemcpyDeviceToDevice));
    return LATER_OK;
}

// Cancels a request: a waiting one leaves the queue at once (state
// CANCELLED); a running one finishes the stage already on its stream, is not
// advanced further and becomes CANCELLED at the next poll.  Either way its
// memory returns to the pool at later_server_release.  A finished request
// cannot be cancelled (LATER_ERR_SHAPE).
LATER_API int later_server_cancel(void *srv, uint64_t id)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return LATER_ERR_SHAPE;
    Req *r = it->second;
    if (r->state == WAITING) {
        S->waiting.erase(std::find(S->waiting.begin(), S->waiting.end(), r));
        r->state = CANCELLED;
        r->status = LATER_ERR_CANCELLED;
        r->t_done = now_s();
        progress(*S);
        return LATER_OK;
    }
    if (r->state == RUNNING_QR || r->state == RUNNING_QR2 || r->state == RUNNING_SOLVE) {
        r->cancel = true;
        progress(*S);
        return LATER_OK;
    }
    return LATER_ERR_SHAPE;
}

LATER_API int later_server_release(void *srv, uint64_t id)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return LATER_ERR_SHAPE;
    Req *r = it->second;
    if (r->state != DONE && r->state != FAILED && r->state != CANCELLED) return LATER_ERR_SHAPE;
    if (r->off >= 0) pool_free(*S, r->off, r->bytes);
    if (r->ev) hipEventDestroy(r->ev);
    S->reqs.erase(it);
    delete r;
    progress(*S);
    return LATER_OK;
}

LATER_API int later_server_stats(void *srv, int64_t *pool_bytes, int64_t *used_bytes, int64_t *peak_bytes,
                                 int *completed, int *rejected)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    std::lock_guard<std::mutex> lk(S->mu);
    if (pool_bytes) *pool_bytes = S->pool_bytes;
    if (used_bytes) *used_bytes = S->used;
    if (peak_bytes) *peak_bytes = S->peak_used;
    if (completed) *completed = S->completed;
    if (rejected) *rejected = S->rejected;
    return LATER_OK;
}

// Waits for every in-flight request, then frees the pool and the streams.
LATER_API int later_server_destroy(void *srv)
{
    Server *S = (Server *)srv;
    if (!S) return LATER_ERR_SHAPE;
    for (;;) {
        int active = 0;
        later_server_poll(srv, &active);
        bool running = false;
        for (Req *r : S->slot_req) running |= (r != nullptr);
        if (!running) break;
    }
    for (hipStream_t s : S->streams) {
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
    }
    for (auto &kv : S->reqs) {
        if (kv.second->ev) hipEventDestroy(kv.second->ev);
        delete kv.second;
    }
    hipFree(S->pool);
    hipHostFree(S->cond_pinned);
    delete S;
    return LATER_OK;
}
