// Graph-captured factorisation plans.
//
// A plan records every kernel that later_rgsqrf_ex enqueues for one fixed set
// of buffers (A, R, work) and options into a HIP graph, once; executing the
// plan replays the whole recursion (thousands of dependent launches at large
// n) with a single graph launch, removing the per-launch host cost.  The
// recorded path clears memory only with fill kernels (no memset nodes).
#include "common.h"

namespace {
struct Plan {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int kernels = 0;
};
} // namespace

LATER_API int later_plan_create(int m, int n, float *A, int lda, float *R, int ldr, float *work, int64_t lwork,
                                uint16_t *hwork, int64_t lhwork, const later_opts *opts, void **plan)
{
    if (!plan) return LATER_ERR_SHAPE;
    *plan = nullptr;
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    Plan *p = new Plan();
    int rc = LATER_OK;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = LATER_ERR_DEVICE;
    later_info info;
    if (!rc) {
        laterCtxt c{(void *)s, nullptr};
        rc = later_rgsqrf_ex(c, m, n, A, lda, R, ldr, work, lwork, hwork, lhwork, opts, &info);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (!rc && (e != hipSuccess || !g)) rc = LATER_ERR_DEVICE;
        p->graph = g;
    }
    if (!rc && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) rc = LATER_ERR_DEVICE;
    hipStreamDestroy(s);
    if (rc) {
        if (p->exec) hipGraphExecDestroy(p->exec);
        if (p->graph) hipGraphDestroy(p->graph);
        delete p;
        return rc;
    }
    p->kernels = info.kernels;
    *plan = p;
    return LATER_OK;
}

LATER_API int later_plan_execute(void *plan, void *stream)
{
    Plan *p = (Plan *)plan;
    if (!p || !p->exec) return LATER_ERR_SHAPE;
    HIP_OK(hipGraphLaunch(p->exec, (hipStream_t)stream));
    return LATER_OK;
}

LATER_API int later_plan_kernels(void *plan)
{
    Plan *p = (Plan *)plan;
    return p ? p->kernels : -1;
}

LATER_API int later_plan_destroy(void *plan)
{
    Plan *p = (Plan *)plan;
    if (!p) return LATER_ERR_SHAPE;
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    delete p;
    return LATER_OK;
}
