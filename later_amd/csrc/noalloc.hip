// later_rgsqrf_noalloc: the status form of the QR with its FP32 workspace in
// a file-scope device array instead of a caller buffer, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).
//
// The array `scratch` is LATER_NOALLOC_SCRATCH_FLOATS floats; every GPU the
// library runs this entry point on holds its own copy until the process
// ends.  Host code never passes the array's name to a kernel: the first call
// on a device (with that device current) asks the runtime for the copy's
// device address with hipGetSymbolAddress, keeps it per device, and every
// later call hands that address to later_rgsqrf_ex as `work`.  The kernels
// clear what must start at zero (the status header, R's lower triangle, the
// second pass's R2), so the previous call's contents do not matter.  Calls
// that overlap on one GPU share the array and race.
#include <mutex>

#include "common.h"

namespace {

constexpr int kMaxDevices = 64;

__device__ __align__(256) float scratch[LATER_NOALLOC_SCRATCH_FLOATS];

std::mutex g_mu;
float *g_ptr[kMaxDevices];  // device address of `scratch` per device (nullptr: not set up)

int scratch_ptr(float **out)
{
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return LATER_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LATER_ERR_DEVICE;
        g_ptr[dev] = (float *)p;
    }
    *out = g_ptr[dev];
    return LATER_OK;
}

} // namespace

LATER_API int later_rgsqrf_noalloc(laterCtxt ctxt, int m, int n, float *A, int lda, float *R, int ldr,
                                   const later_opts *opts, later_info *info)
{
    if (info) info->status = LATER_OK;
    int64_t lwork = 0;
    int rc = later_rgsqrf_workspace(m, n, opts, &lwork, nullptr);
    if (!rc && lwork > (int64_t)LATER_NOALLOC_SCRATCH_FLOATS) rc = LATER_ERR_WORKSPACE;
    float *work = nullptr;
    if (!rc) rc = scratch_ptr(&work);
    if (rc) {
        if (info) info->status = rc;
        return rc;
    }
    // No 16-bit arena: the node GEMMs convert their operands in the kernel
    // (the same result bits as staged pieces).
    return later_rgsqrf_ex(ctxt, m, n, A, lda, R, ldr, work, lwork, nullptr, 0, opts, info);
}
