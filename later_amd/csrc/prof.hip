// Kernel-class timing (see prof.h).  Host-side state, guarded by a mutex;
// only active between later_prof_begin and later_prof_end.
#include <mutex>
#include <vector>

#include "prof.h"

namespace {
struct Rec {
    hipEvent_t a, b;
    int cls;
    double flops, bytes;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
} // namespace

int prof_start(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return -1;
    Rec r{};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
    hipEventRecord(r.a, s);
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

void prof_stop(int token, hipStream_t s, int cls, double flops, double bytes)
{
    if (token < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    Rec &r = g_recs[token];
    hipEventRecord(r.b, s);
    r.cls = cls;
    r.flops = flops;
    r.bytes = bytes;
}

LATER_API void later_prof_begin(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_recs.clear();
    g_on = true;
}

// Synchronises the recorded events and fills, per class, launches, total ms,
// algorithmic flops and bytes.  Arrays are indexed by LATER_PROF_* class.
LATER_API int later_prof_end(int *launches, double *ms, double *flops, double *bytes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = false;
    for (int c = 0; c < PROF_NCLASS; ++c) {
        launches[c] = 0;
        ms[c] = flops[c] = bytes[c] = 0.0;
    }
    int rc = LATER_OK;
    for (Rec &r : g_recs) {
        float t = 0.0f;
        if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&t, r.a, r.b) != hipSuccess)
            rc = LATER_ERR_DEVICE;
        launches[r.cls] += 1;
        ms[r.cls] += t;
        flops[r.cls] += r.flops;
        bytes[r.cls] += r.bytes;
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    g_recs.clear();
    return rc;
}
