// Optional per-kernel-class timing with HIP events on the launch stream.
// Off unless later_prof_begin() was called; used by the benchmark to time the
// dominant kernel live (kernel-launch granularity, eager launches only).
#pragma once
#include "common.h"

enum { PROF_GEMM_PROJ = 0, PROF_GEMM_UPDATE = 1, PROF_PANEL = 2, PROF_OTHER = 3, PROF_NCLASS = 4 };

// Returns a token for prof_stop, or -1 when profiling is off.
int prof_start(hipStream_t s);
void prof_stop(int token, hipStream_t s, int cls, double flops, double bytes);
