// Optional per-kernel-class timing with HIP events on the launch stream.
// Off unless later_prof_begin() was called; used by the benchmark to time the
// dominant kernel live (kernel-launch granularity, eager launches only).
#pragma once
#include "common.h"

// PROF_UPDATE_X3: every launch of the streaming FP16 hi/lo kernel on 256-row
// tiles (pre-split operands, no FP64 folds; the mixed form's largest): the
// node updates, the R2 R1 composition and the first pass's projections of
// enough tiles; PROF_PROJ_I8: the int8 four-slice projections (the mixed
// form's second pass); PROF_GEMM_UPDATE and PROF_GEMM_PROJ hold the other
// products.
enum { PROF_GEMM_PROJ = 0, PROF_GEMM_UPDATE = 1, PROF_PANEL = 2, PROF_OTHER = 3, PROF_UPDATE_X3 = 4,
       PROF_PROJ_I8 = 5, PROF_NCLASS = 6 };

// Returns a token for prof_stop, or -1 when profiling is off.
int prof_start(hipStream_t s);
void prof_stop(int token, hipStream_t s, int cls, double flops, double bytes);
