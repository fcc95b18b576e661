// Internal interface of the TSQR panel (panel.hip).
#pragma once
#include "common.h"

constexpr int kPanelBlockRows = 256;  // row block of the tree (QR/panel.cu:87-133)

// Floats of `work` the tree needs for an m x w panel.
int64_t panel_work_floats(int m, int w);

// QR of the m x w panel A (w <= 32) in place: A <- Q, R (w x w, ldr) upper.
// fp64 selects FP64 column norms and dot products.  *status gets bit 0 set
// on a zero or non-finite column norm.
int panel_tsqr(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work,
               int64_t lwork, bool fp64, int *status, int *launches);
