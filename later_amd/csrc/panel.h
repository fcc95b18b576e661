// Internal interface of the TSQR panel (panel.hip).
#pragma once
#include "common.h"

constexpr int kPanelBlockRows = 256;  // row block of the tree (QR/panel.cu:87-133)

// One level of the tree for one panel (device-side argument block of the
// grouped launches): blocks of kPanelBlockRows rows of the m x w panel A
// (the last block takes the remainder); the MGS writes each block's R into W
// (rows b*w, ld ldw) and the apply multiplies block b by those rows of W.
struct PanelArgs {
    int m, w, nblk;
    float *A;
    int64_t lda;
    float *W;
    int64_t ldw;
    int *status;
};

// Number of row blocks of an m x w panel level (1: a single block).
int panel_blocks(int m, int w);

// Grouped launches: task b = (args index, block).
int mgs_grouped(hipStream_t s, bool fp64, const PanelArgs *args, const int2 *tasks, int ntasks);
int apply_grouped(hipStream_t s, bool fp64, const PanelArgs *args, const int2 *tasks, int ntasks);

// Floats of `work` the tree needs for an m x w panel, and the largest over
// every width w <= 32.
int64_t panel_work_floats(int m, int w);
int64_t panel_work_floats_any(int m);

// QR of the m x w panel A (w <= 32) in place: A <- Q, R (w x w, ldr) upper.
// fp64 selects FP64 column norms and dot products.  *status gets bit 0 set
// on a zero or non-finite column norm.
int panel_tsqr(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work,
               int64_t lwork, bool fp64, int *status, int *launches);

// Cholesky QR twice for an m x w panel (w <= 64) in three launches (once:
// one pass, two launches, for panels with condition number near 1); *status
// gets bit 0 set on a numerically rank-deficient or non-finite Gram pivot.
// The workspace depends on whether w is above 32.
int64_t panel_chol_work_floats(int w);
int panel_cholqr2(hipStream_t s, int m, int w, float *A, int lda, float *R, int ldr, float *work, int64_t lwork,
                  int *status, int *launches, bool once = false);
