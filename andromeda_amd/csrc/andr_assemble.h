// Resistance-system assembly job description shared by the assembly and driver units.
#pragma once
#include "andr_internal.h"

// right-hand sides carried per body (6 rigid-body problems, padded to 8)
#define ANDR_NRHS 8

// One body of a batch: geometry, output system A (3ne x lda, blocked rows c ne + r and columns
// en ne + e, element type float or double) and right-hand sides B (3ne x ANDR_NRHS) of the six
// rigid-body problems.
struct SysJob {
  BodyGeom g;
  void *A_;
  int64_t lda;
  double *B;                // 3 ne x ANDR_NRHS right-hand sides (problems 0..5, the rest 0)
  const uint8_t *moving;    // per node: 1 if it carries the rigid-body velocity (null: all nodes)
  double center[3];         // rotation centre of problems 3..5
  double mu, extra;
  int qoff, nq;        // the fixed-order rule
  double *err_rows;    // ne (may be null)
  uint8_t *cls;        // ne x ne rule class per pair (learned form), null: fixed-order rule
  const double *defl;  // per wall id (1..nwalls): scale of the rank-one null-space term, null: none
#ifdef __HIPCC__
  template <class ST>
  __device__ ST *Aptr() const { return (ST *)A_; }
#endif
};

int andr_launch_stokes_system(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows, bool fp32_quad,
                              bool fp32_store, double alg_bytes, double alg_flops);
// algorithmic work of a pair evaluation: Stokes G + T kernels 107 flops, Laplace G + H 50 flops
#define ANDR_FLOPS_PER_STOKES_EVAL 107.0
#define ANDR_FLOPS_PER_LAPLACE_EVAL 50.0
// Dunavant rules (classes 0..6 of the learned form) in the quadrature table
#define ANDR_DUN_OFF0 297
