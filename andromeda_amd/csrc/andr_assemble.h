// Resistance-system assembly job description shared by the assembly and driver units.
#pragma once
#include "andr_internal.h"

// right-hand sides carried per body (6 rigid-body problems, padded to 8)
#define ANDR_NRHS 8

template <class ST>
struct SysJobT {
  BodyGeom g;
  ST *A;                  // 3ne x lda, blocked rows (c ne + r) and columns (en ne + e)
  int64_t lda;
  double *B;              // 3ne x ANDR_NRHS
  const double *U;        // nn x 3 x ANDR_NRHS prescribed velocities
  double mu, extra;
  int qoff, nq;
  double *err_rows;       // ne (may be null)
};

struct SysJob {
  BodyGeom g;
  void *A_;
  int64_t lda;
  double *B;
  const double *U;
  double mu, extra;
  int qoff, nq;
  double *err_rows;
#ifdef __HIPCC__
  template <class ST>
  __device__ ST *Aptr() const { return (ST *)A_; }
#endif
};

int andr_launch_stokes_system(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows, bool fp32_quad,
                              bool fp32_store, double alg_bytes, double alg_flops);
// algorithmic work of a Stokes pair evaluation: G + T kernels, 107 flops per quadrature point
#define ANDR_FLOPS_PER_STOKES_EVAL 107.0
#define ANDR_FLOPS_PER_LAPLACE_EVAL 50.0
