// Device-resident dense solvers: multi-RHS LSQR (the plain FP64 path, restating src/lib_lsqr.f)
// and restarted GMRES with FP32 matrix storage (the mixed path). Vectors are stored interleaved,
// [row][ANDR_NB] with ANDR_NB = 8 right-hand sides per block.
#pragma once
#include "andr_internal.h"

#define ANDR_NB 8

// LSQR scalar state of one right-hand side (the locals of lsqr_LSQR, src/lib_lsqr.f:470-1207)
struct LsqrState {
  double alpha, beta, rhobar, phibar, bnorm, rnorm, anorm, dnorm, xnorm, xnorm1, res2, cs2, sn2, z;
  double atol, btol, ctol;
  double t1, t2;       // update coefficients of the current iteration
  double sx, sy;       // scale factors handed to the next product
  int itn, istop, itnlim, active;
};

// One dense system of a batch: A is m x n (row-major, leading dimension lda) of type ST.
struct DenseSys {
  const void *A;
  int64_t lda;
  int m, n;
};

int andr_colscale(andr_ctx *ctx, void *A, bool fp32, int64_t lda, int m, int n, double *q);

// LSQR on one dense FP64/FP32 system with up to ANDR_NB right-hand sides.
// B: m x ANDR_NB (device), X: n x ANDR_NB (device) hot start on entry, solution on exit;
// q: column scale (device, may be null for none). nrhs active columns; maxit per column
// (0 -> 10(m+n+50)); nit/istop (host, nrhs entries) receive the results.
int andr_lsqr_device(andr_ctx *ctx, const void *A, bool fp32, int64_t lda, int m, int n, const double *q,
                     const double *B, double *X, int nrhs, int maxit, double eps, int *nit, int *istop);

// Restarted GMRES on a batch of square systems (FP32 or FP64 storage, FP64 vectors); restart
// cycles recompute the FP64 residual b - A x. Stops per system when every active column reaches
// ||r|| <= tol ||b||. X holds the result (zero start). iters[s*nrhs + k] gets the iteration count.
// Fast-path system matrix layout: FP32 in 16 x 16 tiles, tile (rt, kb) holding rows 16 rt..16 rt+15
// and columns 16 kb..16 kb+15 row-major, tiles of a row band contiguous, KB = ceil(n / 16) tiles
// per band and ceil(n / 64) * 4 bands (the product reads 64-row bands). One wave instruction of the
// matrix-core product then reads one whole 1 KB tile.
__host__ __device__ inline size_t andr_tiled_off(int row, int col, int KB) {
  return (((size_t)(row >> 4) * KB + (col >> 4)) * 16 + (row & 15)) * 16 + (col & 15);
}
__host__ __device__ inline size_t andr_tiled_elems(int n) {
  return (size_t)((n + 63) / 64 * 64) * (size_t)((n + 15) / 16 * 16);
}

struct GmresSys {
  const void *A;    // FP32: tiled layout (andr_tiled_off); FP64: row-major with leading dimension lda
  int64_t lda;
  int n;
  const double *B;  // n x ANDR_NB
  double *X;        // n x ANDR_NB
};
int andr_gmres_batch(andr_ctx *ctx, const GmresSys *sys, int nsys, bool fp32, int nrhs, int restart, int maxit,
                     double tol, int *iters);
