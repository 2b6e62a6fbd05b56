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
// The same on a sparse FP64 system: A (m x n) and its transpose in CSR with 0-based row pointers and
// columns, all on the device; each product sums a row's entries in stored order.
int andr_lsqr_device_csr(andr_ctx *ctx, int m, int n, int64_t nnz, const int64_t *ia, const int32_t *ja,
                         const double *a, const int64_t *tia, const int32_t *tja, const double *ta, const double *q,
                         const double *B, double *X, int nrhs, int maxit, double eps, int *nit, int *istop);
// Several independent LSQR solves (as andr_lsqr_device each), advanced together on up to
// LSQR_STREAMS streams so that the small ones overlap on the device. Waits for the context's stream
// first; returns when every solve has finished.
struct LsqrJob {
  const void *A;
  bool fp32;
  int64_t lda;
  int m, n;
  const double *q, *B;
  double *X;
  int nrhs, maxit;
  double eps;
  int *nit, *istop;
};
#define LSQR_STREAMS 4
// BiCGSTAB with a Jacobi preconditioner on a square FP64 system (device arrays, A row-major with
// leading dimension lda, B and X n x ANDR_NB interleaved, X the initial guess and the result):
// per column nit = iterations, status 0 converged (||r|| <= tol ||b||), 1 iteration limit,
// 2 breakdown. maxit <= 0: 1000.
// ws: null -> the workspace comes from the context's memory cache; else a device region of ws_bytes
// (at least andr_bicgstab_ws_bytes(n), 256-byte aligned) that the workspace is carved from, with no
// allocation, no memset and no profiling records in the call.
int andr_bicgstab_device(andr_ctx *ctx, const double *A, int64_t lda, int n, const double *B, double *X, int nrhs,
                         int maxit, double tol, int *nit, int *status, double *ws = nullptr, size_t ws_bytes = 0);
size_t andr_bicgstab_ws_bytes(int n);
int andr_lsqr_device_batch(andr_ctx *ctx, int count, const LsqrJob *jobs);

// Restarted GMRES on a batch of square systems (FP32 or FP64 storage, FP64 vectors); restart
// cycles recompute the FP64 residual b - A x. Stops per system when every active column reaches
// ||r|| <= tol ||b||. X holds the result (zero start). iters[s*nrhs + k] gets the iteration count.
// Fast-path system matrix: the six distinct Stokeslet blocks S_q(r, e) = -(mul_e / mu) G_q(r, e),
// q = {xx, xy, xz, yy, yz, zz}, of the centroid-collocation system (the full matrix has 3 x 3 blocks
// with G_{c,en} = G_{en,c}). FP32 in 16 x 16 tiles: tile (band, kb) of block q holds element rows
// 16 band.. and columns 16 kb.. row-major; for one (band, kb) the six q tiles are contiguous, and
// the bands of a 32-row pair follow each other. KBe = ceil(ne / 16); ceil(ne / 32) * 2 bands.
__host__ __device__ inline size_t andr_sym_off(int q, int r, int e, int KBe) {
  return ((((size_t)(r >> 4) * KBe + (e >> 4)) * 6 + q) * 16 + (r & 15)) * 16 + (e & 15);
}
__host__ __device__ inline size_t andr_sym_elems(int ne) {
  return (size_t)((ne + 31) / 32 * 32) * (size_t)((ne + 15) / 16 * 16) * 6;
}

// The structured operator of the fast resistance path, in GMRES's variable y (x = M^-1 y):
//   A M^-1 y = S z + sum_w d_w (A_r n_r) (sum_{e in w} n_e . z_e),   z = M^-1 y,
// with S the symmetric Stokeslet blocks above, M the 3 x 3 self blocks (block-Jacobi) and the
// rank-one null-space term of every closed wall w.
struct StokesOp {
  const float *S;            // andr_sym_off layout
  int ne;
  const double *minv;        // ne x 9
  const double *area, *nrm;  // ne, 3 ne
  const int *bcid;           // ne, wall ids 1..nwalls
  const double *defl;        // per wall id: d_w (0: no term); walls above ANDR_OP_MAXW carry none
};
#define ANDR_OP_MAXW 4

struct GmresSys {
  const void *A;    // unused (the operator is op)
  int64_t lda;      // unused
  int n;
  StokesOp op;      // the structured Stokes operator, n = 3 op.ne
  const double *B;  // n x ANDR_NB
  double *X;        // n x ANDR_NB
};
// Batched restarted GMRES (right-preconditioned, FP64 vectors) on structured Stokes systems; fp32
// must be true (the FP32 symmetric-block operator). iters: per system and rhs, the Krylov steps;
// status (may be null): per system and rhs, 0 converged (||r|| <= tol ||b||), ANDR_GMRES_FLOOR
// stagnated at a relative residual <= 1e-6 above tol (the FP32 floor of the near-null direction;
// accepted), ANDR_GMRES_MAXIT stopped at the iteration limit. The codes follow LSQR's istop
// (4: ill-conditioned, 5: iteration limit), so a resistance report's status is > 3 for both.
#define ANDR_GMRES_FLOOR 4
#define ANDR_GMRES_MAXIT 5
int andr_gmres_batch(andr_ctx *ctx, const GmresSys *sys, int nsys, bool fp32, int nrhs, int restart, int maxit,
                     double tol, int *iters, int *status);
