// Resistance-system assembly job description shared by the assembly and driver units.
#pragma once
#include "andr_internal.h"
#include "andr_solve.h"

// right-hand sides carried per body (6 rigid-body problems, padded to 8)
#define ANDR_NRHS 8

// One body of a batch: geometry, output system A (3ne x lda, blocked rows c ne + r and columns
// en ne + e, element type float or double) and right-hand sides B (3ne x ANDR_NRHS) of the six
// rigid-body problems.
struct SysJob {
  BodyGeom g;
  void *A_;
  int64_t lda;         // row-major leading dimension; 0: the six symmetric blocks (andr_sym_off)
  double *B;                // 3 ne x ANDR_NRHS right-hand sides (problems 0..5, the rest 0)
  const uint8_t *moving;    // per node: 1 if it carries the rigid-body velocity (null: all nodes)
  double center[3];         // rotation centre of problems 3..5
  int strain;               // 0: the six rigid-body problems; 1: columns 0..2 are the strain fields
                            // (0, z, y), (z, 0, x), (y, x, 0) about the centre (TorquePiCRS)
  double mu, extra;
  int qoff, nq;        // the fixed-order rule
  double *err_rows;    // ne (may be null)
  uint8_t *cls;        // ne x ne rule class per pair (learned form), null: fixed-order rule
  const double *defl;  // per wall id (1..nwalls): scale of the rank-one null-space term, null: none
  double *minv;        // ne x 9: inverse 3x3 self blocks (right block-Jacobi preconditioner), null: none
  double *X;           // 3 ne x ANDR_NRHS solution (for andr_launch_unprecondition)
#ifdef __HIPCC__
  template <class ST>
  __device__ ST *Aptr() const { return (ST *)A_; }
#endif
};

// vertex_sums (FP32 quadrature and storage only): rigid-body problems on bodies whose elements each
// move as a whole (all three nodes moving or all at rest) take the vertex-summed stresslet moments
int andr_launch_stokes_system(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows, bool fp32_quad,
                              bool fp32_store, double alg_bytes, double alg_flops, bool vertex_sums = false);
// Right block-Jacobi preconditioner of the iterative forms: minv[e] = inverse of the 3x3 self block
// (singular Stokeslet block of element e at its own centroid, scaled as the system, plus the
// null-space term). The assembly then stores A M^-1 and the solve returns y with x = M^-1 y.
// In the symmetric-block layout (lda 0) it also writes the blocks' diagonal entries, which the fast
// assembly (FP32 quadrature and storage) leaves to it: launch it before andr_launch_stokes_system.
int andr_launch_self_blocks(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows);
int andr_launch_unprecondition(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows);
// Force and torque of the six problems on wall 1 of every job, and its largest row self-check error:
// d_out[37 j + 6 k + (0..5)] = (F, T) of problem k, d_out[37 j + 36] = error.
int andr_launch_force_torque(andr_ctx *ctx, const SysJob *d_jobs, int njobs, double *d_out);
// Velocity, pressure and closure error at npts domain points (device arrays pts[3 npts], u[3 nn],
// q[3 ne]; out[5 npts]); triangle meshes.
int andr_launch_stokes_domain(andr_ctx *ctx, DevBody &b, int npts, const double *pts, const double *u, const double *q,
                              double diff, double *out);
// Surface pressure at the element centres (device arrays u[3 nn], q[3 ne], p[ne]); triangle meshes.
int andr_launch_stokes_surface_pressure(andr_ctx *ctx, DevBody &b, const double *u, const double *q, double diff,
                                        double *p);
// algorithmic work of a pair evaluation: Stokes G + T kernels 107 flops (per-vertex stresslet
// moments), 60 flops with the vertex-summed moments of the fast form's rigid-body problems
// (stokes_pair_sum); Laplace G + H 50 flops
#define ANDR_FLOPS_PER_STOKES_EVAL 107.0
#define ANDR_FLOPS_PER_STOKES_EVAL_SUM 60.0
#define ANDR_FLOPS_PER_LAPLACE_EVAL 50.0
// Dunavant rules (classes 0..6 of the learned form) in the quadrature table
#define ANDR_DUN_OFF0 297
