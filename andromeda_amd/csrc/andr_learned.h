// Learned Gauss-order selection for near-field pairs (mixed path): device network, training
// kernels and the assembly variant that consults it.
#pragma once
#include "andr_internal.h"
#include "andr_assemble.h"

struct LearnedStats {
  int64_t points;    // regular quadrature points evaluated
  int64_t learned;   // near-field pairs integrated with the rule chosen by the network
  int64_t fallback;  // near-field pairs sent to the fixed-order rule
};

int andr_launch_stokes_system_learned(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows,
                                      LearnedStats *stats);
