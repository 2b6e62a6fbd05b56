// Learned Gauss-order selection for near-field pairs (mixed path): device network, training
// kernels and the classification pass that feeds the assembly.
#pragma once
#include "andr_internal.h"
#include "andr_assemble.h"

// algorithmic flops of one pass of the 6 -> 64 -> 64 -> 64 -> 8 network: 2 x (6*64 + 2*64*64 + 64*8)
#define ANDR_MLP_FLOPS_PER_PASS 18176.0

struct LearnedStats {
  int64_t points;    // regular quadrature points evaluated
  int64_t learned;   // near-field pairs integrated with the rule chosen by the network
  int64_t fallback;  // near-field pairs sent to the fixed-order rule
};

// classifies every pair of the listed rows (writes SysJob::cls), then assembles with those rules;
// stats[njobs] (may be null) gets the counts of each job
int andr_launch_stokes_system_learned(andr_ctx *ctx, SysJob *h_jobs, const SysJob *d_jobs, const int2 *d_rowmap,
                                      int nrows, int njobs, bool vertex_sums, LearnedStats *stats);
