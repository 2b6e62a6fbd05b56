// Regular (non-singular) pair integrals shared by the assembly kernels: the Stokeslet and
// stresslet moments of one (collocation point, element) pair with one quadrature rule.
#pragma once
#include "andr_internal.h"

// quadrature table of this translation unit: l1, l2, l3, w per point (see andr_rule)
static __constant__ double c_qd[ANDR_QTAB_N][4];
static __constant__ float c_qf[ANDR_QTAB_N][4];

static inline int andr_upload_rules_here(andr_ctx *ctx, const double *tab, const float *tabf) {
  ANDR_HIP(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_qd), tab, sizeof(double) * 4 * ANDR_QTAB_N));
  ANDR_HIP(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_qf), tabf, sizeof(float) * 4 * ANDR_QTAB_N));
  return ANDR_OK;
}


// Gs: Stokeslet block (FP64); Tv: stresslet moments per vertex, in the quadrature type CT
template <class CT>
__device__ inline void stokes_pair_regular(const double s[3], const double *x1, const double *x2, const double *x3,
                                           double nx, double ny, double nz, double area, int qoff, int nq,
                                           double Gs[6], CT Tv[6][3]) {
  // relative geometry formed in FP64, quadrature in CT
  const CT d0 = (CT)(x3[0] - s[0]), d1 = (CT)(x3[1] - s[1]), d2 = (CT)(x3[2] - s[2]);
  const CT a0 = (CT)(x1[0] - x3[0]), a1 = (CT)(x1[1] - x3[1]), a2 = (CT)(x1[2] - x3[2]);
  const CT b0 = (CT)(x2[0] - x3[0]), b1 = (CT)(x2[1] - x3[1]), b2 = (CT)(x2[2] - x3[2]);
  const CT n0 = (CT)nx, n1 = (CT)ny, n2 = (CT)nz;
  CT g[6] = {0, 0, 0, 0, 0, 0};
  CT t[6][3] = {{0}};
  for (int q = 0; q < nq; ++q) {
    CT l1, l2, l3, w;
    if constexpr (sizeof(CT) == 8) {
      l1 = c_qd[qoff + q][0]; l2 = c_qd[qoff + q][1]; l3 = c_qd[qoff + q][2]; w = c_qd[qoff + q][3];
    } else {
      l1 = c_qf[qoff + q][0]; l2 = c_qf[qoff + q][1]; l3 = c_qf[qoff + q][2]; w = c_qf[qoff + q][3];
    }
    const CT rx = d0 + l1 * a0 + l2 * b0;
    const CT ry = d1 + l1 * a1 + l2 * b1;
    const CT rz = d2 + l1 * a2 + l2 * b2;
    const CT rxx = rx * rx, rxy = rx * ry, rxz = rx * rz, ryy = ry * ry, ryz = ry * rz, rzz = rz * rz;
    CT ir;
    if constexpr (sizeof(CT) == 8) ir = rsqrt(rxx + ryy + rzz);
    else ir = __frsqrt_rn(rxx + ryy + rzz);
    const CT ir2 = ir * ir, ir3 = ir2 * ir;
    const CT wir = w * ir, wir3 = w * ir3;
    g[0] += wir + rxx * wir3;
    g[1] += rxy * wir3;
    g[2] += rxz * wir3;
    g[3] += wir + ryy * wir3;
    g[4] += ryz * wir3;
    g[5] += wir + rzz * wir3;
    const CT ct = (CT)(-6.0) * ir2 * ir3 * (rx * n0 + ry * n1 + rz * n2);
    const CT wc[3] = {w * l1 * ct, w * l2 * ct, w * l3 * ct};
    const CT rr[6] = {rxx, rxy, rxz, ryy, ryz, rzz};
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int v = 0; v < 3; ++v) t[k][v] += wc[v] * rr[k];
  }
  const CT ca = (CT)area;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    Gs[k] = (double)g[k] * area;
#pragma unroll
    for (int v = 0; v < 3; ++v) Tv[k][v] = t[k][v] * ca;
  }
}

