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


// The regular integrals of one triangle seen from s, in FP64 with pressure (Triangle_3DStokesInt,
// src/lib_Stokes.f90:2341-2446, with StokesKernel :2450-2544): Stokeslet Gs, pressure vector pG,
// stresslet moments Tv and pressure moments pT per vertex, for the normal (nx, ny, nz) and area A.
__device__ inline void stokes_tri_full(const double s[3], const double *x1, const double *x2, const double *x3,
                                       double nx, double ny, double nz, double A, int qoff, int nq, double Gs[6],
                                       double pG[3], double Tv[6][3], double pT[3][3]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    Gs[k] = 0.0;
#pragma unroll
    for (int v = 0; v < 3; ++v) Tv[k][v] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pG[k] = 0.0;
#pragma unroll
    for (int v = 0; v < 3; ++v) pT[k][v] = 0.0;
  }
  for (int q = 0; q < nq; ++q) {
    const double l1 = c_qd[qoff + q][0], l2 = c_qd[qoff + q][1], l3 = c_qd[qoff + q][2];
    const double w = c_qd[qoff + q][3];
    const double rx = (l1 * x1[0] + l2 * x2[0] + l3 * x3[0]) - s[0];
    const double ry = (l1 * x1[1] + l2 * x2[1] + l3 * x3[1]) - s[1];
    const double rz = (l1 * x1[2] + l2 * x2[2] + l3 * x3[2]) - s[2];
    const double rxx = rx * rx, rxy = rx * ry, rxz = rx * rz, ryy = ry * ry, ryz = ry * rz, rzz = rz * rz;
    const double ir2 = 1.0 / (rxx + ryy + rzz);
    const double ir = sqrt(ir2);
    const double ir3 = ir2 * ir;
    Gs[0] += w * (ir * (1.0 + rxx * ir2));
    Gs[1] += w * (rxy * ir3);
    Gs[2] += w * (rxz * ir3);
    Gs[3] += w * (ir * (1.0 + ryy * ir2));
    Gs[4] += w * (ryz * ir3);
    Gs[5] += w * (ir * (1.0 + rzz * ir2));
    const double c2 = 2.0 * ir3;
    pG[0] += w * (c2 * rx);
    pG[1] += w * (c2 * ry);
    pG[2] += w * (c2 * rz);
    const double ct = -6.0 * ir2 * ir3 * (rx * nx + ry * ny + rz * nz);
    const double tc[6] = {ct * rxx, ct * rxy, ct * rxz, ct * ryy, ct * ryz, ct * rzz};
    const double c4 = 4.0 * ir3;
    const double pt[3] = {c4 * ((-1.0 + 3.0 * rxx * ir2) * nx + 3.0 * rxy * ir2 * ny + 3.0 * rxz * ir2 * nz),
                          c4 * (3.0 * rxy * ir2 * nx + (-1.0 + 3.0 * ryy * ir2) * ny + 3.0 * ryz * ir2 * nz),
                          c4 * (3.0 * rxz * ir2 * nx + 3.0 * ryz * ir2 * ny + (-1.0 + 3.0 * rzz * ir2) * nz)};
    const double wl[3] = {w * l1, w * l2, w * l3};
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int v = 0; v < 3; ++v) Tv[k][v] += wl[v] * tc[k];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int v = 0; v < 3; ++v) pT[k][v] += wl[v] * pt[k];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    Gs[k] *= A;
#pragma unroll
    for (int v = 0; v < 3; ++v) Tv[k][v] *= A;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pG[k] *= A;
#pragma unroll
    for (int v = 0; v < 3; ++v) pT[k][v] *= A;
  }
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

// As stokes_pair_regular, with the stresslet moments summed over the vertices: Ts[k] = sum_v Tv[k][v]
// (the shape functions sum to one, so the vertex weights drop out of the quadrature loop). Enough for
// a velocity field that is affine over the element (a rigid motion of all three vertices): then
// sum_v Tv[k][v] u(x_v) = int T_k u(x), and the rotation terms reduce to Ts and the collocation point
// (the third moments int T_{c,i} r_j enter the rotations antisymmetrised in (i, j), where
// T_{c,i} r_j = t r_c r_i r_j cancels at every quadrature point).
template <class CT>
__device__ inline void stokes_pair_sum(const double s[3], const double *x1, const double *x2, const double *x3,
                                       double nx, double ny, double nz, double area, int qoff, int nq, double Gs[6],
                                       CT Ts[6]) {
  const CT d0 = (CT)(x3[0] - s[0]), d1 = (CT)(x3[1] - s[1]), d2 = (CT)(x3[2] - s[2]);
  const CT a0 = (CT)(x1[0] - x3[0]), a1 = (CT)(x1[1] - x3[1]), a2 = (CT)(x1[2] - x3[2]);
  const CT b0 = (CT)(x2[0] - x3[0]), b1 = (CT)(x2[1] - x3[1]), b2 = (CT)(x2[2] - x3[2]);
  const CT n0 = (CT)nx, n1 = (CT)ny, n2 = (CT)nz;
  CT g[6] = {0, 0, 0, 0, 0, 0};
  CT t[6] = {0, 0, 0, 0, 0, 0};
  for (int q = 0; q < nq; ++q) {
    CT l1, l2, w;
    if constexpr (sizeof(CT) == 8) {
      l1 = c_qd[qoff + q][0]; l2 = c_qd[qoff + q][1]; w = c_qd[qoff + q][3];
    } else {
      l1 = c_qf[qoff + q][0]; l2 = c_qf[qoff + q][1]; w = c_qf[qoff + q][3];
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
    const CT wc = (CT)(-6.0) * wir3 * ir2 * (rx * n0 + ry * n1 + rz * n2);
    t[0] += wc * rxx;
    t[1] += wc * rxy;
    t[2] += wc * rxz;
    t[3] += wc * ryy;
    t[4] += wc * ryz;
    t[5] += wc * rzz;
  }
  const CT ca = (CT)area;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    Gs[k] = (double)g[k] * area;
    Ts[k] = t[k] * ca;
  }
}

