// Analytic singular integrals over a flat triangle whose source point P lies in the triangle
// (a vertex, isrc = 1..3, or the centroid, isrc = 4). They replace si_tri_lap / si_tri_stk
// (src/mod_singInt.f90:34-167), which use the Ren & Chan (2015) reference-triangle transform;
// here the same integrals are evaluated in polar coordinates about P, edge by edge:
//
//   the triangle is split into sub-triangles (P, A, B) over the edges AB not containing P.
//   With t the unit vector A->B, F the foot of the perpendicular from P on AB, h = |F - P|,
//   nu = (F - P)/h and s_A = (A - F).t, s_B = (B - F).t, rho = sqrt(s^2 + h^2):
//     int 1/r dA                = h [asinh(s/h)]_A^B                               (= h I0)
//     int e_i e_j / r dA        = h ( S nu nu + C (nu t + t nu) + (I0 - S) t t )
//   with S = [s/rho]_A^B and C = -[h/rho]_A^B (e the in-plane unit vector from P).
//   Laplace:  G = (1/4pi) sum h I0                                   (H = 0 since r.n = 0)
//   Stokeslet (delta_ij/r + r_i r_j/r^3, no 1/8pi):  G_ij = sum [ h I0 delta_ij + int e_i e_j/r ]
//
// All arithmetic in FP64 (singular integrals stay FP64 in every precision form).
#pragma once

#include <math.h>

#ifndef ANDR_HD
#if defined(__HIPCC__)
#define ANDR_HD __host__ __device__ inline
#else
#define ANDR_HD static inline
#endif
#endif

// Adds the contribution of sub-triangle (P, A, B). M (6: xx xy xz yy yz zz) may be null.
ANDR_HD void andr_sing_edge(const double P[3], const double A[3], const double B[3], double *I0sum,
                            double *M) {
  double ab[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  double L = sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  double t[3] = {ab[0] / L, ab[1] / L, ab[2] / L};
  double ap[3] = {P[0] - A[0], P[1] - A[1], P[2] - A[2]};
  double sp = ap[0] * t[0] + ap[1] * t[1] + ap[2] * t[2];  // position of the foot from A
  double fp[3] = {A[0] + sp * t[0] - P[0], A[1] + sp * t[1] - P[1], A[2] + sp * t[2] - P[2]};
  double h = sqrt(fp[0] * fp[0] + fp[1] * fp[1] + fp[2] * fp[2]);
  double nu[3] = {fp[0] / h, fp[1] / h, fp[2] / h};
  double sA = -sp, sB = L - sp;
  double I0 = asinh(sB / h) - asinh(sA / h);
  *I0sum += h * I0;
  if (M) {
    double rA = sqrt(sA * sA + h * h), rB = sqrt(sB * sB + h * h);
    double S = sB / rB - sA / rA;
    double C = -(h / rB - h / rA);
    double k1 = h * S, k2 = h * C, k3 = h * (I0 - S);
    M[0] += k1 * nu[0] * nu[0] + 2.0 * k2 * nu[0] * t[0] + k3 * t[0] * t[0];
    M[1] += k1 * nu[0] * nu[1] + k2 * (nu[0] * t[1] + t[0] * nu[1]) + k3 * t[0] * t[1];
    M[2] += k1 * nu[0] * nu[2] + k2 * (nu[0] * t[2] + t[0] * nu[2]) + k3 * t[0] * t[2];
    M[3] += k1 * nu[1] * nu[1] + 2.0 * k2 * nu[1] * t[1] + k3 * t[1] * t[1];
    M[4] += k1 * nu[1] * nu[2] + k2 * (nu[1] * t[2] + t[1] * nu[2]) + k3 * t[1] * t[2];
    M[5] += k1 * nu[2] * nu[2] + 2.0 * k2 * nu[2] * t[2] + k3 * t[2] * t[2];
  }
}

// Source point of a singular pair: vertex isrc (1..3) or the centroid (isrc = 4), given as the
// caller computed it (the centroid must be the collocation point itself).
ANDR_HD void andr_sing_tri(const double x1[3], const double x2[3], const double x3[3], const double P[3],
                           int isrc, double *I0, double *M) {
  *I0 = 0.0;
  if (M)
    for (int i = 0; i < 6; ++i) M[i] = 0.0;
  // a vertex source sees only the opposite edge; the centroid sees all three
  if (isrc == 4 || isrc == 1) andr_sing_edge(P, x2, x3, I0, M);
  if (isrc == 4 || isrc == 2) andr_sing_edge(P, x3, x1, I0, M);
  if (isrc == 4 || isrc == 3) andr_sing_edge(P, x1, x2, I0, M);
}

// Laplace singular G (si_tri_lap): 1/(4 pi) int 1/r dA.
ANDR_HD double andr_sing_lap(const double x1[3], const double x2[3], const double x3[3], const double P[3],
                             int isrc) {
  double I0;
  andr_sing_tri(x1, x2, x3, P, isrc, &I0, 0);
  return I0 * (0.25 / 3.141592653589793238462643383);
}

// Stokeslet singular block (si_tri_stk): G6 = {xx, xy, xz, yy, yz, zz}.
ANDR_HD void andr_sing_stk(const double x1[3], const double x2[3], const double x3[3], const double P[3],
                           int isrc, double G6[6]) {
  double I0, M[6];
  andr_sing_tri(x1, x2, x3, P, isrc, &I0, M);
  G6[0] = I0 + M[0];
  G6[1] = M[1];
  G6[2] = M[2];
  G6[3] = I0 + M[3];
  G6[4] = M[4];
  G6[5] = I0 + M[5];
}
