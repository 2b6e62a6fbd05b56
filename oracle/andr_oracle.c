/*
 * CPU oracle for the Andromeda BEM hot path -- TEST INFRASTRUCTURE ONLY (see andr_oracle.h).
 * Plain C, FP64, OpenMP over collocation rows. Restates the Fortran reference routine by
 * routine; each function cites the reference file:line it follows.
 */
#include "andr_oracle.h"
#include "../andromeda_amd/csrc/andr_wx_rules.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793238462643383

/* ------------------------------------------------------------------------------------------
 * Quadrature table: Triangle_IntegrationInit (src/mod_trikotnik.f90:282-677) lists orbit
 * generators; WandzuratXiaoBary (:682-781) expands each orbit by the 3 rotations (and the
 * reflection for 6-point orbits) and converts to barycentric coordinates.
 * ---------------------------------------------------------------------------------------- */
int orc_tri_rule(int itri, double *l1, double *l2, double *l3, double *w) {
  const struct andr_wx_orbit *orb;
  int norb;
  switch (itri) {
    case 1: orb = andr_wx_deg5; norb = 3; break;
    case 2: orb = andr_wx_deg10; norb = 7; break;
    case 3: orb = andr_wx_deg15; norb = 12; break;
    case 4: orb = andr_wx_deg20; norb = 19; break;
    case 5: orb = andr_wx_deg25; norb = 26; break;
    default: return -1;
  }
  int k = 0;
  for (int o = 0; o < norb; ++o) {
    const double a = orb[o].a, b = orb[o].b, c = orb[o].c, ww = orb[o].w;
    double P[6][3] = {{a, b, c}, {c, a, b}, {b, c, a}, {b, a, c}, {c, b, a}, {a, c, b}};
    for (int p = 0; p < orb[o].m; ++p) {
      l1[k] = P[p][0]; l2[k] = P[p][1]; l3[k] = P[p][2]; w[k] = ww; ++k;
    }
  }
  return k;
}

/* Heron's formula, Triangle_Area (src/mod_trikotnik.f90:786-802) == tri_area (src/mod_singInt.f90:394-413) */
static double heron(const double x1[3], const double x2[3], const double x3[3]) {
  double a = sqrt(pow(x1[0] - x2[0], 2.0) + pow(x1[1] - x2[1], 2.0) + pow(x1[2] - x2[2], 2.0));
  double b = sqrt(pow(x3[0] - x2[0], 2.0) + pow(x3[1] - x2[1], 2.0) + pow(x3[2] - x2[2], 2.0));
  double c = sqrt(pow(x1[0] - x3[0], 2.0) + pow(x1[1] - x3[1], 2.0) + pow(x1[2] - x3[2], 2.0));
  double p = 0.5 * (a + b + c);
  return sqrt(p * (p - a) * (p - b) * (p - c));
}

/* ------------------------------------------------------------------------------------------
 * Gauss-Legendre rules: GaussPosWeights (src/mod_gauss.f90:33-391) holds N = 2,4,6,8,16,32,40,48
 * as literals; here they are computed (Newton iteration on the three-term Legendre recurrence,
 * nodes ascending as in the reference tables).
 * ---------------------------------------------------------------------------------------- */
int orc_gauss_rule(int iq, double *x, double *w) {
  static const int Ns[8] = {2, 4, 6, 8, 16, 32, 40, 48};
  if (iq < 1 || iq > 8) return -1;
  const int N = Ns[iq - 1];
  for (int i = 0; i < N / 2; ++i) {
    double z = cos(ORC_PI * (i + 0.75) / (N + 0.5)), pp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= N; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = N * (z * p1 - p2) / (z * z - 1.0);
      const double dz = p1 / pp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    x[N - 1 - i] = z;
    w[i] = w[N - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
  return N;
}

/* bilinear quadrilateral at local coordinates (gi, gj) in [-1, 1]^2 (Quad_BEMIntNum,
 * src/lib_intAna.f90:300-325): shape functions fig4, point f, unnormalised normal an = x_ksi x x_eta */
static void quad_point(const double *x1, const double *x2, const double *x3, const double *x4, double gi, double gj,
                       double fig4[4], double f[3], double an[3]) {
  const double e1m = 1.0 - gi, e1p = 1.0 + gi, e2m = 1.0 - gj, e2p = 1.0 + gj;
  fig4[0] = 0.25 * e1m * e2m;
  fig4[1] = 0.25 * e1p * e2m;
  fig4[2] = 0.25 * e1p * e2p;
  fig4[3] = 0.25 * e1m * e2p;
  double ks[3], et[3];
  for (int j = 0; j < 3; ++j) {
    f[j] = fig4[0] * x1[j] + fig4[1] * x2[j] + fig4[2] * x3[j] + fig4[3] * x4[j];
    ks[j] = 0.25 * (-e2m * x1[j] + e2m * x2[j] + e2p * x3[j] - e2p * x4[j]);
    et[j] = 0.25 * (-e1m * x1[j] - e1p * x2[j] + e1p * x3[j] + e1m * x4[j]);
  }
  an[0] = ks[1] * et[2] - et[1] * ks[2];
  an[1] = et[0] * ks[2] - ks[0] * et[2];
  an[2] = ks[0] * et[1] - et[0] * ks[1];
}

void orc_geometry_n(int nnodes, int nelem, const double *xyz, const int32_t *con, int nvert,
                    double *normal, double *area, double *centroid) {
  (void)nnodes;
  double gx[8], gw[8];
  orc_gauss_rule(4, gx, gw); /* CalElementArea uses gaus%ng1(4):ng2(4) (src/lib_int.f90:648-649) */
  for (int e = 0; e < nelem; ++e) {
    const int32_t *cn = con + (size_t)nvert * e;
    const int quad = nvert == 4 && cn[3] >= 0;
    const double *t1 = xyz + 3 * cn[0], *t2 = xyz + 3 * cn[1], *t3 = xyz + 3 * cn[2];
    /* CalElementNormal (src/lib_mesh.f90:716-769): first three nodes for either type */
    double v1[3] = {t2[0] - t1[0], t2[1] - t1[1], t2[2] - t1[2]};
    double v2[3] = {t3[0] - t1[0], t3[1] - t1[1], t3[2] - t1[2]};
    double n0 = v1[1] * v2[2] - v1[2] * v2[1];
    double n1 = -v1[0] * v2[2] + v1[2] * v2[0];
    double n2 = v1[0] * v2[1] - v1[1] * v2[0];
    double d = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    normal[3 * e] = n0 / d; normal[3 * e + 1] = n1 / d; normal[3 * e + 2] = n2 / d;
    if (!quad) {
      area[e] = heron(t1, t2, t3);
      /* CalQMesh (src/lib_mesh.f90:577-629): barycentre */
      for (int j = 0; j < 3; ++j) centroid[3 * e + j] = (1.0 / 3.0) * t1[j] + (1.0 / 3.0) * t2[j] + (1.0 / 3.0) * t3[j];
    } else {
      const double *t4 = xyz + 3 * cn[3];
      double A = 0.0;
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
          double fig4[4], f[3], an[3];
          quad_point(t1, t2, t3, t4, gx[i], gx[j], fig4, f, an);
          const double aj = gw[i] * gw[j] * sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
          for (int k = 0; k < 4; ++k) A += aj * fig4[k];
        }
      area[e] = A;
      for (int j = 0; j < 3; ++j) centroid[3 * e + j] = 0.25 * t1[j] + 0.25 * t2[j] + 0.25 * t3[j] + 0.25 * t4[j];
    }
  }
}

void orc_geometry(int nnodes, int nelem, const double *xyz, const int32_t *con,
                  double *normal, double *area, double *centroid) {
  orc_geometry_n(nnodes, nelem, xyz, con, 3, normal, area, centroid);
}

void orc_bem_kernel(const double s[3], const double f[3], const double n[3], double *G, double *H) {
  double rx = -f[0] + s[0], ry = -f[1] + s[1], rz = -f[2] + s[2];
  double ra = sqrt(rx * rx + ry * ry + rz * rz);
  double ira2 = 1.0 / (ra * ra);
  double g = 0.25 / (ORC_PI * ra);
  *G = g;
  *H = (rx * n[0] + ry * n[1] + rz * n[2]) * ira2 * g;
}

static int g_itri = 5;
static int g_nq = 0;
static double g_l1[126], g_l2[126], g_l3[126], g_w[126];

static void ensure_rule(int itri) {
  if (g_nq == 0 || itri != g_itri) {
    g_nq = orc_tri_rule(itri, g_l1, g_l2, g_l3, g_w);
    g_itri = itri;
  }
}

/* Triangle_3DLapInt (src/lib_sdBem.f90:859-905) */
void orc_tri_lap_int(const double x1[3], const double x2[3], const double x3[3], const double s[3],
                     const double n[3], double area, double out[4]) {
  if (g_nq == 0) ensure_rule(5);
  for (int i = 0; i < g_nq; ++i) {
    double f[3];
    for (int j = 0; j < 3; ++j) f[j] = g_l1[i] * x1[j] + g_l2[i] * x2[j] + g_l3[i] * x3[j];
    double G, H;
    orc_bem_kernel(s, f, n, &G, &H);
    out[0] += g_w[i] * G * area;
    out[1] += g_w[i] * H * area * g_l1[i];
    out[2] += g_w[i] * H * area * g_l2[i];
    out[3] += g_w[i] * H * area * g_l3[i];
  }
}

/* ------------------------------------------------------------------------------------------
 * Ren & Chan (2015) analytic singular integrals, src/mod_singInt.f90
 * ---------------------------------------------------------------------------------------- */
/* getXi100 (:418-430) */
static double xi100(double a, double b, double c) {
  double d = 2.0 * sqrt(b) * sqrt(a + b - c) + 2.0 * b - c;
  double e = 2.0 * sqrt(a) * sqrt(b) - c;
  return log(d / e) / sqrt(b);
}
/* getXi010 (:434-447) */
static double xi010(double a, double b, double c) {
  double d = 2.0 * sqrt(a) * sqrt(a + b - c) + 2.0 * a - c;
  double e = 2.0 * sqrt(a) * sqrt(b) - c;
  return log(d / e) / sqrt(a);
}
/* getXi000 (:453-463) */
static double xi000(double a, double b, double c) {
  double d = 2.0 * sqrt(a) * sqrt(a + b - c) + 2.0 * a - c;
  double e = 2.0 * sqrt(b) * sqrt(a + b - c) - 2.0 * b + c;
  return log(d / e) / sqrt(a + b - c);
}
/* getXi13130, FF2 of Eq. (24) (:827-899) */
static double xi13130(double a, double b, double c) {
  double d = 4.0 * a + 2.0 * b + 2.0 * sqrt(4.0 * a + b - 2.0 * c) * sqrt(a + b - c) - 3.0 * c;
  double e = -2.0 * a - 4.0 * b + 2.0 * sqrt(a + 4.0 * b - 2.0 * c) * sqrt(a + b - c) + 3.0 * c;
  double t1 = log(d / e) / sqrt(a + b - c);
  d = 4.0 * a + 2.0 * sqrt(4.0 * a + b - 2.0 * c) * sqrt(a) - c;
  e = -2.0 * a - c + 2.0 * sqrt(a + b + c) * sqrt(a);
  double t2 = log(d / e) / sqrt(a);
  d = (4.0 * b + 2.0 * sqrt(a + 4.0 * b - 2.0 * c) * sqrt(b) - c) * (2.0 * b + 2.0 * sqrt(a + b + c) * sqrt(b) + c);
  e = (b + sqrt(4.0 * a + b - 2.0 * c) * sqrt(b) - c) * (-b + sqrt(4.0 * a + b - 2.0 * c) * sqrt(b) + c);
  double t3 = log(d / e) / sqrt(b);
  return (t1 + t2 + t3) / 3.0;
}
/* getXi13130G, Eq. (34) (:468-521) */
static double xi13130G(double a, double b, double c, double d, double e, double f) {
  double G = 0.0, st, im;
  st = 2.0 * sqrt(b) * (b * c * d + a * c * e - 2.0 * a * b * f) * (sqrt(a + b + c) - sqrt(4.0 * a + b - 2.0 * c));
  im = a * (4.0 * a * b - c * c);
  G += st / im;
  st = 2.0 * sqrt(b) * (b * (2.0 * b - c) * d - a * (2.0 * b * e + c * e - 2.0 * b * f) + c * (c * e - b * f)) *
       (sqrt(a + 4.0 * b - 2.0 * c) - sqrt(4.0 * a + b - 2.0 * c));
  im = (a + b - c) * (4.0 * a * b - c * c);
  G += st / im;
  st = 4.0 * a + 2.0 * b + 2.0 * sqrt(4.0 * a + b - 2.0 * c) * sqrt(a + b - c) - 3.0 * c;
  im = -2.0 * a - 4.0 * b + 2.0 * sqrt(a + 4.0 * b - 2.0 * c) * sqrt(a + b - c) + 3.0 * c;
  G += e * sqrt(b) * log(st / im) / sqrt(a + b - c);
  G += sqrt(b) * (b * d - a * e + c * e - b * f) * log(st / im) / pow(a + b - c, 1.5);
  st = 4.0 * a + 2.0 * sqrt(a) * sqrt(4.0 * a + b - 2.0 * c) - c;
  im = -2.0 * a - c + 2.0 * sqrt(a) * sqrt(a + b + c);
  G += e * sqrt(b) * log(st / im) / sqrt(a);
  G += sqrt(b) * (b * d - a * e) * log(st / im) / pow(a, 1.5);
  st = (4.0 * b + 2.0 * sqrt(b) * sqrt(a + 4.0 * b - 2.0 * c) - c) * (2.0 * b + 2.0 * sqrt(b) * sqrt(a + b + c) + c);
  im = 4.0 * (b + sqrt(b) * sqrt(4.0 * a + b - 2.0 * c) - c) * (-b + sqrt(b) * sqrt(4.0 * a + b - 2.0 * c) + c);
  G += e * log(st / im);
  G += e * log(4.0);
  return G / (3.0 * pow(b, 1.5));
}
/* getXi000G (:760-776) */
static double xi000G(double a, double b, double c, double d, double e, double f) {
  double sa = sqrt(a), sb = sqrt(b), sabc = sqrt(a + b - c);
  double b15 = pow(b, 1.5), abc15 = pow(a + b - c, 1.5), k = 4.0 * a * b - c * c;
  double num = -4.0 * a * b * e + c * c * e +
               (2.0 * b * (2.0 * b * b * d + c * (-a + c) * e - b * (2.0 * a * (e - f) + c * (d + f)))) / (a + b - c) +
               (2.0 * sa * sb * (-2.0 * b * b * d + (a - c) * c * e + b * (2.0 * a * (e - f) + c * (d + f)))) / (a + b - c) +
               (b15 * k * (d + e - f) * log(2.0 * a + 2.0 * sa * sabc - c)) / abc15 +
               (-2.0 * b15 * c * d - 2.0 * a * sb * c * e + 4.0 * a * b15 * f +
                sa * (-4.0 * a * b + c * c) * e * (-1.0 + log(2.0 * sa * sb + c))) / sa +
               4.0 * a * b * e * log(2.0 * sa * sb + c) - c * c * e * log(2.0 * sa * sb + c) -
               (b15 * k * (d + e - f) * log(-2.0 * b + 2.0 * sb * sabc + c)) / abc15;
  return num / (b15 * k);
}
/* getXi100G (:782-795) */
static double xi100G(double a, double b, double c, double d, double e, double f) {
  double sa = sqrt(a), sb = sqrt(b), sabc = sqrt(a + b - c), k = 4.0 * a * b - c * c;
  double num = 2.0 * sb * (b * sabc * c * d - pow(a, 1.5) * (c * e + 2.0 * b * (e - f)) + a * sabc * (c * e - 2.0 * b * f) +
                           sa * (2.0 * b * b * d + c * c * e - b * c * (d + f))) -
               sa * sabc * k * e * (log(2.0 * sa * sb - c) - log(2.0 * b + 2.0 * sb * sabc - c));
  return num / (sa * pow(b, 1.5) * sabc * k);
}
/* getXi010G (:801-822) */
static double xi010G(double a, double b, double c, double d, double e, double f) {
  double sa = sqrt(a), sb = sqrt(b), sabc = sqrt(a + b - c), k = 4.0 * a * b - c * c;
  double b15 = pow(b, 1.5), b25 = pow(b, 2.5);
  double t1 = (2.0 * sb * (-2.0 * b * b * d + (a - c) * c * e + b * (2.0 * a * (e - f) + c * (d + f))) +
               sabc * k * e * (-1.0 + log(-2.0 * b + 2.0 * sb * sabc + c))) / sabc;
  double t2 = (b15 * sabc * (-k) * d * log(2.0 * sa * sb - c) + b15 * sabc * k * d * log(2.0 * a + 2.0 * sa * sabc - c) +
               sa * (2.0 * b15 * c * (-b + sb * sabc + c) * d +
                     a * a * (4.0 * b * sabc * e - 2.0 * sb * c * e + 4.0 * b15 * f) +
                     a * (2.0 * b * sabc * c * e + 2.0 * sb * c * c * e - sabc * c * c * e + 4.0 * b25 * f -
                          4.0 * b * b * sabc * f - 2.0 * b15 * c * (d + e + 2.0 * f)) +
                     a * sabc * (-k) * e * log(-2.0 * b + 2.0 * sb * sabc + c))) /
              (pow(a, 1.5) * sabc);
  return (t1 + t2) / (b15 * k);
}

static void cross(double a[3], const double b[3], const double c[3]) { /* a = b x c (siCrossProduct) */
  a[0] = b[1] * c[2] - b[2] * c[1];
  a[1] = -b[0] * c[2] + b[2] * c[0];
  a[2] = b[0] * c[1] - b[1] * c[0];
}
static void normv(double a[3]) { /* siNormVector (:1141-1155), accumulator started at zero */
  double nrm = 0.0;
  for (int i = 0; i < 3; ++i) nrm += a[i] * a[i];
  nrm = sqrt(nrm);
  for (int i = 0; i < 3; ++i) a[i] /= nrm;
}
/* getMrot (:905-950) */
static void get_mrot(const double rr1[3], const double rr2[3], const double rr3[3], double M[3][3]) {
  double ov1[3], ov2[3], ov3[3], ov4[3];
  for (int i = 0; i < 3; ++i) { ov1[i] = rr2[i] - rr1[i]; ov4[i] = rr3[i] - rr1[i]; }
  cross(ov3, ov4, ov1);
  cross(ov2, ov3, ov1);
  normv(ov1); normv(ov2); normv(ov3);
  for (int j = 0; j < 3; ++j) { M[0][j] = ov1[j]; M[1][j] = ov2[j]; M[2][j] = ov3[j]; }
}
/* getrcABC / getABCDEF (:1023-1108) with getABC (:955-977) and getDEF (:982-1017) */
static void get_frame(const double x1[3], const double x2[3], const double x3[3], double *a, double *b, double *c,
                      double def[9], double M[3][3]) {
  double cen[3], rr[3][3], rrr[3][3];
  for (int j = 0; j < 3; ++j) cen[j] = (x1[j] + x2[j] + x3[j]) / 3.0;
  for (int j = 0; j < 3; ++j) { rr[0][j] = x1[j] - cen[j]; rr[1][j] = x2[j] - cen[j]; rr[2][j] = x3[j] - cen[j]; }
  get_mrot(rr[0], rr[1], rr[2], M);
  for (int v = 0; v < 3; ++v)
    for (int i = 0; i < 3; ++i) rrr[v][i] = M[i][0] * rr[v][0] + M[i][1] * rr[v][1] + M[i][2] * rr[v][2];
  double A11 = rrr[0][0] - rrr[2][0], A12 = rrr[1][0] - rrr[2][0];
  double A21 = rrr[0][1] - rrr[2][1], A22 = rrr[1][1] - rrr[2][1];
  *a = A11 * A11 + A21 * A21;
  *b = A12 * A12 + A22 * A22;
  *c = 2.0 * (A11 * A12 + A21 * A22);
  if (def) {
    def[0] = A11 * A11; def[1] = A12 * A12; def[2] = 2.0 * A11 * A12;          /* dxx exx fxx */
    def[3] = A11 * A21; def[4] = A12 * A22; def[5] = A11 * A22 + A21 * A12;    /* dxy exy fxy */
    def[6] = A21 * A21; def[7] = A22 * A22; def[8] = 2.0 * A21 * A22;          /* dyy eyy fyy */
  }
}

/* si_tri_lap (:127-167) */
double orc_si_tri_lap(const double x1[3], const double x2[3], const double x3[3], double area, int isrc) {
  double a, b, c, M[3][3], f = 0.0;
  get_frame(x1, x2, x3, &a, &b, &c, NULL, M);
  switch (isrc) {
    case 1: f = xi100(a, b, c); break;
    case 2: f = xi010(a, b, c); break;
    case 3: f = xi000(a, b, c); break;
    case 4: f = xi13130(a, b, c); break;
    default: return NAN;
  }
  double jac = 2.0 * area;
  return 0.25 / ORC_PI * jac * f;
}

/* si_tri_stk (:34-123) */
void orc_si_tri_stk(const double x1[3], const double x2[3], const double x3[3], double area, int isrc,
                    double Gout[6]) {
  double a, b, c, def[9], M[3][3];
  get_frame(x1, x2, x3, &a, &b, &c, def, M);
  double ff, gxx, gxy, gyy;
  double (*fG)(double, double, double, double, double, double);
  switch (isrc) {
    case 1: ff = xi100(a, b, c); fG = xi100G; break;
    case 2: ff = xi010(a, b, c); fG = xi010G; break;
    case 3: ff = xi000(a, b, c); fG = xi000G; break;
    case 4: ff = xi13130(a, b, c); fG = xi13130G; break;
    default: for (int i = 0; i < 6; ++i) Gout[i] = NAN; return;
  }
  gxx = fG(a, b, c, def[0], def[1], def[2]);
  gxy = fG(a, b, c, def[3], def[4], def[5]);
  gyy = fG(a, b, c, def[6], def[7], def[8]);
  double mat[3][3] = {{0}};
  mat[0][0] = 2.0 * area * (ff + gxx);
  mat[1][1] = 2.0 * area * (ff + gyy);
  mat[2][2] = 2.0 * area * ff;
  mat[0][1] = mat[1][0] = 2.0 * area * gxy;
  /* mat = Mrot^T mat Mrot */
  double t[3][3], r[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += M[k][i] * mat[k][j];
      t[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += t[i][k] * M[k][j];
      r[i][j] = s;
    }
  Gout[0] = r[0][0]; Gout[1] = r[0][1]; Gout[2] = r[0][2];
  Gout[3] = r[1][1]; Gout[4] = r[1][2]; Gout[5] = r[2][2];
}

/* StokesKernel (src/lib_Stokes.f90:2450-2544): r = f - s, no 1/(8 pi) */
void orc_stokes_kernel(const double s[3], const double f[3], const double n[3],
                       double G[9], double T[9], double prG[3], double prT[3]) {
  double rx = f[0] - s[0], ry = f[1] - s[1], rz = f[2] - s[2];
  double rxx = rx * rx, rxy = rx * ry, rxz = rx * rz, ryy = ry * ry, ryz = ry * rz, rzz = rz * rz;
  double ir2 = 1.0 / (rxx + ryy + rzz);
  double ir = sqrt(ir2);
  double ir3 = ir2 * ir;
  G[0] = ir * (1.0 + rxx * ir2); G[1] = rxy * ir3; G[2] = rxz * ir3;
  G[3] = G[1]; G[4] = ir * (1.0 + ryy * ir2); G[5] = ryz * ir3;
  G[6] = G[2]; G[7] = G[5]; G[8] = ir * (1.0 + rzz * ir2);
  double c = -6.0 * ir2 * ir3 * (rx * n[0] + ry * n[1] + rz * n[2]);
  T[0] = c * rxx; T[1] = c * rxy; T[2] = c * rxz;
  T[3] = T[1]; T[4] = c * ryy; T[5] = c * ryz;
  T[6] = T[2]; T[7] = T[5]; T[8] = c * rzz;
  c = 2.0 * ir3;
  prG[0] = c * rx; prG[1] = c * ry; prG[2] = c * rz;
  c = 4.0 * ir3;
  prT[0] = c * ((-1.0 + 3.0 * rxx * ir2) * n[0] + 3.0 * rxy * ir2 * n[1] + 3.0 * rxz * ir2 * n[2]);
  prT[1] = c * (3.0 * rxy * ir2 * n[0] + (-1.0 + 3.0 * ryy * ir2) * n[1] + 3.0 * ryz * ir2 * n[2]);
  prT[2] = c * (3.0 * rxz * ir2 * n[0] + 3.0 * ryz * ir2 * n[1] + (-1.0 + 3.0 * rzz * ir2) * n[2]);
}

/* Triangle_3DStokesInt (src/lib_Stokes.f90:2341-2446) */
void orc_tri_stokes_int(const double x1[3], const double x2[3], const double x3[3], const double s[3],
                        const double n[3], double area, double acc[36]) {
  if (g_nq == 0) ensure_rule(5);
  static const int Tidx[6] = {0, 1, 2, 4, 5, 8}; /* xx xy xz yy yz zz */
  for (int i = 0; i < g_nq; ++i) {
    double f[3];
    for (int j = 0; j < 3; ++j) f[j] = g_l1[i] * x1[j] + g_l2[i] * x2[j] + g_l3[i] * x3[j];
    double G[9], T[9], prG[3], prT[3];
    orc_stokes_kernel(s, f, n, G, T, prG, prT);
    double w = g_w[i] * area;
    double wl[3] = {w * g_l1[i], w * g_l2[i], w * g_l3[i]};
    for (int k = 0; k < 6; ++k) acc[k] += w * G[Tidx[k]];
    for (int k = 0; k < 3; ++k) acc[6 + k] += w * prG[k];
    for (int k = 0; k < 6; ++k)
      for (int v = 0; v < 3; ++v) acc[9 + 3 * k + v] += wl[v] * T[Tidx[k]];
    for (int k = 0; k < 3; ++k)
      for (int v = 0; v < 3; ++v) acc[27 + 3 * k + v] += wl[v] * prT[k];
  }
}

/* ------------------------------------------------------------------------------------------
 * Quadrilaterals: regular integrals over bilinear sub-cells, singular ones split in triangles
 * ---------------------------------------------------------------------------------------- */
static double dist3(const double *a, const double *b) {
  return sqrt(pow(a[0] - b[0], 2.0) + pow(a[1] - b[1], 2.0) + pow(a[2] - b[2], 2.0));
}

/* ndiv sub-cells x ndiv sub-cells of N x N points; calls fn(fig4, f, n_unit * mul, weight) */
typedef void (*quad_fn)(void *ctx, const double fig4[4], const double f[3], const double n[3], double w);
static void quad_loop(const double *x1, const double *x2, const double *x3, const double *x4, double mul, int iq,
                      int ndiv_xi, int ndiv_et, quad_fn fn, void *ctx) {
  double gx[48], gw[48];
  const int N = orc_gauss_rule(iq, gx, gw);
  for (int ix = 1; ix <= ndiv_xi; ++ix)
    for (int ie = 1; ie <= ndiv_et; ++ie) {
      const double a = -1.0 + (ix - 1) * 2.0 / ndiv_xi, b = -1.0 + ix * 2.0 / ndiv_xi;
      const double c = -1.0 + (ie - 1) * 2.0 / ndiv_et, d = -1.0 + ie * 2.0 / ndiv_et;
      const double dex = 0.25 * (b - a) * (d - c);
      for (int i = 0; i < N; ++i) {
        const double gii = a + 0.5 * (gx[i] + 1.0) * (b - a);
        for (int j = 0; j < N; ++j) {
          const double gij = c + 0.5 * (gx[j] + 1.0) * (d - c);
          double fig4[4], f[3], an[3];
          quad_point(x1, x2, x3, x4, gii, gij, fig4, f, an);
          const double inv = 1.0 / sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
          const double n[3] = {mul * an[0] * inv, mul * an[1] * inv, mul * an[2] * inv};
          fn(ctx, fig4, f, n, dex * gw[i] * gw[j] / inv);
        }
      }
    }
}

typedef struct { const double *s; double *out; } lapq_ctx;
static void lapq_fn(void *vc, const double fig4[4], const double f[3], const double n[3], double w) {
  lapq_ctx *c = (lapq_ctx *)vc;
  double G, H;
  orc_bem_kernel(c->s, f, n, &G, &H);
  for (int k = 0; k < 4; ++k) c->out[1 + k] += H * w * fig4[k];
  c->out[0] += G * w;
}

void orc_quad_lap_int(const double *x1, const double *x2, const double *x3, const double *x4,
                      const double s[3], double mul, int iq, double out[5]) {
  for (int k = 0; k < 5; ++k) out[k] = 0.0;
  const double minedge = 1.0; /* fixed in the reference (src/lib_intAna.f90:263) */
  const double d13 = fmax(dist3(x1, x2), dist3(x3, x4)), d24 = fmax(dist3(x2, x3), dist3(x4, x1));
  int nxi = (int)(d13 / minedge + 0.5), net = (int)(d24 / minedge + 0.5);
  nxi = nxi > 2 ? nxi : 2;
  net = net > 2 ? net : 2;
  lapq_ctx c = {s, out};
  quad_loop(x1, x2, x3, x4, mul, iq, nxi, net, lapq_fn, &c);
}

/* two (vertex source) or four (centre source) triangles with the source at their first vertex */
double orc_si_qua_lap(const double *x1, const double *x2, const double *x3, const double *x4, int isrc) {
  const double *v[4] = {x1, x2, x3, x4};
  if (isrc >= 1 && isrc <= 4) {
    const double *a = v[isrc - 1], *b = v[isrc % 4], *c = v[(isrc + 1) % 4], *d = v[(isrc + 2) % 4];
    return orc_si_tri_lap(a, b, c, heron(a, b, c), 1) + orc_si_tri_lap(a, c, d, heron(a, c, d), 1);
  }
  if (isrc == 5) {
    double x5[3];
    for (int j = 0; j < 3; ++j) x5[j] = 0.25 * (x1[j] + x2[j] + x3[j] + x4[j]);
    double G = 0.0;
    for (int k = 0; k < 4; ++k) G += orc_si_tri_lap(x5, v[k], v[(k + 1) % 4], heron(x5, v[k], v[(k + 1) % 4]), 1);
    return G;
  }
  return 0.0;
}

typedef struct { const double *s; double *acc; } stkq_ctx;
static void stkq_fn(void *vc, const double fig4[4], const double f[3], const double n[3], double w) {
  stkq_ctx *c = (stkq_ctx *)vc;
  double G[9], T[9], prG[3], prT[3];
  orc_stokes_kernel(c->s, f, n, G, T, prG, prT);
  static const int idx[6] = {0, 1, 2, 4, 5, 8};
  for (int k = 0; k < 6; ++k) c->acc[k] += w * G[idx[k]];
  for (int k = 0; k < 3; ++k) c->acc[6 + k] += w * prG[k];
  for (int v = 0; v < 4; ++v) {
    const double wv = w * fig4[v];
    for (int k = 0; k < 6; ++k) c->acc[9 + 4 * k + v] += wv * T[idx[k]];
    for (int k = 0; k < 3; ++k) c->acc[33 + 4 * k + v] += wv * prT[k];
  }
}

void orc_quad_stokes_int(const double *x1, const double *x2, const double *x3, const double *x4,
                         const double s[3], double mul, int iq, double acc[48]) {
  for (int k = 0; k < 48; ++k) acc[k] = 0.0;
  stkq_ctx c = {s, acc};
  quad_loop(x1, x2, x3, x4, mul, iq, 3, 3, stkq_fn, &c); /* ndivXi = ndivEt = 3 (:3845-3846) */
}

void orc_si_qua_stk(const double *x1, const double *x2, const double *x3, const double *x4, int isrc,
                    double Gout[6]) {
  const double *v[4] = {x1, x2, x3, x4};
  double g[6];
  for (int k = 0; k < 6; ++k) Gout[k] = 0.0;
  if (isrc >= 1 && isrc <= 4) {
    const double *a = v[isrc - 1], *b = v[isrc % 4], *c = v[(isrc + 1) % 4], *d = v[(isrc + 2) % 4];
    orc_si_tri_stk(a, b, c, heron(a, b, c), 1, g);
    for (int k = 0; k < 6; ++k) Gout[k] += g[k];
    orc_si_tri_stk(a, c, d, heron(a, c, d), 1, g);
    for (int k = 0; k < 6; ++k) Gout[k] += g[k];
  } else if (isrc == 5) {
    double x5[3];
    for (int j = 0; j < 3; ++j) x5[j] = 0.25 * (x1[j] + x2[j] + x3[j] + x4[j]);
    for (int k = 0; k < 4; ++k) {
      orc_si_tri_stk(x5, v[k], v[(k + 1) % 4], heron(x5, v[k], v[(k + 1) % 4]), 1, g);
      for (int q = 0; q < 6; ++q) Gout[q] += g[q];
    }
  }
}

/* ------------------------------------------------------------------------------------------
 * Row integration and matrix assembly
 * ---------------------------------------------------------------------------------------- */
typedef struct {
  double *normal, *area, *cen;
} geom_t;

static void make_geom(const orc_body *b, geom_t *g) {
  g->normal = (double *)malloc(sizeof(double) * 3 * b->nelem);
  g->area = (double *)malloc(sizeof(double) * b->nelem);
  g->cen = (double *)malloc(sizeof(double) * 3 * b->nelem);
  orc_geometry_n(b->nnodes, b->nelem, b->xyz, b->con, b->nvert == 4 ? 4 : 3, g->normal, g->area, g->cen);
}
static void free_geom(geom_t *g) { free(g->normal); free(g->area); free(g->cen); }

/* sdIntRowU (src/lib_sdBem.f90:438-584) / sdIntRowQ (:589-733): one Laplace row.
 * qrow < 0 -> node row for node `src`, else centroid row for element qrow. */
static double lap_row(const orc_body *b, const geom_t *g, int node_src, int qrow, double *rowH, double *rowG) {
  const int nn = b->nnodes, ne = b->nelem;
  memset(rowH, 0, sizeof(double) * nn);
  memset(rowG, 0, sizeof(double) * ne);
  double s[3];
  if (qrow < 0) { for (int j = 0; j < 3; ++j) s[j] = b->xyz[3 * node_src + j]; }
  else { for (int j = 0; j < 3; ++j) s[j] = g->cen[3 * qrow + j]; }
  for (int wall = 1; wall <= b->nwalls; ++wall) {
    const double mul = b->normmul[wall - 1];
    for (int e = 0; e < ne; ++e) {
      if (b->bcid[e] != wall) continue;
      const int nv = b->nvert == 4 ? 4 : 3;
      const int32_t *cn = b->con + (size_t)nv * e;
      const int nno = (nv == 4 && cn[3] >= 0) ? 4 : 3;
      int isrc = 0;
      if (qrow < 0) { for (int j = 0; j < nno; ++j) if (cn[j] == node_src) isrc = j + 1; }
      else if (e == qrow) isrc = nno == 4 ? 5 : 4;
      double out[5] = {0, 0, 0, 0, 0};
      const double *x1 = b->xyz + 3 * cn[0], *x2 = b->xyz + 3 * cn[1], *x3 = b->xyz + 3 * cn[2];
      if (nno == 4) { /* Quad_BEMIntNumAna (src/lib_intAna.f90:153-176) */
        const double *x4 = b->xyz + 3 * cn[3];
        if (isrc == 0) orc_quad_lap_int(x1, x2, x3, x4, s, mul, b->iquad, out);
        else out[0] = orc_si_qua_lap(x1, x2, x3, x4, isrc); /* Quad_BEMIntAna: H = 0 */
      } else if (isrc == 0) {
        double n[3] = {mul * g->normal[3 * e], mul * g->normal[3 * e + 1], mul * g->normal[3 * e + 2]};
        orc_tri_lap_int(x1, x2, x3, s, n, g->area[e], out);
      } else {
        out[0] = orc_si_tri_lap(x1, x2, x3, g->area[e], isrc); /* Triangle_3DSingLapInt, H = 0 */
      }
      for (int j = 0; j < nno; ++j) rowH[cn[j]] += out[1 + j];
      rowG[e] = out[0];
    }
  }
  double err = 0.0, c = 0.0;
  for (int i = 0; i < nn; ++i) c -= rowH[i];
  if (qrow < 0) {
    rowH[node_src] += c;
    /* check u = x, q = n_x (:566-581) */
    for (int j = 0; j < 3; ++j) {
      double cc = 0.0;
      for (int i = 0; i < nn; ++i) cc += rowH[i] * b->xyz[3 * i + j];
      for (int i = 0; i < ne; ++i) cc -= rowG[i] * g->normal[3 * i + j] * b->normmul[b->bcid[i] - 1];
      err = fmax(err, fabs(cc));
    }
  } else {
    err = fabs(0.5 - c);
    const int nv = b->nvert == 4 ? 4 : 3;
    const int32_t *cn = b->con + (size_t)nv * qrow;
    const int nno = (nv == 4 && cn[3] >= 0) ? 4 : 3;
    const double shpf = nno == 4 ? 0.25 : 1.0 / 3.0;
    for (int j = 0; j < nno; ++j) rowH[cn[j]] += shpf * c;
  }
  return err;
}

void orc_assemble_laplace(const orc_body *b, double *H, double *G, double err[2]) {
  ensure_rule(b->itri);
  geom_t g;
  make_geom(b, &g);
  const int nn = b->nnodes, ne = b->nelem, nsp = nn + ne;
  double eu = 0.0, eq = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(max : eu, eq)
  for (int r = 0; r < nsp; ++r) {
    double e = (r < nn) ? lap_row(b, &g, r, -1, H + (size_t)r * nn, G + (size_t)r * ne)
                        : lap_row(b, &g, -1, r - nn, H + (size_t)r * nn, G + (size_t)r * ne);
    if (r < nn) eu = fmax(eu, e); else eq = fmax(eq, e);
  }
  err[0] = eu;
  err[1] = fmax(eu, eq); /* the reference keeps accumulating into the same err (lib_int.f90:252,298) */
  free_geom(&g);
}

/* sdIntRowUStokes (src/lib_Stokes.f90:1904-2245) / sdIntRowQStokes (:1562-1897).
 * rT[6]/rprT[3] have nnodes entries, rG[6]/rprG[3] nelem entries. */
static double stokes_row(const orc_body *b, const geom_t *g, int node_src, int qrow, double extra,
                         double *rT[6], double *rG[6], double *rprT[3], double *rprG[3]) {
  const int nn = b->nnodes, ne = b->nelem;
  for (int k = 0; k < 6; ++k) { memset(rT[k], 0, sizeof(double) * nn); memset(rG[k], 0, sizeof(double) * ne); }
  for (int k = 0; k < 3; ++k) { memset(rprT[k], 0, sizeof(double) * nn); memset(rprG[k], 0, sizeof(double) * ne); }
  const double osemPi = (double)(atanf(1.0f)) * 4.0 * 8.0; /* REAL(4) ATAN(1.0) (:1596) */
  double s[3];
  if (qrow < 0) { for (int j = 0; j < 3; ++j) s[j] = b->xyz[3 * node_src + j]; }
  else { for (int j = 0; j < 3; ++j) s[j] = g->cen[3 * qrow + j]; }
  for (int wall = 1; wall <= b->nwalls; ++wall) {
    const double mul = b->normmul[wall - 1];
    for (int e = 0; e < ne; ++e) {
      if (b->bcid[e] != wall) continue;
      const int nv = b->nvert == 4 ? 4 : 3;
      const int32_t *cn = b->con + (size_t)nv * e;
      const int nno = (nv == 4 && cn[3] >= 0) ? 4 : 3;
      int isrc = 0;
      if (qrow < 0) { for (int j = 0; j < nno; ++j) if (cn[j] == node_src) isrc = j + 1; }
      else if (e == qrow) isrc = nno == 4 ? 5 : 4;
      /* unified layout: G6, prG3, T[6][4], prT[3][4] */
      double acc[48];
      memset(acc, 0, sizeof(acc));
      const double *x1 = b->xyz + 3 * cn[0], *x2 = b->xyz + 3 * cn[1], *x3 = b->xyz + 3 * cn[2];
      if (nno == 4) { /* Quad_StokesIntNumAna (src/lib_intAna.f90:5-37) */
        const double *x4 = b->xyz + 3 * cn[3];
        if (isrc == 0) orc_quad_stokes_int(x1, x2, x3, x4, s, mul, b->iquad, acc);
        else orc_si_qua_stk(x1, x2, x3, x4, isrc, acc); /* Quad_StokesIntAna: T = prT = prG = 0 */
      } else {
        double t36[36];
        memset(t36, 0, sizeof(t36));
        if (isrc == 0) {
          double n[3] = {mul * g->normal[3 * e], mul * g->normal[3 * e + 1], mul * g->normal[3 * e + 2]};
          orc_tri_stokes_int(x1, x2, x3, s, n, g->area[e], t36);
        } else {
          orc_si_tri_stk(x1, x2, x3, g->area[e], isrc, t36); /* Triangle_StokesIntAna: T = prT = prG = 0 */
        }
        for (int k = 0; k < 9; ++k) acc[k] = t36[k];
        for (int k = 0; k < 6; ++k)
          for (int v = 0; v < 3; ++v) acc[9 + 4 * k + v] = t36[9 + 3 * k + v];
        for (int k = 0; k < 3; ++k)
          for (int v = 0; v < 3; ++v) acc[33 + 4 * k + v] = t36[27 + 3 * k + v];
      }
      for (int k = 0; k < 6; ++k)
        for (int v = 0; v < nno; ++v) rT[k][cn[v]] += acc[9 + 4 * k + v];
      for (int k = 0; k < 6; ++k) rG[k][e] = acc[k];
      if (qrow < 0 || isrc == 0) {
        for (int k = 0; k < 3; ++k)
          for (int v = 0; v < nno; ++v) rprT[k][cn[v]] += acc[33 + 4 * k + v];
        for (int k = 0; k < 3; ++k) rprG[k][e] = acc[6 + k];
      }
    }
  }
  double err = -9.99;
  double cx = 0, cy = 0, cz = 0;
  if (qrow < 0) {
    for (int i = 0; i < nn; ++i) { cx -= rT[0][i]; cy -= rT[3][i]; cz -= rT[5][i]; }
    rT[0][node_src] += cx + extra;
    rT[3][node_src] += cy + extra;
    rT[5][node_src] += cz + extra;
    /* check sum n.G = 0 (:2152-2178) */
    cx = cy = cz = 0.0;
    for (int i = 0; i < ne; ++i) {
      double m = b->normmul[b->bcid[i] - 1];
      const double *n = g->normal + 3 * i;
      cx = cx - rG[0][i] * n[0] * m - rG[1][i] * n[1] * m - rG[2][i] * n[2] * m;
      cy = cy - rG[1][i] * n[0] * m - rG[3][i] * n[1] * m - rG[4][i] * n[2] * m;
      cz = cz - rG[2][i] * n[0] * m - rG[4][i] * n[1] * m - rG[5][i] * n[2] * m;
    }
    err = fmax(err, fabs(cx)); err = fmax(err, fabs(cy)); err = fmax(err, fabs(cz));
  } else {
    for (int i = 0; i < nn; ++i) { cx -= rT[1][i]; cy -= rT[4][i]; cz -= rT[2][i]; }
    err = fmax(err, fabs(cx)); err = fmax(err, fabs(cy)); err = fmax(err, fabs(cz));
    cx = cy = cz = 0.0;
    for (int i = 0; i < nn; ++i) { cx -= rT[0][i]; cy -= rT[3][i]; cz -= rT[5][i]; }
    err = fmax(err, fabs(cx - 0.5 * osemPi));
    err = fmax(err, fabs(cy - 0.5 * osemPi));
    err = fmax(err, fabs(cz - 0.5 * osemPi));
    cx += extra; cy += extra; cz += extra;
    const int nv = b->nvert == 4 ? 4 : 3;
    const int32_t *cn = b->con + (size_t)nv * qrow;
    const int nno = (nv == 4 && cn[3] >= 0) ? 4 : 3;
    const double shpf = nno == 4 ? 0.25 : 1.0 / 3.0;
    for (int v = 0; v < nno; ++v) {
      rT[0][cn[v]] += shpf * cx;
      rT[3][cn[v]] += shpf * cy;
      rT[5][cn[v]] += shpf * cz;
    }
    /* pressure correction (:1851-1863) */
    double px = 0, py = 0, pz = 0;
    for (int j = 0; j < nn; ++j) { px += rprT[0][j]; py += rprT[1][j]; pz += rprT[2][j]; }
    for (int v = 0; v < nno; ++v) {
      rprT[0][cn[v]] -= px / nno; rprT[1][cn[v]] -= py / nno; rprT[2][cn[v]] -= pz / nno;
    }
  }
  return err;
}

void orc_assemble_stokes(const orc_body *b, int nrows, const int32_t *rows, double extra_free_term,
                         double *T[6], double *G[6], double *prT[3], double *prG[3], double err[2]) {
  ensure_rule(b->itri);
  geom_t g;
  make_geom(b, &g);
  const int nn = b->nnodes, ne = b->nelem, nsp = nn + ne;
  const int nr = rows ? nrows : nsp;
  double eu = -9.99, eq = -9.99;
#pragma omp parallel reduction(max : eu, eq)
  {
    double *buf = (double *)malloc(sizeof(double) * (size_t)(9 * nn + 9 * ne));
    double *rT[6], *rG[6], *rprT[3], *rprG[3];
    for (int k = 0; k < 6; ++k) { rT[k] = buf + k * nn; rG[k] = buf + 9 * nn + k * ne; }
    for (int k = 0; k < 3; ++k) { rprT[k] = buf + (6 + k) * nn; rprG[k] = buf + 9 * nn + (6 + k) * ne; }
#pragma omp for schedule(dynamic, 2)
    for (int i = 0; i < nr; ++i) {
      int r = rows ? rows[i] : i;
      double e = (r < nn) ? stokes_row(b, &g, r, -1, extra_free_term, rT, rG, rprT, rprG)
                          : stokes_row(b, &g, -1, r - nn, extra_free_term, rT, rG, rprT, rprG);
      if (r < nn) eu = fmax(eu, e); else eq = fmax(eq, e);
      for (int k = 0; k < 6; ++k) {
        memcpy(T[k] + (size_t)i * nn, rT[k], sizeof(double) * nn);
        memcpy(G[k] + (size_t)i * ne, rG[k], sizeof(double) * ne);
      }
      for (int k = 0; k < 3; ++k) {
        if (prT && prT[k]) memcpy(prT[k] + (size_t)i * nn, rprT[k], sizeof(double) * nn);
        if (prG && prG[k]) memcpy(prG[k] + (size_t)i * ne, rprG[k], sizeof(double) * ne);
      }
    }
    free(buf);
  }
  err[0] = eu;
  err[1] = eq;
  free_geom(&g);
}

void orc_stokes_domain_values(const orc_body *b, int npts, const double *pts, const double *u, const double *q,
                              double diff, double *vel, double *p, double *err) {
  ensure_rule(b->itri);
  geom_t g;
  make_geom(b, &g);
  const int nn = b->nnodes, ne = b->nelem;
  const double osemPi = (double)(atanf(1.0f)) * 4.0 * 8.0;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * (size_t)(9 * nn + 9 * ne));
    double *rT[6], *rG[6], *rprT[3], *rprG[3];
    for (int k = 0; k < 6; ++k) { rT[k] = buf + k * nn; rG[k] = buf + 9 * nn + k * ne; }
    for (int k = 0; k < 3; ++k) { rprT[k] = buf + (6 + k) * nn; rprG[k] = buf + 9 * nn + (6 + k) * ne; }
#pragma omp for schedule(dynamic, 2)
    for (int ip = 0; ip < npts; ++ip) {
      memset(buf, 0, sizeof(double) * (size_t)(9 * nn + 9 * ne));
      const double *s = pts + 3 * (size_t)ip;
      /* walls of the subdomain, elements of each wall (:3123-3216) */
      for (int wall = 1; wall <= b->nwalls; ++wall) {
        const double mul = b->normmul[wall - 1];
        for (int e = 0; e < ne; ++e) {
          if (b->bcid[e] != wall) continue;
          const int32_t *cn = b->con + 3 * (size_t)e;
          const double n[3] = {mul * g.normal[3 * e], mul * g.normal[3 * e + 1], mul * g.normal[3 * e + 2]};
          double acc[36];
          memset(acc, 0, sizeof(acc));
          orc_tri_stokes_int(b->xyz + 3 * cn[0], b->xyz + 3 * cn[1], b->xyz + 3 * cn[2], s, n, g.area[e], acc);
          for (int v = 0; v < 3; ++v) {
            for (int k = 0; k < 6; ++k) rT[k][cn[v]] += acc[9 + 3 * k + v];
            for (int k = 0; k < 3; ++k) rprT[k][cn[v]] += acc[27 + 3 * k + v];
          }
          for (int k = 0; k < 6; ++k) rG[k][e] = acc[k];
          for (int k = 0; k < 3; ++k) rprG[k][e] = acc[6 + k];
        }
      }
      /* closure checks (:3255-3300) */
      double er = -9.99, cx = 0.0, cy = 0.0, cz = 0.0;
      for (int i = 0; i < nn; ++i) { cx -= rT[0][i]; cy -= rT[3][i]; cz -= rT[5][i]; }
      er = fmax(er, fabs(cx / osemPi - 1.0));
      er = fmax(er, fabs(cy / osemPi - 1.0));
      er = fmax(er, fabs(cz / osemPi - 1.0));
      cx = cy = cz = 0.0;
      for (int i = 0; i < ne; ++i) {
        const double multi = b->normmul[b->bcid[i] - 1];
        const double *nr = g.normal + 3 * (size_t)i;
        cx = cx - rG[0][i] * nr[0] * multi - rG[1][i] * nr[1] * multi - rG[2][i] * nr[2] * multi;
        cy = cy - rG[1][i] * nr[0] * multi - rG[3][i] * nr[1] * multi - rG[4][i] * nr[2] * multi;
        cz = cz - rG[2][i] * nr[0] * multi - rG[4][i] * nr[1] * multi - rG[5][i] * nr[2] * multi;
      }
      er = fmax(er, fabs(cx));
      er = fmax(er, fabs(cy));
      er = fmax(er, fabs(cz));
      /* velocity and pressure (:3337-3381) */
      double rx = 0.0, ry = 0.0, rz = 0.0, pp = 0.0;
      for (int i = 0; i < nn; ++i) {
        const double ux = u[3 * i], uy = u[3 * i + 1], uz = u[3 * i + 2];
        rx = rx + rT[0][i] * ux + rT[1][i] * uy + rT[2][i] * uz;
        ry = ry + rT[1][i] * ux + rT[3][i] * uy + rT[4][i] * uz;
        rz = rz + rT[2][i] * ux + rT[4][i] * uy + rT[5][i] * uz;
        pp = pp + diff * (rprT[0][i] * ux + rprT[1][i] * uy + rprT[2][i] * uz);
      }
      for (int i = 0; i < ne; ++i) {
        const double qx = q[3 * i], qy = q[3 * i + 1], qz = q[3 * i + 2];
        rx = rx + rG[0][i] * qx / diff + rG[1][i] * qy / diff + rG[2][i] * qz / diff;
        ry = ry + rG[1][i] * qx / diff + rG[3][i] * qy / diff + rG[4][i] * qz / diff;
        rz = rz + rG[2][i] * qx / diff + rG[4][i] * qy / diff + rG[5][i] * qz / diff;
        pp = pp - (rprG[0][i] * qx + rprG[1][i] * qy + rprG[2][i] * qz);
      }
      vel[3 * (size_t)ip] = -rx / osemPi;
      vel[3 * (size_t)ip + 1] = -ry / osemPi;
      vel[3 * (size_t)ip + 2] = -rz / osemPi;
      p[ip] = -pp / osemPi;
      if (err) err[ip] = er;
    }
    free(buf);
  }
  free_geom(&g);
}

/* ------------------------------------------------------------------------------------------
 * LSQR (src/lib_lsqr.f)
 * ---------------------------------------------------------------------------------------- */
static double dnrm2(int64_t n, const double *x) { /* BLAS dnrm2 definition: scaled sum of squares */
  double scale = 0.0, ssq = 1.0;
  for (int64_t i = 0; i < n; ++i) {
    if (x[i] != 0.0) {
      double a = fabs(x[i]);
      if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
      else ssq += (a / scale) * (a / scale);
    }
  }
  return scale * sqrt(ssq);
}
/* lsqr_d2norm (:1236-1263) */
static double d2norm(double a, double b) {
  double scale = fabs(a) + fabs(b);
  if (scale == 0.0) return 0.0;
  return scale * sqrt((a / scale) * (a / scale) + (b / scale) * (b / scale));
}
/* lsqr_aprod modes 1 and 2 (:172-265): y += A x ; x += A^T y */
static void aprod1(const orc_csr *A, const double *x, double *y) {
  for (int64_t i = 0; i < A->m; ++i) {
    double s = 0.0;
    for (int64_t j = A->ia[i]; j < A->ia[i + 1]; ++j) s += A->a[j] * x[A->ja[j]];
    y[i] += s;
  }
}
static void aprod2(const orc_csr *A, double *x, const double *y) {
  for (int64_t j = 0; j < A->m; ++j)
    for (int64_t i = A->ia[j]; i < A->ia[j + 1]; ++i) x[A->ja[i]] += A->a[i] * y[j];
}

void orc_lsqr_precon(orc_csr *A, double *q) {
  for (int64_t i = 0; i < A->n; ++i) q[i] = 0.0; /* lsqr_tmatvec3 (:267-297) */
  for (int64_t j = 0; j < A->m; ++j)
    for (int64_t i = A->ia[j]; i < A->ia[j + 1]; ++i) q[A->ja[i]] += A->a[i] * A->a[i];
  for (int64_t i = 0; i < A->n; ++i) q[i] = 1.0 / sqrt(q[i]);
  for (int64_t i = 0; i < A->m; ++i) /* lsqr_matvec4 (:301-329) */
    for (int64_t ij = A->ia[i]; ij < A->ia[i + 1]; ++ij) A->a[ij] *= q[A->ja[ij]];
}

/* lsqr_LSQR (:470-1207), damp = 0, wantse = false */
static void lsqr_core(const orc_csr *A, double *u, double *x, double *v, double *w, double atol, double btol,
                      double conlim, int itnlim, int *istop_out, int *itn_out) {
  const int64_t m = A->m, n = A->n;
  int itn = 0, istop = 0, nstop = 0;
  double ctol = conlim > 0 ? 1.0 / conlim : 0.0;
  double anorm = 0, acond = 0, dnorm = 0, res2 = 0, psi = 0, xnorm = 0, xnorm1 = 0, cs2 = -1, sn2 = 0, z = 0;
  for (int64_t i = 0; i < n; ++i) { v[i] = 0; x[i] = 0; }
  double alpha = 0.0, beta = dnrm2(m, u);
  if (beta > 0) {
    for (int64_t i = 0; i < m; ++i) u[i] *= 1.0 / beta;
    aprod2(A, v, u);
    alpha = dnrm2(n, v);
  }
  if (alpha > 0) {
    for (int64_t i = 0; i < n; ++i) { v[i] *= 1.0 / alpha; w[i] = v[i]; }
  }
  double arnorm = alpha * beta;
  if (arnorm == 0.0) { *istop_out = 0; *itn_out = 0; return; }
  double rhobar = alpha, phibar = beta, bnorm = beta, rnorm = beta;
  for (;;) {
    ++itn;
    for (int64_t i = 0; i < m; ++i) u[i] *= -alpha;
    aprod1(A, v, u);
    beta = dnrm2(m, u);
    double temp = d2norm(alpha, beta);
    temp = d2norm(temp, 0.0);
    anorm = d2norm(anorm, temp);
    if (beta > 0) {
      for (int64_t i = 0; i < m; ++i) u[i] *= 1.0 / beta;
      for (int64_t i = 0; i < n; ++i) v[i] *= -beta;
      aprod2(A, v, u);
      alpha = dnrm2(n, v);
      if (alpha > 0) for (int64_t i = 0; i < n; ++i) v[i] *= 1.0 / alpha;
    }
    double rhbar1 = rhobar;
    double rho = d2norm(rhbar1, beta);
    double cs = rhbar1 / rho, sn = beta / rho;
    double theta = sn * alpha;
    rhobar = -cs * alpha;
    double phi = cs * phibar;
    phibar = sn * phibar;
    double tau = sn * phi;
    double t1 = phi / rho, t2 = -theta / rho, t3 = 1.0 / rho;
    double dknorm = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double t = w[i];
      x[i] = t1 * t + x[i];
      w[i] = t2 * t + v[i];
      dknorm = (t3 * t) * (t3 * t) + dknorm;
    }
    dknorm = sqrt(dknorm);
    dnorm = d2norm(dnorm, dknorm);
    double delta = sn2 * rho, gambar = -cs2 * rho;
    double rhs = phi - delta * z, zbar = rhs / gambar;
    xnorm = d2norm(xnorm1, zbar);
    double gamma = d2norm(gambar, theta);
    cs2 = gambar / gamma;
    sn2 = theta / gamma;
    z = rhs / gamma;
    xnorm1 = d2norm(xnorm1, z);
    acond = anorm * dnorm;
    res2 = d2norm(res2, psi);
    rnorm = d2norm(res2, phibar);
    arnorm = alpha * fabs(tau);
    double test1 = rnorm / bnorm, test2 = 0.0;
    if (rnorm > 0) test2 = arnorm / (anorm * rnorm);
    double test3 = 1.0 / acond;
    t1 = test1 / (1.0 + anorm * xnorm / bnorm);
    double rtol = btol + atol * anorm * xnorm / bnorm;
    t3 = 1.0 + test3; t2 = 1.0 + test2; t1 = 1.0 + t1;
    if (itn >= itnlim) istop = 5;
    if (t3 <= 1.0) istop = 4;
    if (t2 <= 1.0) istop = 2;
    if (t1 <= 1.0) istop = 1;
    if (test3 <= ctol) istop = 4;
    if (test2 <= atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    if (istop == 0) nstop = 0;
    else { nstop = nstop + 1; if (nstop < 1 && itn < itnlim) istop = 0; }
    if (istop != 0) break;
  }
  *istop_out = istop;
  *itn_out = itn;
}

int orc_lsqr_solve(const orc_csr *A, int *maxit, double eps, int *nit, int *istop_out,
                   const double *q, const double *b, double *x) {
  const int64_t m = A->m, n = A->n;
  if (*maxit == 0) *maxit = (int)(10 * (m + n + 50));
  double *x0 = (double *)malloc(sizeof(double) * n), *wm = (double *)calloc(m, sizeof(double));
  double *u = (double *)malloc(sizeof(double) * m), *v = (double *)malloc(sizeof(double) * n);
  double *w = (double *)malloc(sizeof(double) * n);
  memcpy(x0, x, sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) x[i] = x[i] / q[i];
  aprod1(A, x, wm);
  for (int64_t i = 0; i < m; ++i) u[i] = b[i] - wm[i];
  double nrm_r = dnrm2(m, u), nrm_b = dnrm2(m, b);
  double atol = eps, btol = eps * nrm_b / nrm_r;
  int istop = 0, itn = 0;
  lsqr_core(A, u, x, v, w, atol, btol, 1.0e8, *maxit, &istop, &itn);
  for (int64_t i = 0; i < n; ++i) x[i] = q[i] * x[i] + x0[i];
  free(x0); free(wm); free(u); free(v); free(w);
  *nit = itn;
  if (istop_out) *istop_out = istop;
  return istop > 3 ? istop : 0;
}

/* ------------------------------------------------------------------------------------------
 * Dirichlet systems
 * ---------------------------------------------------------------------------------------- */
void orc_laplace_dirichlet_system(const orc_body *b, const double *H, const double *G, double diff,
                                  double *sys, double *rhs) {
  const int nn = b->nnodes, ne = b->nelem;
  for (int e = 0; e < ne; ++e) {
    const double *Hr = H + (size_t)(nn + e) * nn, *Gr = G + (size_t)(nn + e) * ne;
    for (int j = 0; j < nn; ++j) rhs[(size_t)e * nn + j] = Hr[j];
    for (int j = 0; j < ne; ++j) sys[(size_t)e * ne + j] = b->normmul[b->bcid[j] - 1] / diff * Gr[j];
  }
}

void orc_stokes_dirichlet_system(const orc_body *b, double *const T[6], double *const G[6], int row_offset,
                                 double diff, double *sys, double *rhs) {
  const int nn = b->nnodes, ne = b->nelem;
  /* component (een, en) -> stored symmetric index, formQrow (src/lib_StokesFull.f90:4-85) */
  static const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  const size_t ncol = 3 * (size_t)ne, ncolr = 3 * (size_t)nn;
  for (int e = 0; e < ne; ++e) {
    size_t r = (size_t)(row_offset + e);
    for (int een = 0; een < 3; ++een) {
      size_t row = 3 * (size_t)e + een;
      for (int en = 0; en < 3; ++en) {
        const double *Tr = T[sym[een][en]] + r * nn;
        const double *Gr = G[sym[een][en]] + r * ne;
        for (int j = 0; j < nn; ++j) rhs[row * ncolr + (size_t)en * nn + j] = Tr[j];
        for (int j = 0; j < ne; ++j)
          sys[row * ncol + (size_t)en * ne + j] = -(b->normmul[b->bcid[j] - 1] / diff * Gr[j]);
      }
    }
  }
}

static void dense_to_csr(int64_t m, int64_t n, double *a, orc_csr *A, int64_t *ia, int32_t *ja) {
  for (int64_t i = 0; i <= m; ++i) ia[i] = i * n;
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) ja[i * n + j] = (int32_t)j;
  A->m = m; A->n = n; A->nnz = m * n; A->ia = ia; A->ja = ja; A->a = a;
}

/* StokesInlet3sidesForceCRS / TorqueOmegaCRS (problems = 0: translations e_k, rotations e_k x r) and
 * StokesInlet3sidesTorquePiCRS (problems = 1: the strain fields (0, z, y), (z, 0, x), (y, x, 0) of
 * src/lib_particle.f90:72-85), src/lib_particle.f90:6-476; bounded form: the velocity on wall 2, the
 * body (wall 1) at rest; exterior form: the velocity on the body. FT[6 k + (0..5)] = (F, T) on wall 1. */
int orc_resistance_problems(const orc_body *b, int exterior, double mu, double eps, int maxit, int problems,
                            double FT[36], int nits[6]) {
  const int nn = b->nnodes, ne = b->nelem;
  /* exterior: only the body wall (wall 1) with normals into the body */
  double *nm = (double *)malloc(sizeof(double) * b->nwalls);
  for (int w = 0; w < b->nwalls; ++w) nm[w] = b->normmul[w];
  orc_body bb = *b;
  if (exterior) { for (int w = 0; w < b->nwalls; ++w) nm[w] = -fabs(nm[w]); }
  bb.normmul = nm;
  int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * ne);
  for (int e = 0; e < ne; ++e) rows[e] = nn + e;
  double *T[6], *G[6];
  for (int k = 0; k < 6; ++k) {
    T[k] = (double *)malloc(sizeof(double) * (size_t)ne * nn);
    G[k] = (double *)malloc(sizeof(double) * (size_t)ne * ne);
  }
  double err[2];
  orc_assemble_stokes(&bb, ne, rows, exterior ? 8.0 * ORC_PI : 0.0, T, G, NULL, NULL, err);
  const size_t n3 = 3 * (size_t)ne, nr3 = 3 * (size_t)nn;
  double *sys = (double *)malloc(sizeof(double) * n3 * n3);
  double *rhs = (double *)malloc(sizeof(double) * n3 * nr3);
  orc_stokes_dirichlet_system(&bb, T, G, 0, mu, sys, rhs);
  for (int k = 0; k < 6; ++k) { free(T[k]); free(G[k]); }
  orc_csr A;
  int64_t *ia = (int64_t *)malloc(sizeof(int64_t) * (n3 + 1));
  int32_t *ja = (int32_t *)malloc(sizeof(int32_t) * n3 * n3);
  dense_to_csr((int64_t)n3, (int64_t)n3, sys, &A, ia, ja);
  double *q = (double *)malloc(sizeof(double) * n3);
  orc_lsqr_precon(&A, q);
  double *geo_n = (double *)malloc(sizeof(double) * 3 * ne), *area = (double *)malloc(sizeof(double) * ne);
  double *cen = (double *)malloc(sizeof(double) * 3 * ne);
  const int nvr = b->nvert == 4 ? 4 : 3;
  orc_geometry_n(nn, ne, b->xyz, b->con, nvr, geo_n, area, cen);
  int32_t *node_wall = (int32_t *)calloc(nn, sizeof(int32_t));
  for (int e = 0; e < ne; ++e)
    for (int v = 0; v < nvr; ++v) {
      const int nd = b->con[(size_t)nvr * e + v];
      if (nd >= 0 && b->bcid[e] > node_wall[nd]) node_wall[nd] = b->bcid[e];
    }
  double *ub = (double *)malloc(sizeof(double) * nr3), *rb = (double *)malloc(sizeof(double) * n3);
  double *x = (double *)malloc(sizeof(double) * n3);
  int worst = 0;
  const int nprob = problems == 1 ? 3 : 6;
  for (int k = 0; k < 36; ++k) FT[k] = 0.0;
  for (int k = 0; k < nprob; ++k) {
    /* velocity of every node: translation e_k, rotation e_(k-3) x r, or strain field k */
    for (int i = 0; i < nn; ++i) {
      const double *p = b->xyz + 3 * i;
      double u[3] = {0, 0, 0};
      if (!exterior && node_wall[i] != 2) { for (int c = 0; c < 3; ++c) ub[(size_t)c * nn + i] = 0.0; continue; }
      if (problems == 1) {
        if (k == 0) { u[1] = p[2]; u[2] = p[1]; }
        if (k == 1) { u[0] = p[2]; u[2] = p[0]; }
        if (k == 2) { u[1] = p[0]; u[0] = p[1]; }
      } else if (k < 3) u[k] = 1.0;
      else {
        double om[3] = {0, 0, 0};
        om[k - 3] = 1.0;
        cross(u, om, p);
      }
      for (int c = 0; c < 3; ++c) ub[(size_t)c * nn + i] = u[c];
    }
    for (size_t r = 0; r < n3; ++r) {
      double s = 0.0;
      for (size_t j = 0; j < nr3; ++j) s += rhs[r * nr3 + j] * ub[j];
      rb[r] = s;
    }
    for (size_t i = 0; i < n3; ++i) x[i] = 0.0;
    int mi = maxit, nit = 0, istop = 0;
    int ierr = orc_lsqr_solve(&A, &mi, eps, &nit, &istop, q, rb, x);
    if (ierr > worst) worst = ierr;
    if (nits) nits[k] = nit;
    double F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
    for (int e = 0; e < ne; ++e) {
      if (b->bcid[e] != 1) continue;
      double qe[3] = {x[e], x[ne + e], x[2 * ne + e]}, cp[3];
      for (int c = 0; c < 3; ++c) F[c] += area[e] * qe[c];
      cross(cp, cen + 3 * e, qe);
      for (int c = 0; c < 3; ++c) Tq[c] += area[e] * cp[c];
    }
    for (int c = 0; c < 3; ++c) { FT[6 * k + c] = F[c]; FT[6 * k + 3 + c] = Tq[c]; }
  }
  free(nm); free(rows); free(sys); free(rhs); free(ia); free(ja); free(q);
  free(geo_n); free(area); free(cen); free(ub); free(rb); free(x); free(node_wall);
  return worst;
}

int orc_resistance(const orc_body *b, int exterior, double mu, double eps, int maxit, double FT[36], int nits[6]) {
  return orc_resistance_problems(b, exterior, mu, eps, maxit, 0, FT, nits);
}
