// BEM integral assembly on MI355X: element geometry, Laplace H/G rows, Stokes T/G rows and the
// direct emission of the Stokes resistance system (matrix in its final solver layout plus the
// right-hand sides of the rigid-body problems), FP64 or FP32 regular quadrature.
//
// Work decomposition: one 256-thread workgroup per collocation row. The quadrature rule sits in
// constant memory (every lane of a wave reads the same point, a scalar load). Node-indexed rows
// (H, T, prT) are accumulated element colour by element colour: no two elements of a colour share
// a node, so the scatter needs neither atomics nor a fixed thread-to-element map, and the sums are
// reproducible run to run. The row buffer lives in LDS when it fits, else in the output row.
#include "andr_internal.h"
#include "andr_singular.h"
#include "andr_wx_rules.h"
#include "andr_dunavant_rules.h"
#include "andr_assemble.h"
#include "andr_prof.h"
#include "andr_pairs.h"

#include <algorithm>
#include <cmath>
#include <bitset>
#include <cstdarg>
#include <chrono>
#include <cstring>


// rule id -> (offset, npts): 1..5 Wandzura-Xiao (ITRI), 6..12 Dunavant 3..19 points
static const int k_rule_off[14] = {0, 0, 7, 32, 86, 171, 297, 300, 304, 310, 317, 329, 342, 361};

QuadRule andr_rule(int id) {
  if (id < 1 || id > 12) return {0, 0};
  return {k_rule_off[id], k_rule_off[id + 1] - k_rule_off[id]};
}

static void expand_wx(const andr_wx_orbit *orb, int norb, std::vector<double> &tab) {
  for (int o = 0; o < norb; ++o) {
    const double a = orb[o].a, b = orb[o].b, c = orb[o].c, w = orb[o].w;
    const double P[6][3] = {{a, b, c}, {c, a, b}, {b, c, a}, {b, a, c}, {c, b, a}, {a, c, b}};
    for (int p = 0; p < orb[o].m; ++p) {
      tab.push_back(P[p][0]);
      tab.push_back(P[p][1]);
      tab.push_back(P[p][2]);
      tab.push_back(w);
    }
  }
}

// Gauss-Legendre rules of quadrilaterals (GaussPosWeights, src/mod_gauss.f90:33-391): IQUD 1..8 ->
// N = 2,4,6,8,16,32,40,48 points per direction, stored one after the other (156 nodes), nodes
// ascending. Computed at start-up by Newton iteration on the Legendre three-term recurrence.
static const int k_gl_n[9] = {0, 2, 4, 6, 8, 16, 32, 40, 48};
static const int k_gl_off[9] = {0, 0, 2, 6, 12, 20, 36, 68, 108};
static __constant__ double c_gl[156][2];  // node, weight
static __constant__ int c_gl_n[9], c_gl_off[9];

static void gauss_legendre(int N, double *x, double *w) {
  for (int i = 0; i < N / 2; ++i) {
    double z = std::cos(ANDR_PI * (i + 0.75) / (N + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;  // P_0, P_1
      for (int j = 2; j <= N; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = N * (z * p1 - p0) / (z * z - 1.0);  // P_N'(z)
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    x[N - 1 - i] = z;
    w[i] = w[N - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

int andr_upload_rules(andr_ctx *ctx) {
  {
    double gl[156][2];
    for (int iq = 1; iq <= 8; ++iq) {
      double x[48], w[48];
      gauss_legendre(k_gl_n[iq], x, w);
      for (int i = 0; i < k_gl_n[iq]; ++i) {
        gl[k_gl_off[iq] + i][0] = x[i];
        gl[k_gl_off[iq] + i][1] = w[i];
      }
    }
    ANDR_HIP(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_gl), gl, sizeof(gl)));
    ANDR_HIP(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_gl_n), k_gl_n, sizeof(k_gl_n)));
    ANDR_HIP(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_gl_off), k_gl_off, sizeof(k_gl_off)));
  }
  std::vector<double> tab;
  expand_wx(andr_wx_deg5, 3, tab);
  expand_wx(andr_wx_deg10, 7, tab);
  expand_wx(andr_wx_deg15, 12, tab);
  expand_wx(andr_wx_deg20, 19, tab);
  expand_wx(andr_wx_deg25, 26, tab);
  for (int i = 0; i < 64; ++i)
    for (int j = 0; j < 4; ++j) tab.push_back(andr_dun_points[i][j]);
  if (tab.size() != 4 * ANDR_QTAB_N) return andr_fail(ctx, ANDR_ERR_STATE, "quadrature table size %zu", tab.size());
  std::vector<float> tf(tab.begin(), tab.end());
  int rc = andr_upload_rules_here(ctx, tab.data(), tf.data());
  if (rc) return rc;
  return andr_upload_rules_learned(ctx, tab.data(), tf.data());
}

// bilinear quadrilateral at local coordinates (gi, gj) in [-1, 1]^2 (src/lib_intAna.f90:300-325):
// shape functions, point, and the unnormalised normal x_ksi x x_eta
__device__ inline void quad_point(const double *x1, const double *x2, const double *x3, const double *x4, double gi,
                                  double gj, double fig4[4], double f[3], double an[3]) {
  const double e1m = 1.0 - gi, e1p = 1.0 + gi, e2m = 1.0 - gj, e2p = 1.0 + gj;
  fig4[0] = 0.25 * e1m * e2m;
  fig4[1] = 0.25 * e1p * e2m;
  fig4[2] = 0.25 * e1p * e2p;
  fig4[3] = 0.25 * e1m * e2p;
  double ks[3], et[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    f[j] = fig4[0] * x1[j] + fig4[1] * x2[j] + fig4[2] * x3[j] + fig4[3] * x4[j];
    ks[j] = 0.25 * (-e2m * x1[j] + e2m * x2[j] + e2p * x3[j] - e2p * x4[j]);
    et[j] = 0.25 * (-e1m * x1[j] - e1p * x2[j] + e1p * x3[j] + e1m * x4[j]);
  }
  an[0] = ks[1] * et[2] - et[1] * ks[2];
  an[1] = et[0] * ks[2] - ks[0] * et[2];
  an[2] = ks[0] * et[1] - et[0] * ks[1];
}

// ------------------------------------------------------------------------------------------------
// geometry: CalElementNormal (src/lib_mesh.f90:716-769, first three nodes for either element type),
// Heron area (src/mod_trikotnik.f90:786-802) or, for quadrilaterals, the integral of |J| with the
// N = 8 rule (CalElementArea, src/lib_int.f90:606-721), CalQMesh centre (src/lib_mesh.f90:577-629)
// ------------------------------------------------------------------------------------------------
__global__ void k_geometry(int ne, int nv, const double *__restrict__ xyz, const int *__restrict__ con,
                           const int *__restrict__ bcid, const double *__restrict__ wallmul, double *__restrict__ nrm,
                           double *__restrict__ area, double *__restrict__ cen, double *__restrict__ mul,
                           float4 *__restrict__ cenl) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int *cn = con + (size_t)nv * e;
  const bool quad = nv == 4 && cn[3] >= 0;
  const double *t1 = xyz + 3 * cn[0], *t2 = xyz + 3 * cn[1], *t3 = xyz + 3 * cn[2];
  double v1[3] = {t2[0] - t1[0], t2[1] - t1[1], t2[2] - t1[2]};
  double v2[3] = {t3[0] - t1[0], t3[1] - t1[1], t3[2] - t1[2]};
  double n0 = v1[1] * v2[2] - v1[2] * v2[1];
  double n1 = -v1[0] * v2[2] + v1[2] * v2[0];
  double n2 = v1[0] * v2[1] - v1[1] * v2[0];
  double d = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  nrm[3 * e] = n0 / d;
  nrm[3 * e + 1] = n1 / d;
  nrm[3 * e + 2] = n2 / d;
  if (!quad) {
    auto sq = [](double x) { return x * x; };
    double a = sqrt(sq(t1[0] - t2[0]) + sq(t1[1] - t2[1]) + sq(t1[2] - t2[2]));
    double b = sqrt(sq(t3[0] - t2[0]) + sq(t3[1] - t2[1]) + sq(t3[2] - t2[2]));
    double c = sqrt(sq(t1[0] - t3[0]) + sq(t1[1] - t3[1]) + sq(t1[2] - t3[2]));
    double p = 0.5 * (a + b + c);
    area[e] = sqrt(p * (p - a) * (p - b) * (p - c));
    const double third = 1.0 / 3.0;
    for (int j = 0; j < 3; ++j) cen[3 * e + j] = third * t1[j] + third * t2[j] + third * t3[j];
  } else {
    const double *t4 = xyz + 3 * cn[3];
    const int off = c_gl_off[4], N = c_gl_n[4];
    double A = 0.0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        double fig4[4], f[3], an[3];
        quad_point(t1, t2, t3, t4, c_gl[off + i][0], c_gl[off + j][0], fig4, f, an);
        const double aj = c_gl[off + i][1] * c_gl[off + j][1] * sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
        for (int k = 0; k < 4; ++k) A += aj * fig4[k];
      }
    area[e] = A;
    for (int j = 0; j < 3; ++j) cen[3 * e + j] = 0.25 * t1[j] + 0.25 * t2[j] + 0.25 * t3[j] + 0.25 * t4[j];
  }
  mul[e] = wallmul[bcid[e] - 1];
  {  // centroid and longest edge (the pair-distance scale of the learned rule selection), FP32
    const int nno = quad ? 4 : 3;
    double l2 = 0.0;
    for (int v = 0; v < nno; ++v) {
      const double *a = xyz + 3 * cn[v], *b = xyz + 3 * cn[(v + 1) % nno];
      const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
      l2 = fmax(l2, dx * dx + dy * dy + dz * dz);
    }
    cenl[e] = make_float4((float)cen[3 * e], (float)cen[3 * e + 1], (float)cen[3 * e + 2], (float)(1.0 / l2));
  }
}

// Greedy colouring of elements so that no two elements of a colour share a node.
static int color_elements(int nn, int ne, int nv, const int *con, std::vector<int> &order, std::vector<int> &off) {
  constexpr int MAXC = 256;  // an element sees at most sum over its nodes of (valence - 1) colours
  std::vector<std::bitset<MAXC>> used(nn);
  std::vector<int> col(ne);
  int ncol = 0;
  for (int e = 0; e < ne; ++e) {
    const int *cn = con + (size_t)nv * e;
    const int nno = (nv == 4 && cn[3] >= 0) ? 4 : 3;
    std::bitset<MAXC> m;
    for (int v = 0; v < nno; ++v) m |= used[cn[v]];
    int c = 0;
    while (c < MAXC && m[c]) ++c;
    if (c >= MAXC) return -1;
    col[e] = c;
    ncol = std::max(ncol, c + 1);
    for (int v = 0; v < nno; ++v) used[cn[v]].set(c);
  }
  off.assign(ncol + 1, 0);
  for (int e = 0; e < ne; ++e) off[col[e] + 1]++;
  for (int c = 0; c < ncol; ++c) off[c + 1] += off[c];
  order.resize(ne);
  std::vector<int> pos(off.begin(), off.end() - 1);
  for (int e = 0; e < ne; ++e) order[pos[col[e]]++] = e;
  return 0;
}

// the mesh checks with the message in msg (host only: no context, no device)
static int check_mesh_msg(const andr_mesh *m, char *msg, size_t len) {
  if (!m || m->nnodes <= 0 || m->nelem <= 0 || !m->xyz || !m->con) {
    snprintf(msg, len, "invalid mesh (sizes or null arrays)");
    return ANDR_ERR_ARG;
  }
  if (m->nvert != 0 && m->nvert != 3 && m->nvert != 4) {
    snprintf(msg, len, "nvert must be 0, 3 or 4");
    return ANDR_ERR_ARG;
  }
  const int nn = m->nnodes, ne = m->nelem, nw = m->nwalls > 0 ? m->nwalls : 1, nv = m->nvert == 4 ? 4 : 3;
  for (size_t i = 0; i < (size_t)nv * ne; ++i) {
    const int c = m->con[i] - m->index_base;
    if (nv == 4 && i % 4 == 3 && c < 0) continue;  // a triangle in a quadrilateral-capable mesh
    if (c < 0 || c >= nn) {
      snprintf(msg, len, "connectivity entry %zu out of range", i);
      return ANDR_ERR_ARG;
    }
  }
  if (m->bcid)
    for (int e = 0; e < ne; ++e)
      if (m->bcid[e] < 1 || m->bcid[e] > nw) {
        snprintf(msg, len, "bcid of element %d is %d, outside 1..%d", e, m->bcid[e], nw);
        return ANDR_ERR_ARG;
      }
  if (m->iquad < 0 || m->iquad > 8) {
    snprintf(msg, len, "iquad must be 1..8 (0: default)");
    return ANDR_ERR_ARG;
  }
  if (m->itri < 0 || m->itri > 5) {
    snprintf(msg, len, "itri must be 1..5 (0: default)");
    return ANDR_ERR_ARG;
  }
  return ANDR_OK;
}

int andr_check_mesh(andr_ctx *ctx, const andr_mesh *m) {
  char msg[256];
  const int rc = check_mesh_msg(m, msg, sizeof(msg));
  return rc ? andr_fail(ctx, rc, "%s", msg) : ANDR_OK;
}

extern "C" int andr_mesh_validate(const andr_mesh *mesh, char *msg, int32_t msglen) {
  char buf[256] = "";
  const int rc = check_mesh_msg(mesh, buf, sizeof(buf));
  if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", buf);
  return rc;
}

int andr_body_upload(andr_ctx *ctx, const andr_mesh *m, DevBody &b, bool with_colors) {
  int rc = andr_check_mesh(ctx, m);
  if (rc) return rc;
  const int nn = m->nnodes, ne = m->nelem;
  const int nw = m->nwalls > 0 ? m->nwalls : 1;
  const int nv = m->nvert == 4 ? 4 : 3;
  std::vector<int> con((size_t)nv * ne), bcid(ne, 1);
  for (size_t i = 0; i < (size_t)nv * ne; ++i) {
    con[i] = m->con[i] - m->index_base;
    if (nv == 4 && i % 4 == 3 && con[i] < 0) con[i] = -1;  // a triangle in a quadrilateral-capable mesh
  }
  b.iquad = m->iquad == 0 ? 4 : m->iquad;
  if (m->bcid)
    for (int e = 0; e < ne; ++e) bcid[e] = m->bcid[e];
  std::vector<double> wm(nw, 1.0);
  if (m->normmul)
    for (int w = 0; w < nw; ++w) wm[w] = m->normmul[w];
  b.itri = m->itri == 0 ? 5 : m->itri;
  DevBuf<double> d_wm;
  if (b.xyz.alloc(3 * (size_t)nn, ctx) || b.con.alloc((size_t)nv * ne, ctx) || b.bcid.alloc(ne, ctx) ||
      b.nrm.alloc(3 * (size_t)ne, ctx) || b.area.alloc(ne, ctx) || b.cen.alloc(3 * (size_t)ne, ctx) ||
      b.mul.alloc(ne, ctx) || b.cenl.alloc(ne, ctx) || d_wm.alloc(nw, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "device allocation for mesh");
  ANDR_HIP(ctx, hipMemcpyAsync(b.xyz.p, m->xyz, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(b.con.p, con.data(), sizeof(int) * nv * ne, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(b.bcid.p, bcid.data(), sizeof(int) * ne, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(d_wm.p, wm.data(), sizeof(double) * nw, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_geometry, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream, ne, nv, b.xyz.p, b.con.p,
                     b.bcid.p, d_wm.p, b.nrm.p, b.area.p, b.cen.p, b.mul.p, b.cenl.p);
  ANDR_HIP(ctx, hipGetLastError());
  if (with_colors) {
    std::vector<int> order;
    if (color_elements(nn, ne, nv, con.data(), order, b.color_off))
      return andr_fail(ctx, ANDR_ERR_ARG, "mesh node valence too high for the element colouring");
    if (b.color_elems.alloc(ne) || b.d_color_off.alloc(b.color_off.size()))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "device allocation for colours");
    ANDR_HIP(ctx, hipMemcpyAsync(b.color_elems.p, order.data(), sizeof(int) * ne, hipMemcpyHostToDevice, ctx->stream));
    ANDR_HIP(ctx, hipMemcpyAsync(b.d_color_off.p, b.color_off.data(), sizeof(int) * b.color_off.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
  }
  // No synchronisation here: d_wm returns to the context's cache, and any later use of that block
  // is ordered after k_geometry on the same stream; the host vectors were staged by the copies.
  b.g.nn = nn;
  b.g.ne = ne;
  b.g.nwalls = nw;
  b.g.nv = nv;
  b.g.iquad = b.iquad;
  b.g.xyz = b.xyz.p;
  b.g.con = b.con.p;
  b.g.bcid = b.bcid.p;
  b.g.nrm = b.nrm.p;
  b.g.area = b.area.p;
  b.g.cen = b.cen.p;
  b.g.mul = b.mul.p;
  b.g.cenl = b.cenl.p;
  return ANDR_OK;
}

// ------------------------------------------------------------------------------------------------
// block reductions
// ------------------------------------------------------------------------------------------------
template <int N>
__device__ inline void block_sum(double (&v)[N], double *sh) {
  // sh: N * (blockDim/64) doubles
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double x = v[i];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    v[i] = x;
  }
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) sh[i * nw + wid] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double x = 0.0;
    for (int w = 0; w < nw; ++w) x += sh[i * nw + w];
    v[i] = x;
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// Quadrilateral pairs. Regular: ndx x ndy sub-cells of the bilinear element, N x N Gauss-Legendre
// points each, the normal evaluated pointwise (Quad_BEMIntNum, src/lib_intAna.f90:219-344, with
// max(2, nint(longest side / 1)) sub-cells per direction, :263-276; Quad_StokesIntNum,
// src/lib_Stokes.f90:3754-3940, 3 x 3 sub-cells). Singular: the element split into two triangles
// with the source vertex first, or four about the centre (si_qua_lap / si_qua_stk,
// src/mod_singInt.f90:172-387); the analytic triangle integrals of andr_singular.h do the rest.
// ------------------------------------------------------------------------------------------------
template <class Fn>
__device__ inline void quad_cells(const double *x1, const double *x2, const double *x3, const double *x4, double mul,
                                  int iq, int ndx, int ndy, Fn &&fn) {
  const int off = c_gl_off[iq], N = c_gl_n[iq];
  for (int ix = 1; ix <= ndx; ++ix)
    for (int ie = 1; ie <= ndy; ++ie) {
      const double a = -1.0 + (ix - 1) * 2.0 / ndx, b = -1.0 + ix * 2.0 / ndx;
      const double c = -1.0 + (ie - 1) * 2.0 / ndy, d = -1.0 + ie * 2.0 / ndy;
      const double dex = 0.25 * (b - a) * (d - c);
      for (int i = 0; i < N; ++i) {
        const double gi = a + 0.5 * (c_gl[off + i][0] + 1.0) * (b - a);
        for (int j = 0; j < N; ++j) {
          const double gj = c + 0.5 * (c_gl[off + j][0] + 1.0) * (d - c);
          double fig4[4], f[3], an[3];
          quad_point(x1, x2, x3, x4, gi, gj, fig4, f, an);
          const double len = sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
          const double n[3] = {mul * an[0] / len, mul * an[1] / len, mul * an[2] / len};
          fn(fig4, f, n, dex * c_gl[off + i][1] * c_gl[off + j][1] * len);
        }
      }
    }
}

__device__ inline double dist3(const double *a, const double *b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Laplace pair of a quadrilateral: out = {G, H at nodes 1..4}; isrc 1..4 vertex, 5 centre source
__device__ __noinline__ void quad_laplace_pair(const double *x1, const double *x2, const double *x3,
                                               const double *x4, const double *s, double mul, int iq, int isrc,
                                               double *out) {
  for (int k = 0; k < 5; ++k) out[k] = 0.0;
  if (isrc == 0) {
    const double d13 = fmax(dist3(x1, x2), dist3(x3, x4)), d24 = fmax(dist3(x2, x3), dist3(x4, x1));
    const int ndx = max(2, (int)(d13 + 0.5)), ndy = max(2, (int)(d24 + 0.5));  // minedge = 1
    const double k4pi = 0.25 / ANDR_PI;
    quad_cells(x1, x2, x3, x4, mul, iq, ndx, ndy, [&](const double *fig4, const double *f, const double *n, double w) {
      const double rx = s[0] - f[0], ry = s[1] - f[1], rz = s[2] - f[2];  // r = s - f
      const double ira = rsqrt(rx * rx + ry * ry + rz * rz);
      const double fg = k4pi * ira;
      const double fh = (rx * n[0] + ry * n[1] + rz * n[2]) * ira * ira * fg;
      out[0] += w * fg;
      for (int k = 0; k < 4; ++k) out[1 + k] += w * fh * fig4[k];
    });
    return;
  }
  const double *v[4] = {x1, x2, x3, x4};
  if (isrc <= 4) {
    const double *a = v[isrc - 1], *b = v[isrc % 4], *c = v[(isrc + 1) % 4], *d = v[(isrc + 2) % 4];
    out[0] = andr_sing_lap(a, b, c, a, 1) + andr_sing_lap(a, c, d, a, 1);
  } else {
    for (int k = 0; k < 4; ++k) out[0] += andr_sing_lap(s, v[k], v[(k + 1) % 4], s, 1);
  }
}

// Stokes pair of a quadrilateral: acc = {G6, prG3, T[6][4], prT[3][4]} (T, prT, prG of a singular
// pair are zero)
__device__ __noinline__ void quad_stokes_pair(const double *x1, const double *x2, const double *x3, const double *x4,
                                              const double *s, double mul, int iq, int isrc, double *acc) {
  for (int k = 0; k < 45; ++k) acc[k] = 0.0;
  if (isrc == 0) {
    quad_cells(x1, x2, x3, x4, mul, iq, 3, 3, [&](const double *fig4, const double *f, const double *n, double w) {
      const double rx = f[0] - s[0], ry = f[1] - s[1], rz = f[2] - s[2];  // r = f - s
      const double rxx = rx * rx, rxy = rx * ry, rxz = rx * rz, ryy = ry * ry, ryz = ry * rz, rzz = rz * rz;
      const double ir2 = 1.0 / (rxx + ryy + rzz);
      const double ir = sqrt(ir2);
      const double ir3 = ir2 * ir;
      acc[0] += w * (ir * (1.0 + rxx * ir2));
      acc[1] += w * (rxy * ir3);
      acc[2] += w * (rxz * ir3);
      acc[3] += w * (ir * (1.0 + ryy * ir2));
      acc[4] += w * (ryz * ir3);
      acc[5] += w * (ir * (1.0 + rzz * ir2));
      const double c2 = 2.0 * ir3;
      acc[6] += w * (c2 * rx);
      acc[7] += w * (c2 * ry);
      acc[8] += w * (c2 * rz);
      const double ct = -6.0 * ir2 * ir3 * (rx * n[0] + ry * n[1] + rz * n[2]);
      const double tc[6] = {ct * rxx, ct * rxy, ct * rxz, ct * ryy, ct * ryz, ct * rzz};
      const double c4 = 4.0 * ir3;
      const double pt[3] = {c4 * ((-1.0 + 3.0 * rxx * ir2) * n[0] + 3.0 * rxy * ir2 * n[1] + 3.0 * rxz * ir2 * n[2]),
                            c4 * (3.0 * rxy * ir2 * n[0] + (-1.0 + 3.0 * ryy * ir2) * n[1] + 3.0 * ryz * ir2 * n[2]),
                            c4 * (3.0 * rxz * ir2 * n[0] + 3.0 * ryz * ir2 * n[1] + (-1.0 + 3.0 * rzz * ir2) * n[2])};
      for (int v = 0; v < 4; ++v) {
        const double wv = w * fig4[v];
        for (int k = 0; k < 6; ++k) acc[9 + 4 * k + v] += wv * tc[k];
        for (int k = 0; k < 3; ++k) acc[33 + 4 * k + v] += wv * pt[k];
      }
    });
    return;
  }
  const double *v[4] = {x1, x2, x3, x4};
  double g6[6];
  if (isrc <= 4) {
    const double *a = v[isrc - 1], *b = v[isrc % 4], *c = v[(isrc + 1) % 4], *d = v[(isrc + 2) % 4];
    andr_sing_stk(a, b, c, a, 1, g6);
    for (int k = 0; k < 6; ++k) acc[k] += g6[k];
    andr_sing_stk(a, c, d, a, 1, g6);
    for (int k = 0; k < 6; ++k) acc[k] += g6[k];
  } else {
    for (int q = 0; q < 4; ++q) {
      andr_sing_stk(s, v[q], v[(q + 1) % 4], s, 1, g6);
      for (int k = 0; k < 6; ++k) acc[k] += g6[k];
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Laplace rows: sdIntRowU / sdIntRowQ (src/lib_sdBem.f90:438-733), regular integrals by
// Triangle_3DLapInt (:859-905) with BemKernel (:911-936), singular by the analytic formula.
// ------------------------------------------------------------------------------------------------
// F32Q: the regular integrals in FP32 (relative geometry formed in FP64, quadrature and sums in FP32,
// scaled by the area in FP64), the singular ones in FP64 (the mixed-precision Laplace form)
template <bool ROW_IN_LDS, bool F32Q>
__global__ __launch_bounds__(256) void k_laplace_rows(BodyGeom g, int qoff, int nq, const int *__restrict__ celems,
                                                      const int *__restrict__ coff, int ncolors,
                                                      double *__restrict__ H, double *__restrict__ G,
                                                      double *__restrict__ err_rows, int r0) {
  extern __shared__ double lds[];
  __shared__ double red[16];
  const int nn = g.nn, ne = g.ne;
  const int r = r0 + blockIdx.x;  // source point; its outputs go to row blockIdx.x of H, G
  const int ro = blockIdx.x;
  const bool qrow = r >= nn;
  const int erow = r - nn;
  double s[3];
  if (!qrow)
    for (int j = 0; j < 3; ++j) s[j] = g.xyz[3 * r + j];
  else
    for (int j = 0; j < 3; ++j) s[j] = g.cen[3 * erow + j];
  double *hrow = ROW_IN_LDS ? lds : H + (size_t)ro * nn;
  double *grow = G + (size_t)ro * ne;
  for (int i = threadIdx.x; i < nn; i += blockDim.x) hrow[i] = 0.0;
  __syncthreads();
  const double k4pi = 0.25 / ANDR_PI;
  for (int c = 0; c < ncolors; ++c) {
    for (int idx = coff[c] + threadIdx.x; idx < coff[c + 1]; idx += blockDim.x) {
      const int e = celems[idx];
      const int *cn = g.con + (size_t)g.nv * e;
      const int v0 = cn[0], v1 = cn[1], v2 = cn[2];
      const double *x1 = g.xyz + 3 * v0, *x2 = g.xyz + 3 * v1, *x3 = g.xyz + 3 * v2;
      if (g.nv == 4 && cn[3] >= 0) {  // quadrilateral (Quad_BEMIntNumAna, src/lib_intAna.f90:153-176)
        const int v3 = cn[3];
        int qs = 0;
        if (!qrow) qs = (v0 == r) ? 1 : (v1 == r) ? 2 : (v2 == r) ? 3 : (v3 == r) ? 4 : 0;
        else if (e == erow) qs = 5;
        double out[5];
        quad_laplace_pair(x1, x2, x3, g.xyz + 3 * v3, s, g.mul[e], g.iquad, qs, out);
        grow[e] = out[0];
        hrow[v0] += out[1];
        hrow[v1] += out[2];
        hrow[v2] += out[3];
        hrow[v3] += out[4];
        continue;
      }
      int isrc = 0;
      if (!qrow) isrc = (v0 == r) ? 1 : (v1 == r) ? 2 : (v2 == r) ? 3 : 0;
      else if (e == erow) isrc = 4;
      double gv, h1 = 0.0, h2 = 0.0, h3 = 0.0;
      if (isrc == 0 && F32Q) {
        const double m = g.mul[e];
        const float nx = (float)(m * g.nrm[3 * e]), ny = (float)(m * g.nrm[3 * e + 1]), nz = (float)(m * g.nrm[3 * e + 2]);
        // r = s - f = d - l1 a - l2 b with d = s - x3, a = x1 - x3, b = x2 - x3 (l3 = 1 - l1 - l2)
        const float d0 = (float)(s[0] - x3[0]), d1 = (float)(s[1] - x3[1]), d2 = (float)(s[2] - x3[2]);
        const float a0 = (float)(x1[0] - x3[0]), a1 = (float)(x1[1] - x3[1]), a2 = (float)(x1[2] - x3[2]);
        const float b0 = (float)(x2[0] - x3[0]), b1 = (float)(x2[1] - x3[1]), b2 = (float)(x2[2] - x3[2]);
        const float k4 = (float)k4pi;
        float gs = 0.f, f1 = 0.f, f2 = 0.f, f3 = 0.f;
        for (int q = 0; q < nq; ++q) {
          const float l1 = c_qf[qoff + q][0], l2 = c_qf[qoff + q][1], l3 = c_qf[qoff + q][2], w = c_qf[qoff + q][3];
          const float rx = d0 - l1 * a0 - l2 * b0;
          const float ry = d1 - l1 * a1 - l2 * b1;
          const float rz = d2 - l1 * a2 - l2 * b2;
          const float ira = __frsqrt_rn(rx * rx + ry * ry + rz * rz);
          const float fg = k4 * ira;
          const float wh = w * (rx * nx + ry * ny + rz * nz) * ira * ira * fg;
          gs += w * fg;
          f1 += wh * l1;
          f2 += wh * l2;
          f3 += wh * l3;
        }
        const double A = g.area[e];
        gv = (double)gs * A;
        h1 = (double)f1 * A;
        h2 = (double)f2 * A;
        h3 = (double)f3 * A;
      } else if (isrc == 0) {
        const double m = g.mul[e];
        const double nx = m * g.nrm[3 * e], ny = m * g.nrm[3 * e + 1], nz = m * g.nrm[3 * e + 2];
        // r = s - f with f = l1 x1 + l2 x2 + l3 x3
        double gs = 0.0;
        for (int q = 0; q < nq; ++q) {
          const double l1 = c_qd[qoff + q][0], l2 = c_qd[qoff + q][1], l3 = c_qd[qoff + q][2], w = c_qd[qoff + q][3];
          const double rx = s[0] - (l1 * x1[0] + l2 * x2[0] + l3 * x3[0]);
          const double ry = s[1] - (l1 * x1[1] + l2 * x2[1] + l3 * x3[1]);
          const double rz = s[2] - (l1 * x1[2] + l2 * x2[2] + l3 * x3[2]);
          const double ira = rsqrt(rx * rx + ry * ry + rz * rz);
          const double fg = k4pi * ira;
          const double fh = (rx * nx + ry * ny + rz * nz) * ira * ira * fg;
          gs += w * fg;
          const double wh = w * fh;
          h1 += wh * l1;
          h2 += wh * l2;
          h3 += wh * l3;
        }
        const double A = g.area[e];
        gv = gs * A;
        h1 *= A;
        h2 *= A;
        h3 *= A;
      } else {
        gv = andr_sing_lap(x1, x2, x3, s, isrc);
      }
      grow[e] = gv;
      hrow[v0] += h1;
      hrow[v1] += h2;
      hrow[v2] += h3;
    }
    __syncthreads();
  }
  // free term c = -sum_row H (rigid-body trick, src/lib_sdBem.f90:552-562, 714-731)
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < nn; i += blockDim.x) acc[0] -= hrow[i];
  block_sum<1>(acc, red);
  const double cfree = acc[0];
  if (threadIdx.x == 0) {
    if (!qrow) hrow[r] += cfree;
    else {  // spread 1/3 (triangle) or 1/4 (quadrilateral) on the element's nodes
      const int *cn = g.con + (size_t)g.nv * erow;
      const int nno = (g.nv == 4 && cn[3] >= 0) ? 4 : 3;
      const double shpf = nno == 4 ? 0.25 : 1.0 / 3.0;
      for (int v = 0; v < nno; ++v) hrow[cn[v]] += shpf * cfree;
    }
  }
  __syncthreads();
  double err = 0.0;
  if (!qrow) {
    // check u = x, q = n_x (src/lib_sdBem.f90:566-581)
    double ck[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nn; i += blockDim.x)
      for (int j = 0; j < 3; ++j) ck[j] += hrow[i] * g.xyz[3 * i + j];
    for (int i = threadIdx.x; i < ne; i += blockDim.x)
      for (int j = 0; j < 3; ++j) ck[j] -= grow[i] * g.nrm[3 * i + j] * g.mul[i];
    block_sum<3>(ck, red);
    err = fmax(fabs(ck[0]), fmax(fabs(ck[1]), fabs(ck[2])));
  } else {
    err = fabs(0.5 - cfree);
  }
  if (ROW_IN_LDS) {
    double *out = H + (size_t)ro * nn;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) out[i] = hrow[i];
  }
  if (threadIdx.x == 0) err_rows[ro] = err;
}

int andr_launch_laplace_rows(andr_ctx *ctx, DevBody &b, double *H, double *G, double *err_rows, int r0, int nrows,
                             bool fp32_quad) {
  QuadRule qr = andr_rule(b.itri);
  const int nsp = nrows < 0 ? b.g.nn + b.g.ne - r0 : nrows;  // rows r0 .. r0 + nsp - 1
  const int ncol = (int)b.color_off.size() - 1;
  const size_t lds = sizeof(double) * b.g.nn;
  const double evals = (double)nsp * b.g.ne * qr.npts;
  const int tok = andr_prof_begin(ctx, fp32_quad ? "laplace_rows_fp32" : "laplace_rows",
                                  8.0 * nsp * ((double)b.g.nn + b.g.ne), evals * ANDR_FLOPS_PER_LAPLACE_EVAL);
  const bool in_lds = lds <= 128 * 1024;
  auto kern = in_lds ? (fp32_quad ? k_laplace_rows<true, true> : k_laplace_rows<true, false>)
                     : (fp32_quad ? k_laplace_rows<false, true> : k_laplace_rows<false, false>);
  if (in_lds) ANDR_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(nsp), dim3(256), in_lds ? lds : 0, ctx->stream, b.g, qr.offset, qr.npts,
                     b.color_elems.p, b.d_color_off.p, ncol, H, G, err_rows, r0);
  andr_prof_end(ctx, tok);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

// ------------------------------------------------------------------------------------------------
// Stokes rows: sdIntRowUStokes / sdIntRowQStokes (src/lib_Stokes.f90:1562-2245), regular integrals
// by Triangle_3DStokesInt (:2341-2446) with StokesKernel (:2450-2544).
// Outputs: T[6][nsp][nn], G[6][nsp][ne], prT[3][nsp][nn], prG[3][nsp][ne].
// ------------------------------------------------------------------------------------------------
template <bool ROW_IN_LDS>
__global__ __launch_bounds__(256) void k_stokes_rows(BodyGeom g, int qoff, int nq, const int *__restrict__ celems,
                                                     const int *__restrict__ coff, int ncolors, double extra,
                                                     double *__restrict__ T, double *__restrict__ G,
                                                     double *__restrict__ prT, double *__restrict__ prG,
                                                     double *__restrict__ err_rows, int r0, int nrows) {
  extern __shared__ double lds[];
  __shared__ double red[6 * 4];
  const int nn = g.nn, ne = g.ne;
  const int r = r0 + blockIdx.x;  // source point; its outputs go to row blockIdx.x of the matrices
  const int ro = blockIdx.x;
  const bool qrow = r >= nn;
  const int erow = r - nn;
  const size_t mstrT = (size_t)nrows * nn, mstrG = (size_t)nrows * ne;
  double s[3];
  if (!qrow)
    for (int j = 0; j < 3; ++j) s[j] = g.xyz[3 * r + j];
  else
    for (int j = 0; j < 3; ++j) s[j] = g.cen[3 * erow + j];
  // node rows: 6 T components then 3 prT components
  auto trow = [&](int k) -> double * {
    if (ROW_IN_LDS) return lds + (size_t)k * nn;
    return k < 6 ? T + k * mstrT + (size_t)ro * nn : prT + (k - 6) * mstrT + (size_t)ro * nn;
  };
  for (int k = 0; k < 9; ++k) {
    double *p = trow(k);
    for (int i = threadIdx.x; i < nn; i += blockDim.x) p[i] = 0.0;
  }
  __syncthreads();
  for (int c = 0; c < ncolors; ++c) {
    for (int idx = coff[c] + threadIdx.x; idx < coff[c + 1]; idx += blockDim.x) {
      const int e = celems[idx];
      const int *cn = g.con + (size_t)g.nv * e;
      const int vv[3] = {cn[0], cn[1], cn[2]};
      const double *x1 = g.xyz + 3 * vv[0], *x2 = g.xyz + 3 * vv[1], *x3 = g.xyz + 3 * vv[2];
      if (g.nv == 4 && cn[3] >= 0) {  // quadrilateral (Quad_StokesIntNumAna, src/lib_intAna.f90:5-37)
        const int v4[4] = {cn[0], cn[1], cn[2], cn[3]};
        int qs = 0;
        if (!qrow) {
          for (int j = 0; j < 4; ++j)
            if (v4[j] == r) qs = j + 1;
        } else if (e == erow) {
          qs = 5;
        }
        double acc[45];
        quad_stokes_pair(x1, x2, x3, g.xyz + 3 * v4[3], s, g.mul[e], g.iquad, qs, acc);
        for (int k = 0; k < 6; ++k) G[k * mstrG + (size_t)ro * ne + e] = acc[k];
        for (int k = 0; k < 3; ++k) prG[k * mstrG + (size_t)ro * ne + e] = acc[6 + k];
        for (int k = 0; k < 6; ++k) {
          double *p = trow(k);
          for (int v = 0; v < 4; ++v) p[v4[v]] += acc[9 + 4 * k + v];
        }
        for (int k = 0; k < 3; ++k) {
          double *p = trow(6 + k);
          for (int v = 0; v < 4; ++v) p[v4[v]] += acc[33 + 4 * k + v];
        }
        continue;
      }
      int isrc = 0;
      if (!qrow) isrc = (vv[0] == r) ? 1 : (vv[1] == r) ? 2 : (vv[2] == r) ? 3 : 0;
      else if (e == erow) isrc = 4;
      double Gs[6] = {0, 0, 0, 0, 0, 0}, pG[3] = {0, 0, 0};
      double Tv[6][3] = {{0}}, pT[3][3] = {{0}};
      if (isrc == 0) {
        const double m = g.mul[e];
        stokes_tri_full(s, x1, x2, x3, m * g.nrm[3 * e], m * g.nrm[3 * e + 1], m * g.nrm[3 * e + 2], g.area[e], qoff,
                        nq, Gs, pG, Tv, pT);
      } else {
        andr_sing_stk(x1, x2, x3, s, isrc, Gs);  // T, prT, prG of a singular pair are zero
      }
      for (int k = 0; k < 6; ++k) G[k * mstrG + (size_t)ro * ne + e] = Gs[k];
      for (int k = 0; k < 3; ++k) prG[k * mstrG + (size_t)ro * ne + e] = pG[k];
      for (int k = 0; k < 6; ++k) {
        double *p = trow(k);
        for (int v = 0; v < 3; ++v) p[vv[v]] += Tv[k][v];
      }
      for (int k = 0; k < 3; ++k) {
        double *p = trow(6 + k);
        for (int v = 0; v < 3; ++v) p[vv[v]] += pT[k][v];
      }
    }
    __syncthreads();
  }
  // row sums of the six T components and of the three prT components
  double sums[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 6; ++k) {
    const double *p = trow(k);
    for (int i = threadIdx.x; i < nn; i += blockDim.x) sums[k] += p[i];
  }
  block_sum<6>(sums, red);
  double psum[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    const double *p = trow(6 + k);
    for (int i = threadIdx.x; i < nn; i += blockDim.x) psum[k] += p[i];
  }
  block_sum<3>(psum, red);
  const double osemPi = (double)(atanf(1.0f)) * 4.0 * 8.0;  // REAL(4) ATAN(1.0) (src/lib_Stokes.f90:1596)
  double err = -9.99;
  const double cx = -sums[0], cy = -sums[3], cz = -sums[5];
  if (!qrow) {
    if (threadIdx.x == 0) {
      trow(0)[r] += cx + extra;
      trow(3)[r] += cy + extra;
      trow(5)[r] += cz + extra;
    }
    // sum n.G check (src/lib_Stokes.f90:2152-2178)
    double ck[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
      const double m = g.mul[i];
      const double n0 = g.nrm[3 * i] * m, n1 = g.nrm[3 * i + 1] * m, n2 = g.nrm[3 * i + 2] * m;
      const size_t o = (size_t)ro * ne + i;
      const double gxx = G[o], gxy = G[mstrG + o], gxz = G[2 * mstrG + o], gyy = G[3 * mstrG + o],
                   gyz = G[4 * mstrG + o], gzz = G[5 * mstrG + o];
      ck[0] -= gxx * n0 + gxy * n1 + gxz * n2;
      ck[1] -= gxy * n0 + gyy * n1 + gyz * n2;
      ck[2] -= gxz * n0 + gyz * n1 + gzz * n2;
    }
    block_sum<3>(ck, red);
    err = fmax(err, fmax(fabs(ck[0]), fmax(fabs(ck[1]), fabs(ck[2]))));
  } else {
    // checks of src/lib_Stokes.f90:1781-1812, then the free term spread 1/3 on the element nodes
    err = fmax(err, fmax(fabs(sums[1]), fmax(fabs(sums[4]), fabs(sums[2]))));
    err = fmax(err, fabs(cx - 0.5 * osemPi));
    err = fmax(err, fabs(cy - 0.5 * osemPi));
    err = fmax(err, fabs(cz - 0.5 * osemPi));
    if (threadIdx.x == 0) {
      const int *cn = g.con + (size_t)g.nv * erow;
      const int nno = (g.nv == 4 && cn[3] >= 0) ? 4 : 3;
      const double shpf = nno == 4 ? 0.25 : 1.0 / 3.0;
      for (int v = 0; v < nno; ++v) {
        trow(0)[cn[v]] += shpf * (cx + extra);
        trow(3)[cn[v]] += shpf * (cy + extra);
        trow(5)[cn[v]] += shpf * (cz + extra);
      }
      // pressure correction (src/lib_Stokes.f90:1851-1863)
      for (int v = 0; v < nno; ++v)
        for (int k = 0; k < 3; ++k) trow(6 + k)[cn[v]] -= psum[k] / nno;
    }
  }
  __syncthreads();
  if (ROW_IN_LDS) {
    for (int k = 0; k < 9; ++k) {
      double *out = k < 6 ? T + k * mstrT + (size_t)ro * nn : prT + (k - 6) * mstrT + (size_t)ro * nn;
      const double *p = lds + (size_t)k * nn;
      for (int i = threadIdx.x; i < nn; i += blockDim.x) out[i] = p[i];
    }
  }
  if (threadIdx.x == 0) err_rows[ro] = err;
}

// ------------------------------------------------------------------------------------------------
// Velocity and pressure at points of the flow domain: GetStokesDomainValue (src/lib_Stokes.f90:
// 3043-3390). One workgroup per point; the threads take elements, integrate each with the mesh's
// rule (the point is off the boundary, so every pair is regular) and contract the integrals with
// the nodal velocities u (3 per node) and the element tractions q (3 per element) at once, so that
// no integral row is stored. out[5 p + 0..2] = velocity, out[5 p + 3] = pressure, out[5 p + 4] =
// the closure error of the point's integrals (:3255-3300): the T closure -sum_i T_ii = 8 pi and
// sum_e G n_e = 0 hold inside the domain. osem = 8 pi as the reference forms it.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stokes_domain(BodyGeom g, int qoff, int nq, const double *__restrict__ pts,
                                                       const double *__restrict__ u, const double *__restrict__ q,
                                                       double diff, double osem, double *__restrict__ out) {
  constexpr int NV = 14;  // T u (3), G q (3), prT u, prG q, T closure (3), G n (3)
  __shared__ double red[NV][4];
  const int ip = blockIdx.x;
  const double s[3] = {pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2]};
  double a[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) a[i] = 0.0;
  for (int e = threadIdx.x; e < g.ne; e += blockDim.x) {
    const int vv[3] = {g.con[3 * e], g.con[3 * e + 1], g.con[3 * e + 2]};
    const double *x1 = g.xyz + 3 * vv[0], *x2 = g.xyz + 3 * vv[1], *x3 = g.xyz + 3 * vv[2];
    const double m = g.mul[e];
    const double nx = m * g.nrm[3 * e], ny = m * g.nrm[3 * e + 1], nz = m * g.nrm[3 * e + 2];
    double Gs[6], pG[3], Tv[6][3], pT[3][3];
    stokes_tri_full(s, x1, x2, x3, nx, ny, nz, g.area[e], qoff, nq, Gs, pG, Tv, pT);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double ux = u[3 * vv[v]], uy = u[3 * vv[v] + 1], uz = u[3 * vv[v] + 2];
      a[0] += Tv[0][v] * ux + Tv[1][v] * uy + Tv[2][v] * uz;
      a[1] += Tv[1][v] * ux + Tv[3][v] * uy + Tv[4][v] * uz;
      a[2] += Tv[2][v] * ux + Tv[4][v] * uy + Tv[5][v] * uz;
      a[6] += pT[0][v] * ux + pT[1][v] * uy + pT[2][v] * uz;
      a[8] += Tv[0][v];
      a[9] += Tv[3][v];
      a[10] += Tv[5][v];
    }
    const double qx = q[3 * e], qy = q[3 * e + 1], qz = q[3 * e + 2];
    a[3] += Gs[0] * qx + Gs[1] * qy + Gs[2] * qz;
    a[4] += Gs[1] * qx + Gs[3] * qy + Gs[4] * qz;
    a[5] += Gs[2] * qx + Gs[4] * qy + Gs[5] * qz;
    a[7] += pG[0] * qx + pG[1] * qy + pG[2] * qz;
    a[11] += Gs[0] * nx + Gs[1] * ny + Gs[2] * nz;
    a[12] += Gs[1] * nx + Gs[3] * ny + Gs[4] * nz;
    a[13] += Gs[2] * nx + Gs[4] * ny + Gs[5] * nz;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double v = a[i];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[i][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = (red[i][0] + red[i][1]) + (red[i][2] + red[i][3]);
    double *o = out + 5 * (size_t)ip;
    for (int c = 0; c < 3; ++c) o[c] = -(t[c] + t[3 + c] / diff) / osem;
    o[3] = -(diff * t[6] - t[7]) / osem;
    double err = -9.99;
    for (int c = 0; c < 3; ++c) err = fmax(err, fabs(-t[8 + c] / osem - 1.0));
    for (int c = 0; c < 3; ++c) err = fmax(err, fabs(t[11 + c]));
    o[4] = err;
  }
}

// Surface pressure at the element centres: pressureStokes (src/lib_Stokes.f90:4-94) with the centroid
// rows of prT and prG (sdIntRowQStokes, :1562-1897). One workgroup per element r; the pair (r, r)
// is singular and has no pressure integrals. The row-sum correction of prT (:1851-1863) moves the
// row sum onto the three nodes of element r; with u_c = the mean of those nodes' velocities
// (InterpolateToQMesh) the row contracts to sum prT u - rowsum . u_c, so it is not formed:
// p_r = -2 (diff (sum prT u - rowsum . u_c) - sum prG q) / 8 pi.
__global__ __launch_bounds__(256) void k_stokes_surface_pressure(BodyGeom g, int qoff, int nq,
                                                                 const double *__restrict__ u,
                                                                 const double *__restrict__ q, double diff,
                                                                 double osem, double *__restrict__ p) {
  constexpr int NV = 5;  // prT u, prT row sums (3), prG q
  __shared__ double red[NV][4];
  const int r = blockIdx.x;
  const double s[3] = {g.cen[3 * r], g.cen[3 * r + 1], g.cen[3 * r + 2]};
  double a[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = threadIdx.x; e < g.ne; e += blockDim.x) {
    if (e == r) continue;
    const int vv[3] = {g.con[3 * e], g.con[3 * e + 1], g.con[3 * e + 2]};
    const double m = g.mul[e];
    double Gs[6], pG[3], Tv[6][3], pT[3][3];
    stokes_tri_full(s, g.xyz + 3 * vv[0], g.xyz + 3 * vv[1], g.xyz + 3 * vv[2], m * g.nrm[3 * e],
                    m * g.nrm[3 * e + 1], m * g.nrm[3 * e + 2], g.area[e], qoff, nq, Gs, pG, Tv, pT);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double *uv = u + 3 * vv[v];
      a[0] += pT[0][v] * uv[0] + pT[1][v] * uv[1] + pT[2][v] * uv[2];
      a[1] += pT[0][v];
      a[2] += pT[1][v];
      a[3] += pT[2][v];
    }
    a[4] += pG[0] * q[3 * e] + pG[1] * q[3 * e + 1] + pG[2] * q[3 * e + 2];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double v = a[i];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[i][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = (red[i][0] + red[i][1]) + (red[i][2] + red[i][3]);
    const int c0 = g.con[3 * r], c1 = g.con[3 * r + 1], c2 = g.con[3 * r + 2];
    double pu = t[0];
    for (int k = 0; k < 3; ++k) pu -= t[1 + k] * ((u[3 * c0 + k] + u[3 * c1 + k] + u[3 * c2 + k]) / 3.0);
    p[r] = -(diff * pu - t[4]) / osem * 2.0;
  }
}

int andr_launch_stokes_surface_pressure(andr_ctx *ctx, DevBody &b, const double *u, const double *q, double diff,
                                        double *p) {
  const QuadRule qr = andr_rule(b.itri);
  const double osem = (double)atanf(1.0f) * 4.0 * 8.0;  // REAL(4) ATAN(1.0) (src/lib_Stokes.f90:17)
  const double evals = (double)b.g.ne * (b.g.ne - 1.0) * qr.npts;
  const int tok = andr_prof_begin(ctx, "stokes_surface_pressure", 8.0 * (3.0 * b.g.nn + 4.0 * b.g.ne), evals * 175.0);
  hipLaunchKernelGGL(k_stokes_surface_pressure, dim3(b.g.ne), dim3(256), 0, ctx->stream, b.g, qr.offset, qr.npts, u,
                     q, diff, osem, p);
  andr_prof_end(ctx, tok);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

int andr_launch_stokes_domain(andr_ctx *ctx, DevBody &b, int npts, const double *pts, const double *u, const double *q,
                              double diff, double *out) {
  if (npts <= 0) return ANDR_OK;
  const QuadRule qr = andr_rule(b.itri);
  const double osem = (double)atanf(1.0f) * 4.0 * 8.0;  // REAL(4) ATAN(1.0) (src/lib_Stokes.f90:3089)
  const double evals = (double)npts * b.g.ne * qr.npts;
  const int tok = andr_prof_begin(ctx, "stokes_domain", 8.0 * (3.0 * npts + 5.0 * npts), evals * 175.0);
  hipLaunchKernelGGL(k_stokes_domain, dim3(npts), dim3(256), 0, ctx->stream, b.g, qr.offset, qr.npts, pts, u, q, diff,
                     osem, out);
  andr_prof_end(ctx, tok);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

int andr_launch_stokes_rows(andr_ctx *ctx, DevBody &b, double extra, double *T, double *G, double *prT, double *prG,
                            double *err_rows, int r0, int nrows) {
  QuadRule qr = andr_rule(b.itri);
  const int nsp = nrows < 0 ? b.g.nn + b.g.ne - r0 : nrows;  // rows r0 .. r0 + nsp - 1
  const int ncol = (int)b.color_off.size() - 1;
  const size_t lds = sizeof(double) * 9 * (size_t)b.g.nn;
  const double evals = (double)nsp * b.g.ne * qr.npts;
  const int tok = andr_prof_begin(ctx, "stokes_rows", 8.0 * 9 * nsp * ((double)b.g.nn + b.g.ne),
                                  evals * ANDR_FLOPS_PER_STOKES_EVAL);
  if (lds <= 144 * 1024) {
    ANDR_HIP(ctx, hipFuncSetAttribute((const void *)k_stokes_rows<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_stokes_rows<true>, dim3(nsp), dim3(256), lds, ctx->stream, b.g, qr.offset, qr.npts,
                       b.color_elems.p, b.d_color_off.p, ncol, extra, T, G, prT, prG, err_rows, r0, nsp);
  } else {
    hipLaunchKernelGGL(k_stokes_rows<false>, dim3(nsp), dim3(256), 0, ctx->stream, b.g, qr.offset, qr.npts,
                       b.color_elems.p, b.d_color_off.p, ncol, extra, T, G, prT, prG, err_rows, r0, nsp);
  }
  andr_prof_end(ctx, tok);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

// ------------------------------------------------------------------------------------------------
// Stokes resistance system, emitted directly in the solver layout (blocked rows and columns):
//   A[(c ne + r), (en ne + e)] = -(mul_e / mu) G_{c,en}(r, e)    (formQrow, src/lib_StokesFull.f90:4-85)
//   B[(c ne + r), k] = sum_{e,v,en} T_{c,en}(r, e, v) u_k(node(e,v))_en + free term
// with u_k the rigid-body velocity of problem k (translation e_k, k < 3; rotation
// e_{k-3} x (x - center)) on the moving nodes and 0 elsewhere. Since u_k is affine in the node
// position, the contraction needs only the moments sum_v T_v and sum_v T_v (x_v - center) of each
// pair (24 sums per row) instead of the nodal velocities. Only centroid rows (the equations of an
// all-Dirichlet problem) are formed; the T matrix is never stored. One workgroup per (job, row); a
// job is one body of a batch.
// ------------------------------------------------------------------------------------------------
template <class CT, class ST>
__global__ __launch_bounds__(256) void k_stokes_system(const SysJob *__restrict__ jobs,
                                                       const int2 *__restrict__ rowmap) {
  constexpr int NACC = 6 + 18 + 6;  // sum T (moving), sum T r (moving), sum T (all nodes)
  // Per-thread partial sums of the T moments in the quadrature type (a few dozen terms each; the
  // FP32 form's terms carry FP32 quadrature error already), summed over the block in FP64.
  using AT = CT;
  __shared__ double red[NACC * 4];
  const int2 jr = rowmap[blockIdx.x];
  const SysJob &J = jobs[jr.x];
  const BodyGeom &g = J.g;
  const int r = jr.y;
  const int ne = g.ne;
  const double s[3] = {g.cen[3 * r], g.cen[3 * r + 1], g.cen[3 * r + 2]};
  AT acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = 0;
  const double inv_mu = 1.0 / J.mu;
  ST *A = J.template Aptr<ST>();
  const bool tiled = J.lda == 0;  // symmetric-block layout: S_q = sc G_q only (no null-space term, no M^-1)
  const int KB = (ne + 15) / 16;
  ST *arow0 = A + (tiled ? 0 : (size_t)(0 * ne + r) * J.lda);
  ST *arow1 = A + (tiled ? 0 : (size_t)(1 * ne + r) * J.lda);
  ST *arow2 = A + (tiled ? 0 : (size_t)(2 * ne + r) * J.lda);
  // one (row, element) pair: matrix entries and the T moments of the right-hand sides
  auto pair = [&](int e) {
    const int vv[3] = {g.con[3 * e], g.con[3 * e + 1], g.con[3 * e + 2]};
    const double *x1 = g.xyz + 3 * vv[0], *x2 = g.xyz + 3 * vv[1], *x3 = g.xyz + 3 * vv[2];
    double Gs[6];
    AT Tv[6][3];
    const double m = g.mul[e];
    if (e == r) {
      andr_sing_stk(x1, x2, x3, s, 4, Gs);
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int v = 0; v < 3; ++v) Tv[k][v] = 0.0;
    } else {
      int qoff = J.qoff, nq = J.nq;
      if (J.cls) {  // learned form: Dunavant class 0..6 or the fixed-order rule (class 7)
        const int c = J.cls[(size_t)r * ne + e];
        if (c < 7) {  // Dunavant offsets {0,3,7,13,20,32,45} and sizes {3,4,6,7,12,13,19}, 6 bits each
          qoff = ANDR_DUN_OFF0 + (int)((0x2D8143470C0ull >> (6 * c)) & 63);
          nq = (int)((0x1334C1C6103ull >> (6 * c)) & 63);
        }
      }
      stokes_pair_regular<CT>(s, x1, x2, x3, m * g.nrm[3 * e], m * g.nrm[3 * e + 1], m * g.nrm[3 * e + 2],
                              g.area[e], qoff, nq, Gs, Tv);
    }
    const double sc = -m * inv_mu;
    if (tiled) {
#pragma unroll
      for (int q = 0; q < 6; ++q) A[andr_sym_off(q, r, e, KB)] = (ST)(sc * Gs[q]);
    } else {
      // full 3 x 3 block, optionally with the rank-one null-space term d_w A_r n_r n_e^T between
      // elements of one closed wall and the right preconditioner M^-1 (block (r, e) times minv[e])
      double dfl = 0.0;
      if (J.defl && g.bcid[e] == g.bcid[r]) dfl = J.defl[g.bcid[r]] * g.area[r];
      const double ne3[3] = {g.nrm[3 * e], g.nrm[3 * e + 1], g.nrm[3 * e + 2]};
      const double nr3[3] = {dfl * g.nrm[3 * r], dfl * g.nrm[3 * r + 1], dfl * g.nrm[3 * r + 2]};
      ST *ar[3] = {arow0, arow1, arow2};
      double blk[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int en = 0; en < 3; ++en) blk[c][en] = sc * Gs[sym3(c, en)] + nr3[c] * ne3[en];
      if (J.minv) {
        const double *mi = J.minv + 9 * (size_t)e;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double b0 = blk[c][0], b1 = blk[c][1], b2 = blk[c][2];
#pragma unroll
          for (int en = 0; en < 3; ++en) blk[c][en] = b0 * mi[en] + b1 * mi[3 + en] + b2 * mi[6 + en];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int en = 0; en < 3; ++en) ar[c][(size_t)en * ne + e] = (ST)blk[c][en];
    }
    // moments of T over the element's moving nodes
    const double *xv[3] = {x1, x2, x3};
    AT mv[3], rv[3][3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      mv[v] = J.moving ? (AT)J.moving[vv[v]] : (AT)1;
#pragma unroll
      for (int d = 0; d < 3; ++d) rv[v][d] = mv[v] * (AT)(xv[v][d] - J.center[d]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      acc[k] += mv[0] * Tv[k][0] + mv[1] * Tv[k][1] + mv[2] * Tv[k][2];
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[6 + 3 * k + d] += Tv[k][0] * rv[0][d] + Tv[k][1] * rv[1][d] + Tv[k][2] * rv[2][d];
      acc[24 + k] += Tv[k][0] + Tv[k][1] + Tv[k][2];
    }
  };
  // Learned form: a pair on the fixed-order rule (class 7, the full table) costs ~20 times a
  // Dunavant pair, and one such lane would hold its whole wave for that long. Those pairs are
  // queued per wave (in element order, so the result does not depend on timing) and integrated
  // after the sweep, all lanes on the same rule. A full queue integrates in place.
  constexpr int QW = 256;
  __shared__ int s_q[4][QW];
  __shared__ int s_qn[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int qn = 0;  // uniform over the wave
  for (int e0 = 0; e0 < ne; e0 += blockDim.x) {
    const int e = e0 + threadIdx.x;
    bool defer = J.cls && e < ne && e != r && J.cls[(size_t)r * ne + e] >= 7;
    const unsigned long long bm = __ballot(defer);
    const int cnt = __popcll(bm);
    if (qn + cnt <= QW) {
      if (defer) s_q[wid][qn + __popcll(bm & ((1ull << lane) - 1ull))] = e;
      qn += cnt;
    } else {
      defer = false;
    }
    if (e < ne && !defer) pair(e);
  }
  if (lane == 0) s_qn[wid] = qn;
  __syncthreads();
  {
    const int n0 = s_qn[0], n1 = s_qn[1], n2 = s_qn[2], n3 = s_qn[3];
    const int total = n0 + n1 + n2 + n3;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const int w = i < n0 ? 0 : (i < n0 + n1 ? 1 : (i < n0 + n1 + n2 ? 2 : 3));
      const int base = w == 0 ? 0 : (w == 1 ? n0 : (w == 2 ? n0 + n1 : n0 + n1 + n2));
      pair(s_q[w][i - base]);
    }
  }
  double accd[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) accd[i] = (double)acc[i];
  block_sum<NACC>(accd, red);
  if (threadIdx.x == 0) {
    const double *S = accd, *M = accd + 6, *tsum = accd + 24;
    const double cfree[3] = {-tsum[0] + J.extra, -tsum[3] + J.extra, -tsum[5] + J.extra};
    // rigid velocities at the row element's nodes, for the free term spread over its vertices
    const int *cn = g.con + 3 * r;
    double ubar[3][ANDR_NRHS];  // (1/3) sum_v u_k(node v), component c
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < ANDR_NRHS; ++k) ubar[c][k] = 0.0;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double w = (J.moving ? (double)J.moving[cn[v]] : 1.0) / 3.0;
      const double *p = g.xyz + 3 * cn[v];
      const double q[3] = {p[0] - J.center[0], p[1] - J.center[1], p[2] - J.center[2]};
      if (J.strain) {  // strain fields (0, qz, qy), (qz, 0, qx), (qy, qx, 0) in columns 0..2
        ubar[1][0] += w * q[2];
        ubar[2][0] += w * q[1];
        ubar[0][1] += w * q[2];
        ubar[2][1] += w * q[0];
        ubar[0][2] += w * q[1];
        ubar[1][2] += w * q[0];
        continue;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) ubar[c][c] += w;
      // e_x x q = (0, -qz, qy); e_y x q = (qz, 0, -qx); e_z x q = (-qy, qx, 0)
      ubar[1][3] -= w * q[2];
      ubar[2][3] += w * q[1];
      ubar[0][4] += w * q[2];
      ubar[2][4] -= w * q[0];
      ubar[0][5] -= w * q[1];
      ubar[1][5] += w * q[0];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // sum_v T_{c,i} u_k,i: translations pick column k; rotations (e_a x r)_i = eps_{i a j} r_j
      const double *Mc[3] = {M + 3 * sym3(c, 0), M + 3 * sym3(c, 1), M + 3 * sym3(c, 2)};
      double b[ANDR_NRHS];
#pragma unroll
      for (int k = 0; k < 3; ++k) b[k] = S[sym3(c, k)];
      b[3] = -Mc[1][2] + Mc[2][1];
      b[4] = Mc[0][2] - Mc[2][0];
      b[5] = -Mc[0][1] + Mc[1][0];
#pragma unroll
      for (int k = 6; k < ANDR_NRHS; ++k) b[k] = 0.0;
      if (J.strain) {  // sum_v T_{c,i} (A_k q_v)_i for the symmetric strain matrices A_k
        b[0] = Mc[1][2] + Mc[2][1];
        b[1] = Mc[0][2] + Mc[2][0];
        b[2] = Mc[0][1] + Mc[1][0];
        b[3] = b[4] = b[5] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < ANDR_NRHS; ++k) J.B[(size_t)(c * ne + r) * ANDR_NRHS + k] = b[k] + cfree[c] * ubar[c][k];
    }
    // self-checks as sdIntRowQStokes: off-diagonal sums ~ 0, diagonal sums ~ 4 pi (bounded form)
    const double osemPi = (double)(atanf(1.0f)) * 4.0 * 8.0;
    double err = fmax(fabs(tsum[1]), fmax(fabs(tsum[4]), fabs(tsum[2])));
    for (int c = 0; c < 3; ++c) err = fmax(err, fabs(cfree[c] - J.extra - 0.5 * osemPi));
    if (J.err_rows) J.err_rows[r] = err;
  }
}

// ------------------------------------------------------------------------------------------------
// Fast form of the resistance system (FP32 quadrature into the six symmetric FP32 blocks; the
// singular diagonal entries come from k_self_blocks in FP64). The rows are those of
// k_stokes_system, with two changes to how the work maps onto the wave:
//  * learned form: the elements of a row are counting-sorted by rule class in LDS before they are
//    integrated, so that a wave runs one rule at a time (the rule's table loads are then scalar and
//    no lane idles through a longer rule of a neighbour);
//  * SUM: rigid-body problems where the three nodes of every element move alike take the
//    vertex-summed stresslet moments (stokes_pair_sum: 6 sums per quadrature point instead of 18);
//    otherwise (strain problems, elements with nodes of both kinds) the per-vertex moments.
// ------------------------------------------------------------------------------------------------
constexpr int FS_CHUNK = 8192;  // elements sorted at a time (16-bit indices in LDS)
constexpr int FS_NCLS = 9;      // Dunavant classes 0..6, the fixed-order rule 7, the row's own element 8

template <bool SUM>
__global__ __launch_bounds__(256) void k_stokes_system_fast(const SysJob *__restrict__ jobs,
                                                            const int2 *__restrict__ rowmap) {
  // S: T sums times the node velocity flag; tsum: all T sums; M (per-vertex form): rotation /
  // strain moments sum_v T_k(e, v) flag_v (x_v - centre)_d
  constexpr int NACC = SUM ? 12 : 30;
  __shared__ uint16_t s_list[FS_CHUNK];
  __shared__ int s_cnt[4][FS_NCLS];
  __shared__ double red[NACC * 4];
  const int2 jr = rowmap[blockIdx.x];
  const SysJob &J = jobs[jr.x];
  const BodyGeom &g = J.g;
  const int r = jr.y, ne = g.ne;
  const double s[3] = {g.cen[3 * r], g.cen[3 * r + 1], g.cen[3 * r + 2]};
  float acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = 0.f;
  const double inv_mu = 1.0 / J.mu;
  float *A = J.template Aptr<float>();
  const int KB = (ne + 15) / 16;
  // qoff, nq: one rule for the whole wave (its table entries are then scalar loads)
  auto pair = [&](int e, int qoff, int nq) {
    qoff = __builtin_amdgcn_readfirstlane(qoff);
    nq = __builtin_amdgcn_readfirstlane(nq);
    const int vv[3] = {g.con[3 * e], g.con[3 * e + 1], g.con[3 * e + 2]};
    const double *x1 = g.xyz + 3 * vv[0], *x2 = g.xyz + 3 * vv[1], *x3 = g.xyz + 3 * vv[2];
    const double m = g.mul[e];
    double Gs[6];
    if constexpr (SUM) {
      const float m0 = J.moving ? (float)J.moving[vv[0]] : 1.f;
      float Ts[6];
      stokes_pair_sum<float>(s, x1, x2, x3, m * g.nrm[3 * e], m * g.nrm[3 * e + 1], m * g.nrm[3 * e + 2], g.area[e],
                             qoff, nq, Gs, Ts);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        acc[k] += m0 * Ts[k];
        acc[6 + k] += Ts[k];
      }
    } else {
      const float mv[3] = {J.moving ? (float)J.moving[vv[0]] : 1.f, J.moving ? (float)J.moving[vv[1]] : 1.f,
                           J.moving ? (float)J.moving[vv[2]] : 1.f};
      float Tv[6][3];
      stokes_pair_regular<float>(s, x1, x2, x3, m * g.nrm[3 * e], m * g.nrm[3 * e + 1], m * g.nrm[3 * e + 2],
                                 g.area[e], qoff, nq, Gs, Tv);
      const double *xv[3] = {x1, x2, x3};
      float rv[3][3];
#pragma unroll
      for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) rv[v][d] = mv[v] * (float)(xv[v][d] - J.center[d]);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        acc[k] += mv[0] * Tv[k][0] + mv[1] * Tv[k][1] + mv[2] * Tv[k][2];
        acc[6 + k] += Tv[k][0] + Tv[k][1] + Tv[k][2];
#pragma unroll
        for (int d = 0; d < 3; ++d)
          acc[12 + 3 * k + d] += Tv[k][0] * rv[0][d] + Tv[k][1] * rv[1][d] + Tv[k][2] * rv[2][d];
      }
    }
    const double sc = -m * inv_mu;
#pragma unroll
    for (int q = 0; q < 6; ++q) A[andr_sym_off(q, r, e, KB)] = (float)(sc * Gs[q]);
  };
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  if (!J.cls) {  // one rule for every pair
    for (int e = threadIdx.x; e < ne; e += blockDim.x)
      if (e != r) pair(e, J.qoff, J.nq);
  } else {
    const uint8_t *crow = J.cls + (size_t)r * ne;
    for (int c0 = 0; c0 < ne; c0 += FS_CHUNK) {
      const int n = min(ne - c0, FS_CHUNK);
      auto cls_of = [&](int i) { return c0 + i == r ? 8 : min((int)crow[c0 + i], 7); };
      // count the chunk's elements per class and wave
      int cnt[FS_NCLS];
#pragma unroll
      for (int k = 0; k < FS_NCLS; ++k) cnt[k] = 0;
      for (int i0 = 0; i0 < n; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        const int k = i < n ? cls_of(i) : FS_NCLS;
#pragma unroll
        for (int kk = 0; kk < FS_NCLS; ++kk) cnt[kk] += __popcll(__ballot(k == kk));
      }
      if (lane == 0)
#pragma unroll
        for (int k = 0; k < FS_NCLS; ++k) s_cnt[wid][k] = cnt[k];
      __syncthreads();
      // class k occupies [off[k], off[k + 1]) of the sorted list; wave w's elements of class k start
      // at off[k] + (class-k count of waves < w)
      int off[FS_NCLS + 1], pos0[FS_NCLS];
      off[0] = 0;
#pragma unroll
      for (int k = 0; k < FS_NCLS; ++k) {
        int before = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const int c = s_cnt[w][k];
          before += w < wid ? c : 0;
          tot += c;
        }
        pos0[k] = __builtin_amdgcn_readfirstlane(off[k] + before);  // uniform over the wave
        off[k + 1] = __builtin_amdgcn_readfirstlane(off[k] + tot);
      }
      for (int i0 = 0; i0 < n; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        const int k = i < n ? cls_of(i) : FS_NCLS;
        int pos = 0;
#pragma unroll
        for (int kk = 0; kk < FS_NCLS; ++kk) {
          const unsigned long long bm = __ballot(k == kk);
          if (k == kk) pos = pos0[kk] + __popcll(bm & below);
          pos0[kk] += __popcll(bm);
        }
        if (i < n) s_list[pos] = (uint16_t)i;
      }
      __syncthreads();
      // integrate in sorted order; a wave whose lanes straddle a class boundary runs one class at a time
      for (int p0 = 0; p0 < n; p0 += blockDim.x) {
        const int p = p0 + threadIdx.x;
        int k = FS_NCLS - 1, e = 0;
        if (p < n) {
          e = c0 + s_list[p];
          k = 0;
#pragma unroll
          for (int kk = 1; kk < FS_NCLS; ++kk) k += p >= off[kk] ? 1 : 0;
        }
        bool todo = k < 8;  // class 8, the row's own element, is k_self_blocks' singular block
        for (;;) {
          const unsigned long long bm = __ballot(todo);
          if (!bm) break;
          const int kk = __builtin_amdgcn_readfirstlane(__shfl(k, (int)__builtin_ctzll(bm), 64));
          if (todo && k == kk) {
            int qoff = J.qoff, nq = J.nq;
            if (kk < 7) {  // Dunavant offsets {0,3,7,13,20,32,45} and sizes {3,4,6,7,12,13,19}, 6 bits each
              qoff = ANDR_DUN_OFF0 + (int)((0x2D8143470C0ull >> (6 * kk)) & 63);
              nq = (int)((0x1334C1C6103ull >> (6 * kk)) & 63);
            }
            pair(e, qoff, nq);
            todo = false;
          }
        }
      }
      __syncthreads();  // s_list and s_cnt are rewritten by the next chunk
    }
  }
  double accd[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) accd[i] = (double)acc[i];
  block_sum<NACC>(accd, red);
  if (threadIdx.x == 0) {
    const double *S = accd, *tsum = accd + 6;
    // rotation moments sum_e sum_v T_k(e, v) flag_v (x_v - centre)_d: as integrated (per-vertex
    // form), or S_k (s - centre)_d (vertex sums: x_v - centre = (x_v - s) + (s - centre), and the
    // first part only adds third moments that the rotations cancel)
    double M[18];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) M[3 * k + d] = SUM ? S[k] * (s[d] - J.center[d]) : accd[12 + 3 * k + d];
    const double cfree[3] = {-tsum[0] + J.extra, -tsum[3] + J.extra, -tsum[5] + J.extra};
    const int *cn = g.con + 3 * r;
    double ubar[3][ANDR_NRHS];  // (1/3) sum_v u_k(node v), component c
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < ANDR_NRHS; ++k) ubar[c][k] = 0.0;
    for (int v = 0; v < 3; ++v) {
      const double w = (J.moving ? (double)J.moving[cn[v]] : 1.0) / 3.0;
      const double *p = g.xyz + 3 * cn[v];
      const double q[3] = {p[0] - J.center[0], p[1] - J.center[1], p[2] - J.center[2]};
      if (J.strain) {
        ubar[1][0] += w * q[2];
        ubar[2][0] += w * q[1];
        ubar[0][1] += w * q[2];
        ubar[2][1] += w * q[0];
        ubar[0][2] += w * q[1];
        ubar[1][2] += w * q[0];
        continue;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) ubar[c][c] += w;
      ubar[1][3] -= w * q[2];
      ubar[2][3] += w * q[1];
      ubar[0][4] += w * q[2];
      ubar[2][4] -= w * q[0];
      ubar[0][5] -= w * q[1];
      ubar[1][5] += w * q[0];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double *Mc[3] = {M + 3 * sym3(c, 0), M + 3 * sym3(c, 1), M + 3 * sym3(c, 2)};
      double b[ANDR_NRHS];
#pragma unroll
      for (int k = 0; k < 3; ++k) b[k] = S[sym3(c, k)];
      b[3] = -Mc[1][2] + Mc[2][1];
      b[4] = Mc[0][2] - Mc[2][0];
      b[5] = -Mc[0][1] + Mc[1][0];
#pragma unroll
      for (int k = 6; k < ANDR_NRHS; ++k) b[k] = 0.0;
      if (J.strain) {
        b[0] = Mc[1][2] + Mc[2][1];
        b[1] = Mc[0][2] + Mc[2][0];
        b[2] = Mc[0][1] + Mc[1][0];
        b[3] = b[4] = b[5] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < ANDR_NRHS; ++k) J.B[(size_t)(c * ne + r) * ANDR_NRHS + k] = b[k] + cfree[c] * ubar[c][k];
    }
    const double osemPi = (double)(atanf(1.0f)) * 4.0 * 8.0;
    double err = fmax(fabs(tsum[1]), fmax(fabs(tsum[4]), fabs(tsum[2])));
    for (int c = 0; c < 3; ++c) err = fmax(err, fabs(cfree[c] - J.extra - 0.5 * osemPi));
    if (J.err_rows) J.err_rows[r] = err;
  }
}

int andr_launch_stokes_system(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows, bool fp32_quad,
                              bool fp32_store, double alg_bytes, double alg_flops, bool vertex_sums) {
  if (nrows <= 0) return ANDR_OK;
  const int tok = andr_prof_begin(ctx, fp32_quad ? "stokes_system_fp32" : "stokes_system_fp64", alg_bytes, alg_flops);
  if (!fp32_quad && !fp32_store)
    hipLaunchKernelGGL((k_stokes_system<double, double>), dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else if (fp32_quad && fp32_store && vertex_sums)  // symmetric blocks, diagonal from k_self_blocks
    hipLaunchKernelGGL(k_stokes_system_fast<true>, dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else if (fp32_quad && fp32_store)
    hipLaunchKernelGGL(k_stokes_system_fast<false>, dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else if (!fp32_quad && fp32_store)
    hipLaunchKernelGGL((k_stokes_system<double, float>), dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else
    return andr_fail(ctx, ANDR_ERR_ARG, "FP32 quadrature with FP64 storage is not offered");
  andr_prof_end(ctx, tok);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

// block per body: force and torque of the six problems on wall 1, F = sum_e A_e q_e and
// T = sum_e A_e (c_e - centre) x q_e (the reductions of src/lib_particle.f90:1751-1831), and the
// largest row self-check error. out[37 b + 6 k + (0..5)] = (F, T) of problem k, out[37 b + 36] = error.
__global__ __launch_bounds__(256) void k_force_torque(const SysJob *__restrict__ jobs, double *__restrict__ out) {
  __shared__ double red[36 * 4];
  __shared__ double emax[4];
  const SysJob &J = jobs[blockIdx.x];
  const BodyGeom &g = J.g;
  const int ne = g.ne;
  double v[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) v[i] = 0.0;
  double err = 0.0;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    if (J.err_rows) err = fmax(err, J.err_rows[e]);
    if (g.bcid[e] != 1) continue;
    const double a = g.area[e];
    const double r0 = g.cen[3 * e] - J.center[0], r1 = g.cen[3 * e + 1] - J.center[1],
                 r2 = g.cen[3 * e + 2] - J.center[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double q0 = J.X[(size_t)e * ANDR_NRHS + k], q1 = J.X[(size_t)(ne + e) * ANDR_NRHS + k],
                   q2 = J.X[(size_t)(2 * ne + e) * ANDR_NRHS + k];
      v[6 * k] += a * q0;
      v[6 * k + 1] += a * q1;
      v[6 * k + 2] += a * q2;
      v[6 * k + 3] += a * (r1 * q2 - r2 * q1);
      v[6 * k + 4] += a * (r2 * q0 - r0 * q2);
      v[6 * k + 5] += a * (r0 * q1 - r1 * q0);
    }
  }
  for (int o = 32; o > 0; o >>= 1) err = fmax(err, __shfl_xor(err, o, 64));
  if ((threadIdx.x & 63) == 0) emax[threadIdx.x >> 6] = err;
  block_sum<36>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 36; ++i) out[(size_t)blockIdx.x * 37 + i] = v[i];
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, emax[w]);
    out[(size_t)blockIdx.x * 37 + 36] = m;
  }
}

int andr_launch_force_torque(andr_ctx *ctx, const SysJob *d_jobs, int njobs, double *d_out) {
  if (njobs <= 0) return ANDR_OK;
  hipLaunchKernelGGL(k_force_torque, dim3(njobs), dim3(256), 0, ctx->stream, d_jobs, d_out);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

// thread per (job, element): inverse of the element's own 3x3 system block
__global__ void k_self_blocks(const SysJob *__restrict__ jobs, const int2 *__restrict__ rowmap, int nrows) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const int2 jr = rowmap[t];
  const SysJob &J = jobs[jr.x];
  const BodyGeom &g = J.g;
  const int e = jr.y;
  const double *x1 = g.xyz + 3 * g.con[3 * e], *x2 = g.xyz + 3 * g.con[3 * e + 1], *x3 = g.xyz + 3 * g.con[3 * e + 2];
  const double s[3] = {g.cen[3 * e], g.cen[3 * e + 1], g.cen[3 * e + 2]};
  double Gs[6];
  andr_sing_stk(x1, x2, x3, s, 4, Gs);
  const double sc = -g.mul[e] / J.mu;
  if (J.lda == 0) {  // symmetric FP32 blocks: their diagonal entries (k_stokes_system_fast skips them)
    float *A = J.template Aptr<float>();
    const int KB = (g.ne + 15) / 16;
    const double scs = -g.mul[e] * (1.0 / J.mu);
#pragma unroll
    for (int q = 0; q < 6; ++q) A[andr_sym_off(q, e, e, KB)] = (float)(scs * Gs[q]);
  }
  const double dfl = J.defl ? J.defl[g.bcid[e]] * g.area[e] : 0.0;
  const double *n = g.nrm + 3 * e;
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = sc * Gs[sym3(i, j)] + dfl * n[i] * n[j];
  const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1], c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2],
               c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  double *mi = J.minv + 9 * (size_t)e;
  if (!(fabs(det) > 0.0)) {  // degenerate element: no scaling
    for (int i = 0; i < 9; ++i) mi[i] = (i % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  const double id = 1.0 / det;
  mi[0] = c00 * id;
  mi[3] = c01 * id;
  mi[6] = c02 * id;
  mi[1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  mi[4] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  mi[7] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  mi[2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  mi[5] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  mi[8] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
}

// thread per (job, element): x_e = minv[e] y_e for every right-hand side (in place on J.X)
__global__ void k_unprecondition(const SysJob *__restrict__ jobs, const int2 *__restrict__ rowmap, int nrows) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const int2 jr = rowmap[t];
  const SysJob &J = jobs[jr.x];
  const int e = jr.y, ne = J.g.ne;
  const double *mi = J.minv + 9 * (size_t)e;
  for (int k = 0; k < ANDR_NRHS; ++k) {
    const double y0 = J.X[(size_t)(0 * ne + e) * ANDR_NRHS + k], y1 = J.X[(size_t)(1 * ne + e) * ANDR_NRHS + k],
                 y2 = J.X[(size_t)(2 * ne + e) * ANDR_NRHS + k];
    for (int c = 0; c < 3; ++c) J.X[(size_t)(c * ne + e) * ANDR_NRHS + k] = mi[3 * c] * y0 + mi[3 * c + 1] * y1 + mi[3 * c + 2] * y2;
  }
}

int andr_launch_self_blocks(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows) {
  if (nrows <= 0) return ANDR_OK;
  hipLaunchKernelGGL(k_self_blocks, dim3((nrows + 255) / 256), dim3(256), 0, ctx->stream, d_jobs, d_rowmap, nrows);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

int andr_launch_unprecondition(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows) {
  if (nrows <= 0) return ANDR_OK;
  hipLaunchKernelGGL(k_unprecondition, dim3((nrows + 255) / 256), dim3(256), 0, ctx->stream, d_jobs, d_rowmap, nrows);
  ANDR_HIP(ctx, hipGetLastError());
  return ANDR_OK;
}

// ------------------------------------------------------------------------------------------------
// row-major device matrix -> host (row- or column-major)
// ------------------------------------------------------------------------------------------------
__global__ void k_transpose(const double *__restrict__ in, double *__restrict__ out, int64_t m, int64_t n) {
  __shared__ double tile[32][33];
  int64_t bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int j = threadIdx.y; j < 32; j += 8) {
    int64_t r = by + j, c = bx + threadIdx.x;
    if (r < m && c < n) tile[j][threadIdx.x] = in[r * n + c];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    int64_t c = bx + j, r = by + threadIdx.x;
    if (r < m && c < n) out[c * m + r] = tile[threadIdx.x][j];
  }
}

int andr_copy_out(andr_ctx *ctx, const double *d, int64_t m, int64_t n, int32_t layout, double *h) {
  if (!h) return ANDR_OK;
  if (layout == ANDR_ROW_MAJOR) {
    ANDR_HIP(ctx, hipMemcpyAsync(h, d, sizeof(double) * m * n, hipMemcpyDeviceToHost, ctx->stream));
    ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ANDR_OK;
  }
  DevBuf<double> t;
  if (t.alloc((size_t)m * n)) return andr_fail(ctx, ANDR_ERR_ALLOC, "transpose buffer");
  dim3 grid((unsigned)((n + 31) / 32), (unsigned)((m + 31) / 32));
  hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, ctx->stream, d, t.p, m, n);
  ANDR_HIP(ctx, hipGetLastError());
  ANDR_HIP(ctx, hipMemcpyAsync(h, t.p, sizeof(double) * m * n, hipMemcpyDeviceToHost, ctx->stream));
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ANDR_OK;
}
