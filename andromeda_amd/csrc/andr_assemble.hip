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

int andr_upload_rules(andr_ctx *ctx) {
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

// ------------------------------------------------------------------------------------------------
// geometry: CalElementNormal (src/lib_mesh.f90:716-769), Heron area (src/mod_trikotnik.f90:786-802),
// CalQMesh centroid (src/lib_mesh.f90:577-629)
// ------------------------------------------------------------------------------------------------
__global__ void k_geometry(int ne, const double *__restrict__ xyz, const int *__restrict__ con,
                           const int *__restrict__ bcid, const double *__restrict__ wallmul, double *__restrict__ nrm,
                           double *__restrict__ area, double *__restrict__ cen, double *__restrict__ mul) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const double *t1 = xyz + 3 * con[3 * e], *t2 = xyz + 3 * con[3 * e + 1], *t3 = xyz + 3 * con[3 * e + 2];
  double v1[3] = {t2[0] - t1[0], t2[1] - t1[1], t2[2] - t1[2]};
  double v2[3] = {t3[0] - t1[0], t3[1] - t1[1], t3[2] - t1[2]};
  double n0 = v1[1] * v2[2] - v1[2] * v2[1];
  double n1 = -v1[0] * v2[2] + v1[2] * v2[0];
  double n2 = v1[0] * v2[1] - v1[1] * v2[0];
  double d = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  nrm[3 * e] = n0 / d;
  nrm[3 * e + 1] = n1 / d;
  nrm[3 * e + 2] = n2 / d;
  auto sq = [](double x) { return x * x; };
  double a = sqrt(sq(t1[0] - t2[0]) + sq(t1[1] - t2[1]) + sq(t1[2] - t2[2]));
  double b = sqrt(sq(t3[0] - t2[0]) + sq(t3[1] - t2[1]) + sq(t3[2] - t2[2]));
  double c = sqrt(sq(t1[0] - t3[0]) + sq(t1[1] - t3[1]) + sq(t1[2] - t3[2]));
  double p = 0.5 * (a + b + c);
  area[e] = sqrt(p * (p - a) * (p - b) * (p - c));
  const double third = 1.0 / 3.0;
  for (int j = 0; j < 3; ++j) cen[3 * e + j] = third * t1[j] + third * t2[j] + third * t3[j];
  mul[e] = wallmul[bcid[e] - 1];
}

// Greedy colouring of elements so that no two elements of a colour share a node.
static void color_elements(int nn, int ne, const int *con, std::vector<int> &order, std::vector<int> &off) {
  std::vector<uint64_t> used(nn, 0);
  std::vector<int> col(ne);
  int ncol = 0;
  for (int e = 0; e < ne; ++e) {
    uint64_t m = used[con[3 * e]] | used[con[3 * e + 1]] | used[con[3 * e + 2]];
    int c = 0;
    while (c < 64 && (m >> c) & 1ull) ++c;
    if (c >= 64) c = 63;  // cannot happen for valence < 21; kept in range
    col[e] = c;
    ncol = std::max(ncol, c + 1);
    for (int v = 0; v < 3; ++v) used[con[3 * e + v]] |= (1ull << c);
  }
  off.assign(ncol + 1, 0);
  for (int e = 0; e < ne; ++e) off[col[e] + 1]++;
  for (int c = 0; c < ncol; ++c) off[c + 1] += off[c];
  order.resize(ne);
  std::vector<int> pos(off.begin(), off.end() - 1);
  for (int e = 0; e < ne; ++e) order[pos[col[e]]++] = e;
}

int andr_body_upload(andr_ctx *ctx, const andr_mesh *m, DevBody &b, bool with_colors) {
  if (!m || m->nnodes <= 0 || m->nelem <= 0 || !m->xyz || !m->con)
    return andr_fail(ctx, ANDR_ERR_ARG, "invalid mesh");
  const int nn = m->nnodes, ne = m->nelem;
  const int nw = m->nwalls > 0 ? m->nwalls : 1;
  std::vector<int> con(3 * (size_t)ne), bcid(ne, 1);
  for (size_t i = 0; i < 3 * (size_t)ne; ++i) {
    con[i] = m->con[i] - m->index_base;
    if (con[i] < 0 || con[i] >= nn) return andr_fail(ctx, ANDR_ERR_ARG, "connectivity entry %zu out of range", i);
  }
  if (m->bcid)
    for (int e = 0; e < ne; ++e) {
      bcid[e] = m->bcid[e];
      if (bcid[e] < 1 || bcid[e] > nw) return andr_fail(ctx, ANDR_ERR_ARG, "bcid of element %d out of range", e);
    }
  std::vector<double> wm(nw, 1.0);
  if (m->normmul)
    for (int w = 0; w < nw; ++w) wm[w] = m->normmul[w];
  b.itri = m->itri == 0 ? 5 : m->itri;
  if (b.itri < 1 || b.itri > 5) return andr_fail(ctx, ANDR_ERR_ARG, "itri must be 1..5");
  DevBuf<double> d_wm;
  if (b.xyz.alloc(3 * (size_t)nn, ctx) || b.con.alloc(3 * (size_t)ne, ctx) || b.bcid.alloc(ne, ctx) ||
      b.nrm.alloc(3 * (size_t)ne, ctx) || b.area.alloc(ne, ctx) || b.cen.alloc(3 * (size_t)ne, ctx) ||
      b.mul.alloc(ne, ctx) || d_wm.alloc(nw, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "device allocation for mesh");
  ANDR_HIP(ctx, hipMemcpyAsync(b.xyz.p, m->xyz, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(b.con.p, con.data(), sizeof(int) * 3 * ne, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(b.bcid.p, bcid.data(), sizeof(int) * ne, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(d_wm.p, wm.data(), sizeof(double) * nw, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_geometry, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream, ne, b.xyz.p, b.con.p, b.bcid.p,
                     d_wm.p, b.nrm.p, b.area.p, b.cen.p, b.mul.p);
  ANDR_HIP(ctx, hipGetLastError());
  if (with_colors) {
    std::vector<int> order;
    color_elements(nn, ne, con.data(), order, b.color_off);
    if (b.color_elems.alloc(ne) || b.d_color_off.alloc(b.color_off.size()))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "device allocation for colours");
    ANDR_HIP(ctx, hipMemcpyAsync(b.color_elems.p, order.data(), sizeof(int) * ne, hipMemcpyHostToDevice, ctx->stream));
    ANDR_HIP(ctx, hipMemcpyAsync(b.d_color_off.p, b.color_off.data(), sizeof(int) * b.color_off.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
  }
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // keeps d_wm alive until the kernel ran
  b.g.nn = nn;
  b.g.ne = ne;
  b.g.nwalls = nw;
  b.g.xyz = b.xyz.p;
  b.g.con = b.con.p;
  b.g.bcid = b.bcid.p;
  b.g.nrm = b.nrm.p;
  b.g.area = b.area.p;
  b.g.cen = b.cen.p;
  b.g.mul = b.mul.p;
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
// Laplace rows: sdIntRowU / sdIntRowQ (src/lib_sdBem.f90:438-733), regular integrals by
// Triangle_3DLapInt (:859-905) with BemKernel (:911-936), singular by the analytic formula.
// ------------------------------------------------------------------------------------------------
template <bool ROW_IN_LDS>
__global__ __launch_bounds__(256) void k_laplace_rows(BodyGeom g, int qoff, int nq, const int *__restrict__ celems,
                                                      const int *__restrict__ coff, int ncolors,
                                                      double *__restrict__ H, double *__restrict__ G,
                                                      double *__restrict__ err_rows) {
  extern __shared__ double lds[];
  __shared__ double red[16];
  const int nn = g.nn, ne = g.ne;
  const int r = blockIdx.x;
  const bool qrow = r >= nn;
  const int erow = r - nn;
  double s[3];
  if (!qrow)
    for (int j = 0; j < 3; ++j) s[j] = g.xyz[3 * r + j];
  else
    for (int j = 0; j < 3; ++j) s[j] = g.cen[3 * erow + j];
  double *hrow = ROW_IN_LDS ? lds : H + (size_t)r * nn;
  double *grow = G + (size_t)r * ne;
  for (int i = threadIdx.x; i < nn; i += blockDim.x) hrow[i] = 0.0;
  __syncthreads();
  const double k4pi = 0.25 / ANDR_PI;
  for (int c = 0; c < ncolors; ++c) {
    for (int idx = coff[c] + threadIdx.x; idx < coff[c + 1]; idx += blockDim.x) {
      const int e = celems[idx];
      const int v0 = g.con[3 * e], v1 = g.con[3 * e + 1], v2 = g.con[3 * e + 2];
      const double *x1 = g.xyz + 3 * v0, *x2 = g.xyz + 3 * v1, *x3 = g.xyz + 3 * v2;
      int isrc = 0;
      if (!qrow) isrc = (v0 == r) ? 1 : (v1 == r) ? 2 : (v2 == r) ? 3 : 0;
      else if (e == erow) isrc = 4;
      double gv, h1 = 0.0, h2 = 0.0, h3 = 0.0;
      if (isrc == 0) {
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
    else {
      const int *cn = g.con + 3 * erow;
      for (int v = 0; v < 3; ++v) hrow[cn[v]] += (1.0 / 3.0) * cfree;
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
    double *out = H + (size_t)r * nn;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) out[i] = hrow[i];
  }
  if (threadIdx.x == 0) err_rows[r] = err;
}

int andr_launch_laplace_rows(andr_ctx *ctx, DevBody &b, double *H, double *G, double *err_rows) {
  QuadRule qr = andr_rule(b.itri);
  const int nsp = b.g.nn + b.g.ne;
  const int ncol = (int)b.color_off.size() - 1;
  const size_t lds = sizeof(double) * b.g.nn;
  const double evals = (double)nsp * b.g.ne * qr.npts;
  const int tok = andr_prof_begin(ctx, "laplace_rows", 8.0 * nsp * ((double)b.g.nn + b.g.ne),
                                  evals * ANDR_FLOPS_PER_LAPLACE_EVAL);
  if (lds <= 128 * 1024) {
    ANDR_HIP(ctx, hipFuncSetAttribute((const void *)k_laplace_rows<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_laplace_rows<true>, dim3(nsp), dim3(256), lds, ctx->stream, b.g, qr.offset, qr.npts,
                       b.color_elems.p, b.d_color_off.p, ncol, H, G, err_rows);
  } else {
    hipLaunchKernelGGL(k_laplace_rows<false>, dim3(nsp), dim3(256), 0, ctx->stream, b.g, qr.offset, qr.npts,
                       b.color_elems.p, b.d_color_off.p, ncol, H, G, err_rows);
  }
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
                                                     double *__restrict__ err_rows) {
  extern __shared__ double lds[];
  __shared__ double red[6 * 4];
  const int nn = g.nn, ne = g.ne;
  const int nsp = nn + ne;
  const int r = blockIdx.x;
  const bool qrow = r >= nn;
  const int erow = r - nn;
  const size_t mstrT = (size_t)nsp * nn, mstrG = (size_t)nsp * ne;
  double s[3];
  if (!qrow)
    for (int j = 0; j < 3; ++j) s[j] = g.xyz[3 * r + j];
  else
    for (int j = 0; j < 3; ++j) s[j] = g.cen[3 * erow + j];
  // node rows: 6 T components then 3 prT components
  auto trow = [&](int k) -> double * {
    if (ROW_IN_LDS) return lds + (size_t)k * nn;
    return k < 6 ? T + k * mstrT + (size_t)r * nn : prT + (k - 6) * mstrT + (size_t)r * nn;
  };
  for (int k = 0; k < 9; ++k) {
    double *p = trow(k);
    for (int i = threadIdx.x; i < nn; i += blockDim.x) p[i] = 0.0;
  }
  __syncthreads();
  for (int c = 0; c < ncolors; ++c) {
    for (int idx = coff[c] + threadIdx.x; idx < coff[c + 1]; idx += blockDim.x) {
      const int e = celems[idx];
      const int vv[3] = {g.con[3 * e], g.con[3 * e + 1], g.con[3 * e + 2]};
      const double *x1 = g.xyz + 3 * vv[0], *x2 = g.xyz + 3 * vv[1], *x3 = g.xyz + 3 * vv[2];
      int isrc = 0;
      if (!qrow) isrc = (vv[0] == r) ? 1 : (vv[1] == r) ? 2 : (vv[2] == r) ? 3 : 0;
      else if (e == erow) isrc = 4;
      double Gs[6] = {0, 0, 0, 0, 0, 0}, pG[3] = {0, 0, 0};
      double Tv[6][3] = {{0}}, pT[3][3] = {{0}};
      if (isrc == 0) {
        const double m = g.mul[e];
        const double nx = m * g.nrm[3 * e], ny = m * g.nrm[3 * e + 1], nz = m * g.nrm[3 * e + 2];
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
        const double A = g.area[e];
        for (int k = 0; k < 6; ++k) {
          Gs[k] *= A;
          for (int v = 0; v < 3; ++v) Tv[k][v] *= A;
        }
        for (int k = 0; k < 3; ++k) {
          pG[k] *= A;
          for (int v = 0; v < 3; ++v) pT[k][v] *= A;
        }
      } else {
        andr_sing_stk(x1, x2, x3, s, isrc, Gs);  // T, prT, prG of a singular pair are zero
      }
      for (int k = 0; k < 6; ++k) G[k * mstrG + (size_t)r * ne + e] = Gs[k];
      for (int k = 0; k < 3; ++k) prG[k * mstrG + (size_t)r * ne + e] = pG[k];
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
      const size_t o = (size_t)r * ne + i;
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
      const int *cn = g.con + 3 * erow;
      const double shpf = 1.0 / 3.0;
      for (int v = 0; v < 3; ++v) {
        trow(0)[cn[v]] += shpf * (cx + extra);
        trow(3)[cn[v]] += shpf * (cy + extra);
        trow(5)[cn[v]] += shpf * (cz + extra);
      }
      // pressure correction (src/lib_Stokes.f90:1851-1863)
      for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) trow(6 + k)[cn[v]] -= psum[k] / 3;
    }
  }
  __syncthreads();
  if (ROW_IN_LDS) {
    for (int k = 0; k < 9; ++k) {
      double *out = k < 6 ? T + k * mstrT + (size_t)r * nn : prT + (k - 6) * mstrT + (size_t)r * nn;
      const double *p = lds + (size_t)k * nn;
      for (int i = threadIdx.x; i < nn; i += blockDim.x) out[i] = p[i];
    }
  }
  if (threadIdx.x == 0) err_rows[r] = err;
}

int andr_launch_stokes_rows(andr_ctx *ctx, DevBody &b, double extra, double *T, double *G, double *prT, double *prG,
                            double *err_rows) {
  QuadRule qr = andr_rule(b.itri);
  const int nsp = b.g.nn + b.g.ne;
  const int ncol = (int)b.color_off.size() - 1;
  const size_t lds = sizeof(double) * 9 * (size_t)b.g.nn;
  const double evals = (double)nsp * b.g.ne * qr.npts;
  const int tok = andr_prof_begin(ctx, "stokes_rows", 8.0 * 9 * nsp * ((double)b.g.nn + b.g.ne),
                                  evals * ANDR_FLOPS_PER_STOKES_EVAL);
  if (lds <= 144 * 1024) {
    ANDR_HIP(ctx, hipFuncSetAttribute((const void *)k_stokes_rows<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_stokes_rows<true>, dim3(nsp), dim3(256), lds, ctx->stream, b.g, qr.offset, qr.npts,
                       b.color_elems.p, b.d_color_off.p, ncol, extra, T, G, prT, prG, err_rows);
  } else {
    hipLaunchKernelGGL(k_stokes_rows<false>, dim3(nsp), dim3(256), 0, ctx->stream, b.g, qr.offset, qr.npts,
                       b.color_elems.p, b.d_color_off.p, ncol, extra, T, G, prT, prG, err_rows);
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
  __shared__ double red[NACC * 4];
  const int2 jr = rowmap[blockIdx.x];
  const SysJob &J = jobs[jr.x];
  const BodyGeom &g = J.g;
  const int r = jr.y;
  const int ne = g.ne;
  const double s[3] = {g.cen[3 * r], g.cen[3 * r + 1], g.cen[3 * r + 2]};
  double acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = 0.0;
  const double inv_mu = 1.0 / J.mu;
  ST *A = J.template Aptr<ST>();
  const bool tiled = J.lda == 0;  // symmetric-block layout: S_q = sc G_q only (no null-space term, no M^-1)
  const int KB = (ne + 15) / 16;
  ST *arow0 = A + (tiled ? 0 : (size_t)(0 * ne + r) * J.lda);
  ST *arow1 = A + (tiled ? 0 : (size_t)(1 * ne + r) * J.lda);
  ST *arow2 = A + (tiled ? 0 : (size_t)(2 * ne + r) * J.lda);
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int vv[3] = {g.con[3 * e], g.con[3 * e + 1], g.con[3 * e + 2]};
    const double *x1 = g.xyz + 3 * vv[0], *x2 = g.xyz + 3 * vv[1], *x3 = g.xyz + 3 * vv[2];
    double Gs[6], Tv[6][3];
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
    double mv[3], rv[3][3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      mv[v] = J.moving ? (double)J.moving[vv[v]] : 1.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) rv[v][d] = mv[v] * (xv[v][d] - J.center[d]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      acc[k] += mv[0] * Tv[k][0] + mv[1] * Tv[k][1] + mv[2] * Tv[k][2];
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[6 + 3 * k + d] += Tv[k][0] * rv[0][d] + Tv[k][1] * rv[1][d] + Tv[k][2] * rv[2][d];
      acc[24 + k] += Tv[k][0] + Tv[k][1] + Tv[k][2];
    }
  }
  block_sum<NACC>(acc, red);
  if (threadIdx.x == 0) {
    const double *S = acc, *M = acc + 6, *tsum = acc + 24;
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

int andr_launch_stokes_system(andr_ctx *ctx, const SysJob *d_jobs, const int2 *d_rowmap, int nrows, bool fp32_quad,
                              bool fp32_store, double alg_bytes, double alg_flops) {
  if (nrows <= 0) return ANDR_OK;
  const int tok = andr_prof_begin(ctx, fp32_quad ? "stokes_system_fp32" : "stokes_system_fp64", alg_bytes, alg_flops);
  if (!fp32_quad && !fp32_store)
    hipLaunchKernelGGL((k_stokes_system<double, double>), dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else if (fp32_quad && fp32_store)
    hipLaunchKernelGGL((k_stokes_system<float, float>), dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else if (!fp32_quad && fp32_store)
    hipLaunchKernelGGL((k_stokes_system<double, float>), dim3(nrows), dim3(256), 0, ctx->stream, d_jobs, d_rowmap);
  else
    return andr_fail(ctx, ANDR_ERR_ARG, "FP32 quadrature with FP64 storage is not offered");
  andr_prof_end(ctx, tok);
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
