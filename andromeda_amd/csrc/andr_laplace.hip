// The Laplace boundary-value problem of the reference's Laplace driver (solveLaplace,
// src/laplace.f90) and the general Stokes problem of its Stokes driver (one subdomain; velocity or
// traction known per wall and component, src/lib_Stokes.f90:417-637, src/lib_StokesFull.f90) on the
// device. Laplace: subdomains of walls, per-wall boundary conditions (function,
// flux or contact), the integral matrices of every subdomain, the system and its right-hand side
// formed from them, the column-scaled LSQR solve, the unknowns scattered back to the boundary values
// and the flux integral of every wall.
//
//   subdomain lists      SetUpSubdomains       src/lib_sdBem.f90:319-430
//   integral matrices    sdFormIntegralMatrices src/lib_int.f90:225-323 (rows sdIntRowU / sdIntRowQ)
//   unknowns and knowns  sdSetUpXandB          src/lib_sdBem.f90:4-107
//   system, rhs matrix   sdSetUpSysMrhsM_CRS   src/lib_CRS.f90:148-343, b = rhsM (-b) by CRSxV :764-787
//   solve and scatter    sdSolveCRS            src/lib_CRS.f90:4-142
//   wall fluxes          CalFluxIntegralWalls  src/lib_mesh.f90:851-882
//
// A subdomain is integrated as the mesh of its own walls: its nodes in ascending global order and its
// elements wall by wall in the subdomain's wall order (the reference's node and q-node lists), wall
// ids renumbered in that order with the subdomain's NORMMUL factors. Its H and G stay on the device;
// the system kernel reads them row by row, writes the unknowns' coefficients into a dense row-major
// system (the reference's CSR holds every entry of these dense blocks) and sums the known columns
// into the right-hand side in the same pass, so the rhs matrix is never stored.
#include "andr_internal.h"
#include "andr_solve.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int NOT_SET = -999;

struct SdLists {
  std::vector<int> elems, qwall;  // element list and the wall of each (q-node list)
  std::vector<int> nodes, nwall;  // node list and the wall each node takes its condition from
};

// the lists of every subdomain; msg receives the reason of a failure
int build_lists(const andr_mesh *m, int nsd, const andr_subdomain *sds, std::vector<SdLists> &L, char *msg,
                size_t len) {
  const int nn = m->nnodes, ne = m->nelem, nw = m->nwalls > 0 ? m->nwalls : 1, nv = m->nvert == 4 ? 4 : 3;
  if (nsd <= 0 || !sds) {
    snprintf(msg, len, "no subdomains");
    return ANDR_ERR_ARG;
  }
  L.assign(nsd, SdLists());
  std::vector<int> tmp(nn);
  for (int s = 0; s < nsd; ++s) {
    const andr_subdomain &sd = sds[s];
    if (sd.nwalls <= 0 || !sd.walls || !(sd.diff > 0.0)) {
      snprintf(msg, len, "subdomain %d: walls and a positive diff are required", s);
      return ANDR_ERR_ARG;
    }
    for (int k = 0; k < sd.nwalls; ++k)
      if (sd.walls[k] < 1 || sd.walls[k] > nw) {
        snprintf(msg, len, "subdomain %d: wall id %d outside 1..%d", s, sd.walls[k], nw);
        return ANDR_ERR_ARG;
      }
    SdLists &l = L[s];
    for (int k = 0; k < sd.nwalls; ++k)
      for (int e = 0; e < ne; ++e) {
        const int b = m->bcid ? m->bcid[e] : 1;
        if (b == sd.walls[k]) {
          l.elems.push_back(e);
          l.qwall.push_back(b);
        }
      }
    if (l.elems.empty()) {
      snprintf(msg, len, "subdomain %d has no elements", s);
      return ANDR_ERR_ARG;
    }
    std::fill(tmp.begin(), tmp.end(), 0);
    for (int e : l.elems)  // the last listed element touching a node gives its wall
      for (int v = 0; v < nv; ++v) {
        const int c = m->con[(size_t)nv * e + v] - m->index_base;
        if (nv == 4 && v == 3 && c < 0) continue;
        tmp[c] = m->bcid ? m->bcid[e] : 1;
      }
    for (int i = 0; i < nn; ++i)
      if (tmp[i]) {
        int w = tmp[i];
        if (sd.node_wall && sd.node_wall[i] != 0) {  // EDGE / CORNER rule of the input
          w = sd.node_wall[i];
          if (w < 1 || w > nw) {
            snprintf(msg, len, "subdomain %d: node_wall of node %d is %d, outside 1..%d", s, i, w, nw);
            return ANDR_ERR_ARG;
          }
        }
        l.nodes.push_back(i);
        l.nwall.push_back(w);
      }
  }
  return ANDR_OK;
}

// One row block per system row of a subdomain: sys[row][col] = H[r][j] or (normmul / diff) G[r][j]
// for the unknowns' columns, rhs[row] = -(sum of the known columns times the known values).
template <class ST>
__global__ __launch_bounds__(256) void k_lap_system(int nn_sd, int ne_sd, const double *__restrict__ H,
                                                    const double *__restrict__ G, const int *__restrict__ rows_r,
                                                    int row0, const int *__restrict__ ncol,
                                                    const uint8_t *__restrict__ nknown, const int *__restrict__ qcol,
                                                    const uint8_t *__restrict__ qknown,
                                                    const double *__restrict__ gmul, const double *__restrict__ b,
                                                    ST *__restrict__ sys, int64_t lda, double *__restrict__ B) {
  __shared__ double red[4];
  const int row = row0 + blockIdx.x;
  const int r = rows_r[blockIdx.x];
  const double *hr = H + (size_t)r * nn_sd;
  const double *gr = G + (size_t)r * ne_sd;
  ST *srow = sys + (size_t)row * lda;
  double acc = 0.0;
  for (int j = threadIdx.x; j < nn_sd; j += blockDim.x) {
    const double v = hr[j];
    if (nknown[j]) acc -= v * b[ncol[j]];
    else srow[ncol[j]] = (ST)v;
  }
  for (int j = threadIdx.x; j < ne_sd; j += blockDim.x) {
    const double v = gmul[j] * gr[j];
    if (qknown[j]) acc -= v * b[qcol[j]];
    else srow[qcol[j]] = (ST)v;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) B[(size_t)row * ANDR_NB] = (red[0] + red[1]) + (red[2] + red[3]);
}

template <class T>
__global__ void k_lap_fill(size_t n, T *__restrict__ p) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = (T)0;
}

// the solution back to the boundary values: u[i] = x[src] where src >= 0 (column 0 of the packed X)
__global__ void k_lap_scatter(int n, const int *__restrict__ src, const double *__restrict__ X,
                              double *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && src[i] >= 0) v[i] = X[(size_t)src[i] * ANDR_NB];
}

// flux integral of wall blockIdx.x + 1: sum of area * q over its elements
__global__ __launch_bounds__(256) void k_wall_flux(int ne, const int *__restrict__ bcid, const double *__restrict__ area,
                                                   const double *__restrict__ q, double *__restrict__ out) {
  __shared__ double red[4];
  const int w = blockIdx.x + 1;
  double acc = 0.0;
  for (int e = threadIdx.x; e < ne; e += blockDim.x)
    if (bcid[e] == w) acc += area[e] * q[e];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// The mesh of one subdomain's walls: nodes in node-list order, elements in element-list order, wall
// ids renumbered 1..nwalls in the subdomain's order with its NORMMUL factors.
struct SubMesh {
  std::vector<double> xyz, nm;
  std::vector<int32_t> con, bc;
  andr_mesh m;
};
void make_submesh(const andr_mesh *mesh, const andr_subdomain &sd, const SdLists &l, SubMesh &S) {
  const int nn = mesh->nnodes, nv = mesh->nvert == 4 ? 4 : 3;
  std::vector<int> remap(nn, -1);
  for (size_t k = 0; k < l.nodes.size(); ++k) remap[l.nodes[k]] = (int)k;
  S.xyz.resize(3 * l.nodes.size());
  for (size_t k = 0; k < l.nodes.size(); ++k)
    for (int d = 0; d < 3; ++d) S.xyz[3 * k + d] = mesh->xyz[3 * (size_t)l.nodes[k] + d];
  S.con.resize((size_t)nv * l.elems.size());
  S.bc.resize(l.elems.size());
  for (size_t k = 0; k < l.elems.size(); ++k) {
    for (int v = 0; v < nv; ++v) {
      const int c = mesh->con[(size_t)nv * l.elems[k] + v] - mesh->index_base;
      S.con[(size_t)nv * k + v] = (nv == 4 && v == 3 && c < 0) ? -1 : remap[c];
    }
    for (int w = 0; w < sd.nwalls; ++w)
      if (sd.walls[w] == l.qwall[k]) {
        S.bc[k] = w + 1;
        break;
      }
  }
  S.nm.assign(sd.nwalls, 1.0);
  if (sd.normmul)
    for (int w = 0; w < sd.nwalls; ++w) S.nm[w] = sd.normmul[w];
  S.m = *mesh;
  S.m.nnodes = (int)l.nodes.size();
  S.m.nelem = (int)l.elems.size();
  S.m.nwalls = sd.nwalls;
  S.m.xyz = S.xyz.data();
  S.m.con = S.con.data();
  S.m.bcid = S.bc.data();
  S.m.normmul = S.nm.data();
  S.m.index_base = 0;
}

// normmul / diff of every element of a subdomain (the G columns' factor in the system)
std::vector<double> g_factors(const andr_subdomain &sd, const SdLists &l) {
  std::vector<double> gm(l.elems.size());
  for (size_t k = 0; k < l.elems.size(); ++k) {
    double multi = 1.0;
    for (int w = 0; w < sd.nwalls; ++w)
      if (sd.walls[w] == l.qwall[k]) multi = sd.normmul ? sd.normmul[w] : 1.0;
    gm[k] = multi / sd.diff;
  }
  return gm;
}

template <class T>
int upload(andr_ctx *ctx, DevBuf<T> &d, const std::vector<T> &h) {
  if (d.alloc(std::max<size_t>(h.size(), 1), ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "Laplace maps");
  if (!h.empty())
    ANDR_HIP(ctx, hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  return ANDR_OK;
}

}  // namespace

extern "C" int andr_subdomain_node_walls(const andr_mesh *mesh, int32_t nsd, const andr_subdomain *sds,
                                         int32_t *node_wall, char *msg, int32_t msglen) {
  char buf[256] = "";
  int rc = andr_mesh_validate(mesh, buf, sizeof(buf));
  std::vector<SdLists> L;
  if (!rc) rc = build_lists(mesh, nsd, sds, L, buf, sizeof(buf));
  if (!rc && !node_wall) {
    snprintf(buf, sizeof(buf), "null output");
    rc = ANDR_ERR_ARG;
  }
  if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", buf);
  if (rc) return rc;
  for (int s = 0; s < nsd; ++s) {
    int32_t *o = node_wall + (size_t)s * mesh->nnodes;
    std::fill(o, o + mesh->nnodes, 0);
    for (size_t k = 0; k < L[s].nodes.size(); ++k) o[L[s].nodes[k]] = L[s].nwall[k];
  }
  return ANDR_OK;
}

namespace {
// mixed: the mixed-precision form (FP32 regular integrals, FP32 system storage with FP64 products
// and LSQR vectors); otherwise the plain FP64 path
int laplace_solve(andr_ctx *ctx, const andr_mesh *mesh, int32_t nsd, const andr_subdomain *sds, int32_t neqn,
                  const int32_t *wall_bc, const andr_lsqr_opts *opts, double *u, double *q, double *wall_flux,
                  andr_laplace_report *reports, bool mixed) {
  if (!ctx) return ANDR_ERR_ARG;
  if (neqn <= 0 || !wall_bc || !u || !q) return andr_fail(ctx, ANDR_ERR_ARG, "andr_laplace_solve arguments");
  int rc = andr_check_mesh(ctx, mesh);
  if (rc) return rc;
  char msg[256];
  std::vector<SdLists> L;
  if ((rc = build_lists(mesh, nsd, sds, L, msg, sizeof(msg)))) return andr_fail(ctx, rc, "%s", msg);
  const int nn = mesh->nnodes, ne = mesh->nelem, nw = mesh->nwalls > 0 ? mesh->nwalls : 1;
  const int nv = mesh->nvert == 4 ? 4 : 3;
  for (int k = 0; k < neqn * nw; ++k)
    if (wall_bc[k] < ANDR_BC_FUNCTION || wall_bc[k] > ANDR_BC_CONTACT)
      return andr_fail(ctx, ANDR_ERR_ARG, "wall_bc[%d] = %d is not a boundary condition kind", k, wall_bc[k]);
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (reports) std::memset(reports, 0, sizeof(andr_laplace_report) * neqn);
  const double t0 = andr_now();
  // ---- integral matrices of every subdomain (kept on the device) ---------------------------------
  struct SdDev {
    DevBuf<double> H, G, E;
    int nn = 0, ne = 0;
  };
  std::vector<SdDev> D(nsd);
  double errU = 0.0, errQ = 0.0;
  for (int s = 0; s < nsd; ++s) {
    SubMesh sub;
    make_submesh(mesh, sds[s], L[s], sub);
    const andr_mesh &sm = sub.m;
    DevBody body;
    if ((rc = andr_body_upload(ctx, &sm, body, true))) return rc;
    SdDev &d = D[s];
    d.nn = sm.nnodes;
    d.ne = sm.nelem;
    const int nsp = d.nn + d.ne;
    if (d.H.alloc((size_t)nsp * d.nn, ctx) || d.G.alloc((size_t)nsp * d.ne, ctx) || d.E.alloc(nsp, ctx))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "integral matrices of subdomain %d", s);
    if ((rc = andr_launch_laplace_rows(ctx, body, d.H.p, d.G.p, d.E.p, 0, -1, mixed))) return rc;
    std::vector<double> he(nsp);
    ANDR_HIP(ctx, hipMemcpyAsync(he.data(), d.E.p, sizeof(double) * nsp, hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));  // the body's buffers go back to the cache after this
    for (int i = 0; i < nsp; ++i) {
      if (i < d.nn) errU = std::max(errU, he[i]);
      else errQ = std::max(errQ, he[i]);
    }
  }
  const double t1 = andr_now();
  // mesh geometry for the wall fluxes (areas, wall ids)
  DevBody whole;
  if (wall_flux && (rc = andr_body_upload(ctx, mesh, whole, false))) return rc;
  // ---- every equation: unknowns, system, solve, scatter, fluxes --------------------------------------
  for (int en = 0; en < neqn; ++en) {
    const int32_t *bcw = wall_bc + (size_t)en * nw;
    double *ue = u + (size_t)en * nn, *qe = q + (size_t)en * ne;
    const double ts = andr_now();
    // sdSetUpXandB: column of every node (col) and element (qcol) in the unknowns or the knowns
    std::vector<int> col(nn, NOT_SET), qcol(ne, NOT_SET);
    int nx = 0, nb = 0, neq = 0;
    for (int s = 0; s < nsd; ++s) {
      const SdLists &l = L[s];
      for (size_t k = 0; k < l.nodes.size(); ++k) {
        const int isp = l.nodes[k], bc = bcw[l.nwall[k] - 1];
        if (bc == ANDR_BC_FUNCTION) col[isp] = nb++;
        else if (bc == ANDR_BC_FLUX) {
          col[isp] = nx++;
          ++neq;
        } else {
          ++neq;
          if (col[isp] == NOT_SET) col[isp] = nx++;
        }
      }
      for (size_t k = 0; k < l.elems.size(); ++k) {
        const int isp = l.elems[k], bc = bcw[l.qwall[k] - 1];
        if (bc == ANDR_BC_FUNCTION) {
          qcol[isp] = nx++;
          ++neq;
        } else if (bc == ANDR_BC_FLUX) qcol[isp] = nb++;
        else {
          ++neq;
          if (qcol[isp] == NOT_SET) qcol[isp] = nx++;
        }
      }
    }
    if (nx == 0 || neq == 0) return andr_fail(ctx, ANDR_ERR_ARG, "equation %d has no unknowns", en);
    // sdSolveCRS: the knowns b and the hot start x from the current boundary values; where the
    // solution goes back to (the last subdomain listing a point decides, as in the reference)
    std::vector<double> b(std::max(nb, 1), 0.0), x0(nx, 0.0);
    std::vector<int> src_u(nn, -1), src_q(ne, -1);
    for (int s = 0; s < nsd; ++s) {
      const SdLists &l = L[s];
      for (size_t k = 0; k < l.nodes.size(); ++k) {
        const int isp = l.nodes[k], bc = bcw[l.nwall[k] - 1];
        if (bc == ANDR_BC_FUNCTION) {
          b[col[isp]] = ue[isp];
          src_u[isp] = -1;
        } else {
          x0[col[isp]] = ue[isp];
          src_u[isp] = col[isp];
        }
      }
      for (size_t k = 0; k < l.elems.size(); ++k) {
        const int isp = l.elems[k], bc = bcw[l.qwall[k] - 1];
        if (bc == ANDR_BC_FLUX) {
          b[qcol[isp]] = qe[isp];
          src_q[isp] = -1;
        } else {
          x0[qcol[isp]] = qe[isp];
          src_q[isp] = qcol[isp];
        }
      }
    }
    // rows start on 128-byte lines (16 FP64 or 32 FP32 entries)
    const int64_t lda = mixed ? ((int64_t)nx + 31) / 32 * 32 : ((int64_t)nx + 15) / 16 * 16;
    DevBuf<double> sys, B, X, db, qs, du, dq;
    DevBuf<float> sys32;
    if ((mixed ? sys32.alloc((size_t)neq * lda, ctx) : sys.alloc((size_t)neq * lda, ctx)) ||
        B.alloc((size_t)neq * ANDR_NB, ctx) || X.alloc((size_t)nx * ANDR_NB, ctx) || qs.alloc(nx, ctx))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "Laplace system %d x %d", neq, nx);
    if (mixed)
      hipLaunchKernelGGL(k_lap_fill<float>, dim3(((size_t)neq * lda + 255) / 256), dim3(256), 0, st, (size_t)neq * lda,
                         sys32.p);
    else
      hipLaunchKernelGGL(k_lap_fill<double>, dim3(((size_t)neq * lda + 255) / 256), dim3(256), 0, st,
                         (size_t)neq * lda, sys.p);
    hipLaunchKernelGGL(k_lap_fill<double>, dim3(((size_t)neq * ANDR_NB + 255) / 256), dim3(256), 0, st,
                       (size_t)neq * ANDR_NB, B.p);
    {
      std::vector<double> xp((size_t)nx * ANDR_NB, 0.0);
      for (int i = 0; i < nx; ++i) xp[(size_t)i * ANDR_NB] = x0[i];
      ANDR_HIP(ctx, hipMemcpyAsync(X.p, xp.data(), sizeof(double) * xp.size(), hipMemcpyHostToDevice, st));
      if ((rc = upload(ctx, db, b))) return rc;
      // sdSetUpSysMrhsM_CRS: rows of flux / contact nodes, then of function / contact elements
      int row = 0;
      for (int s = 0; s < nsd; ++s) {
        const SdLists &l = L[s];
        const andr_subdomain &sd = sds[s];
        const int nns = (int)l.nodes.size(), nes = (int)l.elems.size();
        std::vector<int> rr, nc(nns), qc(nes);
        std::vector<uint8_t> nk(nns), qk(nes);
        const std::vector<double> gm = g_factors(sd, l);
        for (int k = 0; k < nns; ++k) {
          const int bc = bcw[l.nwall[k] - 1];
          nc[k] = col[l.nodes[k]];
          nk[k] = bc == ANDR_BC_FUNCTION;
          if (bc != ANDR_BC_FUNCTION) rr.push_back(k);
        }
        for (int k = 0; k < nes; ++k) {
          const int bc = bcw[l.qwall[k] - 1];
          qc[k] = qcol[l.elems[k]];
          qk[k] = bc == ANDR_BC_FLUX;
          if (bc != ANDR_BC_FLUX) rr.push_back(nns + k);
        }
        // every column must address the unknowns or the knowns it is written to: a point listed by two
        // subdomains with different kinds of condition would not (the system kernel trusts these maps)
        for (int k = 0; k < nns; ++k)
          if (nc[k] < 0 || nc[k] >= (nk[k] ? nb : nx))
            return andr_fail(ctx, ANDR_ERR_ARG, "node %d: conditions of different kinds in two subdomains", l.nodes[k]);
        for (int k = 0; k < nes; ++k)
          if (qc[k] < 0 || qc[k] >= (qk[k] ? nb : nx))
            return andr_fail(ctx, ANDR_ERR_ARG, "element %d: conditions of different kinds in two subdomains",
                             l.elems[k]);
        if (rr.empty()) continue;
        DevBuf<int> drr, dnc, dqc;
        DevBuf<uint8_t> dnk, dqk;
        DevBuf<double> dgm;
        if ((rc = upload(ctx, drr, rr)) || (rc = upload(ctx, dnc, nc)) || (rc = upload(ctx, dqc, qc)) ||
            (rc = upload(ctx, dnk, nk)) || (rc = upload(ctx, dqk, qk)) || (rc = upload(ctx, dgm, gm)))
          return rc;
        if (mixed)
          hipLaunchKernelGGL(k_lap_system<float>, dim3((int)rr.size()), dim3(256), 0, st, D[s].nn, D[s].ne, D[s].H.p,
                             D[s].G.p, drr.p, row, dnc.p, dnk.p, dqc.p, dqk.p, dgm.p, db.p, sys32.p, lda, B.p);
        else
          hipLaunchKernelGGL(k_lap_system<double>, dim3((int)rr.size()), dim3(256), 0, st, D[s].nn, D[s].ne, D[s].H.p,
                             D[s].G.p, drr.p, row, dnc.p, dnk.p, dqc.p, dqk.p, dgm.p, db.p, sys.p, lda, B.p);
        ANDR_HIP(ctx, hipGetLastError());
        ANDR_HIP(ctx, hipStreamSynchronize(st));  // the maps go back to the cache after this
        row += (int)rr.size();
      }
      if (row != neq) return andr_fail(ctx, ANDR_ERR_STATE, "system rows %d, expected %d", row, neq);
    }
    // lsqr_precon + lsqr_solve (column scaling in place, hot start, the reference's stopping rules)
    const int maxit = opts ? opts[en].maxit : 0;
    const double eps = opts && opts[en].eps > 0.0 ? opts[en].eps : 1e-15;
    void *A = mixed ? (void *)sys32.p : (void *)sys.p;
    if ((rc = andr_colscale(ctx, A, mixed, lda, neq, nx, qs.p))) return rc;
    int nit = 0, istop = 0;
    if ((rc = andr_lsqr_device(ctx, A, mixed, lda, neq, nx, qs.p, B.p, X.p, 1, maxit, eps, &nit, &istop)))
      return rc;
    // the solution back to u and q, then the wall fluxes
    DevBuf<int> dsu, dsq;
    if (du.alloc(nn, ctx) || dq.alloc(ne, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "boundary values");
    ANDR_HIP(ctx, hipMemcpyAsync(du.p, ue, sizeof(double) * nn, hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(dq.p, qe, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    if ((rc = upload(ctx, dsu, src_u)) || (rc = upload(ctx, dsq, src_q))) return rc;
    hipLaunchKernelGGL(k_lap_scatter, dim3((nn + 255) / 256), dim3(256), 0, st, nn, dsu.p, X.p, du.p);
    hipLaunchKernelGGL(k_lap_scatter, dim3((ne + 255) / 256), dim3(256), 0, st, ne, dsq.p, X.p, dq.p);
    DevBuf<double> dflux;
    if (wall_flux) {
      if (dflux.alloc(nw, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "wall fluxes");
      hipLaunchKernelGGL(k_wall_flux, dim3(nw), dim3(256), 0, st, ne, whole.g.bcid, whole.g.area, dq.p, dflux.p);
      ANDR_HIP(ctx, hipMemcpyAsync(wall_flux + (size_t)en * nw, dflux.p, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
    }
    ANDR_HIP(ctx, hipGetLastError());
    ANDR_HIP(ctx, hipMemcpyAsync(ue, du.p, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipMemcpyAsync(qe, dq.p, sizeof(double) * ne, hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));
    if (reports) {
      andr_laplace_report &R = reports[en];
      R.neq = neq;
      R.nx = nx;
      R.nb = nb;
      R.nit = nit;
      R.istop = istop;
      R.err[0] = errU;
      R.err[1] = std::max(errU, errQ);
      R.t_assemble = t1 - t0;
      R.t_solve = andr_now() - ts;
    }
  }
  return ANDR_OK;
}
}  // namespace

extern "C" int andr_laplace_solve(andr_ctx *ctx, const andr_mesh *mesh, int32_t nsd, const andr_subdomain *sds,
                                  int32_t neqn, const int32_t *wall_bc, const andr_lsqr_opts *opts, double *u,
                                  double *q, double *wall_flux, andr_laplace_report *reports) {
  return laplace_solve(ctx, mesh, nsd, sds, neqn, wall_bc, opts, u, q, wall_flux, reports, false);
}

extern "C" int andr_laplace_solve_mixed(andr_ctx *ctx, const andr_mesh *mesh, int32_t nsd,
                                        const andr_subdomain *sds, int32_t neqn, const int32_t *wall_bc,
                                        const andr_lsqr_opts *opts, double *u, double *q, double *wall_flux,
                                        andr_laplace_report *reports) {
  return laplace_solve(ctx, mesh, nsd, sds, neqn, wall_bc, opts, u, q, wall_flux, reports, true);
}

// ---- the general Stokes problem (one subdomain) ------------------------------------------------------
namespace {

// One block per system row (component een of source point r): for every column component en,
// sys[row][col] = -T_{een,en}(r, j) or -(normmul / diff) G_{een,en}(r, j) where the value is
// unknown, and rhs[row] = sum of +T and +(normmul / diff) G times the known values (formUrow /
// formQrow, src/lib_StokesFull.f90:4-160; b = rhsM b, src/lib_Stokes.f90:417-469).
__global__ __launch_bounds__(256) void k_stk_system(int nn_sd, int ne_sd, int nsp, const double *__restrict__ T,
                                                    const double *__restrict__ G, const int2 *__restrict__ rows,
                                                    const int *__restrict__ ncol, const uint8_t *__restrict__ nknown,
                                                    const int *__restrict__ qcol, const uint8_t *__restrict__ qknown,
                                                    const double *__restrict__ gmul, const double *__restrict__ b,
                                                    double *__restrict__ sys, int64_t lda, double *__restrict__ B) {
  __shared__ double red[4];
  const int row = blockIdx.x;
  const int een = rows[row].x, r = rows[row].y;
  double *srow = sys + (size_t)row * lda;
  const size_t sT = (size_t)nsp * nn_sd, sG = (size_t)nsp * ne_sd;
  double acc = 0.0;
  for (int en = 0; en < 3; ++en) {
    const double *tr = T + sym3(een, en) * sT + (size_t)r * nn_sd;
    const double *gr = G + sym3(een, en) * sG + (size_t)r * ne_sd;
    const int *nc = ncol + (size_t)en * nn_sd, *qc = qcol + (size_t)en * ne_sd;
    const uint8_t *nk = nknown + (size_t)en * nn_sd, *qk = qknown + (size_t)en * ne_sd;
    for (int j = threadIdx.x; j < nn_sd; j += blockDim.x) {
      const double v = tr[j];
      if (nk[j]) acc += v * b[nc[j]];
      else srow[nc[j]] = -v;
    }
    for (int j = threadIdx.x; j < ne_sd; j += blockDim.x) {
      const double v = gmul[j] * gr[j];
      if (qk[j]) acc += v * b[qc[j]];
      else srow[qc[j]] = -v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) B[(size_t)row * ANDR_NB] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace

extern "C" int andr_stokes_solve(andr_ctx *ctx, const andr_mesh *mesh, const andr_subdomain *sd,
                                 const int32_t *wall_bc, const andr_lsqr_opts *opts, double *u, double *q,
                                 double *wall_force, andr_laplace_report *report) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!wall_bc || !u || !q) return andr_fail(ctx, ANDR_ERR_ARG, "andr_stokes_solve arguments");
  int rc = andr_check_mesh(ctx, mesh);
  if (rc) return rc;
  if (mesh->nvert == 4) return andr_fail(ctx, ANDR_ERR_ARG, "andr_stokes_solve takes triangle meshes");
  char msg[256];
  std::vector<SdLists> LL;
  if ((rc = build_lists(mesh, 1, sd, LL, msg, sizeof(msg)))) return andr_fail(ctx, rc, "%s", msg);
  const SdLists &l = LL[0];
  const int nn = mesh->nnodes, ne = mesh->nelem, nw = mesh->nwalls > 0 ? mesh->nwalls : 1;
  for (int k = 0; k < 3 * nw; ++k)
    if (wall_bc[k] < ANDR_BC_FUNCTION || wall_bc[k] > ANDR_BC_CONTACT)
      return andr_fail(ctx, ANDR_ERR_ARG, "wall_bc[%d] = %d is not a boundary condition kind", k, wall_bc[k]);
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (report) std::memset(report, 0, sizeof(andr_laplace_report));
  const double t0 = andr_now();
  // ---- the 12 integral matrices of the subdomain (bounded form: no exterior free term) ------------
  SubMesh sub;
  make_submesh(mesh, *sd, l, sub);
  DevBody body;
  if ((rc = andr_body_upload(ctx, &sub.m, body, true))) return rc;
  const int nns = sub.m.nnodes, nes = sub.m.nelem, nsp = nns + nes;
  const size_t sT = (size_t)nsp * nns, sG = (size_t)nsp * nes;
  DevBuf<double> dT, dG, dpT, dpG, dE;
  if (dT.alloc(6 * sT, ctx) || dG.alloc(6 * sG, ctx) || dpT.alloc(3 * sT, ctx) || dpG.alloc(3 * sG, ctx) ||
      dE.alloc(nsp, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "Stokes integral matrices of the subdomain");
  if ((rc = andr_launch_stokes_rows(ctx, body, 0.0, dT.p, dG.p, dpT.p, dpG.p, dE.p))) return rc;
  std::vector<double> he(nsp);
  ANDR_HIP(ctx, hipMemcpyAsync(he.data(), dE.p, sizeof(double) * nsp, hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipStreamSynchronize(st));
  dpT.release();
  dpG.release();
  double errU = 0.0, errQ = 0.0;
  for (int i = 0; i < nsp; ++i) {
    if (i < nns) errU = std::max(errU, he[i]);
    else errQ = std::max(errQ, he[i]);
  }
  const double t1 = andr_now();
  // ---- sdSetUpXandB per component, stacked as StokesBigSystemXB ---------------------------------------
  std::vector<int> ncol(3 * (size_t)nns), qcol(3 * (size_t)nes);
  std::vector<uint8_t> nknown(3 * (size_t)nns), qknown(3 * (size_t)nes);
  int nx = 0, nb = 0;
  for (int en = 0; en < 3; ++en) {
    const int32_t *bcw = wall_bc + (size_t)en * nw;
    int cx = 0, cb = 0;  // this component's unknowns and knowns, in the reference's order
    for (int k = 0; k < nns; ++k) {
      const bool kn = bcw[l.nwall[k] - 1] == ANDR_BC_FUNCTION;
      nknown[(size_t)en * nns + k] = kn;
      ncol[(size_t)en * nns + k] = kn ? nb + cb++ : nx + cx++;
    }
    for (int k = 0; k < nes; ++k) {
      const bool kn = bcw[l.qwall[k] - 1] == ANDR_BC_FLUX;
      qknown[(size_t)en * nes + k] = kn;
      qcol[(size_t)en * nes + k] = kn ? nb + cb++ : nx + cx++;
    }
    nx += cx;
    nb += cb;
  }
  // rows in order: nodes, then elements, the components of a point inside (stokesFormCRSsysMrhsM)
  std::vector<int2> rows;
  for (int k = 0; k < nns; ++k)
    for (int een = 0; een < 3; ++een)
      if (!nknown[(size_t)een * nns + k]) rows.push_back(make_int2(een, k));
  for (int k = 0; k < nes; ++k)
    for (int een = 0; een < 3; ++een)
      if (!qknown[(size_t)een * nes + k]) rows.push_back(make_int2(een, nns + k));
  const int neq = (int)rows.size();
  if (nx == 0 || neq == 0) return andr_fail(ctx, ANDR_ERR_ARG, "the Stokes problem has no unknowns");
  // formStokesLeftRightHandSideVectors: knowns b and the hot start x from u and q
  std::vector<double> b(std::max(nb, 1), 0.0), xp((size_t)nx * ANDR_NB, 0.0);
  std::vector<int> src_u(3 * (size_t)nn, -1), src_q(3 * (size_t)ne, -1);
  for (int en = 0; en < 3; ++en) {
    for (int k = 0; k < nns; ++k) {
      const int isp = l.nodes[k], c = ncol[(size_t)en * nns + k];
      const double v = u[(size_t)en * nn + isp];
      if (nknown[(size_t)en * nns + k]) b[c] = v;
      else {
        xp[(size_t)c * ANDR_NB] = v;
        src_u[(size_t)en * nn + isp] = c;
      }
    }
    for (int k = 0; k < nes; ++k) {
      const int isp = l.elems[k], c = qcol[(size_t)en * nes + k];
      const double v = q[(size_t)en * ne + isp];
      if (qknown[(size_t)en * nes + k]) b[c] = v;
      else {
        xp[(size_t)c * ANDR_NB] = v;
        src_q[(size_t)en * ne + isp] = c;
      }
    }
  }
  const int64_t lda = ((int64_t)nx + 15) / 16 * 16;
  DevBuf<double> sys, B, X, db, qs, dgm, du, dq;
  DevBuf<int2> drows;
  DevBuf<int> dnc, dqc, dsu, dsq;
  DevBuf<uint8_t> dnk, dqk;
  if (sys.alloc((size_t)neq * lda, ctx) || B.alloc((size_t)neq * ANDR_NB, ctx) || X.alloc((size_t)nx * ANDR_NB, ctx) ||
      qs.alloc(nx, ctx) || du.alloc(3 * (size_t)nn, ctx) || dq.alloc(3 * (size_t)ne, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "Stokes system %d x %d", neq, nx);
  if ((rc = upload(ctx, db, b)) || (rc = upload(ctx, dgm, g_factors(*sd, l))) || (rc = upload(ctx, drows, rows)) ||
      (rc = upload(ctx, dnc, ncol)) || (rc = upload(ctx, dqc, qcol)) || (rc = upload(ctx, dnk, nknown)) ||
      (rc = upload(ctx, dqk, qknown)) || (rc = upload(ctx, dsu, src_u)) || (rc = upload(ctx, dsq, src_q)))
    return rc;
  ANDR_HIP(ctx, hipMemcpyAsync(X.p, xp.data(), sizeof(double) * xp.size(), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_lap_fill<double>, dim3(((size_t)neq * lda + 255) / 256), dim3(256), 0, st, (size_t)neq * lda, sys.p);
  hipLaunchKernelGGL(k_lap_fill<double>, dim3(((size_t)neq * ANDR_NB + 255) / 256), dim3(256), 0, st, (size_t)neq * ANDR_NB,
                     B.p);
  hipLaunchKernelGGL(k_stk_system, dim3(neq), dim3(256), 0, st, nns, nes, nsp, dT.p, dG.p, drows.p, dnc.p, dnk.p,
                     dqc.p, dqk.p, dgm.p, db.p, sys.p, lda, B.p);
  ANDR_HIP(ctx, hipGetLastError());
  const int maxit = opts ? opts->maxit : 0;
  const double eps = opts && opts->eps > 0.0 ? opts->eps : 1e-15;
  if ((rc = andr_colscale(ctx, sys.p, false, lda, neq, nx, qs.p))) return rc;
  int nit = 0, istop = 0;
  if ((rc = andr_lsqr_device(ctx, sys.p, false, lda, neq, nx, qs.p, B.p, X.p, 1, maxit, eps, &nit, &istop))) return rc;
  // StokesDistributeUnknowns, then the force on every wall (sum_e A_e q_e per component)
  ANDR_HIP(ctx, hipMemcpyAsync(du.p, u, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(dq.p, q, sizeof(double) * 3 * ne, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_lap_scatter, dim3((3 * nn + 255) / 256), dim3(256), 0, st, 3 * nn, dsu.p, X.p, du.p);
  hipLaunchKernelGGL(k_lap_scatter, dim3((3 * ne + 255) / 256), dim3(256), 0, st, 3 * ne, dsq.p, X.p, dq.p);
  DevBody whole;
  DevBuf<double> dforce;
  if (wall_force) {
    if ((rc = andr_body_upload(ctx, mesh, whole, false))) return rc;
    if (dforce.alloc(3 * (size_t)nw, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "wall forces");
    for (int en = 0; en < 3; ++en)
      hipLaunchKernelGGL(k_wall_flux, dim3(nw), dim3(256), 0, st, ne, whole.g.bcid, whole.g.area, dq.p + (size_t)en * ne,
                         dforce.p + (size_t)en * nw);
    ANDR_HIP(ctx, hipMemcpyAsync(wall_force, dforce.p, sizeof(double) * 3 * nw, hipMemcpyDeviceToHost, st));
  }
  ANDR_HIP(ctx, hipGetLastError());
  ANDR_HIP(ctx, hipMemcpyAsync(u, du.p, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipMemcpyAsync(q, dq.p, sizeof(double) * 3 * ne, hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipStreamSynchronize(st));
  if (report) {
    report->neq = neq;
    report->nx = nx;
    report->nb = nb;
    report->nit = nit;
    report->istop = istop;
    report->err[0] = errU;
    report->err[1] = std::max(errU, errQ);
    report->t_assemble = t1 - t0;
    report->t_solve = andr_now() - t1;
  }
  return ANDR_OK;
}
