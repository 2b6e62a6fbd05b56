// Resistance tensors of rigid bodies: the six rigid-body Stokes problems per body, assembled and
// solved on the device (replaces StokesInlet3sidesForceCRS / TorqueOmegaCRS,
// src/lib_particle.f90:176-476, and the force/torque reductions :1751-1831).
//
// Plain FP64 path: FP64 assembly and storage, column-scaled LSQR (the reference solver) with the
// six right-hand sides in one block. Mixed path: FP32 regular quadrature, FP64 singular blocks,
// FP32 storage of A M^-1 (M: 3x3 self blocks, a right block-Jacobi preconditioner), batched
// restarted GMRES with FP64 vectors and residuals, x = M^-1 y at the end. Bodies are processed in
// waves sized to the free device memory; each wave is one assembly launch for all its bodies.
#include "andr_internal.h"
#include "andr_assemble.h"
#include "andr_solve.h"
#include "andr_learned.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace {

// one body on the device: geometry, node -> wall map and the null-space scales of its walls
struct Prepared {
  DevBody body;
  std::vector<int> node_wall;
  std::vector<double> defl;  // per wall id 0..nwalls (iterative forms)
  int ne = 0, nn = 0;
  bool moving_alike = true;  // bounded form: the three nodes of every element all move or all rest
};

// Null-space deflation for the iterative (mixed and learned) forms. On a closed surface the
// single-layer operator annihilates the normal density, int_S G(x,y) n(y) dS_y = 0, so the
// discrete system has a near-null right vector v (v_e = n_e) and left vector w (w_r = A_r n_r).
// Discretisation error leaves a small component of the right-hand side along w that no
// traction field fits; least squares (the reference's LSQR) leaves that direction unresolved,
// and GMRES stalls on it. Adding d w v^T, with d of the size of the smooth-mode eigenvalues,
// moves the direction into the bulk of the spectrum and makes the system consistent. Force and
// torque are blind to the normal density on a closed polyhedron (sum A_e n_e = 0 and
// sum A_e c_e x n_e = 0 exactly), so they are unchanged. Only walls that are closed, consistently
// oriented surfaces (every edge shared by two of its elements in opposite directions) get the
// term; defl[w] = 0 for the others. defl is indexed by wall id 1..nwalls.
void deflation_scales(const andr_mesh *m, const double *nm, double mu, std::vector<double> &defl) {
  const int nw = std::max(1, m->nwalls), ne = m->nelem, base = m->index_base;
  defl.assign(nw + 1, 0.0);
  std::vector<double> S(nw + 1, 0.0), S2(nw + 1, 0.0);
  std::vector<int> cnt(nw + 1, 0);
  std::vector<std::vector<std::pair<uint64_t, int>>> edges(nw + 1);
  for (int e = 0; e < ne; ++e) {
    const int w = m->bcid ? m->bcid[e] : 1;
    const int v[3] = {m->con[3 * e] - base, m->con[3 * e + 1] - base, m->con[3 * e + 2] - base};
    const double *p0 = m->xyz + 3 * v[0], *p1 = m->xyz + 3 * v[1], *p2 = m->xyz + 3 * v[2];
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    S[w] += area;
    S2[w] += area * area;
    ++cnt[w];
    for (int k = 0; k < 3; ++k) {
      const uint32_t i = (uint32_t)v[k], j = (uint32_t)v[(k + 1) % 3];
      const uint64_t key = ((uint64_t)std::min(i, j) << 32) | std::max(i, j);
      edges[w].push_back({key, i < j ? 1 : -1});
    }
  }
  for (int w = 1; w <= std::min(nw, ANDR_OP_MAXW); ++w) {  // the operator carries the term for walls 1..4
    auto &E = edges[w];
    if (E.empty()) continue;
    std::sort(E.begin(), E.end());
    bool closed = E.size() % 2 == 0;
    for (size_t k = 0; closed && k < E.size(); k += 2)
      closed = E[k].first == E[k + 1].first && E[k].second + E[k + 1].second == 0 &&
               (k + 2 >= E.size() || E[k + 2].first != E[k].first);
    if (!closed) continue;
    // smooth-mode eigenvalue scale: a sphere of radius a has 16 pi a / 3 for uniform traction
    const double a = std::sqrt(S[w] / (4.0 * ANDR_PI));
    const double lam = (8.0 * ANDR_PI / 3.0) * a;
    defl[w] = (-nm[w - 1] / mu) * lam / (std::sqrt(S2[w]) * std::sqrt((double)cnt[w]));
  }
}

}  // namespace

static int check_opts(andr_ctx *ctx, const andr_resistance_opts *o) {
  if (!o) return andr_fail(ctx, ANDR_ERR_ARG, "null options");
  if (o->form != ANDR_FORM_EXTERIOR && o->form != ANDR_FORM_BOUNDED) return andr_fail(ctx, ANDR_ERR_ARG, "bad form");
  if (o->precision < ANDR_PREC_FP64 || o->precision > ANDR_PREC_LEARNED)
    return andr_fail(ctx, ANDR_ERR_ARG, "bad precision");
  if (o->precision == ANDR_PREC_LEARNED && !ctx->has_model)
    return andr_fail(ctx, ANDR_ERR_STATE, "learned precision needs andr_set_quadrature_model first");
  if (!(o->mu > 0.0)) return andr_fail(ctx, ANDR_ERR_ARG, "mu must be positive");
  if (o->problems != ANDR_PROBLEMS_RIGID && o->problems != ANDR_PROBLEMS_STRAIN)
    return andr_fail(ctx, ANDR_ERR_ARG, "problems must be ANDR_PROBLEMS_RIGID or ANDR_PROBLEMS_STRAIN");
  return ANDR_OK;
}

namespace {

// normal orientation factor of every wall: the exterior form takes the normals into the body (out
// of the fluid), whatever the input orientation factor
std::vector<double> wall_normmul(const andr_mesh *m, bool exterior) {
  std::vector<double> nm(std::max(1, m->nwalls), 1.0);
  if (m->normmul)
    for (int w = 0; w < m->nwalls; ++w) nm[w] = m->normmul[w];
  if (exterior)
    for (auto &v : nm) v = -std::fabs(v);
  return nm;
}

// Uploads bodies [b0, b1) of `meshes` into P: the closure tests of the walls (iterative forms) on
// up to 16 host threads, then geometry and node -> wall maps.
int prepare_bodies(andr_ctx *ctx, const andr_mesh *meshes, int b0, int b1, bool exterior, double mu, bool with_defl,
                   std::vector<std::unique_ptr<Prepared>> &P) {
  const int nb = b1 - b0;
  P.clear();
  P.resize(nb);
  for (auto &p : P) p.reset(new Prepared());
  if (with_defl) {
    const int nt = std::max(1, std::min<int>(nb, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&, t]() {
        for (int i = t; i < nb; i += nt) {
          const andr_mesh *m = &meshes[b0 + i];
          deflation_scales(m, wall_normmul(m, exterior).data(), mu, P[i]->defl);
        }
      });
    for (auto &th : pool) th.join();
  }
  for (int i = 0; i < nb; ++i) {
    const andr_mesh *m = &meshes[b0 + i];
    Prepared &p = *P[i];
    const std::vector<double> nm = wall_normmul(m, exterior);
    andr_mesh mm = *m;
    mm.normmul = nm.data();
    if (mm.nwalls <= 0) mm.nwalls = 1;
    int rc = andr_body_upload(ctx, &mm, p.body, false);
    if (rc) return rc;
    p.nn = m->nnodes;
    p.ne = m->nelem;
    p.node_wall.assign(p.nn, 0);
    for (int e = 0; e < p.ne; ++e) {
      const int w = m->bcid ? m->bcid[e] : 1;
      for (int v = 0; v < 3; ++v) {
        int nd = m->con[3 * e + v] - m->index_base;
        p.node_wall[nd] = std::max(p.node_wall[nd], w);
      }
    }
    for (int e = 0; e < p.ne && p.moving_alike; ++e) {  // moving nodes: those of wall 2 (run_wave)
      const int *c = m->con + 3 * e;
      const bool m0 = p.node_wall[c[0] - m->index_base] == 2;
      p.moving_alike = m0 == (p.node_wall[c[1] - m->index_base] == 2) && m0 == (p.node_wall[c[2] - m->index_base] == 2);
    }
  }
  return ANDR_OK;
}

// device bytes of one body's system in a form (matrix, vectors, Krylov basis, rule classes)
size_t body_bytes(int ne, int nn, bool plain, bool learned, int restart) {
  const size_t n = 3 * (size_t)ne;
  const size_t lda = (n + 63) / 64 * 64;
  size_t s = (plain ? n * lda * sizeof(double) : andr_sym_elems(ne) * sizeof(float)) + 4 * n * ANDR_NRHS * sizeof(double);
  if (learned) s += (size_t)ne * ne;  // rule class per pair
  s += plain ? 10 * n * ANDR_NRHS * sizeof(double) : (size_t)(restart + 4) * n * ANDR_NRHS * sizeof(double);
  s += (size_t)nn * 200 + (size_t)ne * 300;
  return s;
}

// The first cycle's length: 128 Krylov steps, longer (up to 192, in steps of 8) while the bodies of
// the call still fit one wave. Slender bodies then converge without a restart: on the 576-body
// workload a 160-step cycle took the fibres from ~290 steps (128 + a long second cycle) to <= 160
// and the step from 2.44 to 2.38 s, while 96 steps made it 2.49 s. Later cycles take the longest
// restart that fits the same buffers.
constexpr int RESTART = 128, RESTART_MAX = 192;

// The resistance tensors of one wave of prepared bodies: assembly, solve, force and torque.
// FT and reports point at the wave's first body; t_up is the wave's upload time (reported).
int run_wave(andr_ctx *ctx, Prepared *const *P, int nb, const andr_resistance_opts *opts, int restart, bool timing,
             double t_up, double *FT, andr_resistance_report *reports) {
  int rc = ANDR_OK;
  const bool exterior = opts->form == ANDR_FORM_EXTERIOR;
  const bool plain = opts->precision == ANDR_PREC_FP64;
  const bool learned = opts->precision == ANDR_PREC_LEARNED;
  const size_t esz = plain ? sizeof(double) : sizeof(float);
  // fast forms: GMRES relative residual 1e-6. On 48 bodies of the batched workload the largest
  // tensor distance to the plain FP64 path is 2.2e-8 at 1e-7 and 3e-7 (the FP32 quadrature's), 8.6e-8
  // at 1e-6 and 5.1e-7 at 3e-6; each factor of ten costs ~10 % more Krylov steps (batch step 2.24 s
  // at 1e-7, 2.00 s at 1e-6). 1e-6 keeps the tensors an order of magnitude inside 1e-6.
  const double eps = opts->eps > 0.0 ? opts->eps : (plain ? 1e-12 : 1e-6);
  const double extra = exterior ? 8.0 * ANDR_PI : 0.0;
  const double ta0 = andr_now();
  std::vector<DevBuf<char>> Abuf(nb);
  std::vector<DevBuf<double>> Bbuf(nb), Xbuf(nb), Ebuf(nb), Dbuf(nb), Pbuf(nb);
  std::vector<DevBuf<uint8_t>> Mbuf(nb);
  std::vector<DevBuf<uint8_t>> Cbuf(nb);
  std::vector<SysJob> jobs(nb);
  std::vector<int2> rowmap;
  for (int i = 0; i < nb; ++i) {
    Prepared &p = *P[i];
    const int n = 3 * p.ne;
    const int64_t lda = plain ? (n + 63) / 64 * 64 : 0;  // fast forms: symmetric blocks, FP32
    if (Abuf[i].alloc(plain ? (size_t)n * lda * esz : andr_sym_elems(p.ne) * sizeof(float), ctx) ||
        Bbuf[i].alloc((size_t)n * ANDR_NRHS, ctx) || Xbuf[i].alloc((size_t)n * ANDR_NRHS, ctx) ||
        Ebuf[i].alloc(p.ne, ctx))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "resistance system of body %d (n = %d)", i, n);
    if (!exterior) {  // bounded form: only the nodes of wall 2 move (src/lib_particle.f90:220-256, 386-409)
      std::vector<uint8_t> mv(p.nn);
      for (int k = 0; k < p.nn; ++k) mv[k] = p.node_wall[k] == 2 ? 1 : 0;
      if (Mbuf[i].alloc(p.nn, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "moving-node flags");
      ANDR_HIP(ctx, hipMemcpyAsync(Mbuf[i].p, mv.data(), p.nn, hipMemcpyHostToDevice, ctx->stream));
      ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // mv goes out of scope
    }
    const QuadRule qr = andr_rule(p.body.itri);
    SysJob &J = jobs[i];
    J.g = p.body.g;
    J.A_ = Abuf[i].p;
    J.lda = lda;
    J.B = Bbuf[i].p;
    J.moving = exterior ? nullptr : Mbuf[i].p;
    for (int d = 0; d < 3; ++d) J.center[d] = opts->center[d];
    J.strain = opts->problems == ANDR_PROBLEMS_STRAIN ? 1 : 0;
    J.mu = opts->mu;
    J.extra = extra;
    J.qoff = qr.offset;
    J.nq = qr.npts;
    J.err_rows = Ebuf[i].p;
    J.cls = nullptr;
    J.defl = nullptr;
    J.minv = nullptr;
    J.X = Xbuf[i].p;
    if (!plain) {
      if (Pbuf[i].alloc(9 * (size_t)p.ne, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "preconditioner blocks");
      J.minv = Pbuf[i].p;
      const std::vector<double> &dh = p.defl;
      if (Dbuf[i].alloc(dh.size(), ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "deflation scales");
      ANDR_HIP(ctx, hipMemcpyAsync(Dbuf[i].p, dh.data(), sizeof(double) * dh.size(), hipMemcpyHostToDevice, ctx->stream));
      J.defl = Dbuf[i].p;
    }
    if (learned) {
      if (Cbuf[i].alloc((size_t)p.ne * p.ne, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "rule classes of body %d", i);
      J.cls = Cbuf[i].p;
    }
    for (int r = 0; r < p.ne; ++r) rowmap.push_back(make_int2(i, r));
  }
  const double tu1 = andr_now();  // job tables
  DevBuf<SysJob> djobs;
  DevBuf<int2> drow;
  if (djobs.alloc(nb, ctx) || drow.alloc(rowmap.size(), ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "job tables");
  ANDR_HIP(ctx, hipMemcpyAsync(djobs.p, jobs.data(), sizeof(SysJob) * nb, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(drow.p, rowmap.data(), sizeof(int2) * rowmap.size(), hipMemcpyHostToDevice, ctx->stream));
  std::vector<LearnedStats> ls(nb, LearnedStats{0, 0, 0});
  double alg_bytes = 0.0, alg_flops = 0.0;
  // fast forms: vertex-summed stresslet moments when every element moves rigidly as a whole
  bool summed = !plain && opts->problems == ANDR_PROBLEMS_RIGID;
  for (int i = 0; i < nb; ++i) summed = summed && (exterior || P[i]->moving_alike);
  for (int i = 0; i < nb; ++i) {
    const double ne = P[i]->ne;
    // plain: the full 3F x 3F FP64 matrix; fast forms: the six symmetric FP32 blocks (6 F^2)
    alg_bytes += (plain ? 9.0 : 6.0) * ne * ne * esz + 3.0 * ne * ANDR_NRHS * 8.0;
    alg_flops += ne * (ne - 1.0) * jobs[i].nq * (summed ? ANDR_FLOPS_PER_STOKES_EVAL_SUM : ANDR_FLOPS_PER_STOKES_EVAL);
  }
  if (!plain && (rc = andr_launch_self_blocks(ctx, djobs.p, drow.p, (int)rowmap.size()))) return rc;
  if (learned)
    rc = andr_launch_stokes_system_learned(ctx, jobs.data(), djobs.p, drow.p, (int)rowmap.size(), nb, summed, ls.data());
  else
    rc = andr_launch_stokes_system(ctx, djobs.p, drow.p, (int)rowmap.size(), !plain, !plain, alg_bytes, alg_flops, summed);
  if (rc) return rc;
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double ta1 = andr_now();
  // ---- solve ----------------------------------------------------------------------------------
  std::vector<int> iters((size_t)nb * 6, 0), stops((size_t)nb * 6, 0);
  const int nprob = opts->problems == ANDR_PROBLEMS_STRAIN ? 3 : 6;  // right-hand sides per body
  if (plain) {
    // column scaling of every body, then the LSQR solves of the wave together on several streams
    std::vector<DevBuf<double>> qs(nb);
    std::vector<LsqrJob> lj(nb);
    for (int i = 0; i < nb; ++i) {
      const int n = 3 * P[i]->ne;
      if (qs[i].alloc(n, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "column scale");
      rc = andr_colscale(ctx, Abuf[i].p, false, jobs[i].lda, n, n, qs[i].p);
      if (rc) return rc;
      ANDR_HIP(ctx, hipMemsetAsync(Xbuf[i].p, 0, sizeof(double) * n * ANDR_NRHS, ctx->stream));
      lj[i] = {Abuf[i].p, false, jobs[i].lda, n, n, qs[i].p, Bbuf[i].p, Xbuf[i].p, nprob, opts->maxit, eps,
               &iters[(size_t)i * 6], &stops[(size_t)i * 6]};
    }
    if ((rc = andr_lsqr_device_batch(ctx, nb, lj.data()))) return rc;
  } else {
    std::vector<GmresSys> gs(nb);
    for (int i = 0; i < nb; ++i) {
      const BodyGeom &g = P[i]->body.g;
      StokesOp op{(const float *)Abuf[i].p, P[i]->ne, jobs[i].minv, g.area, g.nrm, g.bcid, jobs[i].defl};
      gs[i] = {nullptr, 0, 3 * P[i]->ne, op, Bbuf[i].p, Xbuf[i].p};
    }
    std::vector<int> it_g((size_t)nb * nprob), st_g((size_t)nb * nprob);
    rc = andr_gmres_batch(ctx, gs.data(), nb, true, nprob, restart, opts->maxit > 0 ? opts->maxit : 2000, eps,
                          it_g.data(), st_g.data());
    for (int i = 0; i < nb; ++i)
      for (int k = 0; k < nprob; ++k) {
        iters[(size_t)i * 6 + k] = it_g[(size_t)i * nprob + k];
        stops[(size_t)i * 6 + k] = st_g[(size_t)i * nprob + k];
      }
    if (rc) return rc;
    if ((rc = andr_launch_unprecondition(ctx, djobs.p, drow.p, (int)rowmap.size()))) return rc;
  }
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double ts1 = andr_now();
  // ---- force and torque on wall 1 (on the device, one copy back for the wave) ------------------
  std::vector<double> ftbuf;  // per body: 36 force/torque values and the self-check error
  {
    DevBuf<double> dft;
    if (dft.alloc((size_t)nb * 37, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "force/torque results");
    if ((rc = andr_launch_force_torque(ctx, djobs.p, nb, dft.p))) return rc;
    ftbuf.resize((size_t)nb * 37);
    ANDR_HIP(ctx, hipMemcpyAsync(ftbuf.data(), dft.p, sizeof(double) * ftbuf.size(), hipMemcpyDeviceToHost,
                                 ctx->stream));
    ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  for (int i = 0; i < nb; ++i) {
    Prepared &p = *P[i];
    const int ne = p.ne;
    double *ft = FT + (size_t)i * 36;
    for (int k = 0; k < 36; ++k) ft[k] = ftbuf[(size_t)i * 37 + k];
    if (reports) {
      andr_resistance_report &R = reports[i];
      for (int k = 0; k < 6; ++k) {
        R.nit[k] = iters[(size_t)i * 6 + k];
        R.status = std::max(R.status, stops[(size_t)i * 6 + k] > 3 ? stops[(size_t)i * 6 + k] : 0);
      }
      R.err[0] = 0.0;
      R.err[1] = ftbuf[(size_t)i * 37 + 36];
      R.t_assemble = (t_up + ta1 - ta0) / nb;
      R.t_solve = (ts1 - ta1) / nb;
      const QuadRule qr = andr_rule(p.body.itri);
      R.quad_points = learned ? ls[i].points : (int64_t)ne * (ne - 1) * qr.npts;
      R.pairs_learned = learned ? ls[i].learned : 0;
      R.pairs_fallback = learned ? ls[i].fallback : 0;
    }
  }
  if (timing)
    fprintf(stderr, "[andr] wave of %d bodies: upload %.3f s, system set-up %.3f s, assembly %.3f s, solve %.3f s, "
            "force/torque %.3f s\n", nb, t_up, tu1 - ta0, ta1 - tu1, ts1 - ta1, andr_now() - ts1);
  return ANDR_OK;
}

// device memory the systems of a call may take: 80 % of the free memory (with the context's cache)
int wave_budget(andr_ctx *ctx, size_t *budget) {
  size_t free_b = 0, total_b = 0;
  ANDR_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
  free_b += ctx->pool.cached;  // the context's cached blocks are reusable
  *budget = (size_t)(0.80 * (double)free_b);
  return ANDR_OK;
}

// the first cycle's length for bodies (ne, nn): RESTART, or longer while they fit one wave
int first_restart(size_t budget, const std::vector<int> &ne, const std::vector<int> &nn, bool plain, bool learned,
                  int asked) {
  if (plain) return RESTART;
  if (asked > 0) return std::min(asked, 1024);
  int best = RESTART;
  for (int r = RESTART + 8; r <= RESTART_MAX; r += 8) {
    size_t total = 0;
    for (size_t b = 0; b < ne.size(); ++b) total += body_bytes(ne[b], nn[b], plain, learned, r);
    if (total > budget) break;
    best = r;
  }
  return best;
}

// waves of bodies whose systems fit the budget
int plan_waves(andr_ctx *ctx, size_t budget, const std::vector<size_t> &bytes, std::vector<int> &ends) {
  ends.clear();
  const int nb = (int)bytes.size();
  int b0 = 0;
  while (b0 < nb) {
    int b1 = b0;
    size_t used = 0;
    while (b1 < nb && (b1 == b0 || used + bytes[b1] <= budget)) used += bytes[b1++];
    if (used > budget) return andr_fail(ctx, ANDR_ERR_ALLOC, "body %d needs %zu bytes, budget %zu", b0, used, budget);
    ends.push_back(b1);
    b0 = b1;
  }
  return ANDR_OK;
}

}  // namespace

// device-resident bodies (andr_bodies_upload)
struct andr_bodies {
  std::vector<std::unique_ptr<Prepared>> P;
  int32_t form = ANDR_FORM_EXTERIOR;
  double mu = 1.0;
};

static int check_meshes(andr_ctx *ctx, int32_t nbodies, const andr_mesh *meshes) {
  if (nbodies <= 0 || !meshes) return andr_fail(ctx, ANDR_ERR_ARG, "no bodies");
  for (int b = 0; b < nbodies; ++b) {
    // every array is checked here, before the host-side closure tests read the meshes
    int rc = andr_check_mesh(ctx, &meshes[b]);
    if (rc) return andr_fail(ctx, rc, "body %d: %s", b, std::string(ctx->err).c_str());
    if (meshes[b].nvert == 4)
      return andr_fail(ctx, ANDR_ERR_ARG, "body %d: resistance tensors take triangle meshes (nvert 0 or 3)", b);
  }
  return ANDR_OK;
}

extern "C" int andr_resistance_batch(andr_ctx *ctx, int32_t nbodies, const andr_mesh *meshes,
                                     const andr_resistance_opts *opts, double *FT, andr_resistance_report *reports) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!FT) return andr_fail(ctx, ANDR_ERR_ARG, "andr_resistance_batch arguments");
  int rc = check_meshes(ctx, nbodies, meshes);
  if (rc || (rc = check_opts(ctx, opts))) return rc;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const bool plain = opts->precision == ANDR_PREC_FP64, learned = opts->precision == ANDR_PREC_LEARNED;
  if (reports) std::memset(reports, 0, sizeof(andr_resistance_report) * nbodies);
  const bool timing = getenv("ANDR_DEBUG_TIMING") != nullptr;  // per-wave host timing on stderr
  std::vector<int> ne(nbodies), nn(nbodies);
  for (int b = 0; b < nbodies; ++b) {
    ne[b] = meshes[b].nelem;
    nn[b] = meshes[b].nnodes;
  }
  size_t budget = 0;
  if ((rc = wave_budget(ctx, &budget))) return rc;
  const int restart = first_restart(budget, ne, nn, plain, learned, opts->restart);
  std::vector<size_t> bytes(nbodies);
  for (int b = 0; b < nbodies; ++b) bytes[b] = body_bytes(ne[b], nn[b], plain, learned, restart);
  std::vector<int> ends;
  if ((rc = plan_waves(ctx, budget, bytes, ends))) return rc;
  int b0 = 0;
  for (int b1 : ends) {
    const double t0 = andr_now();
    std::vector<std::unique_ptr<Prepared>> P;
    if ((rc = prepare_bodies(ctx, meshes, b0, b1, opts->form == ANDR_FORM_EXTERIOR, opts->mu, !plain, P))) return rc;
    std::vector<Prepared *> pp(P.size());
    for (size_t i = 0; i < P.size(); ++i) pp[i] = P[i].get();
    if ((rc = run_wave(ctx, pp.data(), b1 - b0, opts, restart, timing, andr_now() - t0, FT + (size_t)b0 * 36,
                       reports ? reports + b0 : nullptr)))
      return rc;
    b0 = b1;
  }
  return ANDR_OK;
}

extern "C" int andr_bodies_upload(andr_ctx *ctx, int32_t nbodies, const andr_mesh *meshes,
                                  const andr_resistance_opts *opts, andr_bodies **out) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!out) return andr_fail(ctx, ANDR_ERR_ARG, "andr_bodies_upload: null output");
  *out = nullptr;
  int rc = check_meshes(ctx, nbodies, meshes);
  if (rc) return rc;
  if (!opts || (opts->form != ANDR_FORM_EXTERIOR && opts->form != ANDR_FORM_BOUNDED) || !(opts->mu > 0.0))
    return andr_fail(ctx, ANDR_ERR_ARG, "andr_bodies_upload: form and mu");
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<andr_bodies> B(new andr_bodies());
  B->form = opts->form;
  B->mu = opts->mu;
  if ((rc = prepare_bodies(ctx, meshes, 0, nbodies, opts->form == ANDR_FORM_EXTERIOR, opts->mu, true, B->P))) return rc;
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = B.release();
  return ANDR_OK;
}

extern "C" int andr_resistance_batch_resident(andr_ctx *ctx, const andr_bodies *bodies,
                                              const andr_resistance_opts *opts, double *FT,
                                              andr_resistance_report *reports) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!bodies || !FT) return andr_fail(ctx, ANDR_ERR_ARG, "andr_resistance_batch_resident arguments");
  int rc = check_opts(ctx, opts);
  if (rc) return rc;
  if (opts->form != bodies->form || opts->mu != bodies->mu)
    return andr_fail(ctx, ANDR_ERR_ARG, "form and mu must be those the bodies were uploaded with");
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const int nbodies = (int)bodies->P.size();
  const bool plain = opts->precision == ANDR_PREC_FP64, learned = opts->precision == ANDR_PREC_LEARNED;
  if (reports) std::memset(reports, 0, sizeof(andr_resistance_report) * nbodies);
  const bool timing = getenv("ANDR_DEBUG_TIMING") != nullptr;
  std::vector<int> ne(nbodies), nn(nbodies);
  for (int b = 0; b < nbodies; ++b) {
    ne[b] = bodies->P[b]->ne;
    nn[b] = bodies->P[b]->nn;
  }
  size_t budget = 0;
  if ((rc = wave_budget(ctx, &budget))) return rc;
  const int restart = first_restart(budget, ne, nn, plain, learned, opts->restart);
  std::vector<size_t> bytes(nbodies);
  for (int b = 0; b < nbodies; ++b) bytes[b] = body_bytes(ne[b], nn[b], plain, learned, restart);
  std::vector<int> ends;
  if ((rc = plan_waves(ctx, budget, bytes, ends))) return rc;
  std::vector<Prepared *> pp(nbodies);
  for (int b = 0; b < nbodies; ++b) pp[b] = bodies->P[b].get();
  int b0 = 0;
  for (int b1 : ends) {
    if ((rc = run_wave(ctx, pp.data() + b0, b1 - b0, opts, restart, timing, 0.0, FT + (size_t)b0 * 36,
                       reports ? reports + b0 : nullptr)))
      return rc;
    b0 = b1;
  }
  return ANDR_OK;
}

extern "C" int andr_bodies_count(const andr_bodies *bodies) { return bodies ? (int)bodies->P.size() : 0; }

extern "C" int andr_bodies_free(andr_ctx *ctx, andr_bodies *bodies) {
  if (!ctx) return ANDR_ERR_ARG;
  if (bodies) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete bodies;
  }
  return ANDR_OK;
}

extern "C" int andr_resistance(andr_ctx *ctx, const andr_mesh *mesh, const andr_resistance_opts *opts, double FT[36],
                               andr_resistance_report *report) {
  return andr_resistance_batch(ctx, 1, mesh, opts, FT, report);
}
