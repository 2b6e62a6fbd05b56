// C-ABI entry points of libandromeda_hip: context handling, the three seams (integral matrices,
// least squares, resistance) and the Fortran-linkable LSQR routines.
#include "andr_internal.h"
#include "andr_assemble.h"
#include "andr_solve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

double andr_now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int andr_fail(andr_ctx *ctx, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  else fprintf(stderr, "andromeda_hip: %s\n", buf);
  return code;
}

extern "C" const char *andr_version(void) { return "andromeda-hip 0.1 (gfx950)"; }

extern "C" int andr_create(andr_ctx **out, int32_t device) {
  if (!out) return ANDR_ERR_ARG;
  *out = nullptr;
  andr_ctx *ctx = new andr_ctx();
  ctx->device = device;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    fprintf(stderr, "andromeda_hip: no HIP device available (%s)\n", hipGetErrorString(e));
    delete ctx;
    return ANDR_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    delete ctx;
    return ANDR_ERR_ARG;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return ANDR_ERR_HIP;
  }
  if (andr_upload_rules(ctx) != ANDR_OK || andr_scratch_setup(ctx) != ANDR_OK) {
    fprintf(stderr, "andromeda_hip: %s\n", ctx->err.c_str());
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ANDR_ERR_HIP;
  }
  *out = ctx;
  return ANDR_OK;
}

extern "C" int andr_destroy(andr_ctx *ctx) {
  if (!ctx) return ANDR_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_model) (void)hipFree(ctx->d_model);
  andr_pool_trim(ctx);
  for (auto &kv : ctx->pool.size_of) (void)hipFree(kv.first);  // blocks still held by nobody's buffers
  ctx->pool.size_of.clear();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ANDR_OK;
}

// ---- device memory cache ------------------------------------------------------------------------
static size_t pool_round(size_t bytes) {
  const size_t g = bytes <= (1u << 20) ? 4096 : (1u << 20);
  return (bytes + g - 1) / g * g;
}

void *andr_dev_alloc(andr_ctx *ctx, size_t bytes) {
  andr_pool &P = ctx->pool;
  const size_t want = pool_round(bytes);
  auto it = P.free_blocks.lower_bound(want);
  if (it != P.free_blocks.end() && it->first <= want + want / 4 + (1u << 20)) {  // close enough: reuse
    void *p = it->second;
    P.cached -= it->first;
    P.free_blocks.erase(it);
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    andr_pool_trim(ctx);
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.size_of[p] = want;
  return p;
}

void andr_dev_release(andr_ctx *ctx, void *p) {
  andr_pool &P = ctx->pool;
  auto it = P.size_of.find(p);
  if (it == P.size_of.end()) {
    (void)hipFree(p);
    return;
  }
  P.free_blocks.emplace(it->second, p);
  P.cached += it->second;
}

void andr_pool_trim(andr_ctx *ctx) {
  andr_pool &P = ctx->pool;
  if (P.free_blocks.empty()) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : P.free_blocks) {
    (void)hipFree(kv.second);
    P.size_of.erase(kv.second);
  }
  P.free_blocks.clear();
  P.cached = 0;
}

extern "C" const char *andr_last_error(const andr_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

// ---- seam 1 ----------------------------------------------------------------------------------
// err[0]: largest U-row (node source point) self-check error of rows r0 .. r0 + nrows - 1, err[1]: the
// largest of all those rows (fspCerr / QspCerr as CoR_Integrals logs them, src/lib_int.f90:80-87)
static int max_err(andr_ctx *ctx, const double *d_err, int r0, int nrows, int nn, double err[2]) {
  std::vector<double> h(nrows);
  ANDR_HIP(ctx, hipMemcpyAsync(h.data(), d_err, sizeof(double) * nrows, hipMemcpyDeviceToHost, ctx->stream));
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double eu = 0.0, eq = 0.0;
  for (int i = 0; i < nrows; ++i) {
    if (r0 + i < nn) eu = std::max(eu, h[i]);
    else eq = std::max(eq, h[i]);
  }
  if (err) {
    err[0] = eu;
    err[1] = std::max(eu, eq);
  }
  return ANDR_OK;
}

static int check_rows(andr_ctx *ctx, const andr_mesh *mesh, int32_t row_begin, int32_t row_end, int32_t layout) {
  if (layout != ANDR_ROW_MAJOR && layout != ANDR_COL_MAJOR) return andr_fail(ctx, ANDR_ERR_ARG, "bad layout");
  int rc = andr_check_mesh(ctx, mesh);
  if (rc) return rc;
  const int nsp = mesh->nnodes + mesh->nelem;
  if (row_begin < 0 || row_end > nsp || row_begin >= row_end)
    return andr_fail(ctx, ANDR_ERR_ARG, "row range [%d, %d) outside the %d source points", row_begin, row_end, nsp);
  return ANDR_OK;
}

extern "C" int andr_assemble_laplace_rows(andr_ctx *ctx, const andr_mesh *mesh, int32_t row_begin, int32_t row_end,
                                          int32_t layout, double *H, double *G, double err[2]) {
  if (!ctx) return ANDR_ERR_ARG;
  int rc = check_rows(ctx, mesh, row_begin, row_end, layout);
  if (rc) return rc;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  DevBody b;
  if ((rc = andr_body_upload(ctx, mesh, b, true))) return rc;
  const int nn = b.g.nn, ne = b.g.ne, nr = row_end - row_begin;
  DevBuf<double> dH, dG, dE;
  if (dH.alloc((size_t)nr * nn, ctx) || dG.alloc((size_t)nr * ne, ctx) || dE.alloc(nr, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "integral matrices (%d x %d)", nr, nn + ne);
  if ((rc = andr_launch_laplace_rows(ctx, b, dH.p, dG.p, dE.p, row_begin, nr))) return rc;
  if (H && (rc = andr_copy_out(ctx, dH.p, nr, nn, layout, H))) return rc;
  if (G && (rc = andr_copy_out(ctx, dG.p, nr, ne, layout, G))) return rc;
  return max_err(ctx, dE.p, row_begin, nr, nn, err);
}

extern "C" int andr_assemble_laplace(andr_ctx *ctx, const andr_mesh *mesh, int32_t layout, double *H, double *G,
                                     double err[2]) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!mesh) return andr_fail(ctx, ANDR_ERR_ARG, "null mesh");
  return andr_assemble_laplace_rows(ctx, mesh, 0, mesh->nnodes + mesh->nelem, layout, H, G, err);
}

extern "C" int andr_assemble_stokes_rows(andr_ctx *ctx, const andr_mesh *mesh, int32_t row_begin, int32_t row_end,
                                         int32_t layout, double extra_free_term, double *T[6], double *G[6],
                                         double *prT[3], double *prG[3], double err[2]) {
  if (!ctx) return ANDR_ERR_ARG;
  int rc = check_rows(ctx, mesh, row_begin, row_end, layout);
  if (rc) return rc;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  DevBody b;
  if ((rc = andr_body_upload(ctx, mesh, b, true))) return rc;
  const int nn = b.g.nn, ne = b.g.ne, nr = row_end - row_begin;
  const size_t sT = (size_t)nr * nn, sG = (size_t)nr * ne;
  DevBuf<double> dT, dG, dpT, dpG, dE;
  if (dT.alloc(6 * sT, ctx) || dG.alloc(6 * sG, ctx) || dpT.alloc(3 * sT, ctx) || dpG.alloc(3 * sG, ctx) ||
      dE.alloc(nr, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "Stokes integral matrices (%zu doubles)", 9 * (sT + sG));
  if ((rc = andr_launch_stokes_rows(ctx, b, extra_free_term, dT.p, dG.p, dpT.p, dpG.p, dE.p, row_begin, nr))) return rc;
  for (int k = 0; k < 6; ++k) {
    if (T && T[k] && (rc = andr_copy_out(ctx, dT.p + k * sT, nr, nn, layout, T[k]))) return rc;
    if (G && G[k] && (rc = andr_copy_out(ctx, dG.p + k * sG, nr, ne, layout, G[k]))) return rc;
  }
  for (int k = 0; k < 3; ++k) {
    if (prT && prT[k] && (rc = andr_copy_out(ctx, dpT.p + k * sT, nr, nn, layout, prT[k]))) return rc;
    if (prG && prG[k] && (rc = andr_copy_out(ctx, dpG.p + k * sG, nr, ne, layout, prG[k]))) return rc;
  }
  return max_err(ctx, dE.p, row_begin, nr, nn, err);
}

extern "C" int andr_assemble_stokes(andr_ctx *ctx, const andr_mesh *mesh, int32_t layout, double extra_free_term,
                                    double *T[6], double *G[6], double *prT[3], double *prG[3], double err[2]) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!mesh) return andr_fail(ctx, ANDR_ERR_ARG, "null mesh");
  return andr_assemble_stokes_rows(ctx, mesh, 0, mesh->nnodes + mesh->nelem, layout, extra_free_term, T, G, prT, prG,
                                   err);
}

extern "C" int andr_stokes_domain_values(andr_ctx *ctx, const andr_mesh *mesh, int32_t npts, const double *pts,
                                         const double *u, const double *q, double diff, double *vel, double *p,
                                         double *err) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!mesh || npts < 0 || (npts > 0 && (!pts || !u || !q || !vel || !p)) || !(diff > 0.0))
    return andr_fail(ctx, ANDR_ERR_ARG, "andr_stokes_domain_values arguments");
  if (mesh->nvert == 4) return andr_fail(ctx, ANDR_ERR_ARG, "andr_stokes_domain_values takes triangle meshes");
  if (npts == 0) return ANDR_OK;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  DevBody b;
  int rc = andr_body_upload(ctx, mesh, b, false);
  if (rc) return rc;
  const int nn = b.g.nn, ne = b.g.ne;
  DevBuf<double> dpts, du, dq, dout;
  if (dpts.alloc(3 * (size_t)npts, ctx) || du.alloc(3 * (size_t)nn, ctx) || dq.alloc(3 * (size_t)ne, ctx) ||
      dout.alloc(5 * (size_t)npts, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "domain values (%d points)", npts);
  ANDR_HIP(ctx, hipMemcpyAsync(dpts.p, pts, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(du.p, u, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(dq.p, q, sizeof(double) * 3 * ne, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = andr_launch_stokes_domain(ctx, b, npts, dpts.p, du.p, dq.p, diff, dout.p))) return rc;
  std::vector<double> h(5 * (size_t)npts);
  ANDR_HIP(ctx, hipMemcpyAsync(h.data(), dout.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < npts; ++i) {
    for (int c = 0; c < 3; ++c) vel[3 * i + c] = h[5 * (size_t)i + c];
    p[i] = h[5 * (size_t)i + 3];
    if (err) err[i] = h[5 * (size_t)i + 4];
  }
  return ANDR_OK;
}

extern "C" int andr_stokes_surface_pressure(andr_ctx *ctx, const andr_mesh *mesh, const double *u, const double *q,
                                            double diff, double *p) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!mesh || !u || !q || !p || !(diff > 0.0)) return andr_fail(ctx, ANDR_ERR_ARG, "andr_stokes_surface_pressure arguments");
  if (mesh->nvert == 4) return andr_fail(ctx, ANDR_ERR_ARG, "andr_stokes_surface_pressure takes triangle meshes");
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  DevBody b;
  int rc = andr_body_upload(ctx, mesh, b, false);
  if (rc) return rc;
  const int nn = b.g.nn, ne = b.g.ne;
  DevBuf<double> du, dq, dp;
  if (du.alloc(3 * (size_t)nn, ctx) || dq.alloc(3 * (size_t)ne, ctx) || dp.alloc(ne, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "surface pressure");
  ANDR_HIP(ctx, hipMemcpyAsync(du.p, u, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, ctx->stream));
  ANDR_HIP(ctx, hipMemcpyAsync(dq.p, q, sizeof(double) * 3 * ne, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = andr_launch_stokes_surface_pressure(ctx, b, du.p, dq.p, diff, dp.p))) return rc;
  ANDR_HIP(ctx, hipMemcpyAsync(p, dp.p, sizeof(double) * ne, hipMemcpyDeviceToHost, ctx->stream));
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ANDR_OK;
}

// ---- seam 2 ----------------------------------------------------------------------------------
__global__ void k_csr_to_dense(int m, int64_t ld, const int64_t *__restrict__ ia, const int32_t *__restrict__ ja,
                               const double *__restrict__ a, double *__restrict__ D) {
  const int i = blockIdx.x;
  // duplicate (i, j) entries add up, as they do in the reference's CSR products (lsqr_matvec1,
  // src/lib_lsqr.f:172-205); D is zero on entry, so a single entry is stored exactly
  for (int64_t p = ia[i] - 1 + threadIdx.x; p < ia[i + 1] - 1; p += blockDim.x)
    atomicAdd(&D[(size_t)i * ld + (ja[p] - 1)], a[p]);
}

__global__ void k_pack_col(int n, const double *__restrict__ v, double *__restrict__ V) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ANDR_NB) return;
  V[t] = (t & 7) == 0 ? v[t >> 3] : 0.0;
}
__global__ void k_unpack_col(int n, const double *__restrict__ V, double *__restrict__ v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  v[t] = V[(size_t)t * ANDR_NB];
}

namespace {
// The device copy of the last matrix, keyed on the array pointers, the sizes and a checksum of every
// byte of ia, ja and a: a solve on a matrix whose values or pattern changed since the copy was made
// uploads it again. The checksum is one streaming pass over the host arrays (threads over chunks),
// cheaper than the upload it saves.
struct LsqrCache {
  const double *key_a = nullptr;
  const int64_t *key_ia = nullptr;
  const int32_t *key_ja = nullptr;
  int64_t nnz = 0;
  int m = 0, n = 0;
  uint64_t sum = 0;
  bool sparse = false;
  DevBuf<double> D;  // dense m x n, leading dimension dense_ld(n)
  // sparse: A and A^T in CSR with 0-based row pointers and columns
  DevBuf<int64_t> cia, tia;
  DevBuf<int32_t> cja, tja;
  DevBuf<double> cva, tva;
  void release() {
    D.release();
    cia.release();
    tia.release();
    cja.release();
    tja.release();
    cva.release();
    tva.release();
    key_a = nullptr;
  }
};
std::mutex g_mu;
andr_ctx *g_ctx = nullptr;
LsqrCache g_cache;
bool g_precon_failed = false;  // the last lsqr_precon_ failed: solves are refused until one succeeds

// 64-bit multiply-xor hash of a byte range, 8 bytes at a time (tail zero-padded)
uint64_t hash_bytes(const void *p, size_t bytes, uint64_t h) {
  const unsigned char *c = (const unsigned char *)p;
  size_t i = 0;
  for (; i + 8 <= bytes; i += 8) {
    uint64_t v;
    std::memcpy(&v, c + i, 8);
    h = (h ^ v) * 0x100000001b3ull;
    h ^= h >> 29;
  }
  uint64_t v = 0;
  std::memcpy(&v, c + i, bytes - i);
  return ((h ^ v) * 0x100000001b3ull) ^ bytes;
}
// checksum of (ia, ja, a): chunks hashed on up to 16 threads, combined in chunk order
uint64_t csr_checksum(int m, int64_t nnz, const int64_t *ia, const int32_t *ja, const double *a) {
  struct Part {
    const void *p;
    size_t bytes;
  };
  std::vector<Part> parts;
  const size_t chunk = (size_t)1 << 24;
  auto add = [&](const void *p, size_t bytes) {
    for (size_t o = 0; o < bytes; o += chunk) parts.push_back({(const char *)p + o, std::min(chunk, bytes - o)});
  };
  add(ia, sizeof(int64_t) * ((size_t)m + 1));
  add(ja, sizeof(int32_t) * (size_t)nnz);
  add(a, sizeof(double) * (size_t)nnz);
  std::vector<uint64_t> hs(parts.size());
  const int nt = (int)std::min<size_t>(parts.size(), 16);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (size_t k = t; k < parts.size(); k += nt) hs[k] = hash_bytes(parts[k].p, parts[k].bytes, 0xcbf29ce484222325ull + k);
    });
  for (auto &th : pool) th.join();
  return hash_bytes(hs.data(), sizeof(uint64_t) * hs.size(), 0x84222325cbf29ce4ull);
}
void cache_key(LsqrCache &c, int m, int n, int64_t nnz, const int64_t *ia, const int32_t *ja, const double *a) {
  c.key_a = a;
  c.key_ia = ia;
  c.key_ja = ja;
  c.nnz = nnz;
  c.m = m;
  c.n = n;
  c.sum = csr_checksum(m, nnz, ia, ja, a);
}
bool cache_hit(const LsqrCache &c, int m, int n, int64_t nnz, const int64_t *ia, const int32_t *ja, const double *a) {
  return c.key_a == a && c.key_ia == ia && c.key_ja == ja && c.nnz == nnz && c.m == m && c.n == n &&
         (c.sparse ? c.cva.p != nullptr : c.D.p != nullptr) && c.sum == csr_checksum(m, nnz, ia, ja, a);
}
// A CSR whose entries fill less than a third of the m x n matrix stays sparse on the device (A and its
// transpose, 12 B per entry each); a denser one (the reference's systems are dense matrices stored
// in CSR) is expanded to a dense copy for the matrix-core products.
bool sparse_csr(int m, int n, int64_t nnz) { return 3.0 * (double)nnz < (double)m * (double)n; }
// leading dimension of the dense device copy: rows start on 128-byte lines
int64_t dense_ld(int n) { return ((int64_t)n + 15) / 16 * 16; }

// checks the 1-based CSR on the host (the device kernels trust its indices); colsq (n entries, may be
// null) receives the sum of the squared stored entries of every column in row order, duplicates
// squared apart (lsqr_tmatvec3, src/lib_lsqr.f:267-299)
int check_csr(andr_ctx *ctx, int m, int n, int64_t nnz, const int64_t *ia, const int32_t *ja, const double *a,
              double *colsq) {
  if (m <= 0 || n <= 0 || nnz < 0 || ia[0] != 1 || ia[m] != nnz + 1)
    return andr_fail(ctx, ANDR_ERR_ARG, "CSR row pointers inconsistent (m=%d nnz=%lld)", m, (long long)nnz);
  for (int i = 0; i < m; ++i)
    if (ia[i + 1] < ia[i]) return andr_fail(ctx, ANDR_ERR_ARG, "CSR row pointer decreases at row %d", i + 1);
  if (colsq)
    for (int j = 0; j < n; ++j) colsq[j] = 0.0;
  for (int64_t p = 0; p < nnz; ++p) {
    if (ja[p] < 1 || ja[p] > n) return andr_fail(ctx, ANDR_ERR_ARG, "CSR column %d out of range", ja[p]);
    if (colsq) colsq[ja[p] - 1] += a[p] * a[p];
  }
  return ANDR_OK;
}

// the device copy of a checked CSR: dense, or sparse with its transpose (sparse_csr)
int upload_csr(andr_ctx *ctx, int m, int n, int64_t nnz, const int64_t *ia, const int32_t *ja, const double *a,
               LsqrCache &c) {
  c.release();
  c.sparse = sparse_csr(m, n, nnz);
  hipStream_t st = ctx->stream;
  if (c.sparse) {
    std::vector<int64_t> hia(m + 1), tia(n + 1, 0);
    std::vector<int32_t> hja(nnz), tja(nnz);
    std::vector<double> tva(nnz);
    for (int i = 0; i <= m; ++i) hia[i] = ia[i] - 1;
    for (int64_t p = 0; p < nnz; ++p) {
      hja[p] = ja[p] - 1;
      ++tia[ja[p]];
    }
    for (int j = 0; j < n; ++j) tia[j + 1] += tia[j];
    std::vector<int64_t> fill(tia.begin(), tia.end() - 1);
    for (int i = 0; i < m; ++i)  // row order within every column: the products' sums are fixed
      for (int64_t p = ia[i] - 1; p < ia[i + 1] - 1; ++p) {
        const int64_t o = fill[ja[p] - 1]++;
        tja[o] = i;
        tva[o] = a[p];
      }
    if (c.cia.alloc(m + 1) || c.cja.alloc(std::max<int64_t>(nnz, 1)) || c.cva.alloc(std::max<int64_t>(nnz, 1)) ||
        c.tia.alloc(n + 1) || c.tja.alloc(std::max<int64_t>(nnz, 1)) || c.tva.alloc(std::max<int64_t>(nnz, 1)))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "sparse LSQR matrix %d x %d (%lld entries)", m, n, (long long)nnz);
    ANDR_HIP(ctx, hipMemcpyAsync(c.cia.p, hia.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(c.tia.p, tia.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
    if (nnz > 0) {
      ANDR_HIP(ctx, hipMemcpyAsync(c.cja.p, hja.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
      ANDR_HIP(ctx, hipMemcpyAsync(c.cva.p, a, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
      ANDR_HIP(ctx, hipMemcpyAsync(c.tja.p, tja.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
      ANDR_HIP(ctx, hipMemcpyAsync(c.tva.p, tva.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    }
    ANDR_HIP(ctx, hipStreamSynchronize(st));  // the host vectors go out of scope
    return ANDR_OK;
  }
  DevBuf<int64_t> dia;
  DevBuf<int32_t> dja;
  DevBuf<double> da;
  const int64_t ld = dense_ld(n);
  if (c.D.alloc((size_t)m * ld) || dia.alloc(m + 1) || dja.alloc(nnz) || da.alloc(nnz))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "LSQR matrix %d x %d", m, n);
  ANDR_HIP(ctx, hipMemsetAsync(c.D.p, 0, sizeof(double) * (size_t)m * ld, st));
  ANDR_HIP(ctx, hipMemcpyAsync(dia.p, ia, sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(dja.p, ja, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(da.p, a, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_csr_to_dense, dim3(m), dim3(256), 0, st, m, ld, dia.p, dja.p, da.p, c.D.p);
  ANDR_HIP(ctx, hipGetLastError());
  ANDR_HIP(ctx, hipStreamSynchronize(st));
  return ANDR_OK;
}
}  // namespace

andr_ctx *andr_default_ctx() {
  if (!g_ctx) {
    const char *env = getenv("ANDR_DEVICE");
    int dev = env ? atoi(env) : 0;
    if (andr_create(&g_ctx, dev) != ANDR_OK) g_ctx = nullptr;
  }
  return g_ctx;
}

extern "C" void andr_lsqr_forget(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_cache.release();
  g_precon_failed = false;
}

extern "C" void lsqr_precon_(const int32_t *m, const int32_t *n, const int64_t *nnz, double *q, const int64_t *ia,
                             const int32_t *ja, double *a) {
  std::lock_guard<std::mutex> lk(g_mu);
  // The Fortran interface has no status argument: on failure q is set to NaN (a is left as it was)
  // and every lsqr_solve_ returns ierr = -1 until a later lsqr_precon_ succeeds.
  LsqrCache &c = g_cache;
  c.key_a = nullptr;
  g_precon_failed = true;
  auto fail = [&](const char *msg) {
    fprintf(stderr, "andromeda_hip: lsqr_precon_: %s\n", msg);
    if (q && n && *n > 0)
      for (int j = 0; j < *n; ++j) q[j] = std::numeric_limits<double>::quiet_NaN();
  };
  if (!m || !n || !nnz || !q || !ia || !ja || !a) return fail("null argument");
  andr_ctx *ctx = andr_default_ctx();
  if (!ctx) return fail("no device");
  (void)hipSetDevice(ctx->device);
  // q = 1 / sqrt(column sums of squares) as lsqr_tmatvec3 forms it (host, one pass with the checks),
  // then A <- A diag(q) in place (lsqr_matvec4, src/lib_lsqr.f:301-329) and the device copy of the
  // scaled matrix, which the next lsqr_solve_ finds in the cache
  std::vector<double> colsq(*n > 0 ? *n : 1);
  if (check_csr(ctx, *m, *n, *nnz, ia, ja, a, colsq.data())) return fail(ctx->err.c_str());
  for (int j = 0; j < *n; ++j) q[j] = 1.0 / std::sqrt(colsq[j]);
  for (int i = 0; i < *m; ++i)
    for (int64_t p = ia[i] - 1; p < ia[i + 1] - 1; ++p) a[p] = a[p] * q[ja[p] - 1];
  if (upload_csr(ctx, *m, *n, *nnz, ia, ja, a, c)) return fail(ctx->err.c_str());
  cache_key(c, *m, *n, *nnz, ia, ja, a);
  g_precon_failed = false;
}

extern "C" void lsqr_solve_(const int32_t *m, const int32_t *n, const int64_t *nnz, int32_t *maxit, const double *eps,
                            int32_t *nit, int32_t *ierr, const double *q, const int64_t *ia, const int32_t *ja,
                            const double *a, const double *b, double *x) {
  std::lock_guard<std::mutex> lk(g_mu);
  andr_ctx *ctx = andr_default_ctx();
  if (!ctx || g_precon_failed) {
    if (g_precon_failed) fprintf(stderr, "andromeda_hip: lsqr_solve_: the last lsqr_precon_ failed\n");
    *ierr = -1;
    return;
  }
  (void)hipSetDevice(ctx->device);
  LsqrCache &c = g_cache;
  if (!cache_hit(c, *m, *n, *nnz, ia, ja, a)) {
    if (check_csr(ctx, *m, *n, *nnz, ia, ja, a, nullptr) || upload_csr(ctx, *m, *n, *nnz, ia, ja, a, c)) {
      c.release();
      *ierr = -2;
      return;
    }
    cache_key(c, *m, *n, *nnz, ia, ja, a);
  }
  if (*maxit == 0) *maxit = 10 * (*m + *n + 50);
  const int M = *m, N = *n;
  DevBuf<double> dq, dB, dX, tmp;
  if (dq.alloc(N) || dB.alloc((size_t)M * ANDR_NB) || dX.alloc((size_t)N * ANDR_NB) || tmp.alloc(std::max(M, N))) {
    *ierr = -3;
    return;
  }
  hipStream_t st = ctx->stream;
  (void)hipMemcpyAsync(dq.p, q, sizeof(double) * N, hipMemcpyHostToDevice, st);
  (void)hipMemcpyAsync(tmp.p, b, sizeof(double) * M, hipMemcpyHostToDevice, st);
  hipLaunchKernelGGL(k_pack_col, dim3((M * ANDR_NB + 255) / 256), dim3(256), 0, st, M, tmp.p, dB.p);
  (void)hipMemcpyAsync(tmp.p, x, sizeof(double) * N, hipMemcpyHostToDevice, st);
  hipLaunchKernelGGL(k_pack_col, dim3((N * ANDR_NB + 255) / 256), dim3(256), 0, st, N, tmp.p, dX.p);
  int it = 0, is = 0;
  int rc = c.sparse ? andr_lsqr_device_csr(ctx, M, N, *nnz, c.cia.p, c.cja.p, c.cva.p, c.tia.p, c.tja.p, c.tva.p, dq.p,
                                            dB.p, dX.p, 1, *maxit, *eps, &it, &is)
                    : andr_lsqr_device(ctx, c.D.p, false, dense_ld(N), M, N, dq.p, dB.p, dX.p, 1, *maxit, *eps, &it, &is);
  if (rc) {
    fprintf(stderr, "andromeda_hip: lsqr_solve_: %s\n", ctx->err.c_str());
    *ierr = rc;
    return;
  }
  hipLaunchKernelGGL(k_unpack_col, dim3((N + 255) / 256), dim3(256), 0, st, N, dX.p, tmp.p);
  (void)hipMemcpyAsync(x, tmp.p, sizeof(double) * N, hipMemcpyDeviceToHost, st);
  (void)hipStreamSynchronize(st);
  *nit = it;
  *ierr = is > 3 ? is : 0;
}

extern "C" int andr_bicgstab_dense(andr_ctx *ctx, int64_t n, const double *A, int32_t nrhs, const double *B, double *X,
                                   double tol, int32_t maxit, andr_lsqr_stats *stats) {
  if (!ctx || !A || !B || !X || n <= 0 || nrhs <= 0 || nrhs > 16)
    return ctx ? andr_fail(ctx, ANDR_ERR_ARG, "andr_bicgstab_dense arguments") : ANDR_ERR_ARG;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const double t0 = andr_now();
  const int64_t lda = (n + 15) / 16 * 16;  // rows start on 128-byte lines
  DevBuf<double> dA, dB, dX;
  if (dA.alloc((size_t)n * lda, ctx) || dB.alloc((size_t)n * ANDR_NB, ctx) || dX.alloc((size_t)n * ANDR_NB, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "andr_bicgstab_dense workspace");
  ANDR_HIP(ctx, hipMemcpy2DAsync(dA.p, sizeof(double) * lda, A, sizeof(double) * n, sizeof(double) * n, n,
                                 hipMemcpyHostToDevice, ctx->stream));
  std::vector<double> hb((size_t)n * ANDR_NB), hx((size_t)n * ANDR_NB);
  for (int g0 = 0; g0 < nrhs; g0 += ANDR_NB) {
    const int nr = std::min(ANDR_NB, nrhs - g0);
    std::fill(hb.begin(), hb.end(), 0.0);
    std::fill(hx.begin(), hx.end(), 0.0);
    for (int k = 0; k < nr; ++k)
      for (int64_t i = 0; i < n; ++i) {
        hb[i * ANDR_NB + k] = B[(g0 + k) * n + i];
        hx[i * ANDR_NB + k] = X[(g0 + k) * n + i];
      }
    ANDR_HIP(ctx, hipMemcpyAsync(dB.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, ctx->stream));
    ANDR_HIP(ctx, hipMemcpyAsync(dX.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, ctx->stream));
    int it[ANDR_NB], is[ANDR_NB];
    int rc = andr_bicgstab_device(ctx, dA.p, lda, (int)n, dB.p, dX.p, nr, maxit, tol > 0.0 ? tol : 1e-8, it, is);
    if (rc) return rc;
    ANDR_HIP(ctx, hipMemcpyAsync(hx.data(), dX.p, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, ctx->stream));
    ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nr; ++k) {
      for (int64_t i = 0; i < n; ++i) X[(g0 + k) * n + i] = hx[i * ANDR_NB + k];
      if (stats) {
        stats->nit[g0 + k] = it[k];
        stats->istop[g0 + k] = is[k];
      }
    }
  }
  if (stats) stats->seconds = andr_now() - t0;
  return ANDR_OK;
}

extern "C" int andr_lsqr_dense(andr_ctx *ctx, int64_t m, int64_t n, const double *A, int32_t nrhs, const double *B,
                               double *X, const andr_lsqr_opts *opts, andr_lsqr_stats *stats) {
  if (!ctx || !A || !B || !X || m <= 0 || n <= 0 || nrhs <= 0 || nrhs > 16) return ctx ? andr_fail(ctx, ANDR_ERR_ARG, "andr_lsqr_dense arguments") : ANDR_ERR_ARG;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const double t0 = andr_now();
  andr_lsqr_opts o = opts ? *opts : andr_lsqr_opts{0, 1e-15, 1};
  DevBuf<double> dA, dq, dB, dX;
  const int64_t lda = (n + 15) / 16 * 16;  // rows start on 128-byte lines
  if (dA.alloc((size_t)m * lda) || dq.alloc(n) || dB.alloc((size_t)m * ANDR_NB) || dX.alloc((size_t)n * ANDR_NB))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "andr_lsqr_dense workspace");
  ANDR_HIP(ctx, hipMemcpy2DAsync(dA.p, sizeof(double) * lda, A, sizeof(double) * n, sizeof(double) * n, m,
                                 hipMemcpyHostToDevice, ctx->stream));
  if (o.precon) {
    int rc = andr_colscale(ctx, dA.p, false, lda, (int)m, (int)n, dq.p);
    if (rc) return rc;
  }
  std::vector<double> hb((size_t)m * ANDR_NB), hx((size_t)n * ANDR_NB);
  for (int g0 = 0; g0 < nrhs; g0 += ANDR_NB) {
    const int nr = std::min(ANDR_NB, nrhs - g0);
    std::fill(hb.begin(), hb.end(), 0.0);
    std::fill(hx.begin(), hx.end(), 0.0);
    for (int k = 0; k < nr; ++k) {
      for (int64_t i = 0; i < m; ++i) hb[i * ANDR_NB + k] = B[(g0 + k) * m + i];
      for (int64_t i = 0; i < n; ++i) hx[i * ANDR_NB + k] = X[(g0 + k) * n + i];
    }
    ANDR_HIP(ctx, hipMemcpyAsync(dB.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, ctx->stream));
    ANDR_HIP(ctx, hipMemcpyAsync(dX.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, ctx->stream));
    int it[ANDR_NB], is[ANDR_NB];
    int rc = andr_lsqr_device(ctx, dA.p, false, lda, (int)m, (int)n, o.precon ? dq.p : nullptr, dB.p, dX.p, nr, o.maxit,
                              o.eps, it, is);
    if (rc) return rc;
    ANDR_HIP(ctx, hipMemcpyAsync(hx.data(), dX.p, sizeof(double) * hx.size(), hipMemcpyDeviceToHost, ctx->stream));
    ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nr; ++k) {
      for (int64_t i = 0; i < n; ++i) X[(g0 + k) * n + i] = hx[i * ANDR_NB + k];
      if (stats) {
        stats->nit[g0 + k] = it[k];
        stats->istop[g0 + k] = is[k];
      }
    }
  }
  if (stats) stats->seconds = andr_now() - t0;
  return ANDR_OK;
}
