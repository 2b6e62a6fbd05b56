// Entry point that allocates no device memory: the Jacobi-preconditioned BiCGSTAB solve of
// andr_bicgstab_dense on caller-owned device arrays, with every work array carved from one
// file-scope device array of fixed size (`scratch`) instead of the context's memory cache.
//
// Each device the library runs on holds its own copy of `scratch` from the first use of the
// library on that device until the process ends. Its device address is looked up once per device
// (andr_scratch_setup, called by andr_create with the device current) and kept in g_scratch; host
// code uses only that address, never the array's name (in host code the name is a host-side
// stand-in, not a device address). Calls on one device share its array and may not overlap.
#include "andr_internal.h"
#include "andr_solve.h"

#include <algorithm>
#include <array>
#include <mutex>

#define ANDR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define ANDR_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[ANDR_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
std::array<double *, ANDR_NOALLOC_MAX_DEVICES> g_scratch{};  // device address of `scratch` per device

// column-contiguous n x nrhs (columns g0 .. g0 + nr - 1) -> interleaved n x ANDR_NB, other columns 0
__global__ void k_na_pack(int n, int nrhs_total, int g0, int nr, const double *__restrict__ src,
                          double *__restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ANDR_NB) return;
  const int i = t / ANDR_NB, k = t % ANDR_NB;
  (void)nrhs_total;
  dst[t] = k < nr ? src[(size_t)(g0 + k) * n + i] : 0.0;
}
__global__ void k_na_unpack(int n, int g0, int nr, const double *__restrict__ src, double *__restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nr) return;
  const int k = t / n, i = t % n;
  dst[(size_t)(g0 + k) * n + i] = src[(size_t)i * ANDR_NB + k];
}
}  // namespace

// Looks up the device address of `scratch` on the context's device (which is current), once per
// device. Called by andr_create.
int andr_scratch_setup(andr_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= ANDR_NOALLOC_MAX_DEVICES)
    return andr_fail(ctx, ANDR_ERR_ARG, "device %d above the no-allocation table", ctx->device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[ctx->device]) return ANDR_OK;
  double *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return andr_fail(ctx, ANDR_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch[ctx->device] = ptr;
  return ANDR_OK;
}

size_t andr_noalloc_bicgstab_bytes(int64_t n) {
  const size_t vec = (((size_t)n * ANDR_NB + 31) / 32 * 32) * sizeof(double);  // packed B, packed X
  return 2 * vec + andr_bicgstab_ws_bytes((int)n);
}

extern "C" int andr_bicgstab_dense_noalloc(andr_ctx *ctx, int64_t n, const double *A, int64_t lda, int32_t nrhs,
                                           const double *B, double *X, double tol, int32_t maxit,
                                           andr_lsqr_stats *stats) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!A || !B || !X || n <= 0 || n > (1 << 30) || lda < n || nrhs <= 0 || nrhs > 16)
    return andr_fail(ctx, ANDR_ERR_ARG, "andr_bicgstab_dense_noalloc arguments (n=%lld lda=%lld nrhs=%d)",
                     (long long)n, (long long)lda, nrhs);
  const size_t need = andr_noalloc_bicgstab_bytes(n);
  if (need > ANDR_NOALLOC_SCRATCH_BYTES)
    return andr_fail(ctx, ANDR_ERR_ARG,
                     "andr_bicgstab_dense_noalloc: n=%lld needs %zu bytes of work space, above the fixed %zu "
                     "(use andr_bicgstab_dense)", (long long)n, need, ANDR_NOALLOC_SCRATCH_BYTES);
  if (ctx->device < 0 || ctx->device >= ANDR_NOALLOC_MAX_DEVICES || !g_scratch[ctx->device])
    return andr_fail(ctx, ANDR_ERR_STATE, "no-allocation work space of device %d not set up", ctx->device);
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const double t0 = andr_now();
  double *base = g_scratch[ctx->device];
  const size_t vecd = ((size_t)n * ANDR_NB + 31) / 32 * 32;
  double *dB = base, *dX = base + vecd, *ws = base + 2 * vecd;
  const size_t ws_bytes = ANDR_NOALLOC_SCRATCH_BYTES - 2 * vecd * sizeof(double);
  hipStream_t st = ctx->stream;
  const int nn = (int)n;
  const int gp = (nn * ANDR_NB + 255) / 256;
  for (int g0 = 0; g0 < nrhs; g0 += ANDR_NB) {
    const int nr = std::min(ANDR_NB, nrhs - g0);
    hipLaunchKernelGGL(k_na_pack, dim3(gp), dim3(256), 0, st, nn, nrhs, g0, nr, B, dB);
    hipLaunchKernelGGL(k_na_pack, dim3(gp), dim3(256), 0, st, nn, nrhs, g0, nr, (const double *)X, dX);
    ANDR_HIP(ctx, hipGetLastError());
    int it[ANDR_NB], is[ANDR_NB];
    int rc = andr_bicgstab_device(ctx, A, lda, nn, dB, dX, nr, maxit, tol > 0.0 ? tol : 1e-8, it, is, ws, ws_bytes);
    if (rc) return rc;
    hipLaunchKernelGGL(k_na_unpack, dim3((nn * nr + 255) / 256), dim3(256), 0, st, nn, g0, nr, (const double *)dX, X);
    ANDR_HIP(ctx, hipGetLastError());
    // the next group's packing reuses dB and dX: stream order keeps it after this unpack
    if (stats)
      for (int k = 0; k < nr; ++k) {
        stats->nit[g0 + k] = it[k];
        stats->istop[g0 + k] = is[k];
      }
  }
  ANDR_HIP(ctx, hipStreamSynchronize(st));  // X is final when the call returns
  if (stats) stats->seconds = andr_now() - t0;
  return ANDR_OK;
}
