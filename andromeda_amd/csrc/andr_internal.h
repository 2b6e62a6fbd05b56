// Internal declarations of libandromeda_hip: context, device buffers, launch helpers and the
// quadrature tables in constant memory. gfx950 (MI355X) only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/andromeda_hip.h"

#define ANDR_PI 3.141592653589793238462643383

// index of component (i, j) in the symmetric storage {xx, xy, xz, yy, yz, zz}
__host__ __device__ constexpr int sym3(int i, int j) {
  return (i < j ? i : j) == 0 ? (i > j ? i : j) : ((i < j ? i : j) == 1 ? i + j + 1 : 5);
}

// Device memory cache of a context. Blocks released during a call stay with the context for the
// next call instead of going back to the driver: hipFree synchronises the device and mapping or
// unmapping large blocks costs milliseconds each, which is most of a batched call's host time
// otherwise. Every use of a block is ordered on the context's stream, so a block handed out again
// is only touched by work queued after the work of its previous owner. The cache is emptied when
// an allocation fails and on andr_destroy.
struct andr_pool {
  std::multimap<size_t, void *> free_blocks;  // size -> cached block
  std::unordered_map<void *, size_t> size_of;  // every block the pool owns (live or cached)
  size_t cached = 0;
};

struct andr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  andr_pool pool;
  // learned quadrature model (BF16 copy lives in the device buffer below)
  bool has_model = false;
  float model_conf = 0.9f;
  double model_near = 4.0;
  void *d_model = nullptr;  // packed bf16 weights + fp32 biases
  // autotuner of the plain solve's FP64 product A X: (ceil log2 m, ceil log2 n) -> (16-row tiles
  // per wave, best time in ms); filled on first use of a bucket while autotune is on
  bool autotune = true;
  int gemv_n_tiles = 4;  // with the autotuner off
  std::map<std::pair<int, int>, std::pair<int, float>> tune_gemv_n;
  // autotuner of the fast form's Krylov product: ceil log2 of the live 32-row waves -> (waves the
  // column segments aim for, best time in ms); tuned when few rows are live (the segments matter)
  std::map<int, std::pair<int, float>> tune_krylov;
};

// status helpers ------------------------------------------------------------------------------
int andr_fail(andr_ctx *ctx, int code, const char *fmt, ...);
#define ANDR_HIP(ctx, call)                                                                       \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return andr_fail((ctx), ANDR_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                       hipGetErrorString(e_));                                                    \
  } while (0)

void *andr_dev_alloc(andr_ctx *ctx, size_t bytes);  // from the context's cache or the driver; null on failure
void andr_dev_release(andr_ctx *ctx, void *p);       // back to the context's cache
void andr_pool_trim(andr_ctx *ctx);                  // return every cached block to the driver

// RAII device buffer; with a context it draws on the context's memory cache
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  andr_ctx *owner = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  hipError_t alloc(size_t count, andr_ctx *pool = nullptr) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    if (pool) {
      p = (T *)andr_dev_alloc(pool, sizeof(T) * count);
      if (!p) {
        n = 0;
        return hipErrorOutOfMemory;
      }
      owner = pool;
      return hipSuccess;
    }
    return hipMalloc(&p, sizeof(T) * count);
  }
  void release() {
    if (p) {
      if (owner) andr_dev_release(owner, p);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    owner = nullptr;
  }
};

// Quadrature tables (constant memory). Rule ids 1..5: the Wandzura-Xiao rules (ITRI) with
// 7/25/54/85/126 points; ids 6..12: Dunavant 3/4/6/7/12/13/19 points.
#define ANDR_QTAB_N (297 + 64)
struct QuadRule {
  int offset, npts;
};
// Every translation unit that integrates keeps its own constant-memory copy of the table
// (andr_pairs.h) and an upload function; andr_upload_rules fills them all.
QuadRule andr_rule(int id);  // host side lookup
int andr_upload_rules(andr_ctx *ctx);
int andr_upload_rules_learned(andr_ctx *ctx, const double *tab, const float *tabf);

// Per-body device geometry -------------------------------------------------------------------
struct BodyGeom {
  int nn = 0, ne = 0, nwalls = 0;
  int nv = 3;                    // entries of con per element: 3, or 4 (quadrilaterals; -1 marks a triangle)
  int iquad = 4;                 // Gauss-Legendre rule of quadrilaterals (IQUD 1..8)
  const double *xyz = nullptr;   // 3*nn
  const int *con = nullptr;      // nv*ne, 0-based
  const int *bcid = nullptr;     // ne
  const double *nrm = nullptr;   // 3*ne unit normals (CalElementNormal)
  const double *area = nullptr;  // ne (Heron)
  const double *cen = nullptr;   // 3*ne centroids (CalQMesh)
  const double *mul = nullptr;   // ne: NORMMUL of the element's wall
  const float4 *cenl = nullptr;  // ne: centroid (x, y, z) and 1 / (longest edge)^2 in FP32, for the
                                 // distance classes of the learned rule selection
};

// Host-side owner of a body's device arrays.
struct DevBody {
  DevBuf<double> xyz, nrm, area, cen, mul;
  DevBuf<float4> cenl;
  DevBuf<int> con, bcid, color_elems;
  std::vector<int> color_off;  // host copy of color offsets (size ncolors+1)
  DevBuf<int> d_color_off;
  BodyGeom g;
  int itri = 5;
  int iquad = 4;
};
// Host-side checks of a mesh before anything reads it: sizes, null arrays, nvert, the range of
// every connectivity entry and wall id, itri and iquad. ANDR_ERR_ARG with a message on failure.
int andr_check_mesh(andr_ctx *ctx, const andr_mesh *m);
int andr_body_upload(andr_ctx *ctx, const andr_mesh *m, DevBody &b, bool with_colors);

// kernels exposed between translation units ----------------------------------------------------
// rows r0 .. r0 + nrows - 1 of the nsp source points (node rows first; nrows < 0: to the end); the
// outputs hold those rows only (row k of H is source point r0 + k)
// fp32_quad: regular integrals in FP32 (the mixed-precision form), singular ones in FP64
int andr_launch_laplace_rows(andr_ctx *ctx, DevBody &b, double *H, double *G, double *err_rows, int r0 = 0,
                             int nrows = -1, bool fp32_quad = false);
int andr_launch_stokes_rows(andr_ctx *ctx, DevBody &b, double extra, double *T, double *G, double *prT,
                            double *prG, double *err_rows, int r0 = 0, int nrows = -1);
// transposes an m x n row-major device matrix into a column-major host buffer
int andr_copy_out(andr_ctx *ctx, const double *d, int64_t m, int64_t n, int32_t layout, double *h);

// Device address of the no-allocation work space on the context's device, looked up once per
// device (andr_noalloc.hip); called by andr_create with the device current.
int andr_scratch_setup(andr_ctx *ctx);

// Process-wide default context used by the Fortran-style entry points.
andr_ctx *andr_default_ctx();

double andr_now();
