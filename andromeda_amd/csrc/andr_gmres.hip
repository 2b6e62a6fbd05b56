// Restarted GMRES on a batch of dense square systems with up to ANDR_NB right-hand sides each:
// the solver of the mixed-precision resistance path.
//
// Storage of A may be FP32; products accumulate in FP64 and every vector (Krylov basis, Hessenberg,
// solution) is FP64. Each restart cycle starts from the FP64 residual r = b - A x, so the cycles are
// the iterative-refinement loop around the Krylov solve. All systems of the batch advance together:
// one launch per step covers every system (work lists map waves and row tiles to systems), so the
// launch count does not grow with the number of bodies.
//
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2): both passes are block
// products over the basis, which keeps the basis traffic streaming instead of latency bound.
#include "andr_internal.h"
#include "andr_solve.h"
#include "andr_prof.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr int NB = ANDR_NB;
constexpr int TILE = 64;   // rows per vector tile (256 threads: 32 rows x 8 right-hand sides per pass)
constexpr int ICH = 16;   // basis vectors per block of the projection kernel

struct GSys {  // device view of one system
  const void *A;
  int64_t lda;
  int n;
  const double *B;
  double *X;
  double *R;     // residual / scratch n x NB
  double *W;     // new Krylov vector n x NB
  double *Q;     // (m+1) x n x NB basis
  double *H;     // NB x (m+1) x m Hessenberg (column-major per rhs)
  double *cs, *sn;  // NB x m
  double *g;     // NB x (m+1)
  double *coef;  // NB x (m+1) projection coefficients of the current pass
  double *small; // per rhs: [0] bnorm, [1] rnorm (true), [2] est, [3] inv_h, [4] active, [5] used steps
  int tile0;     // first vector tile of the system
  int ntiles;
  int nseg;      // column segments of the matrix-core product (1: written straight to W)
  double *P;     // nseg x n x NB partial products when nseg > 1
};

struct WaveMap {
  int sys, row0;
};
struct MWave {  // matrix-core product: 64 rows x columns [k0, k1) of one system, column segment seg
  int sys, row0, k0, k1, seg;
};
struct TileMap {
  int sys, row0;
};

constexpr int WROWS = 64;  // rows of A per wave in the products

// Y = sx .* (A X) + sy .* Y for every system; sx/sy per (system, rhs). part[wave][k] = sum Y^2 over
// the wave's rows. FP64 accumulation (the residual product of the refinement loop): the wave
// works through its 64 rows in blocks of 8, each lane covering every 64th column.
template <class ST>
__global__ __launch_bounds__(256) void kb_gemv(const GSys *__restrict__ sys, const WaveMap *__restrict__ wmap, int nwaves,
                                               int which_x, int jq, int which_y, const double *__restrict__ sxy,
                                               double *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= nwaves) return;
  const WaveMap wm = wmap[wave];
  const GSys &S = sys[wm.sys];
  const int m = S.n, n = S.n;
  const ST *A = (const ST *)S.A;
  const double *X = which_x == 0 ? S.X : S.Q + (size_t)jq * S.n * NB;
  double *Y = which_y == 0 ? S.R : S.W;
  const int64_t lda = S.lda;
  double y2 = 0.0;
  for (int row0 = wm.row0; row0 < min(wm.row0 + WROWS, m); row0 += 8) {
    double acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.0;
    const ST *arow[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) arow[r] = A + (size_t)min(row0 + r, m - 1) * lda;
    for (int j = lane; j < n; j += 64) {
      const double4 x0 = *reinterpret_cast<const double4 *>(X + (size_t)j * NB);
      const double4 x1 = *reinterpret_cast<const double4 *>(X + (size_t)j * NB + 4);
      const double xv[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const double a = (double)arow[r][j];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[r * 8 + k] += a * xv[k];
      }
    }
#pragma unroll
    for (int half = 32; half >= 1; half >>= 1) {
      const bool hi = (lane & half) != 0;
#pragma unroll
      for (int i = 0; i < half; ++i) {
        const double keep = hi ? acc[i + half] : acc[i];
        const double send = hi ? acc[i] : acc[i + half];
        acc[i] = keep + __shfl_xor(send, half, 64);
      }
    }
    const int r = lane >> 3, k = lane & 7;
    if (row0 + r < m) {
      double *py = Y + (size_t)(row0 + r) * NB + k;
      const double sx = sxy[(size_t)wm.sys * 2 * NB + k], sy = sxy[(size_t)wm.sys * 2 * NB + NB + k];
      const double y = sx * acc[0] + (sy != 0.0 ? sy * (*py) : 0.0);
      *py = y;
      y2 += y * y;
    }
  }
  y2 += __shfl_xor(y2, 8, 64);
  y2 += __shfl_xor(y2, 16, 64);
  y2 += __shfl_xor(y2, 32, 64);
  if (part && lane < 8) part[(size_t)wave * NB + lane] = y2;
}

// The Krylov product W = A Q_j on the matrix cores: FP32 A (as stored), Q_j rounded to FP32,
// v_mfma_f32_16x16x4f32 with the 8 right-hand sides in the 16 B columns (8 unused). A wave owns 64
// rows as four 16-row tiles; lane (g = lane / 16, c = lane % 16) feeds row 16 T + c of tile T with
// columns k0 + 4 g + t (one float4 per tile and 16 columns), and column c of B with the same k.
// The FP32 accumulators are added into FP64 every 64 columns, which keeps the summation error at
// the size of the FP32 storage error. Output lane layout: D[4 g + v][c] in element v of tile T.
typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void kb_gemv_mfma(const GSys *__restrict__ sys, const MWave *__restrict__ wmap,
                                                    int nwaves, int jq, const double *__restrict__ sxy) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= nwaves) return;
  const MWave wm = wmap[wave];
  const GSys &S = sys[wm.sys];
  const int n = S.n, row0 = wm.row0, kend = min(wm.k1, n);
  const int KB = (n + 15) / 16;
  const float *A = (const float *)S.A;
  const double *X = S.Q + (size_t)jq * n * NB;
  const int g = lane >> 4, c = lane & 15;
  const bool cl = c < NB;
  // tiled layout: the float4 of lane (g, c) in tile (band, kb) sits at 256 (band KB + kb) + 16 c + 4 g;
  // ap[T] + 16 k0 addresses it for the k-block k0 / 16 (16 floats per column of tiles)
  const float *ap[4];
#pragma unroll
  for (int T = 0; T < 4; ++T) ap[T] = A + (size_t)((row0 >> 4) + T) * KB * 256 + 16 * c + 4 * g;
  const double *xp = X + (size_t)(4 * g) * NB + (cl ? c : 0);
  f4v acc[4];
  double acc64[4][4];
#pragma unroll
  for (int T = 0; T < 4; ++T) {
    acc[T] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int v = 0; v < 4; ++v) acc64[T][v] = 0.0;
  }
  auto flush = [&]() {
#pragma unroll
    for (int T = 0; T < 4; ++T) {
#pragma unroll
      for (int v = 0; v < 4; ++v) acc64[T][v] += (double)acc[T][v];
      acc[T] = f4v{0.f, 0.f, 0.f, 0.f};
    }
  };
  int k0 = wm.k0, step = 0;
  for (; k0 + 32 <= kend; k0 += 32) {
    float4 a0[4], a1[4];
#pragma unroll
    for (int T = 0; T < 4; ++T) {
      a0[T] = *reinterpret_cast<const float4 *>(ap[T] + (size_t)16 * k0);
      a1[T] = *reinterpret_cast<const float4 *>(ap[T] + (size_t)16 * (k0 + 16));
    }
    float b0[4], b1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      b0[t] = cl ? (float)xp[(size_t)(k0 + t) * NB] : 0.f;
      b1[t] = cl ? (float)xp[(size_t)(k0 + 16 + t) * NB] : 0.f;
    }
#pragma unroll
    for (int T = 0; T < 4; ++T) {
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[T].x, b0[0], acc[T], 0, 0, 0);
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[T].y, b0[1], acc[T], 0, 0, 0);
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[T].z, b0[2], acc[T], 0, 0, 0);
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[T].w, b0[3], acc[T], 0, 0, 0);
    }
#pragma unroll
    for (int T = 0; T < 4; ++T) {
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[T].x, b1[0], acc[T], 0, 0, 0);
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[T].y, b1[1], acc[T], 0, 0, 0);
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[T].z, b1[2], acc[T], 0, 0, 0);
      acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[T].w, b1[3], acc[T], 0, 0, 0);
    }
    if (++step & 1) continue;
    flush();
  }
  // last columns of the segment (fewer than 32): masked, 16 at a time
  for (; k0 < kend; k0 += 16) {
    float a[4][4], b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool in = k0 + 4 * g + t < kend;
      b[t] = (cl && in) ? (float)xp[(size_t)(k0 + t) * NB] : 0.f;
#pragma unroll
      for (int T = 0; T < 4; ++T) a[T][t] = in ? ap[T][(size_t)16 * k0 + t] : 0.f;
    }
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][t], b[t], acc[T], 0, 0, 0);
  }
  flush();
  if (cl) {
    const bool direct = S.nseg <= 1;
    const double sx = direct ? sxy[(size_t)wm.sys * 2 * NB + c] : 1.0;
    double *Y = direct ? S.W : S.P + (size_t)wm.seg * n * NB;
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + 16 * T + 4 * g + v;
        if (row < n) Y[(size_t)row * NB + c] = sx * acc64[T][v];
      }
  }
}

// Y = sx .* (A X) + sy .* Y on the tiled FP32 layout with FP64 products and sums (the residual of
// the refinement loop): wave per 64-row band, lane (g, c) covers row 16 T + c and columns 4 g..4 g+3
// of every 16-column block; the four column groups are summed across lanes at the end.
// part[wave][k] = sum over the band's rows of Y^2.
__global__ __launch_bounds__(256) void kb_resid_tiled(const GSys *__restrict__ sys, const WaveMap *__restrict__ wmap,
                                                      int nwaves, const double *__restrict__ sxy,
                                                      double *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= nwaves) return;
  const WaveMap wm = wmap[wave];
  const GSys &S = sys[wm.sys];
  const int n = S.n, row0 = wm.row0, KB = (n + 15) / 16;
  const int g = lane >> 4, c = lane & 15;
  const float *A = (const float *)S.A;
  double acc[4][NB];
#pragma unroll
  for (int T = 0; T < 4; ++T)
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[T][k] = 0.0;
  for (int kb = 0; kb < KB; ++kb) {
    double x[4][NB];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int col = 16 * kb + 4 * g + t;
      if (col < n) {
        const double4 x0 = *reinterpret_cast<const double4 *>(S.X + (size_t)col * NB);
        const double4 x1 = *reinterpret_cast<const double4 *>(S.X + (size_t)col * NB + 4);
        x[t][0] = x0.x; x[t][1] = x0.y; x[t][2] = x0.z; x[t][3] = x0.w;
        x[t][4] = x1.x; x[t][5] = x1.y; x[t][6] = x1.z; x[t][7] = x1.w;
      } else {
#pragma unroll
        for (int k = 0; k < NB; ++k) x[t][k] = 0.0;
      }
    }
#pragma unroll
    for (int T = 0; T < 4; ++T) {
      const float4 a = *reinterpret_cast<const float4 *>(A + ((size_t)((row0 >> 4) + T) * KB + kb) * 256 + 16 * c + 4 * g);
      // columns past n are padding of the last block (not written by the assembly): masked
      const int c0 = 16 * kb + 4 * g;
      const double av[4] = {c0 < n ? (double)a.x : 0.0, c0 + 1 < n ? (double)a.y : 0.0,
                            c0 + 2 < n ? (double)a.z : 0.0, c0 + 3 < n ? (double)a.w : 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int k = 0; k < NB; ++k) acc[T][k] += av[t] * x[t][k];
    }
  }
  // sum the four column groups (lanes c, c + 16, c + 32, c + 48)
#pragma unroll
  for (int T = 0; T < 4; ++T)
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      acc[T][k] += __shfl_xor(acc[T][k], 16, 64);
      acc[T][k] += __shfl_xor(acc[T][k], 32, 64);
    }
  // lane (g, c) writes row 16 g + c (tile T = g) for all right-hand sides
  double y2[NB];
  const int row = row0 + 16 * g + c;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    double v = acc[0][k];
#pragma unroll
    for (int T = 1; T < 4; ++T) v = g == T ? acc[T][k] : v;
    y2[k] = 0.0;
    if (row < n) {
      double *py = S.R + (size_t)row * NB + k;
      const double sx = sxy[(size_t)wm.sys * 2 * NB + k], sy = sxy[(size_t)wm.sys * 2 * NB + NB + k];
      const double y = sx * v + (sy != 0.0 ? sy * (*py) : 0.0);
      *py = y;
      y2[k] = y * y;
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k)
    for (int off = 1; off < 64; off <<= 1) y2[k] += __shfl_xor(y2[k], off, 64);
  if (lane < NB) {
    double v = y2[0];
#pragma unroll
    for (int k = 1; k < NB; ++k) v = lane == k ? y2[k] : v;
    part[(size_t)wave * NB + lane] = v;
  }
}

// per tile of a system with nseg > 1: W = sx .* sum over segments of the partial products
__global__ __launch_bounds__(256) void kb_seg_sum(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap,
                                                  const double *__restrict__ sxy) {
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  if (S.nseg <= 1) return;
  const int n = S.n;
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const double sx = sxy[(size_t)tm.sys * 2 * NB + k];
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row >= n) continue;
    double y = 0.0;
    for (int sg = 0; sg < S.nseg; ++sg) y += S.P[((size_t)sg * n + row) * NB + k];
    S.W[(size_t)row * NB + k] = sx * y;
  }
}

// per (tile, chunk of ICH basis vectors): partial dots part[tile][i][k] = sum_rows Q_i[row][k] W[row][k]
__global__ __launch_bounds__(256) void kb_dots(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int j,
                                               int mstride, double *__restrict__ part) {
  __shared__ double red[4][ICH][NB];
  const int i0 = blockIdx.y * ICH;
  if (i0 > j) return;  // uniform over the block
  const int ni = min(ICH, j + 1 - i0);
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n;
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double w[TILE / 32];
  int rows[TILE / 32];
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    rows[t] = row < n ? row : -1;
    w[t] = row < n ? S.W[(size_t)row * NB + k] : 0.0;
  }
  double acc[ICH];
#pragma unroll
  for (int ii = 0; ii < ICH; ++ii) {
    double a = 0.0;
    if (ii < ni) {
      const double *Qi = S.Q + (size_t)(i0 + ii) * n * NB;
#pragma unroll
      for (int t = 0; t < TILE / 32; ++t)
        if (rows[t] >= 0) a += Qi[(size_t)rows[t] * NB + k] * w[t];
    }
    a += __shfl_xor(a, 8, 64);
    a += __shfl_xor(a, 16, 64);
    a += __shfl_xor(a, 32, 64);
    acc[ii] = a;
  }
  if (lane < 8)
#pragma unroll
    for (int ii = 0; ii < ICH; ++ii) red[wid][ii][lane] = acc[ii];
  __syncthreads();
  if (threadIdx.x < ni * NB) {
    const int ii = threadIdx.x >> 3, kk = threadIdx.x & 7;
    part[((size_t)blockIdx.x * mstride + i0 + ii) * NB + kk] =
        red[0][ii][kk] + red[1][ii][kk] + red[2][ii][kk] + red[3][ii][kk];
  }
}

// thread per (system, i, k): coef = sum over the system's tiles; H(i, j) += coef
__global__ void kb_dots_final(const GSys *__restrict__ sys, int nsys, int j, int m, int mstride,
                              const double *__restrict__ part, int first_pass) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = (j + 1) * NB;
  if (t >= nsys * per) return;
  const int s = t / per, rem = t % per, i = rem / NB, k = rem % NB;
  const GSys &S = sys[s];
  double c = 0.0;
  for (int tl = 0; tl < S.ntiles; ++tl) c += part[((size_t)(S.tile0 + tl) * mstride + i) * NB + k];
  S.coef[k * (m + 1) + i] = c;
  double &h = S.H[(size_t)k * (m + 1) * m + (size_t)j * (m + 1) + i];
  h = first_pass ? c : h + c;
}

// per tile: W -= sum_i Q_i coef_i ; part2[tile][k] = sum W^2
__global__ __launch_bounds__(256) void kb_sub(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int j, int m,
                                              double *__restrict__ part2) {
  __shared__ double red[4][NB];
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n;
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double w2 = 0.0;
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row >= n) continue;
    double w = S.W[(size_t)row * NB + k];
    for (int i = 0; i <= j; ++i) w -= S.Q[((size_t)i * n + row) * NB + k] * S.coef[k * (m + 1) + i];
    S.W[(size_t)row * NB + k] = w;
    w2 += w * w;
  }
  w2 += __shfl_xor(w2, 8, 64);
  w2 += __shfl_xor(w2, 16, 64);
  w2 += __shfl_xor(w2, 32, 64);
  if (lane < 8) red[wid][lane] = w2;
  __syncthreads();
  if (threadIdx.x < 8)
    part2[(size_t)blockIdx.x * NB + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// thread per (system, k): h(j+1, j) = ||W||, Givens rotations, residual estimate
__global__ void kb_givens(const GSys *__restrict__ sys, int nsys, int j, int m, const double *__restrict__ part2,
                          double tol) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys * NB) return;
  const int s = t / NB, k = t % NB;
  const GSys &S = sys[s];
  double *sm = S.small + k * 8;
  double hn2 = 0.0;
  for (int tl = 0; tl < S.ntiles; ++tl) hn2 += part2[(size_t)(S.tile0 + tl) * NB + k];
  const double hn = sqrt(hn2);
  double *Hc = S.H + (size_t)k * (m + 1) * m + (size_t)j * (m + 1);
  Hc[j + 1] = hn;
  const bool live = sm[4] > 0.0 && sm[5] < 0.0;  // active and not yet frozen in this cycle
  // apply previous rotations
  double *cs = S.cs + k * m, *sn = S.sn + k * m, *g = S.g + k * (m + 1);
  for (int i = 0; i < j; ++i) {
    const double a = Hc[i], b = Hc[i + 1];
    Hc[i] = cs[i] * a + sn[i] * b;
    Hc[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = Hc[j], b = Hc[j + 1];
  const double rr = sqrt(a * a + b * b);
  const double c = rr > 0.0 ? a / rr : 1.0, sg = rr > 0.0 ? b / rr : 0.0;
  cs[j] = c;
  sn[j] = sg;
  Hc[j] = rr;
  Hc[j + 1] = 0.0;
  g[j + 1] = -sg * g[j];
  g[j] = c * g[j];
  const double est = fabs(g[j + 1]);
  sm[2] = est;
  const double tiny = 1e-300;
  sm[3] = hn > tiny ? 1.0 / hn : 0.0;
  if (live && (est <= 0.1 * tol * sm[0] || hn <= 1e-14 * rr)) sm[5] = (double)(j + 1);  // steps to use
}

// per tile: Q_{j+1} = W * inv_h
__global__ __launch_bounds__(256) void kb_scale_next(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int j) {
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n;
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const double inv = S.small[k * 8 + 3];
  double *Qn = S.Q + (size_t)(j + 1) * n * NB;
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row < n) Qn[(size_t)row * NB + k] = S.W[(size_t)row * NB + k] * inv;
  }
}

// thread per (system, k): back substitution y = R^{-1} g over the steps used; y stored in coef
__global__ void kb_solve_y(const GSys *__restrict__ sys, int nsys, int steps, int m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys * NB) return;
  const int s = t / NB, k = t % NB;
  const GSys &S = sys[s];
  double *sm = S.small + k * 8;
  double *y = S.coef + k * (m + 1);
  int kk = sm[5] > 0.0 ? (int)sm[5] : steps;
  if (sm[4] <= 0.0) kk = 0;
  const double *H = S.H + (size_t)k * (m + 1) * m;
  const double *g = S.g + k * (m + 1);
  for (int i = 0; i <= m; ++i) y[i] = 0.0;
  for (int i = kk - 1; i >= 0; --i) {
    double v = g[i];
    for (int c = i + 1; c < kk; ++c) v -= H[(size_t)c * (m + 1) + i] * y[c];
    const double d = H[(size_t)i * (m + 1) + i];
    y[i] = d != 0.0 ? v / d : 0.0;
  }
}

// per tile: X += sum_j Q_j y_j
__global__ __launch_bounds__(256) void kb_xupdate(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int steps,
                                                  int m) {
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n;
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const double *y = S.coef + k * (m + 1);
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row >= n) continue;
    double x = S.X[(size_t)row * NB + k];
    for (int i = 0; i < steps; ++i) x += S.Q[((size_t)i * n + row) * NB + k] * y[i];
    S.X[(size_t)row * NB + k] = x;
  }
}

// per tile: R = B (start of the residual product)
__global__ __launch_bounds__(256) void kb_copy_b(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap) {
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row < S.n) S.R[(size_t)row * NB + k] = S.B[(size_t)row * NB + k];
  }
}

// thread per (system, k): true residual norm from the gemv partials; restart set-up
__global__ void kb_restart(const GSys *__restrict__ sys, int nsys, int m, const double *__restrict__ part,
                           const int *__restrict__ wave0, const int *__restrict__ nwv, double tol, int first,
                           int nrhs, int *__restrict__ done) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys * NB) return;
  const int s = t / NB, k = t % NB;
  const GSys &S = sys[s];
  double *sm = S.small + k * 8;
  double r2 = 0.0;
  for (int w = 0; w < nwv[s]; ++w) r2 += part[(size_t)(wave0[s] + w) * NB + k];
  const double rn = sqrt(r2);
  const double prev = sm[1];
  if (first) {
    sm[0] = rn;  // ||b|| (x = 0)
    sm[4] = (k < nrhs && rn > 0.0) ? 1.0 : 0.0;
    sm[7] = 0.0;
  }
  sm[1] = rn;
  if (sm[4] > 0.0 && rn <= tol * sm[0]) sm[4] = 0.0;  // converged
  // The discrete single-layer operator has a near-null direction (normal tractions). In FP32
  // storage b keeps a component of size ~1e-7 ||b|| outside the range, a floor no Krylov method
  // passes; a full cycle that gains less than a factor 2 below 1e-6 ||b|| has reached it. That
  // component carries no force or torque (sum_e A_e n_e = 0 on a closed surface).
  if (sm[4] > 0.0 && !first && rn > 0.5 * prev && rn <= 1e-6 * sm[0]) {
    sm[4] = 0.0;
    sm[7] = 1.0;
  }
  sm[5] = -1.0;
  double *g = S.g + k * (m + 1);
  for (int i = 0; i <= m; ++i) g[i] = 0.0;
  g[0] = sm[4] > 0.0 ? rn : 0.0;
  sm[3] = (sm[4] > 0.0 && rn > 0.0) ? 1.0 / rn : 0.0;
  if (sm[4] > 0.0) atomicAdd(done, 1);
}

// per tile: Q_0 = R * (1/||r||)
__global__ __launch_bounds__(256) void kb_q0(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap) {
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const double inv = S.small[k * 8 + 3];
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row < S.n) S.Q[(size_t)row * NB + k] = S.R[(size_t)row * NB + k] * inv;
  }
}

}  // namespace

int andr_gmres_batch(andr_ctx *ctx, const GmresSys *in, int nsys, bool fp32, int nrhs, int restart, int maxit,
                     double tol, int *iters) {
  if (nsys <= 0) return ANDR_OK;
  if (nrhs < 1 || nrhs > NB) return andr_fail(ctx, ANDR_ERR_ARG, "nrhs must be 1..%d", NB);
  const int m = std::max(4, restart);
  hipStream_t st = ctx->stream;
  // full work maps (sizes the partial-sum buffers); subsets are rebuilt as systems converge
  size_t vec_elems = 0, basis_elems = 0;
  int nw_full = 0, nt_full = 0;
  for (int s = 0; s < nsys; ++s) {
    vec_elems += (size_t)in[s].n * NB;
    basis_elems += (size_t)(m + 1) * in[s].n * NB;
    nw_full += (in[s].n + WROWS - 1) / WROWS;
    nt_full += (in[s].n + TILE - 1) / TILE;
  }
  const size_t small_per = (size_t)NB * ((m + 1) * m + 2 * m + 2 * (m + 1) + 8);
  DevBuf<double> R, W, Q, SM, partW, partT, partT2;
  DevBuf<GSys> dsys;
  DevBuf<int> ddone;
  if (R.alloc(vec_elems, ctx) || W.alloc(vec_elems, ctx) || Q.alloc(basis_elems, ctx) ||
      SM.alloc(small_per * nsys, ctx) || partW.alloc((size_t)nw_full * NB, ctx) ||
      partT.alloc((size_t)nt_full * (m + 1) * NB, ctx) || partT2.alloc((size_t)nt_full * NB, ctx) ||
      dsys.alloc(nsys, ctx) || ddone.alloc(1, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES workspace (%zu basis doubles)", basis_elems);
  // a map set over a list of systems: waves of WROWS rows (FP64-accumulating product), tiles of
  // TILE rows (vector kernels) and matrix-core waves (64 rows x a column segment). When the list
  // has few rows (the last, slowest systems of a batch) the columns are split into segments so
  // that the product still fills the chip; the segments are summed by kb_seg_sum.
  constexpr int MF_TARGET_WAVES = 2048, MF_MAX_SEG = 8;
  struct MapSet {
    DevBuf<WaveMap> w;
    DevBuf<TileMap> t;
    DevBuf<MWave> mw;
    DevBuf<int> w0, nw;
    int n_w = 0, n_t = 0, n_mw = 0, max_seg = 1;
    size_t mw_cap = 0;
    std::vector<int> tile0, ntiles, nseg;
    std::vector<size_t> poff;
    size_t p_elems = 0;
  };
  MapSet cyc, live;
  for (MapSet *ms : {&cyc, &live})
    if (ms->w.alloc(nw_full, ctx) || ms->t.alloc(nt_full, ctx) || ms->w0.alloc(nsys, ctx) || ms->nw.alloc(nsys, ctx))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES work maps");
  std::vector<GSys> hs(nsys);
  DevBuf<double> Pseg;
  size_t pseg_cap = 0;
  auto build = [&](const std::vector<int> &list, MapSet &ms) -> int {
    std::vector<WaveMap> wm;
    std::vector<TileMap> tm;
    std::vector<MWave> mw;
    std::vector<int> w0(nsys, 0), nw(nsys, 0);
    ms.tile0.assign(nsys, 0);
    ms.ntiles.assign(nsys, 0);
    ms.nseg.assign(nsys, 1);
    ms.poff.assign(nsys, 0);
    ms.p_elems = 0;
    ms.max_seg = 1;
    int row_waves = 0;
    for (int s : list) row_waves += (in[s].n + WROWS - 1) / WROWS;
    const int want = row_waves > 0 ? std::min(MF_MAX_SEG, (MF_TARGET_WAVES + row_waves - 1) / row_waves) : 1;
    for (int s : list) {
      const int n = in[s].n;
      w0[s] = (int)wm.size();
      for (int r = 0; r < n; r += WROWS) wm.push_back({s, r});
      nw[s] = (int)wm.size() - w0[s];
      ms.tile0[s] = (int)tm.size();
      for (int r = 0; r < n; r += TILE) tm.push_back({s, r});
      ms.ntiles[s] = (int)tm.size() - ms.tile0[s];
      const int seglen = ((n + want - 1) / want + 31) / 32 * 32;
      const int nseg = (n + seglen - 1) / seglen;
      ms.nseg[s] = nseg;
      ms.max_seg = std::max(ms.max_seg, nseg);
      if (nseg > 1) {
        ms.poff[s] = ms.p_elems;
        ms.p_elems += (size_t)nseg * n * NB;
      }
      for (int r = 0; r < n; r += WROWS)
        for (int sg = 0; sg < nseg; ++sg) mw.push_back({s, r, sg * seglen, std::min(n, (sg + 1) * seglen), sg});
    }
    ms.n_w = (int)wm.size();
    ms.n_t = (int)tm.size();
    ms.n_mw = (int)mw.size();
    if (mw.size() > ms.mw_cap) {
      if (ms.mw.alloc(mw.size(), ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES matrix-core map");
      ms.mw_cap = mw.size();
    }
    if (ms.n_w) ANDR_HIP(ctx, hipMemcpyAsync(ms.w.p, wm.data(), sizeof(WaveMap) * wm.size(), hipMemcpyHostToDevice, st));
    if (ms.n_t) ANDR_HIP(ctx, hipMemcpyAsync(ms.t.p, tm.data(), sizeof(TileMap) * tm.size(), hipMemcpyHostToDevice, st));
    if (ms.n_mw) ANDR_HIP(ctx, hipMemcpyAsync(ms.mw.p, mw.data(), sizeof(MWave) * mw.size(), hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(ms.w0.p, w0.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(ms.nw.p, nw.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));  // host vectors go out of scope
    return ANDR_OK;
  };
  // the reductions index partials through S.tile0/ntiles; the product reads S.nseg/P
  auto use_tiles = [&](const MapSet &ms) -> int {
    if (ms.p_elems > pseg_cap) {
      if (Pseg.alloc(ms.p_elems, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES segment partials");
      pseg_cap = ms.p_elems;
    }
    for (int s = 0; s < nsys; ++s) {
      hs[s].tile0 = ms.tile0[s];
      hs[s].ntiles = ms.ntiles[s];
      hs[s].nseg = ms.nseg[s];
      hs[s].P = ms.nseg[s] > 1 ? Pseg.p + ms.poff[s] : nullptr;
    }
    ANDR_HIP(ctx, hipMemcpyAsync(dsys.p, hs.data(), sizeof(GSys) * nsys, hipMemcpyHostToDevice, st));
    return ANDR_OK;
  };
  size_t vo = 0, bo = 0;
  for (int s = 0; s < nsys; ++s) {
    GSys &g = hs[s];
    const int n = in[s].n;
    g.A = in[s].A;
    g.lda = in[s].lda;
    g.n = n;
    g.B = in[s].B;
    g.X = in[s].X;
    g.R = R.p + vo;
    g.W = W.p + vo;
    g.Q = Q.p + bo;
    double *sp = SM.p + small_per * s;
    g.H = sp;
    sp += (size_t)NB * (m + 1) * m;
    g.cs = sp;
    sp += NB * m;
    g.sn = sp;
    sp += NB * m;
    g.g = sp;
    sp += NB * (m + 1);
    g.coef = sp;
    sp += NB * (m + 1);
    g.small = sp;
    vo += (size_t)n * NB;
    bo += (size_t)(m + 1) * n * NB;
  }
  ANDR_HIP(ctx, hipMemsetAsync(SM.p, 0, sizeof(double) * small_per * nsys, st));
  for (int s = 0; s < nsys; ++s) ANDR_HIP(ctx, hipMemsetAsync(in[s].X, 0, sizeof(double) * in[s].n * NB, st));
  // scale factors: residual product (sx = -1, sy = 1) and Arnoldi product (sx = 1, sy = 0)
  std::vector<double> h_res(2 * NB * (size_t)nsys), h_arn(2 * NB * (size_t)nsys);
  for (int s = 0; s < nsys; ++s)
    for (int k = 0; k < NB; ++k) {
      h_res[s * 2 * NB + k] = -1.0;
      h_res[s * 2 * NB + NB + k] = 1.0;
      h_arn[s * 2 * NB + k] = 1.0;
      h_arn[s * 2 * NB + NB + k] = 0.0;
    }
  DevBuf<double> d_res, d_arn;
  if (d_res.alloc(h_res.size(), ctx) || d_arn.alloc(h_arn.size(), ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES scales");
  ANDR_HIP(ctx, hipMemcpyAsync(d_res.p, h_res.data(), sizeof(double) * h_res.size(), hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(d_arn.p, h_arn.data(), sizeof(double) * h_arn.size(), hipMemcpyHostToDevice, st));

  const int gs = (nsys * NB + 63) / 64;
  auto gemv = [&](const MapSet &ms, int which_x, int jq, int which_y, const double *scales) {
    if (ms.n_w == 0) return;
    double n2 = 0.0, n1 = 0.0;
    for (int s = 0; s < nsys; ++s)
      if (ms.ntiles[s]) {
        n2 += (double)in[s].n * in[s].n;
        n1 += in[s].n;
      }
    const bool mfma = fp32 && which_x == 1 && which_y == 1;
    const int tok = andr_prof_begin(ctx, mfma ? "gmres_gemv_mfma" : (fp32 ? "gmres_gemv_fp32" : "gmres_gemv_fp64"),
                                    n2 * (fp32 ? 4.0 : 8.0) + 3.0 * n1 * NB * 8.0, 2.0 * n2 * nrhs);
    const int gw = (ms.n_w + 3) / 4;
    if (mfma) {
      hipLaunchKernelGGL(kb_gemv_mfma, dim3((ms.n_mw + 3) / 4), dim3(256), 0, st, dsys.p, ms.mw.p, ms.n_mw, jq, scales);
      if (ms.max_seg > 1) hipLaunchKernelGGL(kb_seg_sum, dim3(ms.n_t), dim3(256), 0, st, dsys.p, ms.t.p, scales);
    } else if (fp32)  // the residual of the refinement loop: X into R
      hipLaunchKernelGGL(kb_resid_tiled, dim3(gw), dim3(256), 0, st, dsys.p, ms.w.p, ms.n_w, scales, partW.p);
    else
      hipLaunchKernelGGL(kb_gemv<double>, dim3(gw), dim3(256), 0, st, dsys.p, ms.w.p, ms.n_w, which_x, jq, which_y,
                         scales, partW.p);
    andr_prof_end(ctx, tok);
  };
  std::vector<int> it_count((size_t)nsys * NB, 0);
  std::vector<double> hsmall(small_per * nsys);
  auto read_small = [&]() -> int {
    for (int s = 0; s < nsys; ++s) {
      hipError_t e = hipMemcpyAsync(hsmall.data() + small_per * s, hs[s].small, sizeof(double) * NB * 8,
                                    hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return andr_fail(ctx, ANDR_ERR_HIP, "GMRES state copy: %s", hipGetErrorString(e));
    }
    ANDR_HIP(ctx, hipStreamSynchronize(st));
    return ANDR_OK;
  };
  auto sm_of = [&](int s, int k) { return hsmall.data() + small_per * s + k * 8; };
  std::vector<int> cyc_list(nsys);
  for (int s = 0; s < nsys; ++s) cyc_list[s] = s;
  int rc = build(cyc_list, cyc);
  if (rc) return rc;
  if ((rc = use_tiles(cyc))) return rc;
  int total = 0;
  bool first = true;
  const bool debug = getenv("ANDR_DEBUG_GMRES") != nullptr;
  while (true) {
    // residual r = b - A x of every system of the last cycle (x = 0 on the first), FP64 accumulation
    if (cyc.n_t) hipLaunchKernelGGL(kb_copy_b, dim3(cyc.n_t), dim3(256), 0, st, dsys.p, cyc.t.p);
    gemv(cyc, 0, 0, 0, d_res.p);
    ANDR_HIP(ctx, hipMemsetAsync(ddone.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(kb_restart, dim3(gs), dim3(64), 0, st, dsys.p, nsys, m, partW.p, cyc.w0.p, cyc.nw.p, tol,
                       first ? 1 : 0, nrhs, ddone.p);
    ANDR_HIP(ctx, hipGetLastError());
    if ((rc = read_small())) return rc;
    first = false;
    if (debug) {  // per-cycle true residuals of the first systems (ANDR_DEBUG_GMRES=1)
      for (int s = 0; s < std::min(nsys, 4); ++s) {
        fprintf(stderr, "[gmres] it %d sys %d n %d:", total, s, in[s].n);
        for (int k = 0; k < nrhs; ++k)
          fprintf(stderr, " %.2e%s", sm_of(s, k)[1] / std::max(sm_of(s, k)[0], 1e-300),
                  sm_of(s, k)[4] > 0.0 ? "" : (sm_of(s, k)[7] > 0.0 ? "s" : "c"));
        fprintf(stderr, "\n");
      }
    }
    // systems with an unconverged column form the next cycle
    cyc_list.clear();
    for (int s = 0; s < nsys; ++s)
      for (int k = 0; k < NB; ++k)
        if (sm_of(s, k)[4] > 0.0) {
          cyc_list.push_back(s);
          break;
        }
    if (cyc_list.empty() || total >= maxit) break;
    if ((rc = build(cyc_list, cyc))) return rc;
    if ((rc = build(cyc_list, live))) return rc;
    if ((rc = use_tiles(live))) return rc;
    hipLaunchKernelGGL(kb_q0, dim3(cyc.n_t), dim3(256), 0, st, dsys.p, cyc.t.p);
    const int steps_max = std::min(m, maxit - total);
    int steps = 0;
    for (int j = 0; j < steps_max; ++j) {
      gemv(live, 1, j, 1, d_arn.p);  // W = A Q_j
      // CGS2: two passes of (coefficients Q^T W, W -= Q coef); each pass streams the basis twice
      double nlive = 0.0;
      for (int s = 0; s < nsys; ++s)
        if (live.ntiles[s]) nlive += in[s].n;
      const double vbytes = nlive * NB * sizeof(double);
      const int tok_o = andr_prof_begin(ctx, "gmres_orthogonalise", 2.0 * (2.0 * (j + 1) + 3.0) * vbytes,
                                        2.0 * 4.0 * (j + 1) * nlive * NB);
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(kb_dots, dim3(live.n_t, (j + ICH) / ICH), dim3(256), 0, st, dsys.p, live.t.p, j, m + 1,
                           partT.p);
        const int nthr = nsys * (j + 1) * NB;
        hipLaunchKernelGGL(kb_dots_final, dim3((nthr + 255) / 256), dim3(256), 0, st, dsys.p, nsys, j, m, m + 1,
                           partT.p, pass == 0 ? 1 : 0);
        hipLaunchKernelGGL(kb_sub, dim3(live.n_t), dim3(256), 0, st, dsys.p, live.t.p, j, m, partT2.p);
      }
      andr_prof_end(ctx, tok_o);
      const int tok_g = andr_prof_begin(ctx, "gmres_givens_scale", 2.0 * vbytes, 0.0);
      hipLaunchKernelGGL(kb_givens, dim3(gs), dim3(64), 0, st, dsys.p, nsys, j, m, partT2.p, tol);
      hipLaunchKernelGGL(kb_scale_next, dim3(live.n_t), dim3(256), 0, st, dsys.p, live.t.p, j);
      andr_prof_end(ctx, tok_g);
      steps = j + 1;
      if (steps % 8 == 0 && steps < steps_max) {  // drop systems whose columns are all resolved
        ANDR_HIP(ctx, hipGetLastError());
        if ((rc = read_small())) return rc;
        std::vector<int> still;
        for (int s : cyc_list) {
          if (!live.ntiles[s]) continue;
          bool l = false;
          for (int k = 0; k < NB; ++k)
            if (sm_of(s, k)[4] > 0.0 && sm_of(s, k)[5] < 0.0) l = true;
          if (l) still.push_back(s);
        }
        if (still.empty()) break;
        int nlive = 0;
        for (int s = 0; s < nsys; ++s) nlive += live.ntiles[s] ? 1 : 0;
        if ((int)still.size() < nlive) {
          if ((rc = build(still, live))) return rc;
          if ((rc = use_tiles(live))) return rc;
        }
      }
    }
    total += steps;
    hipLaunchKernelGGL(kb_solve_y, dim3(gs), dim3(64), 0, st, dsys.p, nsys, steps, m);
    hipLaunchKernelGGL(kb_xupdate, dim3(cyc.n_t), dim3(256), 0, st, dsys.p, cyc.t.p, steps, m);
    ANDR_HIP(ctx, hipGetLastError());
    if ((rc = read_small())) return rc;
    for (int s : cyc_list)
      for (int k = 0; k < NB; ++k) {
        const double *sm = sm_of(s, k);
        if (sm[4] > 0.0) it_count[(size_t)s * NB + k] += sm[5] > 0.0 ? (int)sm[5] : steps;
      }
    if ((rc = use_tiles(cyc))) return rc;
  }
  if (iters)
    for (int s = 0; s < nsys; ++s)
      for (int k = 0; k < nrhs; ++k) iters[s * nrhs + k] = it_count[(size_t)s * NB + k];
  return ANDR_OK;
}
