// Multi-RHS LSQR on a dense device matrix: the plain FP64 drop-in for src/lib_lsqr.f.
//
// Every right-hand side follows its own Paige-Saunders recurrence (lsqr_LSQR, src/lib_lsqr.f:470-1207,
// damp = 0, conlim = 1e8, nconv = 1), but the two products of an iteration -- A v and A^T u -- are
// formed for all ANDR_NB columns in one pass over A, so the matrix is streamed from HBM twice per
// iteration for the whole block instead of twice per iteration per right-hand side.
//
// The u and v vectors are kept unnormalised together with their norms (true u = U / beta, true
// v = V / alpha); the products fold the normalisation into their scale factors, which removes the
// separate scaling passes of the reference (dscal at :947, :974, :980).
#include "andr_internal.h"
#include "andr_solve.h"
#include "andr_prof.h"

#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr int NB = ANDR_NB;

__device__ inline double d2norm(double a, double b) {  // lsqr_d2norm (src/lib_lsqr.f:1236-1263)
  double scale = fabs(a) + fabs(b);
  if (scale == 0.0) return 0.0;
  double x = a / scale, y = b / scale;
  return scale * sqrt(x * x + y * y);
}

// Butterfly reduce-scatter of 64 per-lane values: afterwards lane l holds the wave sum of v[l].
__device__ inline double wave_reduce_scatter64(double (&v)[64]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int half = 32; half >= 1; half >>= 1) {
    const bool hi = (lane & half) != 0;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      const double keep = hi ? v[i + half] : v[i];
      const double send = hi ? v[i] : v[i + half];
      v[i] = keep + __shfl_xor(send, half, 64);
    }
  }
  return v[0];
}

// Y = sx .* (A X) + sy .* Y for NB interleaved columns; one wave per 8 rows.
// part[wave][k] = sum of the squares of the new Y over the wave's rows.
template <class ST>
__global__ __launch_bounds__(256) void k_gemv_n(const ST *__restrict__ A, int64_t lda, int m, int n,
                                                const double *__restrict__ X, double *__restrict__ Y,
                                                const double *__restrict__ sx, const double *__restrict__ sy,
                                                double *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int row0 = wave * 8;
  if (row0 >= m) return;
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
  double s = wave_reduce_scatter64(acc);  // lane l: row row0 + l/8, column l%8
  const int r = lane >> 3, k = lane & 7;
  double y = 0.0;
  if (row0 + r < m) {
    double *py = Y + (size_t)(row0 + r) * NB + k;
    y = sx[k] * s + sy[k] * (*py);
    *py = y;
  }
  double y2 = y * y;
  y2 += __shfl_xor(y2, 8, 64);
  y2 += __shfl_xor(y2, 16, 64);
  y2 += __shfl_xor(y2, 32, 64);
  if (lane < 8) part[(size_t)wave * NB + lane] = y2;
}

// The same product for FP64 A on the matrix cores (v_mfma_f64_16x16x4f64). A wave owns 16 GN_T rows
// (GN_T 16-row tiles) over the column segment blockIdx.y; lane (g = lane / 16, c = lane % 16) feeds
// row c of tile T with columns 16 kb + 4 g + t (t = 0..3, one instruction each) and column c of
// B = X (right-hand side c; the upper half c >= NB is zero). The result D[g + 4 v][c] sits in
// register v of lane (g, c). With one segment the update Y = sx .* D + sy .* Y and
// out[wave][k] = sum of the new Y^2 over the wave's rows are done here; with several the raw
// segment products go to out = P[segment][row][k] and k_combine_n finishes. The segment length
// depends on the matrix shape only, and every row's sum runs over its segment in the same order for
// every GN_T, so the tile count (chosen by the autotuner) never changes a result bit.
constexpr int GN_SEGLEN = 1024;  // measured on 13500^2 / 24000^2: 4.5 / 4.8 TB/s (1360 / 1104: 4.3 / 4.6)
typedef double d4v __attribute__((ext_vector_type(4)));

template <int GN_T, bool SEG, bool VEC>
__global__ __launch_bounds__(256) void k_gemv_n_mf(const double *__restrict__ A, int64_t lda, int m, int n,
                                                   const double *__restrict__ X, int seglen, double *__restrict__ Y,
                                                   const double *__restrict__ sx, const double *__restrict__ sy,
                                                   double *__restrict__ out) {
  typedef const double __attribute__((address_space(1))) *gdp;
  typedef double d2n __attribute__((ext_vector_type(2)));
  // VEC: the tiles are read in whole 128-byte row segments (lane (rr = lane / 8, p = lane % 8) takes
  // columns 2 p, 2 p + 1 of row rr of an 8-row half tile: a load covers 8 full lines, where the
  // matrix-core lane pattern covers 16 half lines, measured 6.9 against 4.6-5.3 TB/s on a 30 GB
  // matrix) and redistributed to the matrix-core layout through LDS (16 x 16 tile per T, rows padded
  // to 18 doubles)
  constexpr int LR = 18;
  __shared__ __attribute__((aligned(16))) double lt[VEC ? 4 : 1][VEC ? GN_T : 1][16 * LR];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int wave = blockIdx.x * (blockDim.x >> 6) + wid;
  const int row0 = wave * 16 * GN_T;
  if (row0 >= m) return;  // uniform over the wave; the LDS tiles are per wave
  const int g = lane >> 4, c = lane & 15;
  const int rr = lane >> 3, p = lane & 7;
  const bool hi = c >= NB;
  const int k0 = blockIdx.y * seglen, kend = min(n, k0 + seglen);
  gdp arow[GN_T];   // matrix-core pattern: row 16 T + c, columns 4 g ..
  gdp lrow[GN_T][2];  // row-segment pattern: rows 16 T + rr and 16 T + 8 + rr, columns 2 p ..
#pragma unroll
  for (int T = 0; T < GN_T; ++T) {
    arow[T] = (gdp)(A + (size_t)min(row0 + 16 * T + c, m - 1) * lda + 4 * g);
    lrow[T][0] = (gdp)(A + (size_t)min(row0 + 16 * T + rr, m - 1) * lda + 2 * p);
    lrow[T][1] = (gdp)(A + (size_t)min(row0 + 16 * T + 8 + rr, m - 1) * lda + 2 * p);
  }
  const gdp xb = (gdp)(X + (size_t)(4 * g) * NB + (c & (NB - 1)));
  d4v acc[GN_T];
#pragma unroll
  for (int T = 0; T < GN_T; ++T) acc[T] = d4v{0.0, 0.0, 0.0, 0.0};
  // raw operands of one 16-column block: VEC: two row segments per T; otherwise the matrix-core
  // values themselves
  auto load = [&](int kb, double (&a)[GN_T][4], double (&x)[4]) {
#pragma unroll
    for (int T = 0; T < GN_T; ++T) {
      if (VEC) {  // the matrix is streamed once: non-temporal, so that X keeps its place in L2
        const d2n u = __builtin_nontemporal_load((const d2n __attribute__((address_space(1))) *)(lrow[T][0] + 16 * kb));
        const d2n w = __builtin_nontemporal_load((const d2n __attribute__((address_space(1))) *)(lrow[T][1] + 16 * kb));
        a[T][0] = u.x;
        a[T][1] = u.y;
        a[T][2] = w.x;
        a[T][3] = w.y;
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) a[T][t] = __builtin_nontemporal_load(arow[T] + 16 * kb + t);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) x[t] = xb[(size_t)(16 * kb + t) * NB];
  };
  // VEC: row segments -> LDS -> matrix-core values (in place)
  auto stage = [&](double (&a)[GN_T][4]) {
    if (VEC) {
#pragma unroll
      for (int T = 0; T < GN_T; ++T) {
        *(d2n *)&lt[wid][T][rr * LR + 2 * p] = d2n{a[T][0], a[T][1]};
        *(d2n *)&lt[wid][T][(8 + rr) * LR + 2 * p] = d2n{a[T][2], a[T][3]};
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int T = 0; T < GN_T; ++T) {
        const d2n u = *(const d2n *)&lt[wid][T][c * LR + 4 * g];
        const d2n w = *(const d2n *)&lt[wid][T][c * LR + 4 * g + 2];
        a[T][0] = u.x;
        a[T][1] = u.y;
        a[T][2] = w.x;
        a[T][3] = w.y;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  };
  auto step = [&](const double (&a)[GN_T][4], const double (&x)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double b = hi ? 0.0 : x[t];
#pragma unroll
      for (int T = 0; T < GN_T; ++T) acc[T] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[T][t], b, acc[T], 0, 0, 0);
    }
  };
  // full 16-column blocks, two register sets in turn: the operands of the next block are requested
  // before the products of the current one (the last prefetch re-reads the final block)
  int kb = k0 >> 4;
  const int kb_end = kend >> 4;
  if (kb < kb_end) {
    double a0[GN_T][4], a1[GN_T][4], x0[4], x1[4];
    load(kb, a0, x0);
    for (; kb + 1 < kb_end; kb += 2) {
      load(kb + 1, a1, x1);
      __builtin_amdgcn_sched_barrier(0);
      stage(a0);
      step(a0, x0);
      __builtin_amdgcn_sched_barrier(0);
      load(min(kb + 2, kb_end - 1), a0, x0);
      __builtin_amdgcn_sched_barrier(0);
      stage(a1);
      step(a1, x1);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (kb < kb_end) {  // odd count: a0 holds block kb
      stage(a0);
      step(a0, x0);
      ++kb;
    }
  }
  if (16 * kb < kend) {  // last, partial column block (last segment only): masked
    double a[GN_T][4], x[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int col = 16 * kb + 4 * g + t;
      const bool in = col < kend;
      x[t] = in ? X[(size_t)col * NB + (c & (NB - 1))] : 0.0;
#pragma unroll
      for (int T = 0; T < GN_T; ++T) a[T][t] = in ? A[(size_t)min(row0 + 16 * T + c, m - 1) * lda + col] : 0.0;
    }
    step(a, x);
  }
  if (SEG) {
    double *P = out + (size_t)blockIdx.y * m * NB;
#pragma unroll
    for (int T = 0; T < GN_T; ++T)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + 16 * T + g + 4 * v;
        if (!hi && r < m) P[(size_t)r * NB + c] = acc[T][v];
      }
    return;
  }
  double ss = 0.0;
#pragma unroll
  for (int T = 0; T < GN_T; ++T)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = row0 + 16 * T + g + 4 * v;
      if (!hi && r < m) {
        double *py = Y + (size_t)r * NB + c;
        const double y = sx[c] * acc[T][v] + sy[c] * (*py);
        *py = y;
        ss += y * y;
      }
    }
  ss += __shfl_xor(ss, 16, 64);  // lanes of equal c
  ss += __shfl_xor(ss, 32, 64);
  if (lane < NB) out[(size_t)wave * NB + lane] = ss;
}

// Y = sx .* sum_s P[s] + sy .* Y for the segmented product; part[block][k] = sum of the new Y^2.
__global__ __launch_bounds__(256) void k_combine_n(int m, int nseg, const double *__restrict__ P,
                                                   double *__restrict__ Y, const double *__restrict__ sx,
                                                   const double *__restrict__ sy, double *__restrict__ part) {
  __shared__ double red[4 * NB];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // one (row, k) pair per thread
  const int k = t & 7;
  double y2 = 0.0;
  if (t < m * NB) {
    double s = 0.0;
    for (int sg = 0; sg < nseg; ++sg) s += P[(size_t)sg * m * NB + t];
    const double y = sx[k] * s + sy[k] * Y[t];
    Y[t] = y;
    y2 = y * y;
  }
  y2 += __shfl_xor(y2, 8, 64);
  y2 += __shfl_xor(y2, 16, 64);
  y2 += __shfl_xor(y2, 32, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane < 8) red[wid * NB + lane] = y2;
  __syncthreads();
  if (threadIdx.x < 8)
    part[(size_t)blockIdx.x * NB + threadIdx.x] =
        (red[threadIdx.x] + red[NB + threadIdx.x]) + (red[2 * NB + threadIdx.x] + red[3 * NB + threadIdx.x]);
}

// P[s][j][k] = sum_{i in split s} A[i][j] Y[i][k]; one column per thread, splits over blockIdx.y.
template <class ST>
__global__ __launch_bounds__(256) void k_gemv_t(const ST *__restrict__ A, int64_t lda, int m, int n,
                                                const double *__restrict__ Y, int rows_per_split,
                                                double *__restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  const int i0 = s * rows_per_split;
  const int i1 = min(m, i0 + rows_per_split);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (j < n) {
    const ST *col = A + j;
    int i = i0;
    for (; i + 4 <= i1; i += 4) {
      const double a0 = (double)col[(size_t)i * lda], a1 = (double)col[(size_t)(i + 1) * lda];
      const double a2 = (double)col[(size_t)(i + 2) * lda], a3 = (double)col[(size_t)(i + 3) * lda];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        acc[k] += a0 * Y[(size_t)i * NB + k] + a1 * Y[(size_t)(i + 1) * NB + k] + a2 * Y[(size_t)(i + 2) * NB + k] +
                  a3 * Y[(size_t)(i + 3) * NB + k];
    }
    for (; i < i1; ++i) {
      const double a = (double)col[(size_t)i * lda];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += a * Y[(size_t)i * NB + k];
    }
    double4 *p = reinterpret_cast<double4 *>(P + ((size_t)s * n + j) * NB);
    p[0] = make_double4(acc[0], acc[1], acc[2], acc[3]);
    p[1] = make_double4(acc[4], acc[5], acc[6], acc[7]);
  }
}

// The same for FP64 A with even lda on a 16-byte boundary: two columns per thread, so that a load
// covers 1 KB of a row (64 lanes x 16 B) instead of 512 B.
__global__ __launch_bounds__(256) void k_gemv_t2(const double *__restrict__ A, int64_t lda, int m, int n,
                                                 const double *__restrict__ Y, int rows_per_split,
                                                 double *__restrict__ P) {
  typedef double d2n __attribute__((ext_vector_type(2)));
  typedef const d2n __attribute__((address_space(1))) *gd2p;
  const int j = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int s = blockIdx.y;
  const int i0 = s * rows_per_split;
  const int i1 = min(m, i0 + rows_per_split);
  if (j >= n) return;
  d2n acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = d2n{0.0, 0.0};
  // columns j, j + 1 (the second may be the row padding when n is odd: read, never stored)
  const gd2p col = (gd2p)(A + j);
  int i = i0;
  for (; i + 4 <= i1; i += 4) {
    d2n a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = __builtin_nontemporal_load(col + (size_t)(i + u) * (lda / 2));
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += a[u] * Y[(size_t)(i + u) * NB + k];
  }
  for (; i < i1; ++i) {
    const d2n a = __builtin_nontemporal_load(col + (size_t)i * (lda / 2));
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] += a * Y[(size_t)i * NB + k];
  }
  double4 *p = reinterpret_cast<double4 *>(P + ((size_t)s * n + j) * NB);
  p[0] = make_double4(acc[0].x, acc[1].x, acc[2].x, acc[3].x);
  p[1] = make_double4(acc[4].x, acc[5].x, acc[6].x, acc[7].x);
  if (j + 1 < n) {
    p[2] = make_double4(acc[0].y, acc[1].y, acc[2].y, acc[3].y);
    p[3] = make_double4(acc[4].y, acc[5].y, acc[6].y, acc[7].y);
  }
}

// V = sx .* sum_s P[s] + sy .* V; partV[block][k] = sum V^2, partW[block][k] = sum W^2 (old W).
__global__ __launch_bounds__(256) void k_combine_t(int n, int nsplit, const double *__restrict__ P,
                                                   double *__restrict__ V, const double *__restrict__ W,
                                                   const double *__restrict__ sx, const double *__restrict__ sy,
                                                   double *__restrict__ partV, double *__restrict__ partW) {
  __shared__ double red[2 * NB * 4];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // one (j, k) pair per thread
  const int k = t & 7;
  double v2 = 0.0, w2 = 0.0;
  if (t < n * NB) {
    double s = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) s += P[(size_t)sp * n * NB + t];
    const double v = sx[k] * s + sy[k] * V[t];
    V[t] = v;
    v2 = v * v;
    if (W) {
      const double w = W[t];
      w2 = w * w;
    }
  }
  // reduce over lanes with equal k (lane & 7)
  v2 += __shfl_xor(v2, 8, 64);
  v2 += __shfl_xor(v2, 16, 64);
  v2 += __shfl_xor(v2, 32, 64);
  w2 += __shfl_xor(w2, 8, 64);
  w2 += __shfl_xor(w2, 16, 64);
  w2 += __shfl_xor(w2, 32, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane < 8) {
    red[wid * NB + lane] = v2;
    red[4 * NB + wid * NB + lane] = w2;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < 4; ++w) {
      a += red[w * NB + threadIdx.x];
      b += red[4 * NB + w * NB + threadIdx.x];
    }
    partV[(size_t)blockIdx.x * NB + threadIdx.x] = a;
    partW[(size_t)blockIdx.x * NB + threadIdx.x] = b;
  }
}

// Column sums of the partials part[i][k], i < np, by a block of 256 threads (32 partial rows per
// column in flight); thread k < NB receives the sum of column k. Every thread of the block calls it.
__device__ double block_sum_parts(const double *__restrict__ part, int np, double *red) {
  const int t = threadIdx.x, k = t & 7;
  double s = 0.0;
  for (int i = t >> 3; i < np; i += 32) s += part[(size_t)i * NB + k];
  s += __shfl_xor(s, 8, 64);
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  __syncthreads();  // red may still be read from the previous call
  if ((t & 63) < 8) red[(t >> 6) * NB + (t & 63)] = s;
  __syncthreads();
  return t < NB ? (red[t] + red[NB + t]) + (red[2 * NB + t] + red[3 * NB + t]) : 0.0;
}
#define LSQR_SCALAR_THREADS 256

// ---- scalar kernels (one thread per column) -------------------------------------------------
// start: U = b - A x_s is formed; beta = ||U||, btol = eps ||b|| / ||r0||; set scales for A^T u.
__global__ void k_lsqr_start(LsqrState *st, const double *partU, int npU, const double *partB, int npB, double eps,
                             double *sxt, double *syt) {
  __shared__ double red[4 * NB];
  const double sr = block_sum_parts(partU, npU, red);
  const double sb = block_sum_parts(partB, npB, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  LsqrState &S = st[k];
  const double nrm_r = sqrt(sr);
  const double nrm_b = sqrt(sb);
  S.atol = eps;
  S.btol = eps * nrm_b / nrm_r;
  S.ctol = 1.0 / 1.0e8;
  S.anorm = S.dnorm = S.res2 = S.xnorm = S.xnorm1 = S.z = 0.0;
  S.cs2 = -1.0;
  S.sn2 = 0.0;
  S.itn = 0;
  S.istop = 0;
  S.beta = nrm_r;
  S.alpha = 0.0;
  if (S.active && S.beta > 0.0) {
    sxt[k] = 1.0 / S.beta;
    syt[k] = 0.0;
  } else {
    sxt[k] = 0.0;
    syt[k] = 0.0;
  }
}

// after the first A^T u: alpha, w = v, the initial recurrence values (src/lib_lsqr.f:952-1008)
__global__ void k_lsqr_init2(LsqrState *st, const double *partV, int npV, double *sxn, double *syn, double *inv_alpha) {
  __shared__ double red[4 * NB];
  const double sv = block_sum_parts(partV, npV, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  LsqrState &S = st[k];
  S.alpha = S.beta > 0.0 ? sqrt(sv) : 0.0;
  inv_alpha[k] = S.alpha > 0.0 ? 1.0 / S.alpha : 0.0;
  const double arnorm = S.alpha * S.beta;
  S.rhobar = S.alpha;
  S.phibar = S.beta;
  S.bnorm = S.beta;
  S.rnorm = S.beta;
  if (arnorm == 0.0) S.active = 0;  // exact solution x = 0 (istop 0)
  if (S.active) {
    sxn[k] = 1.0 / S.alpha;
    syn[k] = -S.alpha / S.beta;
  } else {
    sxn[k] = 0.0;
    syn[k] = 0.0;
  }
}

// after U = A v - alpha u: beta, anorm; scales of the A^T product
__global__ void k_lsqr_mid(LsqrState *st, const double *partU, int npU, double *sxt, double *syt) {
  __shared__ double red[4 * NB];
  const double su = block_sum_parts(partU, npU, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  LsqrState &S = st[k];
  if (!S.active) {
    sxt[k] = 0.0;
    syt[k] = 0.0;
    return;
  }
  S.itn += 1;
  const double beta = sqrt(su);
  double temp = d2norm(S.alpha, beta);
  temp = d2norm(temp, 0.0);
  S.anorm = d2norm(S.anorm, temp);
  const double alpha_old = S.alpha;
  S.beta = beta;
  if (beta > 0.0) {
    sxt[k] = 1.0 / beta;
    syt[k] = -beta / alpha_old;
  } else {
    sxt[k] = 0.0;
    syt[k] = 1.0;  // v unchanged
  }
}

// after V = A^T u - beta v: alpha, the plane rotations and the convergence tests
// (src/lib_lsqr.f:1015-1152)
__global__ void k_lsqr_end(LsqrState *st, const double *partV, const double *partW, int np, double *sxn, double *syn,
                           double *t1o, double *t2o, double *inv_alpha, int *active_out) {
  __shared__ double red[4 * NB];
  const double sv = block_sum_parts(partV, np, red);
  const double sw = block_sum_parts(partW, np, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  LsqrState &S = st[k];
  if (!S.active) {
    t1o[k] = 0.0;
    t2o[k] = 1.0;
    inv_alpha[k] = 0.0;
    sxn[k] = 0.0;
    syn[k] = 0.0;
    active_out[k] = 0;
    return;
  }
  if (S.beta > 0.0) S.alpha = sqrt(sv);
  const double alpha = S.alpha, beta = S.beta;
  const double wnorm = sqrt(sw);
  const double rhbar1 = S.rhobar;
  const double rho = d2norm(rhbar1, beta);
  const double cs = rhbar1 / rho, sn = beta / rho;
  const double theta = sn * alpha;
  S.rhobar = -cs * alpha;
  const double phi = cs * S.phibar;
  S.phibar = sn * S.phibar;
  const double tau = sn * phi;
  const double t1 = phi / rho, t2 = -theta / rho, t3 = 1.0 / rho;
  const double dknorm = t3 * wnorm;
  S.dnorm = d2norm(S.dnorm, dknorm);
  const double delta = S.sn2 * rho, gambar = -S.cs2 * rho;
  const double rhs = phi - delta * S.z, zbar = rhs / gambar;
  S.xnorm = d2norm(S.xnorm1, zbar);
  const double gamma = d2norm(gambar, theta);
  S.cs2 = gambar / gamma;
  S.sn2 = theta / gamma;
  S.z = rhs / gamma;
  S.xnorm1 = d2norm(S.xnorm1, S.z);
  const double acond = S.anorm * S.dnorm;
  S.res2 = d2norm(S.res2, 0.0);
  S.rnorm = d2norm(S.res2, S.phibar);
  const double arnorm = alpha * fabs(tau);
  const double test1 = S.rnorm / S.bnorm;
  const double test2 = S.rnorm > 0.0 ? arnorm / (S.anorm * S.rnorm) : 0.0;
  const double test3 = 1.0 / acond;
  double tt1 = test1 / (1.0 + S.anorm * S.xnorm / S.bnorm);
  const double rtol = S.btol + S.atol * S.anorm * S.xnorm / S.bnorm;
  const double tt3 = 1.0 + test3, tt2 = 1.0 + test2;
  tt1 = 1.0 + tt1;
  int istop = 0;
  if (S.itn >= S.itnlim) istop = 5;
  if (tt3 <= 1.0) istop = 4;
  if (tt2 <= 1.0) istop = 2;
  if (tt1 <= 1.0) istop = 1;
  if (test3 <= S.ctol) istop = 4;
  if (test2 <= S.atol) istop = 2;
  if (test1 <= rtol) istop = 1;
  S.istop = istop;
  t1o[k] = t1;
  t2o[k] = t2;
  inv_alpha[k] = alpha > 0.0 ? 1.0 / alpha : 1.0;
  if (istop != 0) S.active = 0;
  active_out[k] = 1;  // the update of this iteration still applies (src/lib_lsqr.f:1057-1062)
  if (S.active) {
    sxn[k] = 1.0 / alpha;
    syn[k] = -alpha / beta;
  } else {
    sxn[k] = 0.0;
    syn[k] = 0.0;
  }
}

// x += t1 w ; w = V * inv_alpha + t2 w  (src/lib_lsqr.f:1057-1062); columns with t1 = 0, t2 = 1,
// inv_alpha = 0 stay unchanged.
__global__ void k_lsqr_update(int n, double *__restrict__ X, double *__restrict__ W, const double *__restrict__ V,
                              const double *__restrict__ t1, const double *__restrict__ t2,
                              const double *__restrict__ inv_alpha, const int *__restrict__ was_active) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  const int k = t & 7;
  if (!was_active[k]) return;
  const double w = W[t];
  X[t] += t1[k] * w;
  W[t] = V[t] * inv_alpha[k] + t2[k] * w;
}

__global__ void k_scale_div(int n, const double *__restrict__ X0, const double *__restrict__ q, double *__restrict__ Xs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  Xs[t] = q ? X0[t] / q[t >> 3] : X0[t];
}

__global__ void k_scale_mul_add(int n, const double *__restrict__ q, const double *__restrict__ Xl,
                                const double *__restrict__ X0, double *__restrict__ X) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  X[t] = (q ? q[t >> 3] * Xl[t] : Xl[t]) + X0[t];
}

__global__ void k_sumsq_parts(int n, const double *__restrict__ Y, double *__restrict__ part) {
  __shared__ double red[NB * 4];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double y2 = 0.0;
  if (t < n * NB) y2 = Y[t] * Y[t];
  y2 += __shfl_xor(y2, 8, 64);
  y2 += __shfl_xor(y2, 16, 64);
  y2 += __shfl_xor(y2, 32, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane < 8) red[wid * NB + lane] = y2;
  __syncthreads();
  if (threadIdx.x < 8) {
    double a = 0.0;
    for (int w = 0; w < 4; ++w) a += red[w * NB + threadIdx.x];
    part[(size_t)blockIdx.x * NB + threadIdx.x] = a;
  }
}

// ---- sparse products (CSR) -----------------------------------------------------------------------
// Row-parallel CSR product over NB interleaved columns, 0-based row pointers and columns: thread
// (row r, column k). UPDATE: Y = sx .* (A X) + sy .* Y and part[block][k] = sum of the new Y^2 over
// the block's 32 rows (the layout of k_combine_n); otherwise the raw product goes to Y (A^T U into
// the split partials of k_combine_t, one split). Each row sums its entries in stored order.
template <bool UPDATE>
__global__ __launch_bounds__(256) void k_csr_mv(int rows, const int64_t *__restrict__ ia, const int32_t *__restrict__ ja,
                                                const double *__restrict__ va, const double *__restrict__ X,
                                                double *__restrict__ Y, const double *__restrict__ sx,
                                                const double *__restrict__ sy, double *__restrict__ part) {
  __shared__ double red[4][NB];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t >> 3, k = t & 7;
  double y = 0.0;
  if (r < rows) {
    double acc = 0.0;
    for (int64_t p = ia[r]; p < ia[r + 1]; ++p) acc += va[p] * X[(size_t)ja[p] * NB + k];
    double *py = Y + (size_t)r * NB + k;
    if (UPDATE) {
      y = sx[k] * acc + sy[k] * (*py);
      *py = y;
    } else {
      *py = acc;
    }
  }
  if (UPDATE) {
    double y2 = y * y;
    y2 += __shfl_xor(y2, 8, 64);
    y2 += __shfl_xor(y2, 16, 64);
    y2 += __shfl_xor(y2, 32, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane < NB) red[wid][lane] = y2;
    __syncthreads();
    if (threadIdx.x < NB)
      part[(size_t)blockIdx.x * NB + threadIdx.x] =
          (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
  }
}

// ---- column scaling (lsqr_precon) ------------------------------------------------------------
template <class ST>
__global__ __launch_bounds__(256) void k_colsq(const ST *__restrict__ A, int64_t lda, int m, int n, int rows_per_split,
                                               double *__restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (j >= n) return;
  const int i0 = s * rows_per_split, i1 = min(m, i0 + rows_per_split);
  double acc = 0.0;
  for (int i = i0; i < i1; ++i) {
    const double a = (double)A[(size_t)i * lda + j];
    acc += a * a;
  }
  P[(size_t)s * n + j] = acc;
}

__global__ void k_colq(int n, int nsplit, const double *__restrict__ P, double *__restrict__ q) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int sp = 0; sp < nsplit; ++sp) s += P[(size_t)sp * n + j];
  q[j] = 1.0 / sqrt(s);
}

template <class ST>
__global__ __launch_bounds__(256) void k_colmul(ST *__restrict__ A, int64_t lda, int m, int n, const double *__restrict__ q) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double qj = q[j];
  for (int i = blockIdx.y; i < m; i += gridDim.y) A[(size_t)i * lda + j] = (ST)((double)A[(size_t)i * lda + j] * qj);
}

int splits_for(int m, int n) {
  const int panels = (n + 255) / 256;
  int s = (2048 + panels - 1) / panels;  // aim at >= 2048 workgroups
  s = std::max(1, std::min(s, (m + 63) / 64));
  return s;
}

}  // namespace

int andr_colscale(andr_ctx *ctx, void *A, bool fp32, int64_t lda, int m, int n, double *q) {
  const int ns = splits_for(m, n);
  const int rps = (m + ns - 1) / ns;
  DevBuf<double> P;
  if (P.alloc((size_t)ns * n, ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "column scaling scratch");
  dim3 g1((n + 255) / 256, ns);
  if (fp32)
    hipLaunchKernelGGL(k_colsq<float>, g1, dim3(256), 0, ctx->stream, (const float *)A, lda, m, n, rps, P.p);
  else
    hipLaunchKernelGGL(k_colsq<double>, g1, dim3(256), 0, ctx->stream, (const double *)A, lda, m, n, rps, P.p);
  hipLaunchKernelGGL(k_colq, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ns, P.p, q);
  dim3 g2((n + 255) / 256, std::min(m, 256));
  if (fp32)
    hipLaunchKernelGGL(k_colmul<float>, g2, dim3(256), 0, ctx->stream, (float *)A, lda, m, n, q);
  else
    hipLaunchKernelGGL(k_colmul<double>, g2, dim3(256), 0, ctx->stream, (double *)A, lda, m, n, q);
  ANDR_HIP(ctx, hipGetLastError());
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ANDR_OK;
}

namespace {

// One LSQR solve, advanced in blocks of iterations on a stream of its own (several solves of a batch
// then overlap on the device); the profiling tokens are taken only on the context's stream.
struct LsqrRun {
  andr_ctx *ctx = nullptr;
  hipStream_t st = nullptr;
  const void *A = nullptr;
  bool fp32 = false;
  int64_t lda = 0;
  int m = 0, n = 0, nrhs = 0, itnlim = 0, it = 0;
  const double *q = nullptr;
  double *X = nullptr;
  bool done = false;
  int wavesN = 0, blocksN = 0, ns = 0, rps = 0, blocksC = 0, blocksM = 0;
  int nsegN = 1, seglenN = 0, partsN = 0;  // A X: column segments and the partial sums it leaves
  int tilesN = 4;                           // A X: 16-row tiles per wave (autotuned)
  bool vecN = false;
  bool quiet = false;  // no profiling records (the no-allocation path: event creation is skipped)
  // sparse mode: A and A^T in CSR (0-based, device); A is then unused
  const int64_t *cia = nullptr, *tia = nullptr;
  const int32_t *cja = nullptr, *tja = nullptr;
  const double *cva = nullptr, *tva = nullptr;
  int64_t nnz = 0;
  double *pn = nullptr;  // A X segment partials when they come from a caller's workspace
  DevBuf<double> U, V, W, Xl, X0, Xs, P, PN, partU, partV, partW, partB, sc;
  DevBuf<LsqrState> S;
  DevBuf<int> act;
  LsqrState hs[NB];
  double *sxn = nullptr, *syn = nullptr, *sxt = nullptr, *syt = nullptr, *t1 = nullptr, *t2 = nullptr, *ia = nullptr;
  double abytes = 0, vbytes = 0, aflops = 0;

  int tok(const char *name, double bytes) {
    return st == ctx->stream && !quiet ? andr_prof_begin(ctx, name, bytes, aflops) : -1;
  }
  template <int T, bool SEG>
  void launch_mf_t(const double *Xin, double *Y, const double *sx, const double *sy, double *out) {
    const dim3 grid((m + 64 * T - 1) / (64 * T), nsegN);  // four waves of 16 T rows per block
    if (vecN)
      hipLaunchKernelGGL((k_gemv_n_mf<T, SEG, true>), grid, dim3(256), 0, st, (const double *)A, lda, m, n, Xin,
                         seglenN, Y, sx, sy, out);
    else
      hipLaunchKernelGGL((k_gemv_n_mf<T, SEG, false>), grid, dim3(256), 0, st, (const double *)A, lda, m, n, Xin,
                         seglenN, Y, sx, sy, out);
  }
  void launch_mf(int tiles, bool seg, const double *Xin, double *Y, const double *sx, const double *sy, double *out) {
    switch (tiles * 2 + (seg ? 1 : 0)) {
      case 4: launch_mf_t<2, false>(Xin, Y, sx, sy, out); break;
      case 5: launch_mf_t<2, true>(Xin, Y, sx, sy, out); break;
      case 16: launch_mf_t<8, false>(Xin, Y, sx, sy, out); break;
      case 17: launch_mf_t<8, true>(Xin, Y, sx, sy, out); break;
      case 9: launch_mf_t<4, true>(Xin, Y, sx, sy, out); break;
      default: launch_mf_t<4, false>(Xin, Y, sx, sy, out); break;
    }
  }
  // Autotuner of the FP64 product: on the first large system of a (log2 m, log2 n) bucket, time the
  // segmented product with 2, 4 and 8 tiles per wave on the actual matrix (its output goes to the
  // segment partials only) and keep the fastest for the bucket. The choice does not change results.
  int tune_tiles() {
    const std::pair<int, int> key{ilog2c(m), ilog2c(n)};
    auto it = ctx->tune_gemv_n.find(key);
    if (it != ctx->tune_gemv_n.end()) {
      tilesN = it->second.first;
      return ANDR_OK;
    }
    ANDR_HIP(ctx, hipDeviceSynchronize());  // time the candidates alone on the device
    hipEvent_t e0, e1;
    ANDR_HIP(ctx, hipEventCreate(&e0));
    ANDR_HIP(ctx, hipEventCreate(&e1));
    int best = 4;
    float best_ms = 1e30f;
    for (int tiles : {2, 4, 8}) {
      launch_mf(tiles, true, V.p, U.p, sxn, syn, PN.p);  // warm-up
      float ms = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        (void)hipEventRecord(e0, st);
        launch_mf(tiles, true, V.p, U.p, sxn, syn, PN.p);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms = std::min(ms, t);
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = tiles;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ANDR_HIP(ctx, hipGetLastError());
    ctx->tune_gemv_n[key] = {best, best_ms};
    tilesN = best;
    return ANDR_OK;
  }
  static int ilog2c(int v) {
    int b = 0;
    while ((1 << b) < v) ++b;
    return b;
  }

  void gemv_n(const double *Xin, double *Y, const double *sx, const double *sy, double *part) {
    const int t = tok("lsqr_gemv_n", abytes + vbytes + (double)m * NB * 8.0);
    if (cia) {
      hipLaunchKernelGGL(k_csr_mv<true>, dim3(blocksM), dim3(256), 0, st, m, cia, cja, cva, Xin, Y, sx, sy, part);
    } else if (fp32) {
      hipLaunchKernelGGL(k_gemv_n<float>, dim3(blocksN), dim3(256), 0, st, (const float *)A, lda, m, n, Xin, Y, sx, sy,
                         part);
    } else {
      double *segp = pn ? pn : PN.p;
      launch_mf(tilesN, nsegN > 1, Xin, Y, sx, sy, nsegN > 1 ? segp : part);
      if (nsegN > 1) hipLaunchKernelGGL(k_combine_n, dim3(blocksM), dim3(256), 0, st, m, nsegN, segp, Y, sx, sy, part);
    }
    andr_prof_end(ctx, t);
  }
  void gemv_t(const double *Y) {
    const int t = tok("lsqr_gemv_t", abytes + vbytes + (double)ns * n * NB * 8.0);
    dim3 g((n + 255) / 256, ns);
    if (cia)
      hipLaunchKernelGGL(k_csr_mv<false>, dim3(blocksC), dim3(256), 0, st, n, tia, tja, tva, Y, P.p,
                         (const double *)nullptr, (const double *)nullptr, (double *)nullptr);
    else if (fp32)
      hipLaunchKernelGGL(k_gemv_t<float>, g, dim3(256), 0, st, (const float *)A, lda, m, n, Y, rps, P.p);
    else if (vecN && (n % 2 == 0 || n < lda))  // column pairs: an odd last column reads the row padding
      hipLaunchKernelGGL(k_gemv_t2, dim3((n + 511) / 512, ns), dim3(256), 0, st, (const double *)A, lda, m, n, Y, rps,
                         P.p);
    else
      hipLaunchKernelGGL(k_gemv_t<double>, g, dim3(256), 0, st, (const double *)A, lda, m, n, Y, rps, P.p);
    andr_prof_end(ctx, t);
  }

  // workspace, initial state, x_s = x0 / q, U = b - A x_s, V = A^T U, W = V / alpha
  int begin(andr_ctx *c, hipStream_t stream, const void *A_, bool fp32_, int64_t lda_, int m_, int n_,
            const double *q_, const double *B, double *X_, int nrhs_, int maxit, double eps) {
    ctx = c;
    st = stream;
    A = A_;
    fp32 = fp32_;
    lda = lda_;
    m = m_;
    n = n_;
    q = q_;
    X = X_;
    nrhs = nrhs_;
    if (nrhs < 1 || nrhs > NB) return andr_fail(ctx, ANDR_ERR_ARG, "nrhs must be 1..%d", NB);
    if (m < 1 || n < 1) return andr_fail(ctx, ANDR_ERR_ARG, "empty system");
    if (cia) {  // CSR products: one thread per (row, column); partials per block of 32 rows
      nsegN = 1;
      partsN = (m * NB + 255) / 256;
    } else if (fp32) {  // VALU product: one wave per 8 rows
      wavesN = (m + 7) / 8;
      blocksN = (wavesN + 3) / 4;
      partsN = wavesN;
    } else {  // matrix-core product over column segments; room for the most partials
      // segments of GN_SEGLEN columns, shorter (down to 256) when that leaves fewer than 2048 waves
      // of 64 rows; a function of the shape only, so that the tile count never changes the sums
      const int64_t rw = (m + 63) / 64;
      seglenN = GN_SEGLEN;
      if (rw * ((n + GN_SEGLEN - 1) / GN_SEGLEN) < 2048)
        seglenN = (int)std::max<int64_t>(256, ((int64_t)n * rw / 2048 + 15) / 16 * 16);
      nsegN = (n + seglenN - 1) / seglenN;
      partsN = std::max((m + 31) / 32, (m * NB + 255) / 256);
      vecN = lda % 2 == 0 && ((uintptr_t)A % 16) == 0;
    }
    ns = cia ? 1 : splits_for(m, n);
    rps = (m + ns - 1) / ns;
    blocksC = (n * NB + 255) / 256;
    blocksM = (m * NB + 255) / 256;
    if (U.alloc((size_t)m * NB, ctx) || V.alloc((size_t)n * NB, ctx) || W.alloc((size_t)n * NB, ctx) ||
        Xl.alloc((size_t)n * NB, ctx) || X0.alloc((size_t)n * NB, ctx) || Xs.alloc((size_t)n * NB, ctx) ||
        P.alloc((size_t)ns * n * NB, ctx) || PN.alloc(nsegN > 1 ? (size_t)nsegN * m * NB : 0, ctx) ||
        partU.alloc((size_t)partsN * NB, ctx) ||
        partV.alloc((size_t)blocksC * NB, ctx) || partW.alloc((size_t)blocksC * NB, ctx) ||
        partB.alloc((size_t)blocksM * NB, ctx) || sc.alloc(8 * NB, ctx) || S.alloc(NB, ctx) || act.alloc(NB, ctx))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "LSQR workspace (m=%d n=%d)", m, n);
    sxn = sc.p;
    syn = sc.p + NB;
    sxt = sc.p + 2 * NB;
    syt = sc.p + 3 * NB;
    t1 = sc.p + 4 * NB;
    t2 = sc.p + 5 * NB;
    ia = sc.p + 6 * NB;
    std::memset(hs, 0, sizeof(hs));
    itnlim = maxit > 0 ? maxit : 10 * (m + n + 50);
    for (int k = 0; k < NB; ++k) {
      hs[k].active = k < nrhs ? 1 : 0;
      hs[k].itnlim = itnlim;
    }
    double hsc[8 * NB];
    for (int k = 0; k < NB; ++k) {
      hsc[k] = -1.0;       // sxn: U = -A x_s + U
      hsc[NB + k] = 1.0;   // syn
      hsc[2 * NB + k] = hsc[3 * NB + k] = hsc[4 * NB + k] = hsc[5 * NB + k] = hsc[6 * NB + k] = 0.0;
      hsc[7 * NB + k] = 1.0;
    }
    ANDR_HIP(ctx, hipMemcpyAsync(S.p, hs, sizeof(hs), hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(sc.p, hsc, sizeof(hsc), hipMemcpyHostToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(X0.p, X, sizeof(double) * n * NB, hipMemcpyDeviceToDevice, st));
    ANDR_HIP(ctx, hipMemcpyAsync(U.p, B, sizeof(double) * m * NB, hipMemcpyDeviceToDevice, st));
    ANDR_HIP(ctx, hipMemsetAsync(V.p, 0, sizeof(double) * n * NB, st));
    ANDR_HIP(ctx, hipMemsetAsync(Xl.p, 0, sizeof(double) * n * NB, st));
    ANDR_HIP(ctx, hipMemsetAsync(W.p, 0, sizeof(double) * n * NB, st));
    if (!fp32 && !cia) {
      tilesN = ctx->autotune ? 4 : ctx->gemv_n_tiles;
      if (nsegN > 1 && ctx->autotune && (double)m * n * 8.0 >= 2.56e8) {
        const int rc = tune_tiles();
        if (rc) return rc;
      }
      partsN = nsegN == 1 ? (m + 16 * tilesN - 1) / (16 * tilesN) : blocksM;  // one per wave, or per block
    }
    abytes = cia ? (double)nnz * 12.0 + (double)(m + 1) * 8.0 : (double)m * n * (fp32 ? 4.0 : 8.0);
    vbytes = ((double)m + n) * NB * 8.0;
    aflops = 2.0 * m * n * nrhs;
    // x_s = x0 / q ; U = b - A x_s ; ||U||, ||b||
    hipLaunchKernelGGL(k_scale_div, dim3(blocksC), dim3(256), 0, st, n, X0.p, q, Xs.p);
    gemv_n(Xs.p, U.p, sxn, syn, partU.p);
    hipLaunchKernelGGL(k_sumsq_parts, dim3(blocksM), dim3(256), 0, st, m, B, partB.p);
    hipLaunchKernelGGL(k_lsqr_start, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S.p, partU.p, partsN, partB.p, blocksM, eps, sxt, syt);
    // V = A^T u
    gemv_t(U.p);
    hipLaunchKernelGGL(k_combine_t, dim3(blocksC), dim3(256), 0, st, n, ns, P.p, V.p, (const double *)nullptr, sxt, syt,
                       partV.p, partW.p);
    hipLaunchKernelGGL(k_lsqr_init2, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S.p, partV.p, blocksC, sxn, syn, ia);
    // w = v / alpha  (X update with t1 = 0, t2 = 0, W = V * inv_alpha)
    int allon[NB];
    for (int k = 0; k < NB; ++k) allon[k] = 1;
    ANDR_HIP(ctx, hipMemcpyAsync(act.p, allon, sizeof(allon), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_lsqr_update, dim3(blocksC), dim3(256), 0, st, n, Xl.p, W.p, V.p, t1, t2, ia, act.p);
    ANDR_HIP(ctx, hipGetLastError());
    return ANDR_OK;
  }

  // enqueues iterations up to the next multiple of `poll` (or the iteration limit); no host wait
  void advance(int poll) {
    do {
      gemv_n(V.p, U.p, sxn, syn, partU.p);
      hipLaunchKernelGGL(k_lsqr_mid, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S.p, partU.p, partsN, sxt, syt);
      gemv_t(U.p);
      hipLaunchKernelGGL(k_combine_t, dim3(blocksC), dim3(256), 0, st, n, ns, P.p, V.p, W.p, sxt, syt, partV.p,
                         partW.p);
      hipLaunchKernelGGL(k_lsqr_end, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S.p, partV.p, partW.p, blocksC, sxn, syn, t1, t2, ia,
                         act.p);
      hipLaunchKernelGGL(k_lsqr_update, dim3(blocksC), dim3(256), 0, st, n, Xl.p, W.p, V.p, t1, t2, ia, act.p);
      ++it;
    } while (it % poll != 0 && it < itnlim);
  }

  // reads the per-column states; done when no column is active or the limit is reached
  int poll_state() {
    ANDR_HIP(ctx, hipGetLastError());
    ANDR_HIP(ctx, hipMemcpyAsync(hs, S.p, sizeof(hs), hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));
    bool any = false;
    for (int k = 0; k < nrhs; ++k) any = any || hs[k].active;
    done = !any || it >= itnlim;
    return ANDR_OK;
  }

  // x = q .* x + x0, iteration counts and stop codes
  int finish(int *nit, int *istop) {
    hipLaunchKernelGGL(k_scale_mul_add, dim3(blocksC), dim3(256), 0, st, n, q, Xl.p, X0.p, X);
    ANDR_HIP(ctx, hipGetLastError());
    ANDR_HIP(ctx, hipMemcpyAsync(hs, S.p, sizeof(hs), hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));
    for (int k = 0; k < nrhs; ++k) {
      if (nit) nit[k] = hs[k].itn;
      if (istop) istop[k] = hs[k].istop;
    }
    return ANDR_OK;
  }
};

constexpr int LSQR_POLL = 8;  // iterations between reads of the convergence state

}  // namespace

int andr_lsqr_device(andr_ctx *ctx, const void *A, bool fp32, int64_t lda, int m, int n, const double *q,
                     const double *B, double *X, int nrhs, int maxit, double eps, int *nit, int *istop) {
  LsqrRun R;
  int rc = R.begin(ctx, ctx->stream, A, fp32, lda, m, n, q, B, X, nrhs, maxit, eps);
  if (rc) return rc;
  while (!R.done) {
    R.advance(LSQR_POLL);
    if ((rc = R.poll_state())) return rc;
  }
  return R.finish(nit, istop);
}

int andr_lsqr_device_csr(andr_ctx *ctx, int m, int n, int64_t nnz, const int64_t *ia, const int32_t *ja,
                         const double *a, const int64_t *tia, const int32_t *tja, const double *ta, const double *q,
                         const double *B, double *X, int nrhs, int maxit, double eps, int *nit, int *istop) {
  LsqrRun R;
  R.cia = ia;
  R.cja = ja;
  R.cva = a;
  R.tia = tia;
  R.tja = tja;
  R.tva = ta;
  R.nnz = nnz;
  int rc = R.begin(ctx, ctx->stream, nullptr, false, 0, m, n, q, B, X, nrhs, maxit, eps);
  if (rc) return rc;
  while (!R.done) {
    R.advance(LSQR_POLL);
    if ((rc = R.poll_state())) return rc;
  }
  return R.finish(nit, istop);
}

int andr_lsqr_device_batch(andr_ctx *ctx, int count, const LsqrJob *jobs) {
  if (count <= 0) return ANDR_OK;
  if (count == 1 || andr_prof(ctx).on) {  // profiled: one after another, so that each event pair times one launch
    for (int i = 0; i < count; ++i) {
      const LsqrJob &J = jobs[i];
      const int rc =
          andr_lsqr_device(ctx, J.A, J.fp32, J.lda, J.m, J.n, J.q, J.B, J.X, J.nrhs, J.maxit, J.eps, J.nit, J.istop);
      if (rc) return rc;
    }
    return ANDR_OK;
  }
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the systems are ready
  const int nst = std::min(count, LSQR_STREAMS);
  std::vector<hipStream_t> streams(nst, nullptr);
  int rc = ANDR_OK;
  for (int i = 0; i < nst && !rc; ++i)
    if (hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking) != hipSuccess)
      rc = andr_fail(ctx, ANDR_ERR_HIP, "stream creation for the LSQR batch");
  {
    std::vector<LsqrRun> runs(count);
    for (int i = 0; i < count && !rc; ++i) {
      const LsqrJob &J = jobs[i];
      rc = runs[i].begin(ctx, streams[i % nst], J.A, J.fp32, J.lda, J.m, J.n, J.q, J.B, J.X, J.nrhs, J.maxit, J.eps);
    }
    // all solves advance together, a block of iterations at a time, each on its stream
    for (bool left = !rc; left && !rc;) {
      left = false;
      for (auto &R : runs)
        if (!R.done) R.advance(LSQR_POLL);
      for (auto &R : runs)
        if (!R.done && !rc) {
          rc = R.poll_state();
          left = left || !R.done;
        }
    }
    for (int i = 0; i < count && !rc; ++i) rc = runs[i].finish(jobs[i].nit, jobs[i].istop);
    for (auto s : streams)
      if (s) (void)hipStreamSynchronize(s);  // the workspaces go back to the cache after this
  }
  for (auto s : streams)
    if (s) (void)hipStreamDestroy(s);
  return rc;
}

// ---- BiCGSTAB with a Jacobi preconditioner ---------------------------------------------------
// The Krylov alternative of the reference's LIS path for square systems (parLISslvSet
// "-i bicgstab -p jacobi -tol 1.0e-8 -maxiter 1000", src/mod_par.f90:288). LIS is not part of the
// reference's sources; the algorithm is the published preconditioned BiCGSTAB (van der Vorst 1992,
// in the form of Barrett et al., Templates, 1994, fig. 2.10): the preconditioner M = diag(A) acts
// on the search directions, the residual is the true one, and a column stops when
// ||r||_2 <= tol ||b||_2 (LIS's default criterion). NB columns advance together, each with its own
// scalars; the products are the matrix-core A X above.
namespace {

struct BicgState {
  double rho, alpha, omega, bnorm2, tol2;
  int active, itn, status;  // status: 0 converged, 1 iteration limit, 2 breakdown
};

// per-block partial sums of X.Y (and of Z.W when Z is given) over n x NB column-interleaved arrays
__global__ __launch_bounds__(256) void k_bicg_dots(int n, const double *__restrict__ X, const double *__restrict__ Y,
                                                   const double *__restrict__ Z, const double *__restrict__ W,
                                                   double *__restrict__ p1, double *__restrict__ p2) {
  __shared__ double red[2][4 * NB];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (t < n * NB) {
    a = X[t] * Y[t];
    if (Z) b = Z[t] * W[t];
  }
#pragma unroll
  for (int off = 8; off < 64; off <<= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane < NB) {
    red[0][wid * NB + lane] = a;
    red[1][wid * NB + lane] = b;
  }
  __syncthreads();
  if (threadIdx.x < NB) {
    const int k = threadIdx.x;
    p1[(size_t)blockIdx.x * NB + k] = (red[0][k] + red[0][NB + k]) + (red[0][2 * NB + k] + red[0][3 * NB + k]);
    if (Z) p2[(size_t)blockIdx.x * NB + k] = (red[1][k] + red[1][NB + k]) + (red[1][2 * NB + k] + red[1][3 * NB + k]);
  }
}

// Jacobi: dinv[j] = 1 / A[j][j] (1 where the diagonal is zero, as a Jacobi preconditioner leaves it)
__global__ void k_bicg_diag(const double *__restrict__ A, int64_t lda, int n, double *__restrict__ dinv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double d = A[(size_t)j * lda + j];
  dinv[j] = d != 0.0 ? 1.0 / d : 1.0;
}

// r = b - A x from Y = A x; r_hat = r; p = v = 0
__global__ void k_bicg_init(int n, const double *__restrict__ B, const double *__restrict__ AX, double *__restrict__ R,
                            double *__restrict__ Rh, double *__restrict__ P, double *__restrict__ V) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  const double r = B[t] - AX[t];
  R[t] = r;
  Rh[t] = r;
  P[t] = 0.0;
  V[t] = 0.0;
}

// start: ||b||^2 and ||r0||^2; columns with r0 already small are done
__global__ void k_bicg_start(BicgState *st, const double *pb, const double *pr, int np, double tol, int nrhs,
                             double *sc) {
  __shared__ double red[4 * NB];
  const double b2 = block_sum_parts(pb, np, red);
  const double r2 = block_sum_parts(pr, np, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  BicgState &S = st[k];
  S.rho = S.alpha = S.omega = 1.0;
  S.bnorm2 = b2;
  S.tol2 = tol * tol * b2;
  S.itn = 0;
  S.status = 0;
  S.active = k < nrhs && r2 > S.tol2 ? 1 : 0;
  for (int i = 0; i < 6; ++i) sc[i * NB + k] = 0.0;
}

// scalar slots sc[i * NB + k]: 0 beta, 1 omega (for p), 2 alpha, 3 omega (for x, r)
// rho_new = r_hat . r; beta = (rho_new / rho) (alpha / omega)
__global__ void k_bicg_beta(BicgState *st, const double *part, int np, double *sc) {
  __shared__ double red[4 * NB];
  const double rho = block_sum_parts(part, np, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  BicgState &S = st[k];
  if (!S.active) {
    sc[k] = 0.0;
    sc[NB + k] = 0.0;
    return;
  }
  if (rho == 0.0) {  // breakdown: r_hat orthogonal to r
    S.active = 0;
    S.status = 2;
    sc[k] = sc[NB + k] = 0.0;
    return;
  }
  S.itn += 1;
  const double beta = S.itn == 1 ? 0.0 : (rho / S.rho) * (S.alpha / S.omega);
  S.rho = rho;
  sc[k] = beta;
  sc[NB + k] = S.omega;
}

// p = r + beta (p - omega v); y = dinv .* p
__global__ void k_bicg_p(int n, const double *__restrict__ R, const double *__restrict__ V, double *__restrict__ P,
                         double *__restrict__ Yv, const double *__restrict__ dinv, const double *__restrict__ sc,
                         const BicgState *__restrict__ st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  const int k = t & 7;
  if (!st[k].active) {
    Yv[t] = 0.0;
    return;
  }
  const double p = R[t] + sc[k] * (P[t] - sc[NB + k] * V[t]);
  P[t] = p;
  Yv[t] = dinv[t >> 3] * p;
}

// alpha = rho / (r_hat . v)
__global__ void k_bicg_alpha(BicgState *st, const double *part, int np, double *sc) {
  __shared__ double red[4 * NB];
  const double rv = block_sum_parts(part, np, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  BicgState &S = st[k];
  if (!S.active || rv == 0.0) {
    if (S.active) {
      S.active = 0;
      S.status = 2;
    }
    sc[2 * NB + k] = 0.0;
    return;
  }
  S.alpha = S.rho / rv;
  sc[2 * NB + k] = S.alpha;
}

// x += alpha y; s = r - alpha v (into R); z = dinv .* s
__global__ void k_bicg_s(int n, double *__restrict__ X, const double *__restrict__ Yv, double *__restrict__ R,
                         const double *__restrict__ V, double *__restrict__ Z, const double *__restrict__ dinv,
                         const double *__restrict__ sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  const int k = t & 7;
  const double a = sc[2 * NB + k];
  X[t] += a * Yv[t];
  const double s = R[t] - a * V[t];
  R[t] = s;
  Z[t] = dinv[t >> 3] * s;
}

// omega = (t . s) / (t . t)
__global__ void k_bicg_omega(BicgState *st, const double *pts, const double *ptt, int np, double *sc) {
  __shared__ double red[4 * NB];
  const double ts = block_sum_parts(pts, np, red);
  const double tt = block_sum_parts(ptt, np, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  BicgState &S = st[k];
  if (!S.active || tt == 0.0) {
    sc[3 * NB + k] = 0.0;
    return;
  }
  S.omega = ts / tt;
  sc[3 * NB + k] = S.omega;
}

// x += omega z; r = s - omega t
__global__ void k_bicg_x(int n, double *__restrict__ X, const double *__restrict__ Z, double *__restrict__ R,
                         const double *__restrict__ Tv, const double *__restrict__ sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * NB) return;
  const int k = t & 7;
  const double w = sc[3 * NB + k];
  X[t] += w * Z[t];
  R[t] -= w * Tv[t];
}

// convergence: ||r||^2 <= tol^2 ||b||^2; omega = 0 (with r not small) is a breakdown
__global__ void k_bicg_check(BicgState *st, const double *part, int np, int itnlim) {
  __shared__ double red[4 * NB];
  const double r2 = block_sum_parts(part, np, red);
  const int k = threadIdx.x;
  if (k >= NB) return;
  BicgState &S = st[k];
  if (!S.active) return;
  if (r2 <= S.tol2) {
    S.active = 0;
    S.status = 0;
  } else if (S.omega == 0.0) {
    S.active = 0;
    S.status = 2;
  } else if (S.itn >= itnlim) {
    S.active = 0;
    S.status = 1;
  }
}

}  // namespace

namespace {
// the shape of the BiCGSTAB product A X: column segments and the partial sums they leave (a function
// of n only, so that the workspace size is known before the call)
struct BicgLayout {
  int seglen, nseg, blocksM, partsN;
};
BicgLayout bicg_layout(int n) {
  BicgLayout L;
  L.seglen = 1024;
  const int64_t rw = (n + 63) / 64;
  if (rw * ((n + 1023) / 1024) < 2048) L.seglen = (int)std::max<int64_t>(256, ((int64_t)n * rw / 2048 + 15) / 16 * 16);
  L.nseg = (n + L.seglen - 1) / L.seglen;
  L.blocksM = (n * NB + 255) / 256;
  L.partsN = L.nseg == 1 ? (n + 63) / 64 : L.blocksM;
  return L;
}
// doubles of each workspace array, in the order they are carved from a caller's region
void bicg_sizes(int n, size_t sz[15]) {
  const BicgLayout L = bicg_layout(n);
  for (int i = 0; i < 7; ++i) sz[i] = (size_t)n * NB;  // r, r_hat, p, v, y, z, t
  sz[7] = n;                                           // dinv
  sz[8] = (size_t)std::max(L.blocksM, L.partsN) * NB;  // p1
  sz[9] = (size_t)L.blocksM * NB;                      // p2
  sz[10] = NB;                                         // one
  sz[11] = NB;                                         // zero
  sz[12] = 6 * NB;                                     // scalars
  sz[13] = (sizeof(BicgState) * NB + 7) / 8;           // per-column states
  sz[14] = L.nseg > 1 ? (size_t)L.nseg * n * NB : 0;   // A X segment partials
}
__global__ void k_fill(size_t n, double v, double *__restrict__ p) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}
}  // namespace

size_t andr_bicgstab_ws_bytes(int n) {
  size_t sz[15], tot = 0;
  bicg_sizes(n, sz);
  for (size_t v : sz) tot += (v + 31) / 32 * 32;  // every array on a 256-byte boundary
  return tot * sizeof(double);
}

int andr_bicgstab_device(andr_ctx *ctx, const double *A, int64_t lda, int n, const double *B, double *X, int nrhs,
                         int maxit, double tol, int *nit, int *status, double *ws, size_t ws_bytes) {
  if (nrhs < 1 || nrhs > NB) return andr_fail(ctx, ANDR_ERR_ARG, "nrhs must be 1..%d", NB);
  if (n < 1) return andr_fail(ctx, ANDR_ERR_ARG, "empty system");
  hipStream_t st = ctx->stream;
  // the matrix-core product of the LSQR run object (q unused)
  const BicgLayout L = bicg_layout(n);
  LsqrRun R;
  R.ctx = ctx;
  R.st = st;
  R.A = A;
  R.fp32 = false;
  R.lda = lda;
  R.m = R.n = n;
  R.seglenN = L.seglen;
  R.nsegN = L.nseg;
  R.tilesN = 4;
  R.vecN = lda % 2 == 0 && ((uintptr_t)A % 16) == 0;
  R.blocksM = L.blocksM;
  R.partsN = L.partsN;
  const int nb = R.blocksM;
  // workspace: the context's memory cache, or carved from the caller's region (no allocation)
  DevBuf<double> own[15];
  double *w[15];
  size_t sz[15];
  bicg_sizes(n, sz);
  if (ws) {
    if (ws_bytes < andr_bicgstab_ws_bytes(n))
      return andr_fail(ctx, ANDR_ERR_ARG, "BiCGSTAB workspace of %zu bytes, %zu needed (n=%d)", ws_bytes,
                       andr_bicgstab_ws_bytes(n), n);
    double *cur = ws;
    for (int i = 0; i < 15; ++i) {
      w[i] = sz[i] ? cur : nullptr;
      cur += (sz[i] + 31) / 32 * 32;
    }
    R.quiet = true;
  } else {
    for (int i = 0; i < 15; ++i) {
      if (own[i].alloc(sz[i], ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "BiCGSTAB workspace (n=%d)", n);
      w[i] = own[i].p;
    }
  }
  double *Rr = w[0], *Rh = w[1], *P = w[2], *V = w[3], *Yv = w[4], *Z = w[5], *Tt = w[6], *dinv = w[7], *p1 = w[8],
         *p2 = w[9], *one = w[10], *zero = w[11], *sc = w[12];
  BicgState *S = (BicgState *)w[13];
  R.pn = w[14];
  // constants and cleared outputs by fill kernels (no host copies, no memset nodes): the products
  // write out = 1 * (A in) + 0 * out, so their outputs start finite
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (size_t)NB, 1.0, one);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (size_t)NB, 0.0, zero);
  hipLaunchKernelGGL(k_fill, dim3((n * NB + 255) / 256), dim3(256), 0, st, (size_t)n * NB, 0.0, V);
  hipLaunchKernelGGL(k_fill, dim3((n * NB + 255) / 256), dim3(256), 0, st, (size_t)n * NB, 0.0, Tt);
  R.abytes = (double)n * n * 8.0;
  R.vbytes = 2.0 * n * NB * 8.0;
  R.aflops = 2.0 * n * n * nrhs;
  const int gb = (n * NB + 255) / 256;
  auto product = [&](const double *in, double *out) {  // out = A in (its partial squared norms are not used)
    R.gemv_n(in, out, one, zero, p1);
  };
  hipLaunchKernelGGL(k_bicg_diag, dim3((n + 255) / 256), dim3(256), 0, st, A, lda, n, dinv);
  product(X, V);  // A x0
  hipLaunchKernelGGL(k_bicg_init, dim3(gb), dim3(256), 0, st, n, B, V, Rr, Rh, P, V);
  hipLaunchKernelGGL(k_bicg_dots, dim3(gb), dim3(256), 0, st, n, B, B, Rr, Rr, p1, p2);
  hipLaunchKernelGGL(k_bicg_start, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S, p1, p2, nb, tol, nrhs, sc);
  const int itnlim = maxit > 0 ? maxit : 1000;
  BicgState hs[NB];
  for (int it = 0; it < itnlim;) {
    for (int inner = 0; inner < 8 && it < itnlim; ++inner, ++it) {
      hipLaunchKernelGGL(k_bicg_dots, dim3(gb), dim3(256), 0, st, n, Rh, Rr, (const double *)nullptr,
                         (const double *)nullptr, p1, p2);
      hipLaunchKernelGGL(k_bicg_beta, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S, p1, nb, sc);
      hipLaunchKernelGGL(k_bicg_p, dim3(gb), dim3(256), 0, st, n, Rr, V, P, Yv, dinv, sc, S);
      product(Yv, V);
      hipLaunchKernelGGL(k_bicg_dots, dim3(gb), dim3(256), 0, st, n, Rh, V, (const double *)nullptr,
                         (const double *)nullptr, p1, p2);
      hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S, p1, nb, sc);
      hipLaunchKernelGGL(k_bicg_s, dim3(gb), dim3(256), 0, st, n, X, Yv, Rr, V, Z, dinv, sc);
      product(Z, Tt);
      hipLaunchKernelGGL(k_bicg_dots, dim3(gb), dim3(256), 0, st, n, Tt, Rr, Tt, Tt, p1, p2);
      hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S, p1, p2, nb, sc);
      hipLaunchKernelGGL(k_bicg_x, dim3(gb), dim3(256), 0, st, n, X, Z, Rr, Tt, sc);
      hipLaunchKernelGGL(k_bicg_dots, dim3(gb), dim3(256), 0, st, n, Rr, Rr, (const double *)nullptr,
                         (const double *)nullptr, p1, p2);
      hipLaunchKernelGGL(k_bicg_check, dim3(1), dim3(LSQR_SCALAR_THREADS), 0, st, S, p1, nb, itnlim);
    }
    ANDR_HIP(ctx, hipGetLastError());
    ANDR_HIP(ctx, hipMemcpyAsync(hs, S, sizeof(hs), hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));
    bool any = false;
    for (int k = 0; k < nrhs; ++k) any = any || hs[k].active;
    if (!any) break;
  }
  ANDR_HIP(ctx, hipMemcpyAsync(hs, S, sizeof(hs), hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipStreamSynchronize(st));
  for (int k = 0; k < nrhs; ++k) {
    if (nit) nit[k] = hs[k].itn;
    if (status) status[k] = hs[k].active ? 1 : hs[k].status;
  }
  return ANDR_OK;
}
