// Restarted GMRES on a batch of structured Stokes systems (the six symmetric FP32 Stokeslet blocks
// of a body, right block-Jacobi preconditioner, null-space term) with up to ANDR_NB right-hand sides
// each: the solver of the mixed-precision and learned resistance forms.
//
// The blocks are stored in FP32; products are formed in FP64 and every vector (Krylov basis,
// Hessenberg, solution) is FP64. Each restart cycle starts from the FP64 residual r = b - A x, so the cycles are
// the iterative-refinement loop around the Krylov solve. All systems of the batch advance together:
// one launch per step covers every system (work lists map waves and row tiles to systems), so the
// launch count does not grow with the number of bodies.
//
// Orthogonalisation is classical Gram-Schmidt applied twice with the second pass delayed to the
// next step (one sweep of dot products and one of updates over the basis per step instead of two
// of each); both sweeps stream the basis.
#include "andr_internal.h"
#include "andr_solve.h"
#include "andr_prof.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

constexpr int NB = ANDR_NB;
static_assert(NB == 8, "the matrix-core product packs two 8-column halves into the 16 B columns");
constexpr int TILE = 64;   // rows per vector tile (256 threads: 32 rows x 8 right-hand sides per pass)
constexpr int ICH = 8;   // basis vectors per block of the projection kernel
constexpr int BT = 256;  // rows per tile of the block Gram-Schmidt sweeps
// A cycle stops when the least-squares estimate of every column's residual is below EST_MARGIN tol |b|:
// the other half of the tolerance covers the difference between the estimate and the true FP64
// residual that the FP32 Krylov products leave (the restart test checks the true residual).
constexpr double EST_MARGIN = 0.5;

struct GSys {  // device view of one system
  const void *A;
  int64_t lda;
  int n;
  const double *B;
  double *X;
  double *R;     // residual / scratch n x NB
  double *W;     // new Krylov vector n x NB
  double *Q;     // (m+1) x n x NB basis
  double *H;     // NB x (m+1) x m Hessenberg after the Givens rotations (column-major per rhs)
  double *Hu;    // NB x (m+1) x m Hessenberg as built (column-major per rhs)
  double *cs, *sn;  // NB x m
  double *g;     // NB x (m+1)
  double *cw;    // NB x cw_stride(m): step coefficients (see kb_adots), then the solution y
  double *small; // per rhs: [0] bnorm, [1] rnorm (true), [2] est, [3] inv_h, [4] active, [5] used steps
  int tile0;     // first vector tile of the system
  int ntiles;
  int nseg;      // column segments of the matrix-core product (1: written straight to W)
  double *P;     // nseg x n x NB partial products when nseg > 1
  // structured Stokes operator (fp32 path)
  StokesOp op;
  int nep;       // op.ne rounded up to 32
  double *ZT;    // 3 x NB x nep: z = M^-1 q_j (FP64 products) or M^-1 x (residual), [component][rhs][element]
  float *ZT32;   // the same in FP32 (FP32 products), null: FP64 products
  double *sdef;  // ANDR_OP_MAXW x NB: sum over wall w of n_e . z_e
  int et0, net;  // element tiles (32 elements) of the system in the current map
  int incyc;     // 1 while the system takes part in the current restart cycle
  // block GMRES (one Krylov space for all right-hand sides of the system; see kb_bsmall)
  int t40, nt4;  // tiles of BT rows of the system in the current map
  double *Rp;    // packed upper triangle of the rotated block Hessenberg: column q at q (q + 1) / 2
  double *hv;    // Householder reflector of column q: tau, v_1 .. v_L at 8 q
  double *gB;    // rotated right-hand sides: L (m + 1) rows x NB
  double *C1;    // first-sweep coefficients: 8 x 8 per basis block ([l][k]); the solution blocks Y at the end
  double *C2;    // second-sweep coefficients, then (block j + 1) the Gram matrix of the swept vector
  double *Ri;    // 8 x 8: inverse of the new subdiagonal block (normalises the next basis block)
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


typedef double d4v __attribute__((ext_vector_type(4)));
constexpr int MAXW = ANDR_OP_MAXW;

// a column of the system still iterates in the current block cycle (the sweeps skip a system whose
// columns have all converged: its later rounds are never used). Slot 6 of column 0's state holds it:
// set by kb_binit, cleared by kb_bsmall in the round that freezes every live column at once.
__device__ inline bool sys_iterating(const GSys &S) { return S.small[6] > 0.0; }

// ---- structured Stokes operator --------------------------------------------------------------
// per tile of 32 elements (256 threads = 32 elements x 8 right-hand sides): z_e = M^-1_e v_e for the
// three components of element e, v = Q_j (mode 0, z to ZT or ZT32) or v = X (mode 1, z to ZT in
// FP64); partial sums of n_e . z_e per wall (1..MAXW) and rhs for the null-space term.
__global__ __launch_bounds__(256) void kb_prep(const GSys *__restrict__ sys, const TileMap *__restrict__ emap, int mode,
                                               int jq, double *__restrict__ part_s, int all) {
  __shared__ double red[4][MAXW][NB];
  const TileMap tm = emap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  // block GMRES: a converged system's later products are never used (the product skips its matrix)
  if (!all && !sys_iterating(S)) return;  // uniform over the block
  const StokesOp &O = S.op;
  const int ne = O.ne, n = S.n;
  const int k = threadIdx.x & 7, el = threadIdx.x >> 3;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int e = tm.row0 + el;
  const double *v = mode == 0 ? S.Q + (size_t)jq * n * NB : S.X;
  double dz = 0.0;
  int w = 0;
  if (e < ne) {
    const double q0 = v[(size_t)(0 * ne + e) * NB + k], q1 = v[(size_t)(1 * ne + e) * NB + k],
                 q2 = v[(size_t)(2 * ne + e) * NB + k];
    const double *mi = O.minv + 9 * (size_t)e;
    const double z[3] = {mi[0] * q0 + mi[1] * q1 + mi[2] * q2, mi[3] * q0 + mi[4] * q1 + mi[5] * q2,
                         mi[6] * q0 + mi[7] * q1 + mi[8] * q2};
    if (mode == 0 && S.ZT32) {
#pragma unroll
      for (int c = 0; c < 3; ++c) S.ZT32[((size_t)c * NB + k) * S.nep + e] = (float)z[c];
    } else if (mode == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) S.ZT[((size_t)c * NB + k) * S.nep + e] = z[c];
    } else {  // the residual's operand, FP64 in the product's layout
#pragma unroll
      for (int c = 0; c < 3; ++c) S.ZT[((size_t)c * NB + k) * S.nep + e] = z[c];
    }
    w = O.bcid[e];
    dz = O.nrm[3 * e] * z[0] + O.nrm[3 * e + 1] * z[1] + O.nrm[3 * e + 2] * z[2];
  } else if (e < S.nep) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (mode == 0 && S.ZT32) S.ZT32[((size_t)c * NB + k) * S.nep + e] = 0.f;
      else S.ZT[((size_t)c * NB + k) * S.nep + e] = 0.0;
    }
  }
#pragma unroll
  for (int ww = 0; ww < MAXW; ++ww) {
    double a = w == ww + 1 ? dz : 0.0;
    a += __shfl_xor(a, 8, 64);
    a += __shfl_xor(a, 16, 64);
    a += __shfl_xor(a, 32, 64);
    if (lane < 8) red[wid][ww][lane] = a;
  }
  __syncthreads();
  if (threadIdx.x < MAXW * NB) {
    const int ww = threadIdx.x / NB, kk = threadIdx.x % NB;
    part_s[((size_t)blockIdx.x * MAXW + ww) * NB + kk] = red[0][ww][kk] + red[1][ww][kk] + red[2][ww][kk] + red[3][ww][kk];
  }
}

// block per system, thread per (wall, rhs) and eighth of the element tiles: sdef = sum of the
// system's tile partials
__global__ __launch_bounds__(256) void kb_sred(const GSys *__restrict__ sys, const double *__restrict__ part_s,
                                               int all) {
  static_assert(MAXW * NB == 32, "eight partial sums per (wall, rhs)");
  __shared__ double red[8][32];
  const GSys &S = sys[blockIdx.x];
  if (!all && !sys_iterating(S)) return;  // as kb_prep (uniform over the block)
  const int r = threadIdx.x & 31, p = threadIdx.x >> 5;
  double a = 0.0;
  for (int i = p; i < S.net; i += 8) a += part_s[(size_t)(S.et0 + i) * MAXW * NB + r];
  red[p][r] = a;
  __syncthreads();
  if (threadIdx.x < 32) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) v += red[q][threadIdx.x];
    S.sdef[threadIdx.x] = v;
  }
}

// null-space term of row (element r, component c) for rhs k
__device__ inline double defl_term(const StokesOp &O, const double *sdef, int r, int c, int k) {
  if (!O.defl) return 0.0;
  const int w = O.bcid[r];
  if (w < 1 || w > MAXW) return 0.0;
  const double d = O.defl[w];
  return d != 0.0 ? d * O.area[r] * O.nrm[3 * r + c] * sdef[(w - 1) * NB + k] : 0.0;
}

// The Krylov product W = A M^-1 Q_j on the matrix cores: y_c = sum_en S_{c,en} z_en (+ the
// null-space term), S the six symmetric FP32 blocks, z = M^-1 q_j.
//  * F32 = false: S converted to FP64 in registers (exactly), z in FP64, products and sums in FP64
//    (v_mfma_f64_16x16x4f64), so the Krylov relation holds to FP64 rounding;
//  * F32 = true: z rounded to FP32, products on the FP32 matrix cores (v_mfma_f32_16x16x4f32, twice
//    the FP64 rate) summed in FP32 over one 16-column block and the block sums accumulated in FP64:
//    the product carries a relative error of ~1e-7, which GMRES tolerates at relative residuals of
//    1e-6 and above (the product then no longer bounds the kernel, the matrix stream does).
//
// The B operand has 16 columns for 8 right-hand sides: each block product fills both halves.
// With h = column half, three accumulators per 16-row band take the six blocks:
//   P: xx [z_x | 0] + xy [z_y | z_x] + yy [0 | z_y]  ->  [y_x part | y_y part]
//   Q: xz [z_x | z_z] + zz [z_z | 0]                  ->  [y_z part | y_x part]
//   R: yz [z_z | z_y]                                 ->  [y_y part | y_z part]
// so y_x = P_lo + Q_hi, y_y = R_lo + P_hi, y_z = Q_lo + R_hi: six block products per column step,
// each matrix block read once. A wave owns 32 element rows (two bands) over a column segment; lane
// (g = lane / 16, c = lane % 16) feeds row 16 T + c with columns 16 kb + 4 g + t of every block (one
// float4 per block and band) and column c of B with the same columns. The FP64 form returns
// D[g + 4 v][c] in register v of lane (g, c), the FP32 form D[4 g + v][c].
template <bool F32>
__global__ __launch_bounds__(256, 2) void kb_gemv_sym(const GSys *__restrict__ sys, const MWave *__restrict__ wmap,
                                                   int nwaves, int all) {
  using ZT_t = typename std::conditional<F32, float, double>::type;  // z operand
  typedef float f4n __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= nwaves) return;
  const MWave wm = wmap[wave];
  const GSys &S = sys[wm.sys];
  const StokesOp &O = S.op;
  const int ne = O.ne, row0 = wm.row0, kend = min(wm.k1, ne), KBe = (ne + 15) / 16;
  const int g = lane >> 4, c = lane & 15;
  const bool hi = c >= NB;
  // A system whose every column has converged in this cycle (its step count fixed by the Arnoldi
  // step) needs no more products until the host drops it from the map every few steps: its waves
  // skip the matrix and write the null-space term alone (the later steps of such a column are
  // never used, and stay finite).
  // (all: every system's product, as the residual of the restart test needs)
  bool need = all != 0;
#pragma unroll
  for (int k = 0; k < NB; ++k) need |= S.small[8 * k + 4] > 0.0 && S.small[8 * k + 5] < 0.0;
  const float *ab[2];
#pragma unroll
  for (int T = 0; T < 2; ++T) ab[T] = O.S + (size_t)((row0 >> 4) + T) * KBe * 1536 + 16 * c + 4 * g;
  const ZT_t *zt;
  if constexpr (F32) zt = S.ZT32;
  else zt = S.ZT;
  const ZT_t *zb[3];
#pragma unroll
  for (int en = 0; en < 3; ++en) zb[en] = zt + ((size_t)en * NB + (c & (NB - 1))) * S.nep + 4 * g;
  d4v acc[2][3];
  f4n accf[2][3];  // F32: the current block's sums
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      acc[T][a] = d4v{0.0, 0.0, 0.0, 0.0};
      accf[T][a] = f4n{0.f, 0.f, 0.f, 0.f};
    }
  // one column step t of a 16-column block: operands a[T][q][t], z[en][t]
  auto step = [&](const float (&a)[2][6][4], const ZT_t (&z)[3][4], int t) {
    const ZT_t z0 = z[0][t], z1 = z[1][t], z2 = z[2][t];
    const ZT_t b0 = hi ? (ZT_t)0 : z0, b1 = hi ? z0 : z1, b3 = hi ? z1 : (ZT_t)0;
    const ZT_t b2 = hi ? z2 : z0, b5 = hi ? (ZT_t)0 : z2, b4 = hi ? z1 : z2;
    // bands interleaved so that consecutive instructions never share an accumulator
    if constexpr (F32) {
#pragma unroll
      for (int T = 0; T < 2; ++T) accf[T][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][0][t], b0, accf[T][0], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) accf[T][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][2][t], b2, accf[T][1], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) accf[T][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][4][t], b4, accf[T][2], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) accf[T][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][1][t], b1, accf[T][0], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) accf[T][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][5][t], b5, accf[T][1], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) accf[T][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[T][3][t], b3, accf[T][0], 0, 0, 0);
    } else {
#pragma unroll
      for (int T = 0; T < 2; ++T) acc[T][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[T][0][t], b0, acc[T][0], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) acc[T][1] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[T][2][t], b2, acc[T][1], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) acc[T][2] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[T][4][t], b4, acc[T][2], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) acc[T][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[T][1][t], b1, acc[T][0], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) acc[T][1] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[T][5][t], b5, acc[T][1], 0, 0, 0);
#pragma unroll
      for (int T = 0; T < 2; ++T) acc[T][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[T][3][t], b3, acc[T][0], 0, 0, 0);
    }
  };
  // F32: a block's FP32 sums into the FP64 totals
  auto flush = [&]() {
    if constexpr (F32) {
#pragma unroll
      for (int T = 0; T < 2; ++T)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[T][a][v] += (double)accf[T][a][v];
          accf[T][a] = f4n{0.f, 0.f, 0.f, 0.f};
        }
    }
  };
  // global-address-space views of the operands (plain global loads instead of flat ones)
  typedef const float __attribute__((address_space(1))) *gfp;
  typedef const ZT_t __attribute__((address_space(1))) *gzp;
  typedef double d2n __attribute__((ext_vector_type(2)));
  const gfp ga[2] = {(gfp)ab[0], (gfp)ab[1]};
  const gzp gz[3] = {(gzp)zb[0], (gzp)zb[1], (gzp)zb[2]};
  auto load = [&](int k, float (&a)[2][6][4], ZT_t (&z)[3][4]) {
#pragma unroll
    for (int T = 0; T < 2; ++T)
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        // streamed once: non-temporal, so that the re-read operand z keeps its place in L2
        const f4n v = __builtin_nontemporal_load((const f4n __attribute__((address_space(1))) *)(ga[T] + (size_t)k * 1536 + q * 256));
        a[T][q][0] = v.x;
        a[T][q][1] = v.y;
        a[T][q][2] = v.z;
        a[T][q][3] = v.w;
      }
#pragma unroll
    for (int en = 0; en < 3; ++en) {
      if constexpr (F32) {
        const f4n u = *(const f4n __attribute__((address_space(1))) *)(gz[en] + 16 * k);
        z[en][0] = u.x;
        z[en][1] = u.y;
        z[en][2] = u.z;
        z[en][3] = u.w;
      } else {
        const d2n u = *(const d2n __attribute__((address_space(1))) *)(gz[en] + 16 * k);
        const d2n w = *(const d2n __attribute__((address_space(1))) *)(gz[en] + 16 * k + 2);
        z[en][0] = u.x;
        z[en][1] = u.y;
        z[en][2] = w.x;
        z[en][3] = w.y;
      }
    }
  };
  // full column blocks kb < kb_end, two register sets in turn: the operands of the next block are
  // requested before the products of the current one. The prefetch is unconditional (the last
  // one re-reads the final block) so that no branch sits between a load and its use.
  int kb = wm.k0 >> 4;
  const int kb_end = need ? kend >> 4 : kb;
  if (kb < kb_end) {
    float a0[2][6][4], a1[2][6][4];
    ZT_t z0[3][4], z1[3][4];
    load(kb, a0, z0);
    for (; kb + 1 < kb_end; kb += 2) {
      // the scheduling barriers keep each group of loads ahead of the products it overlaps (left
      // to itself the scheduler moves loads next to their uses and the prefetch is lost)
      load(kb + 1, a1, z1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < 4; ++t) step(a0, z0, t);
      flush();
      __builtin_amdgcn_sched_barrier(0);
      load(min(kb + 2, kb_end - 1), a0, z0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < 4; ++t) step(a1, z1, t);
      flush();
      __builtin_amdgcn_sched_barrier(0);
    }
    if (kb < kb_end) {  // odd count: a0 holds block kb
#pragma unroll
      for (int t = 0; t < 4; ++t) step(a0, z0, t);
      flush();
      ++kb;
    }
  }
  if (need && 16 * kb < kend) {  // last, partial column block: masked (padding columns are not written)
    float a[2][6][4];
    ZT_t z[3][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool in = 16 * kb + 4 * g + t < kend;
#pragma unroll
      for (int en = 0; en < 3; ++en) z[en][t] = in ? zb[en][16 * kb + t] : (ZT_t)0;
#pragma unroll
      for (int T = 0; T < 2; ++T)
#pragma unroll
        for (int q = 0; q < 6; ++q) a[T][q][t] = in ? ab[T][(size_t)kb * 1536 + q * 256 + t] : 0.f;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) step(a, z, t);
    flush();
  }
  const bool direct = S.nseg <= 1;
  double *Y = direct ? S.W : S.P + (size_t)wm.seg * S.n * NB;
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double p = acc[T][0][v], q = acc[T][1][v], rr = acc[T][2][v];
      const double ph = __shfl_xor(p, NB, 64), qh = __shfl_xor(q, NB, 64), rh = __shfl_xor(rr, NB, 64);
      const int r = row0 + 16 * T + (F32 ? 4 * g + v : g + 4 * v);
      if (hi || r >= ne) continue;
      const double y[3] = {p + qh, rr + ph, q + rh};
#pragma unroll
      for (int co = 0; co < 3; ++co)
        Y[(size_t)(co * ne + r) * NB + c] = y[co] + (wm.seg == 0 ? defl_term(O, S.sdef, r, co, c) : 0.0);
    }
}

// R = B - A M^-1 X, the residual of the refinement loop: W = A M^-1 X (+ the null-space term) comes
// from the product on the FP64 matrix cores (kb_gemv_sym<false> on z = M^-1 X in FP64, every
// system); R holds B. Block of 4 waves per 64 element rows of one system (one 16-row band per wave);
// lane (g, c) takes row c of the band, component g (g < 3), all right-hand sides.
// part[4 block + wave][k] = sum over the band's rows of R^2 (rows of all three components).
__global__ __launch_bounds__(256) void kb_resid_fin(const GSys *__restrict__ sys, const WaveMap *__restrict__ bmap,
                                                    double *__restrict__ part) {
  const WaveMap bm = bmap[blockIdx.x];
  const GSys &S = sys[bm.sys];
  const int ne = S.op.ne;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int band = (bm.row0 >> 4) + wid;
  const int r = 16 * band + c;
  double r2[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) r2[k] = 0.0;
  if (g < 3 && r < ne) {
    const size_t o = (size_t)(g * ne + r) * NB;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const double v = S.R[o + k] - S.W[o + k];  // R holds B
      S.R[o + k] = v;
      r2[k] = v * v;
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k)
    for (int off = 1; off < 64; off <<= 1) r2[k] += __shfl_xor(r2[k], off, 64);
  if (lane < NB) {
    double v = r2[0];
#pragma unroll
    for (int k = 1; k < NB; ++k) v = lane == k ? r2[k] : v;
    part[((size_t)blockIdx.x * 4 + wid) * NB + lane] = v;
  }
}

// The residual of the first cycle, x = 0: R = B (already copied) and the same partial sums of R^2
// as kb_resid_fin, over the same bands, without the product (A M^-1 0 contributes exact zeros).
__global__ __launch_bounds__(256) void kb_resid_zero(const GSys *__restrict__ sys, const WaveMap *__restrict__ bmap,
                                                     double *__restrict__ part) {
  const WaveMap bm = bmap[blockIdx.x];
  const GSys &S = sys[bm.sys];
  const int ne = S.op.ne;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int band = (bm.row0 >> 4) + wid;
  const int r = 16 * band + c;
  double r2[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) r2[k] = 0.0;
  if (16 * band < ne && g == 0 && r < ne) {
#pragma unroll
    for (int co = 0; co < 3; ++co)
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        const double v = S.R[(size_t)(co * ne + r) * NB + k];
        r2[k] += v * v;
      }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k)
    for (int off = 1; off < 64; off <<= 1) r2[k] += __shfl_xor(r2[k], off, 64);
  if (lane < NB) {
    double v = r2[0];
#pragma unroll
    for (int k = 1; k < NB; ++k) v = lane == k ? r2[k] : v;
    part[((size_t)blockIdx.x * 4 + wid) * NB + lane] = v;
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

// ---- Arnoldi step with delayed reorthogonalisation (classical Gram-Schmidt twice, one sweep of
// dot products and one of updates per step). At step j the basis slot Q_j holds the candidate u_j
// (last step's vector after one projection, normalised by an estimate of its norm) and W = A u_j.
// One sweep over the basis gives s = V^T u_j, t = V^T W and u.u, u.W, W.W; then
//   rho = |u_j - V s|,  v_j = (u_j - V s) / rho   (the second projection of the previous vector),
//   column j-1 of H gains eta_{j-1} s and its subdiagonal becomes eta_{j-1} rho (it is final now),
//   A v_j = (W - V_{j+1} H s) / rho, so column j is h = (t - (H s)) / rho with h_j from
//   delta = v_j . W = (u.W - s.t) / rho, and the next candidate is
//   u_{j+1} = (W - V t - delta v_j) / (rho eta_j),  eta_j^2 = (W.W - t.t - delta^2) / rho^2.
// The second sweep writes v_j and u_{j+1}. Per right-hand side, the coefficient array cw holds
// s (slots 0..m), t (m+1..2m+1), and 1/rho, delta, 1/eta (2m+2..2m+4); the dot sweep leaves
// u.u, u.W, W.W in those three slots.
__host__ __device__ constexpr int cw_stride(int m) { return 2 * (m + 1) + 6; }
__host__ __device__ constexpr int part_stride(int m) { return 2 * (m + 1) + 3; }

// per (tile, chunk of ICH basis vectors): partial s_i = Q_i . u and t_i = Q_i . W over the tile's
// rows, part[(tile * PS + slot) * NB + k]; the first chunk also adds u.u, u.W, W.W. fin: W = 0
// (closing a cycle: only the reorthogonalisation of the last candidate). Wave w of the block takes
// the basis vectors i0 + w, i0 + w + 4, ... over all 64 rows of the tile; a lane holds two
// right-hand sides of one row (16-byte loads, 16 rows per load), so that each dot product is
// reduced across lanes once per tile.
__global__ __launch_bounds__(256) void kb_adots(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int j,
                                                int m, double *__restrict__ part, int fin) {
  constexpr int PW = ICH / 4;  // basis vectors per wave
  constexpr int RT = TILE / 16;  // rows per lane
  typedef double d2 __attribute__((ext_vector_type(2)));
  const int i0 = blockIdx.y * ICH;
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n, PS = part_stride(m);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int kp = lane & 3, r16 = lane >> 2;  // right-hand sides 2 kp, 2 kp + 1; row within 16
  const d2 *U = (const d2 *)(S.Q + (size_t)j * n * NB) + kp;
  const d2 *Wv = (const d2 *)S.W + kp;
  d2 u[RT], w[RT];
  int rows[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int row = tm.row0 + 16 * t + r16;
    rows[t] = row < n ? row : -1;
    u[t] = row < n ? U[(size_t)row * (NB / 2)] : d2{0.0, 0.0};
    w[t] = (row < n && !fin) ? Wv[(size_t)row * (NB / 2)] : d2{0.0, 0.0};
  }
  d2 as[PW], at[PW];
#pragma unroll
  for (int p = 0; p < PW; ++p) {
    const int i = i0 + wid + 4 * p;
    d2 a = {0.0, 0.0}, b = {0.0, 0.0};
    if (i < j) {
      const d2 *Qi = (const d2 *)(S.Q + (size_t)i * n * NB) + kp;
#pragma unroll
      for (int t = 0; t < RT; ++t)
        if (rows[t] >= 0) {
          const d2 q = Qi[(size_t)rows[t] * (NB / 2)];
          a += q * u[t];
          b += q * w[t];
        }
    }
    as[p] = a;
    at[p] = b;
  }
  auto rsum = [&](double x) {  // over the 16 row groups of the wave
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    x += __shfl_xor(x, 16, 64);
    x += __shfl_xor(x, 32, 64);
    return x;
  };
#pragma unroll
  for (int p = 0; p < PW; ++p) {
    const double a0 = rsum(as[p].x), a1 = rsum(as[p].y), b0 = rsum(at[p].x), b1 = rsum(at[p].y);
    const int i = i0 + wid + 4 * p;
    if (lane < 4 && i < j) {
      double *ps = part + ((size_t)blockIdx.x * PS + i) * NB + 2 * kp;
      double *pt = part + ((size_t)blockIdx.x * PS + (m + 1) + i) * NB + 2 * kp;
      ps[0] = a0;
      ps[1] = a1;
      pt[0] = b0;
      pt[1] = b1;
    }
  }
  if (blockIdx.y == 0 && wid == 0) {
    d2 sc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      sc[0] += u[t] * u[t];
      sc[1] += u[t] * w[t];
      sc[2] += w[t] * w[t];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double v0 = rsum(sc[c].x), v1 = rsum(sc[c].y);
      if (lane < 4) {
        double *pc = part + ((size_t)blockIdx.x * PS + 2 * (m + 1) + c) * NB + 2 * kp;
        pc[0] = v0;
        pc[1] = v1;
      }
    }
  }
}

// block per (system, 8 of the 2j + 3 values): sums over the system's tiles (the four waves take
// every fourth tile) into cw. Systems outside the current map have no tiles and return.
__global__ __launch_bounds__(256) void kb_adots_final(const GSys *__restrict__ sys, int j, int m,
                                                      const double *__restrict__ part) {
  __shared__ double red[4][64];
  const GSys &S = sys[blockIdx.x];
  if (S.ntiles == 0) return;  // uniform over the block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = lane & 7, q = blockIdx.y * 8 + (lane >> 3);
  const int nq = 2 * j + 3, PS = part_stride(m);
  // value q: s_q (q < j), t_{q-j} (q < 2j), then u.u, u.W, W.W
  const int slot = q < j ? q : (q < 2 * j ? (m + 1) + (q - j) : 2 * (m + 1) + (q - 2 * j));
  double c = 0.0;
  if (q < nq)
    for (int tl = wv; tl < S.ntiles; tl += 4) c += part[((size_t)(S.tile0 + tl) * PS + slot) * NB + k];
  red[wv][lane] = c;
  __syncthreads();
  if (wv == 0 && q < nq) S.cw[k * cw_stride(m) + slot] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// wave per (system, rhs) of the current map: the scalar part of step j (see above). Finalises
// column j - 1 of the Hessenberg matrix (unrotated in Hu), rotates it into H with the earlier Givens
// rotations, updates g and the residual estimate (convergence of j steps), and, unless fin, forms
// column j and the coefficients of the update sweep. Dynamic LDS: 5 m + 2 doubles.
__global__ __launch_bounds__(64) void kb_arnoldi(const GSys *__restrict__ sys, int j, int m, double tol, int fin) {
  extern __shared__ double lds[];
  const int s = blockIdx.x / NB, k = blockIdx.x % NB, lane = threadIdx.x;
  const GSys &S = sys[s];
  if (S.ntiles == 0) return;  // uniform over the block
  const int CW = cw_stride(m);
  double *cw = S.cw + (size_t)k * CW;
  double *sm = S.small + k * 8;
  double *Hu = S.Hu + (size_t)k * (m + 1) * m, *Hr = S.H + (size_t)k * (m + 1) * m;
  double *cs = S.cs + k * m, *sn = S.sn + k * m, *g = S.g + k * (m + 1);
  double *ls = lds, *lt = lds + m, *lh = lds + 2 * m, *lc = lds + 3 * m + 2;  // s, t, column, rotations (cs, sn)
  const double alpha = cw[2 * m + 2], beta = cw[2 * m + 3], gamma = cw[2 * m + 4];
  double ss = 0.0, st = 0.0, tt = 0.0;
  for (int i = lane; i < j; i += 64) {
    const double si = cw[i], ti = cw[(m + 1) + i];
    ls[i] = si;
    lt[i] = ti;
    ss += si * si;
    st += si * ti;
    tt += ti * ti;
  }
  for (int o = 32; o > 0; o >>= 1) {
    ss += __shfl_xor(ss, o, 64);
    st += __shfl_xor(st, o, 64);
    tt += __shfl_xor(tt, o, 64);
  }
  const double rho2 = alpha - ss;
  const bool ok = sm[4] > 0.0 && alpha > 0.0 && rho2 > 1e-24 * alpha;
  const double rho = ok ? sqrt(rho2) : 0.0;
  __syncthreads();
  if (j >= 1) {
    // column j - 1 becomes final
    const double eta = Hu[(size_t)(j - 1) * (m + 1) + j];
    double *hc = Hu + (size_t)(j - 1) * (m + 1);
    for (int i = lane; i < j; i += 64) {
      const double v = hc[i] + eta * ls[i];
      hc[i] = v;
      lh[i] = v;
    }
    if (lane == 0) {
      hc[j] = eta * rho;
      lh[j] = eta * rho;
    }
    for (int i = lane; i < j - 1; i += 64) lc[i] = cs[i];
    for (int i = lane; i < j - 1; i += 64) lc[m + i] = sn[i];
    __syncthreads();
    if (lane == 0) {
      const bool live = sm[4] > 0.0 && sm[5] < 0.0;  // active and not yet frozen in this cycle
      double cur = lh[0];
      for (int i = 0; i < j - 1; ++i) {
        const double b = lh[i + 1], c = lc[i], sg = lc[m + i];
        lh[i] = c * cur + sg * b;
        cur = -sg * cur + c * b;
      }
      const double a = cur, b = lh[j];
      const double rr = sqrt(a * a + b * b);
      const double c = rr > 0.0 ? a / rr : 1.0, sg = rr > 0.0 ? b / rr : 0.0;
      cs[j - 1] = c;
      sn[j - 1] = sg;
      lh[j - 1] = rr;
      g[j] = -sg * g[j - 1];
      g[j - 1] = c * g[j - 1];
      const double est = fabs(g[j]);
      sm[2] = est;
      if (live && (est <= EST_MARGIN * tol * sm[0] || fabs(b) <= 1e-14 * rr)) sm[5] = (double)j;  // steps to use
    }
    __syncthreads();
    double *hr = Hr + (size_t)(j - 1) * (m + 1);
    for (int i = lane; i < j; i += 64) hr[i] = lh[i];
    if (lane == 0) hr[j] = 0.0;
  }
  if (fin) return;
  // column j: h_i = (t_i - (H s)_i) / rho, i < j;  h_j = (delta - (H s)_j) / rho
  const double delta = ok ? (beta - st) / rho : 0.0;
  double *hj = Hu + (size_t)j * (m + 1);
  for (int i = lane; i <= j; i += 64) {
    double hs = 0.0;
    for (int c = (i > 0 ? i - 1 : 0); c < j; ++c) hs += Hu[(size_t)c * (m + 1) + i] * ls[c];
    const double num = (i < j ? lt[i] : delta) - hs;
    hj[i] = ok ? num / rho : 0.0;
  }
  if (lane == 0) {
    double eta2 = ok ? (gamma - tt - delta * delta) / rho2 : 0.0;
    const double floor2 = ok ? 1e-28 * gamma / rho2 : 0.0;
    const double eta = eta2 > floor2 ? sqrt(eta2) : 0.0;
    hj[j + 1] = eta;
    cw[2 * m + 2] = ok ? 1.0 / rho : 0.0;
    cw[2 * m + 3] = delta;
    cw[2 * m + 4] = eta > 0.0 ? 1.0 / eta : 0.0;
  }
}

// per tile: v_j = (u_j - V s) / rho into Q_j, u_{j+1} = (W - V t - delta v_j) / (rho eta) into Q_{j+1}.
// The basis vectors are taken in chunks of CCH: the block stages the chunk's step coefficients s, t of
// its system in LDS (lc[i NB + k] = s_(c0+i), lc[(cn + i) NB + k] = t_(c0+i)). A thread holds two
// right-hand sides of one row (16-byte loads); eight basis vectors per iteration are in flight.
__global__ __launch_bounds__(256) void kb_aupdate(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int j,
                                                  int m) {
  constexpr int CCH = 128;
  typedef double d2 __attribute__((ext_vector_type(2)));
  __shared__ double lc[2 * CCH * NB];
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n;
  const int kp = threadIdx.x & 3, rr = threadIdx.x >> 2;
  static_assert(TILE == 64, "one row per thread");
  const int r0 = tm.row0 + rr;
  const bool v0 = r0 < n;
  const int q0 = v0 ? r0 : tm.row0;  // an in-range row for the loads
  const double *cwa = S.cw + (size_t)(2 * kp) * cw_stride(m), *cwb = cwa + cw_stride(m);
  const d2 irho = {cwa[2 * m + 2], cwb[2 * m + 2]}, delta = {cwa[2 * m + 3], cwb[2 * m + 3]},
           ieta = {cwa[2 * m + 4], cwb[2 * m + 4]};
  const size_t vs = (size_t)n * (NB / 2);  // vector stride in d2
  const d2 *Qa = (const d2 *)S.Q + (size_t)q0 * (NB / 2) + kp;
  d2 sa0 = {0.0, 0.0}, sa1 = {0.0, 0.0}, ta0 = {0.0, 0.0}, ta1 = {0.0, 0.0};
  for (int c0 = 0; c0 < j; c0 += CCH) {  // j is uniform over the block: every thread meets the barriers
    const int cn = min(CCH, j - c0);
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * cn * NB; q += 256) {
      const int kk = q % NB, ii = q / NB;  // ii < cn: s_(c0+ii), else t_(c0+ii-cn)
      lc[q] = S.cw[(size_t)kk * cw_stride(m) + (ii < cn ? c0 + ii : (m + 1) + c0 + (ii - cn))];
    }
    __syncthreads();
    const d2 *cs = (const d2 *)lc + kp, *ct = (const d2 *)(lc + (size_t)cn * NB) + kp;
    const d2 *Pa = Qa + (size_t)c0 * vs;
    int i = 0;
    for (; i + 8 <= cn; i += 8) {
      d2 qa[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) qa[u] = Pa[(size_t)(i + u) * vs];
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        sa0 += qa[u] * cs[(i + u) * (NB / 2)];
        sa1 += qa[u + 1] * cs[(i + u + 1) * (NB / 2)];
        ta0 += qa[u] * ct[(i + u) * (NB / 2)];
        ta1 += qa[u + 1] * ct[(i + u + 1) * (NB / 2)];
      }
    }
    for (; i < cn; ++i) {
      const d2 qa = Pa[(size_t)i * vs];
      sa0 += qa * cs[i * (NB / 2)];
      ta0 += qa * ct[i * (NB / 2)];
    }
  }
  d2 *U = (d2 *)(S.Q + (size_t)j * n * NB), *Un = (d2 *)(S.Q + (size_t)(j + 1) * n * NB);
  if (v0) {
    const size_t o = (size_t)r0 * (NB / 2) + kp;
    const d2 v = (U[o] - (sa0 + sa1)) * irho;
    U[o] = v;
    Un[o] = (((const d2 *)S.W)[o] - (ta0 + ta1) - delta * v) * irho * ieta;
  }
}

// wave per (system, rhs): back substitution y = R^{-1} g over the steps used (y stored in cw),
// column by column: y_i = g_i / R_ii, then the lanes update g_l -= R_li y_i for l < i (g in LDS)
__global__ __launch_bounds__(64) void kb_solve_y(const GSys *__restrict__ sys, int steps, int m) {
  extern __shared__ double lg[];
  const int s = blockIdx.x / NB, k = blockIdx.x % NB, lane = threadIdx.x;
  const GSys &S = sys[s];
  if (!S.incyc) return;  // uniform over the block
  const double *sm = S.small + k * 8;
  double *y = S.cw + (size_t)k * cw_stride(m);
  int kk = sm[5] > 0.0 ? (int)sm[5] : steps;
  if (sm[4] <= 0.0) kk = 0;
  const double *H = S.H + (size_t)k * (m + 1) * m;
  const double *g = S.g + k * (m + 1);
  for (int i = lane; i <= m; i += 64) y[i] = 0.0;
  for (int i = lane; i < kk; i += 64) lg[i] = g[i];
  __syncthreads();
  for (int i = kk - 1; i >= 0; --i) {
    const double d = H[(size_t)i * (m + 1) + i];
    const double yi = d != 0.0 ? lg[i] / d : 0.0;
    for (int l = lane; l < i; l += 64) lg[l] -= H[(size_t)i * (m + 1) + l] * yi;
    if (lane == 0) y[i] = yi;
    __syncthreads();
  }
}

// per tile: X += sum_j Q_j y_j over the steps each column used (y_j = 0 beyond them, as in kb_solve_y)
__global__ __launch_bounds__(256) void kb_xupdate(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int steps,
                                                  int m) {
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  const int n = S.n;
  const int k = threadIdx.x & 7, rr = threadIdx.x >> 3;
  const double *y = S.cw + (size_t)k * cw_stride(m);
  const double *sm = S.small + k * 8;
  int kk = sm[5] > 0.0 ? min((int)sm[5], steps) : steps;
  if (sm[4] <= 0.0) kk = 0;
#pragma unroll
  for (int t = 0; t < TILE / 32; ++t) {
    const int row = tm.row0 + rr + 32 * t;
    if (row >= n) continue;
    double x = S.X[(size_t)row * NB + k];
    for (int i = 0; i < kk; ++i) x += S.Q[((size_t)i * n + row) * NB + k] * y[i];
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
__global__ void kb_restart(const GSys *__restrict__ sys, int nsys, const double *__restrict__ part,
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
  sm[3] = (sm[4] > 0.0 && rn > 0.0) ? 1.0 / rn : 0.0;
  if (sm[4] > 0.0) atomicAdd(done, 1);
}

// thread per (system, k) of the cycle: g = (||r||, 0, ..., 0) in the cycle's layout
__global__ void kb_ginit(const GSys *__restrict__ sys, int nsys, int m) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys * NB) return;
  const int s = t / NB, k = t % NB;
  const GSys &S = sys[s];
  if (S.ntiles == 0) return;
  const double *sm = S.small + k * 8;
  double *g = S.g + k * (m + 1);
  for (int i = 0; i <= m; ++i) g[i] = 0.0;
  g[0] = sm[4] > 0.0 ? sm[1] : 0.0;
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

// ---- Block GMRES ------------------------------------------------------------------------------

// One Krylov space for the L right-hand sides of a system: round j multiplies the basis block Q_j
// (n x L, orthonormal columns; a column that takes no part is zero) and orthogonalises W = A M^-1 Q_j
// against Q_0 .. Q_j by block classical Gram-Schmidt applied twice:
//   C1 = Q^T W, W1 = W - Q C1;  C2 = Q^T W1, G1 = W1^T W1,
//   W2 = W1 - Q C2 = Q_(j+1) R_(j+1) with R_(j+1) the Cholesky factor of W2^T W2 = G1 - C2^T C2,
// so that A M^-1 Q_j = Q C + Q_(j+1) R_(j+1) (C = C1 + C2): column block j of the block Hessenberg
// matrix. Its least-squares problem min |E - H Y| (E = the Cholesky factor of the initial residual
// block in the first L rows) is reduced by one Householder reflector per column acting on L + 1
// rows (the Hessenberg matrix has L subdiagonals); the rows L (j + 1) .. L (j + 2) - 1 of the
// rotated right-hand sides are the residual of each column. A column whose product has no new
// direction left (pivot below 1e-20 of its squared norm) ends the cycle: the restart test then takes
// the true residual. Dead columns (converged, or beyond nrhs) stay exact zeros throughout.
//
// Dot sweep: per (tile of BT rows, 16 blocks) partials of C_b = Q_b^T V, b < nb, and with gram V^T V
// as block nb. The tile's rows of V are staged in LDS (zero past the end of the system). Wave w
// takes blocks 16 y + 4 w .. + 3 as two FP64 matrix-core A operands of two blocks each (lane (g, i):
// A[i][g] = Q_(b(i))[row + g][i & 7], B[g][i] = V[row + g][i & 7]); register v of lane (g, i < 8)
// ends as C_b[l][i] with b, l from row g + 4 v of the product. Rows past the end meet zeros in V,
// so the A operands load a valid row unconditionally.
__global__ __launch_bounds__(256) void kb_bdots(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap,
                                                int which, int nb, int gram, double *__restrict__ part, int pstride,
                                                int all) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  __shared__ double vs[BT * NB];
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  if (!all && !sys_iterating(S)) return;  // uniform over the block
  const int n = S.n, nbt = nb + (gram ? 1 : 0);
  const double *V = which ? S.R : S.W;
  {
    const int row = tm.row0 + threadIdx.x;
    d2 *dst = (d2 *)(vs + threadIdx.x * NB);
    const d2 *src = (const d2 *)(V + (size_t)min(row, n - 1) * NB);
#pragma unroll
    for (int p = 0; p < NB / 2; ++p) dst[p] = row < n ? src[p] : d2{0.0, 0.0};
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g = lane >> 4, i = lane & 15, l = i & 7;
  const int bw = blockIdx.y * 16 + 4 * wid;  // the wave's first block
  if (bw >= nbt) return;                      // after the only barrier
  const double *Ap[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int b = bw + 2 * h + (i >> 3);
    const double *Ab = b < nb ? S.Q + (size_t)b * n * NB : (b == nb && gram ? V : nullptr);
    Ap[h] = Ab ? Ab : V;  // lanes without a block compute partials that are not written
  }
  const int rlast = n - 1;
  d4v acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  for (int r = 0; r < BT; r += 64) {
    double a[2][16], bv[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const size_t rr = (size_t)min(tm.row0 + r + 4 * t + g, rlast) * NB + l;
      a[0][t] = Ap[0][rr];
      a[1][t] = Ap[1][rr];
      bv[t] = vs[(r + 4 * t + g) * NB + l];
    }
    // all 48 loads of the iteration go out before its products (left to itself the scheduler
    // moves each load next to its use and a wave keeps few loads in flight)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][t], bv[t], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1][t], bv[t], acc[1], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  if (i < 8) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int mrow = g + 4 * v, blk = bw + 2 * h + (mrow >> 3);
        if (blk < nbt) part[((size_t)blockIdx.x * pstride + blk) * 64 + (mrow & 7) * 8 + i] = acc[h][v];
      }
  }
}

// block per (system, basis block b < nbt): C (dst 0: C1, 1: C2) block b = sum of the system's tile
// partials (the four waves take every fourth tile). Systems outside the map have no tiles.
__global__ __launch_bounds__(256) void kb_bdots_final(const GSys *__restrict__ sys, const double *__restrict__ part,
                                                      int pstride, int dst, int all) {
  __shared__ double red[4][64];
  const GSys &S = sys[blockIdx.x];
  if (S.nt4 == 0 || (!all && !sys_iterating(S))) return;  // uniform over the block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, b = blockIdx.y;
  double c = 0.0;
  for (int tl = wv; tl < S.nt4; tl += 4) c += part[((size_t)(S.t40 + tl) * pstride + b) * 64 + lane];
  red[wv][lane] = c;
  __syncthreads();
  if (wv == 0) (dst ? S.C2 : S.C1)[(size_t)b * 64 + lane] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// thread per row, all NB columns (16-byte loads; the coefficients are uniform over the block):
//   mode 0: W -= sum_(b < nb) Q_b C1_b
//   mode 1: Q_nb = (W - sum_(b < nb) Q_b C2_b) Ri          (the next basis block)
//   mode 2: X += sum_(b < kk) Q_b Y_b, Y in C1, kk the rounds the system used (at most nb)
//   mode 3: Q_0 = R Ri                                      (start of a cycle)
__global__ __launch_bounds__(256) void kb_bupdate(const GSys *__restrict__ sys, const TileMap *__restrict__ tmap, int nb,
                                                  int mode, int all) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const TileMap tm = tmap[blockIdx.x];
  const GSys &S = sys[tm.sys];
  if (mode < 2 && !all && !sys_iterating(S)) return;  // the sweeps of a round; no barriers here
  const int n = S.n;
  const int row = tm.row0 + threadIdx.x;
  if (row >= n) return;  // no barriers in this kernel
  int nbb = mode == 3 ? 0 : nb;
  if (mode == 2) {  // the rounds used: the live columns freeze together (none live: no update)
    int kk = -1;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const double *sm = S.small + 8 * k;
      if (sm[4] > 0.0) kk = max(kk, sm[5] > 0.0 ? min((int)sm[5], nb) : nb);
    }
    nbb = max(kk, 0);
  }
  // the coefficients are uniform over the block and read-only here: loads through the constant
  // address space go to the scalar cache and feed the FMAs as scalar operands
  typedef const __attribute__((address_space(4))) double cdouble;
  cdouble *Cf = (cdouble *)(uintptr_t)(mode == 1 ? S.C2 : S.C1);
  cdouble *Ric = (cdouble *)(uintptr_t)S.Ri;
  const double *src = mode == 3 ? S.R : (mode == 2 ? S.X : S.W);
  d2 x[4], s[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    x[p] = ((const d2 *)(src + (size_t)row * NB))[p];
    s[p] = d2{0.0, 0.0};
  }
  // blocks in groups of four: the group's 16 row loads go out before its products (past the last
  // block the loads repeat it and are not used)
  const size_t vstride = (size_t)n * NB;
  for (int b0 = 0; b0 < nbb; b0 += 4) {
    d2 qv[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const d2 *q = (const d2 *)(S.Q + (size_t)min(b0 + u, nbb - 1) * vstride + (size_t)row * NB);
#pragma unroll
      for (int p = 0; p < 4; ++p) qv[u][p] = q[p];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (b0 + u >= nbb) break;  // uniform
      cdouble *C = Cf + (size_t)(b0 + u) * 64;
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        const double ql = (l & 1) ? qv[u][l >> 1].y : qv[u][l >> 1].x;
#pragma unroll
        for (int p = 0; p < 4; ++p) s[p] += ql * d2{C[l * 8 + 2 * p], C[l * 8 + 2 * p + 1]};
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  if (mode == 0 || mode == 2) {
    d2 *dst = (d2 *)((mode == 0 ? S.W : S.X) + (size_t)row * NB);
#pragma unroll
    for (int p = 0; p < 4; ++p) dst[p] = mode == 0 ? x[p] - s[p] : x[p] + s[p];
    return;
  }
  double w[8];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const d2 v = x[p] - s[p];
    w[2 * p] = v.x;
    w[2 * p + 1] = v.y;
  }
  const size_t blk = mode == 1 ? (size_t)nb : 0;  // mode 3 writes Q_0
  d2 *dst = (d2 *)(S.Q + (blk * n + row) * NB);
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    d2 o = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 8; ++c) o += w[c] * d2{Ric[c * 8 + 2 * p], Ric[c * 8 + 2 * p + 1]};
    dst[p] = o;
  }
}

// Cholesky factor of an L x L Gram matrix G (row-major, stride 8) with dropped columns: column c is
// dropped (zero row, zero diagonal) when it is dead or its pivot is at most drop * nrm2[c]; rows of
// R are written to R (stride 8), the inverse over the kept columns to Ri (full 8 x 8, zero
// elsewhere). Returns 1 when a live column with nrm2 > 0 was dropped. One thread.
__device__ inline int chol_drop(const double *G, const double *nrm2, const bool *live, int L, double drop, double *R,
                                double *Ri) {
  int brk = 0;
  for (int a = 0; a < 64; ++a) {
    R[a] = 0.0;
    Ri[a] = 0.0;
  }
  for (int c = 0; c < L; ++c) {
    double d = G[c * 8 + c];
    for (int t = 0; t < c; ++t) d -= R[t * 8 + c] * R[t * 8 + c];
    if (!live[c] || !(d > drop * nrm2[c])) {
      // a live column whose product still had a norm lost its new direction: breakdown; a zero
      // product (the column has no basis vector, e.g. a right-hand side that repeats another) is not
      if (live[c] && nrm2[c] > 0.0) brk = 1;
      continue;  // row c stays zero
    }
    const double r = sqrt(d);
    R[c * 8 + c] = r;
    for (int c2 = c + 1; c2 < L; ++c2) {
      double v = G[c * 8 + c2];
      for (int t = 0; t < c; ++t) v -= R[t * 8 + c] * R[t * 8 + c2];
      R[c * 8 + c2] = v / r;
    }
  }
  // Ri = R^-1 over the kept columns (upper triangular), column by column
  for (int k = 0; k < L; ++k) {
    if (R[k * 8 + k] == 0.0) continue;
    Ri[k * 8 + k] = 1.0 / R[k * 8 + k];
    for (int a = k - 1; a >= 0; --a) {
      if (R[a * 8 + a] == 0.0) continue;
      double v = 0.0;
      for (int t = a + 1; t <= k; ++t) v += R[a * 8 + t] * Ri[t * 8 + k];
      Ri[a * 8 + k] = -v / R[a * 8 + a];
    }
  }
  return brk;
}

// wave per system of the cycle: Q_0 = R R0^-1 from the Gram matrix of the residual block (C2 block 0),
// g = (R0; 0). Dead columns (converged or beyond nrhs) take no part.
__global__ __launch_bounds__(64) void kb_binit(const GSys *__restrict__ sys, int m, int L) {
  __shared__ double sR[64], sRi[64];
  const GSys &S = sys[blockIdx.x];
  if (!S.incyc) return;  // uniform over the block
  const int lane = threadIdx.x;
  if (lane == 0) {
    bool live[NB], any = false;
    double nrm2[NB];
    for (int c = 0; c < NB; ++c) {
      live[c] = c < L && S.small[8 * c + 4] > 0.0;
      any = any || live[c];
      nrm2[c] = S.C2[c * 9];
    }
    (void)chol_drop(S.C2, nrm2, live, L, 1e-24, sR, sRi);
    S.small[6] = any ? 1.0 : 0.0;  // iterating in this cycle
  }
  __syncthreads();
  S.Ri[lane] = sRi[lane];
  // rows 0 .. L - 1 of g hold R0 (column k of the live k), the rest zero
  for (int idx = lane; idx < L * (m + 1) * NB; idx += 64) {
    const int r = idx / NB, k = idx % NB;
    S.gB[idx] = (r < L && k < L) ? sR[r * 8 + k] : 0.0;
  }
}

// wave per system of the current map: the small dense part of round j. Lane (c, i) = 8 c + i:
// group c handles new Hessenberg column L j + c (and later right-hand side c), sub-lane i the row
// q + i of the reflector window at step q.
__global__ __launch_bounds__(64) void kb_bsmall(const GSys *__restrict__ sys, int j, int L, double tol, int all) {
  __shared__ double sG[64], sR[64], sRi[64], sN[NB];
  __shared__ double sv[NB][NB];  // reflectors of the round's new columns: tau, v_1 .. v_L
  __shared__ double shv[NB][NB], sen[NB][NB];  // a chunk of 8 earlier reflectors, entering rows
  __shared__ double sC1[8][64], sC2[8][64];       // a chunk of 8 coefficient blocks
  __shared__ int sbrk, sok[NB];
  const GSys &S = sys[blockIdx.x];
  if (S.nt4 == 0 || (!all && !sys_iterating(S))) return;  // uniform over the block
  const int lane = threadIdx.x, c = lane >> 3, i = lane & 7;
  const double *C1 = S.C1, *C2 = S.C2;
  // G2 = G1 - sum_b C2_b^T C2_b; squared norms of the product's columns before the sweeps. The
  // coefficient blocks pass through LDS 8 at a time (coalesced loads, all in flight together).
  {
    double gs = C2[(size_t)(j + 1) * 64 + c * 8 + i];
    double w = C2[(size_t)(j + 1) * 64 + c * 9];
    for (int b0 = 0; b0 <= j; b0 += 8) {
      const int nb8 = min(8, j + 1 - b0);
      double v1[8], v2[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        v1[u] = u < nb8 ? C1[(size_t)(b0 + u) * 64 + lane] : 0.0;
        v2[u] = u < nb8 ? C2[(size_t)(b0 + u) * 64 + lane] : 0.0;
      }
      __syncthreads();  // the previous chunk's reads are done
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        sC1[u][lane] = v1[u];
        sC2[u][lane] = v2[u];
      }
      __syncthreads();
      for (int u = 0; u < nb8; ++u)
        for (int l = 0; l < L; ++l) {
          gs -= sC2[u][l * 8 + c] * sC2[u][l * 8 + i];
          w += sC1[u][l * 8 + c] * sC1[u][l * 8 + c];
        }
    }
    sG[lane] = gs;
    if (c == i) sN[c] = w;
  }
  __syncthreads();
  if (lane == 0) {
    bool live[NB];
    for (int k = 0; k < NB; ++k) live[k] = k < L && S.small[8 * k + 4] > 0.0;
    sbrk = chol_drop(sG, sN, live, L, 1e-20, sR, sRi);
  }
  __syncthreads();
  S.Ri[lane] = sRi[lane];
  const int LJ = L * j, rows = L * (j + 2);
  const bool mine = c < L;  // a new column of this round
  // H[r][LJ + c] as built: C1 + C2 rows of the blocks b <= j, then the subdiagonal block R_(j+1)
  auto hval = [&](int r) -> double {
    if (!mine || r >= rows) return 0.0;
    if (r < L * (j + 1)) {
      const int b = r / L, l = r - b * L;
      return C1[(size_t)b * 64 + l * 8 + c] + C2[(size_t)b * 64 + l * 8 + c];
    }
    const int cc = r - L * (j + 1);
    return cc <= c ? sR[cc * 8 + c] : 0.0;
  };
  auto gsum = [&](double x) {  // over the 8 sub-lanes of a group
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    return x;
  };
  double w = i <= L ? hval(i) : 0.0;  // window: sub-lane i holds row q + i
  const int qc = LJ + c;
  double *Rcol = S.Rp + (size_t)qc * (qc + 1) / 2;
  // the reflectors of earlier rounds, 8 steps at a time: lane (c, i) fetches entry i of reflector
  // q0 + c and the row that enters group c's window at step q0 + i, so that the chain of steps
  // reads LDS only
  for (int q0 = 0; q0 < LJ; q0 += 8) {
    const int qn = min(8, LJ - q0);
    const double hvv = c < qn ? S.hv[(size_t)(q0 + c) * 8 + i] : 0.0;
    const double env = i < qn ? hval(q0 + i + 1 + L) : 0.0;
    __syncthreads();
    shv[c][i] = hvv;
    sen[c][i] = env;
    __syncthreads();
    for (int t = 0; t < qn; ++t) {
      const double tau = shv[t][0], vi = i == 0 ? 1.0 : (i <= L ? shv[t][i] : 0.0);
      const double sd = gsum(mine ? vi * w : 0.0);
      if (mine) {
        w -= tau * sd * vi;
        if (i == 0) Rcol[q0 + t] = w;  // row q0 + t of column qc is final
      }
      const double up = __shfl_down(w, 1, 8);
      if (mine) w = i < L ? up : (i == L ? sen[c][t] : 0.0);
    }
  }
  // the round's own columns: column LJ + cq forms its reflector at step LJ + cq, the later ones apply it
  for (int q = LJ; q < LJ + L; ++q) {
    const int cq = q - LJ;  // the new column whose reflector this step forms
    const double x0 = __shfl(w, 8 * c, 64);
    const double s2 = gsum(i >= 1 && i <= L ? w * w : 0.0);
    if (c == cq) {
      double alpha = x0, t = 0.0, sc = 0.0;
      if (s2 > 0.0) {
        const double nx = sqrt(x0 * x0 + s2);
        alpha = x0 > 0.0 ? -nx : nx;
        t = (alpha - x0) / alpha;
        sc = 1.0 / (x0 - alpha);
      }
      if (i == 0) {
        sv[cq][0] = t;
        Rcol[q] = alpha;
      } else if (i <= L) {
        sv[cq][i] = w * sc;
      }
      if (i <= L) S.hv[(size_t)q * 8 + i] = i == 0 ? t : w * sc;
    }
    __syncthreads();
    const double tau = sv[cq][0], vi = i == 0 ? 1.0 : (i <= L ? sv[cq][i] : 0.0);
    const bool act = mine && q < qc;  // group c applies the reflectors of the columns before its own
    const double sd = gsum(act ? vi * w : 0.0);
    if (act) {
      w -= tau * sd * vi;
      if (i == 0) Rcol[q] = w;  // row q of column qc is final
    }
    const double up = __shfl_down(w, 1, 8);
    const double enter = (act && i == L) ? hval(q + 1 + L) : 0.0;
    if (act) w = i < L ? up : enter;
  }
  __syncthreads();
  // the round's reflectors on the right-hand sides: group k, window rows LJ + t + i
  const int k = c;
  const bool klive = k < L && S.small[8 * k + 4] > 0.0;
  double gw = (klive && i <= L) ? S.gB[(size_t)(LJ + i) * NB + k] : 0.0;
  for (int t = 0; t < L; ++t) {
    const double tau = sv[t][0], vi = i == 0 ? 1.0 : (i <= L ? sv[t][i] : 0.0);
    const double s = gsum(vi * gw);
    gw -= tau * s * vi;
    if (klive && i == 0) S.gB[(size_t)(LJ + t) * NB + k] = gw;
    const double up = __shfl_down(gw, 1, 8);
    const int r = LJ + t + 1 + L;
    const double enter = (klive && i == L && r < rows) ? S.gB[(size_t)r * NB + k] : 0.0;
    gw = i < L ? up : (i == L ? enter : 0.0);
  }
  // rows L (j + 1) .. L (j + 2) - 1: the residual of right-hand side k
  if (klive && i < L) S.gB[(size_t)(L * (j + 1) + i) * NB + k] = gw;
  const double est = sqrt(gsum(i < L ? gw * gw : 0.0));
  double *sm = S.small + 8 * k;
  if (i == 0) {
    sok[k] = 1;
    if (klive) {
      sm[2] = est;
      sok[k] = est <= EST_MARGIN * tol * sm[0] ? 1 : 0;
    }
  }
  __syncthreads();
  if (i == 0 && klive && sm[5] < 0.0) {
    bool all = sbrk != 0;
    if (!all) {
      all = true;
      for (int kk = 0; kk < NB; ++kk) all = all && sok[kk];
    }
    if (all) {
      sm[5] = (double)(j + 1);  // rounds to use; the product skips the system from now on
      S.small[6] = 0.0;         // and the sweeps (every live column writes the same value)
    }
  }
}

// wave per system of the cycle: Y = R^-1 g over the rounds used (in place in g), then the solution
// blocks Y_b[l][k] into C1. Lane 8 p + k owns the entries (r, k) with r = p (mod 8) of g, so that
// each element of g is only ever read back by the lane that wrote it.
__global__ __launch_bounds__(64) void kb_bsolve(const GSys *__restrict__ sys, int steps, int L) {
  __shared__ double sy[NB];
  const GSys &S = sys[blockIdx.x];
  if (!S.incyc) return;  // uniform over the block
  const int lane = threadIdx.x, p = lane >> 3, k = lane & 7;
  int kk = -1;
  for (int c = 0; c < L; ++c) {
    const double *sm = S.small + 8 * c;
    if (sm[4] > 0.0) kk = max(kk, sm[5] > 0.0 ? min((int)sm[5], steps) : steps);
  }
  kk = max(kk, 0);
  const int N = L * kk;
  for (int q = N - 1; q >= 0; --q) {
    const double *Rc = S.Rp + (size_t)q * (q + 1) / 2;
    if (p == (q & 7)) {
      double *gq = S.gB + (size_t)q * NB + k;
      const double d = Rc[q];
      const double y = d != 0.0 ? *gq / d : 0.0;
      *gq = y;
      sy[k] = y;
    }
    __syncthreads();
    const double yk = sy[k];
    for (int r = p; r < q; r += 8) S.gB[(size_t)r * NB + k] -= Rc[r] * yk;
    __syncthreads();
  }
  for (int r = p; r < N; r += 8) {
    const int b = r / L, l = r - b * L;
    S.C1[(size_t)b * 64 + l * 8 + k] = S.gB[(size_t)r * NB + k];
  }
  for (int idx = lane; idx < kk * 64; idx += 64)
    if (((idx >> 3) & 7) >= L) S.C1[idx] = 0.0;
}

}  // namespace

int andr_gmres_batch(andr_ctx *ctx, const GmresSys *in, int nsys, bool fp32, int nrhs, int restart, int maxit,
                     double tol, int *iters, int *status) {
  if (nsys <= 0) return ANDR_OK;
  if (nrhs < 1 || nrhs > NB) return andr_fail(ctx, ANDR_ERR_ARG, "nrhs must be 1..%d", NB);
  const int m = std::max(4, restart);
  hipStream_t st = ctx->stream;
  // full work maps (sizes the partial-sum buffers); subsets are rebuilt as systems converge
  // the structured Stokes operator of the fast path (FP32 symmetric blocks) is the only operator
  if (!fp32) return andr_fail(ctx, ANDR_ERR_ARG, "GMRES takes the structured Stokes operator only");
  size_t vec_elems = 0, basis_elems = 0, zt_elems = 0;
  int nw_full = 0, nt_full = 0, net_full = 0, nt4_full = 0;
  for (int s = 0; s < nsys; ++s) {
    const int n = in[s].n, ne = in[s].op.ne;
    if (!in[s].op.S || !in[s].op.minv || 3 * ne != n)
      return andr_fail(ctx, ANDR_ERR_ARG, "structured GMRES system %d is incomplete", s);
    vec_elems += (size_t)n * NB;
    basis_elems += (size_t)(m + 1) * n * NB;
    nw_full += 4 * ((ne + 63) / 64);
    nt_full += (n + TILE - 1) / TILE;
    nt4_full += (n + BT - 1) / BT;
    net_full += (ne + 31) / 32;
    zt_elems += (size_t)3 * NB * ((ne + 31) / 32 * 32);
  }
  // per system: Hessenberg, rotations, g, coefficients; the per-rhs state (8 doubles per rhs) of all
  // systems sits in one separate array so that the host reads it with one copy
  // Block GMRES (one Krylov space per system for its nrhs columns; the reflector window of its
  // least-squares update holds nrhs + 1 rows in 8 lanes, so nrhs <= 7) unless
  // ANDR_GMRES_INDEPENDENT=1, which runs one Krylov space per column.
  const bool blk = getenv("ANDR_GMRES_INDEPENDENT") == nullptr && nrhs <= NB - 1;
  const int L = nrhs;
  auto small_of = [&](int mc) {
    if (blk)  // Rp, hv, gB, C1, C2, Ri
      return (size_t)L * mc * ((size_t)L * mc + 1) / 2 + (size_t)8 * L * mc + (size_t)L * (mc + 1) * NB +
             (size_t)(mc + 1) * 64 + (size_t)(mc + 2) * 64 + 64;
    // H, Hu, cs, sn, g, cw per rhs
    return (size_t)NB * (2 * (size_t)(mc + 1) * mc + 2 * mc + (mc + 1) + cw_stride(mc));
  };
  const size_t small_per = small_of(m);
  const size_t state_per = (size_t)NB * 8;
  DevBuf<double> R, W, Q, SM, ST8, partW, partT, SDEF, partS, partB;
  DevBuf<double> ZT;
  DevBuf<float> ZT32;
  DevBuf<GSys> dsys;
  DevBuf<int> ddone;
  if (R.alloc(vec_elems, ctx) || W.alloc(vec_elems, ctx) || Q.alloc(basis_elems, ctx) ||
      SM.alloc(small_per * nsys, ctx) || ST8.alloc(state_per * nsys, ctx) || partW.alloc((size_t)nw_full * NB, ctx) ||
      (!blk && partT.alloc((size_t)nt_full * part_stride(m) * NB, ctx)) ||
      (blk && partB.alloc((size_t)nt4_full * (m + 2) * 64, ctx)) ||
      dsys.alloc(nsys, ctx) || ddone.alloc(1, ctx))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES workspace (%zu basis doubles)", basis_elems);
  // Krylov products on the FP32 matrix cores at relative residuals of 1e-6 and above (their ~1e-7
  // relative error stays below the tolerance); FP64 products for tighter tolerances.
  // ANDR_GMRES_FP64_PRODUCTS=1 keeps FP64 products at every tolerance.
  const bool f32prod = tol >= 1e-6 && getenv("ANDR_GMRES_FP64_PRODUCTS") == nullptr;
  if ((ZT.alloc(zt_elems, ctx) || (f32prod && ZT32.alloc(zt_elems, ctx)) ||
                     SDEF.alloc((size_t)nsys * MAXW * NB, ctx) || partS.alloc((size_t)net_full * MAXW * NB, ctx)))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES operator workspace");
  // a map set over a list of systems: residual blocks of 64 element rows, tiles of
  // TILE rows (vector kernels) and matrix-core waves (64 rows x a column segment). When the list
  // has few rows (the last, slowest systems of a batch) the columns are split into segments so
  // that the product still fills the chip; the segments are summed by kb_seg_sum.
  constexpr int MF_TARGET_WAVES = 2048, MF_MAX_SEG = 16;
  // structured operator: w = residual blocks of 64 element rows (4 partial entries each), mw =
  // matrix-core waves of 32 element rows, et = element tiles of 32 (z = M^-1 v and wall sums)
  struct MapSet {
    DevBuf<WaveMap> w;
    DevBuf<TileMap> t, et, t4;
    DevBuf<MWave> mw;
    DevBuf<int> w0, nw;
    int n_w = 0, n_t = 0, n_et = 0, n_mw = 0, n_t4 = 0, max_seg = 1;
    size_t mw_cap = 0;
    std::vector<int> tile0, ntiles, et0, net, nseg, t40, nt4;
    std::vector<size_t> poff;
    size_t p_elems = 0;
  };
  MapSet cyc, live;
  for (MapSet *ms : {&cyc, &live})
    if (ms->w.alloc(nw_full, ctx) || ms->t.alloc(nt_full, ctx) || ms->w0.alloc(nsys, ctx) || ms->nw.alloc(nsys, ctx) ||
        ms->et.alloc(net_full, ctx) || (blk && ms->t4.alloc(nt4_full, ctx)))
      return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES work maps");
  std::vector<GSys> hs(nsys);
  DevBuf<double> Pseg;
  size_t pseg_cap = 0;
  auto build = [&](const std::vector<int> &list, MapSet &ms, int target_waves = MF_TARGET_WAVES) -> int {
    std::vector<WaveMap> wm;
    std::vector<TileMap> tm, em, t4;
    std::vector<MWave> mw;
    std::vector<int> w0(nsys, 0), nw(nsys, 0);
    ms.t40.assign(nsys, 0);
    ms.nt4.assign(nsys, 0);
    ms.tile0.assign(nsys, 0);
    ms.ntiles.assign(nsys, 0);
    ms.et0.assign(nsys, 0);
    ms.net.assign(nsys, 0);
    ms.nseg.assign(nsys, 1);
    ms.poff.assign(nsys, 0);
    ms.p_elems = 0;
    ms.max_seg = 1;
    int row_waves = 0;
    for (int s : list) row_waves += (in[s].op.ne + 31) / 32;
    const int want = row_waves > 0 ? std::min(MF_MAX_SEG, (target_waves + row_waves - 1) / row_waves) : 1;
    for (int s : list) {
      const int n = in[s].n, ne = in[s].op.ne;
      w0[s] = 4 * (int)wm.size();
      for (int r = 0; r < ne; r += 64) wm.push_back({s, r});
      nw[s] = 4 * (int)wm.size() - w0[s];
      ms.tile0[s] = (int)tm.size();
      for (int r = 0; r < n; r += TILE) tm.push_back({s, r});
      ms.ntiles[s] = (int)tm.size() - ms.tile0[s];
      if (blk) {
        ms.t40[s] = (int)t4.size();
        for (int r = 0; r < n; r += BT) t4.push_back({s, r});
        ms.nt4[s] = (int)t4.size() - ms.t40[s];
      }
      ms.et0[s] = (int)em.size();
      for (int e = 0; e < ne; e += 32) em.push_back({s, e});
      ms.net[s] = (int)em.size() - ms.et0[s];
      const int seglen = ((ne + want - 1) / want + 31) / 32 * 32;
      const int nseg = (ne + seglen - 1) / seglen;
      ms.nseg[s] = nseg;
      ms.max_seg = std::max(ms.max_seg, nseg);
      if (nseg > 1) {
        ms.poff[s] = ms.p_elems;
        ms.p_elems += (size_t)nseg * n * NB;
      }
      for (int r = 0; r < ne; r += 32)
        for (int sg = 0; sg < nseg; ++sg) mw.push_back({s, r, sg * seglen, std::min(ne, (sg + 1) * seglen), sg});
    }
    ms.n_w = (int)wm.size();
    ms.n_t = (int)tm.size();
    ms.n_et = (int)em.size();
    ms.n_mw = (int)mw.size();
    ms.n_t4 = (int)t4.size();
    if (ms.n_t4) ANDR_HIP(ctx, hipMemcpyAsync(ms.t4.p, t4.data(), sizeof(TileMap) * t4.size(), hipMemcpyHostToDevice, st));
    if (ms.n_et) ANDR_HIP(ctx, hipMemcpyAsync(ms.et.p, em.data(), sizeof(TileMap) * em.size(), hipMemcpyHostToDevice, st));
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
      hs[s].t40 = blk ? ms.t40[s] : 0;
      hs[s].nt4 = blk ? ms.nt4[s] : 0;
      hs[s].et0 = ms.et0[s];
      hs[s].net = ms.net[s];
      hs[s].nseg = ms.nseg[s];
      hs[s].P = ms.nseg[s] > 1 ? Pseg.p + ms.poff[s] : nullptr;
    }
    ANDR_HIP(ctx, hipMemcpyAsync(dsys.p, hs.data(), sizeof(GSys) * nsys, hipMemcpyHostToDevice, st));
    return ANDR_OK;
  };
  size_t vo = 0, zo = 0;
  for (int s = 0; s < nsys; ++s) {
    GSys &g = hs[s];
    const int n = in[s].n;
    g.op = in[s].op;
    g.nep = (g.op.ne + 31) / 32 * 32;
    g.ZT = ZT.p + zo;
    g.ZT32 = f32prod ? ZT32.p + zo : nullptr;
    g.sdef = SDEF.p + (size_t)s * MAXW * NB;
    zo += (size_t)3 * NB * g.nep;
    g.A = in[s].A;
    g.lda = in[s].lda;
    g.n = n;
    g.B = in[s].B;
    g.X = in[s].X;
    g.R = R.p + vo;
    g.W = W.p + vo;
    g.small = ST8.p + state_per * s;
    vo += (size_t)n * NB;
  }
  // Basis and Hessenberg storage of the systems of a cycle, for restart length mc. The first cycle
  // takes every system at length m; the systems left after it (slender bodies whose hard columns
  // converge slowly under restarts) are laid out again with the longest restart that fits the same
  // buffers (Q, SM, partT), up to MMAX.
  constexpr int MMAX = 1024;
  auto layout = [&](const std::vector<int> &list, int mc) {
    size_t bo = 0, so = 0;
    for (int s = 0; s < nsys; ++s) hs[s].incyc = 0;
    for (int s : list) {
      GSys &g = hs[s];
      g.incyc = 1;
      g.Q = Q.p + bo;
      bo += (size_t)(mc + 1) * in[s].n * NB;
      double *sp = SM.p + so;
      so += small_of(mc);
      if (blk) {
        g.Rp = sp;
        sp += (size_t)L * mc * ((size_t)L * mc + 1) / 2;
        g.hv = sp;
        sp += (size_t)8 * L * mc;
        g.gB = sp;
        sp += (size_t)L * (mc + 1) * NB;
        g.C1 = sp;
        sp += (size_t)(mc + 1) * 64;
        g.C2 = sp;
        sp += (size_t)(mc + 2) * 64;
        g.Ri = sp;
        continue;
      }
      g.H = sp;
      sp += (size_t)NB * (mc + 1) * mc;
      g.Hu = sp;
      sp += (size_t)NB * (mc + 1) * mc;
      g.cs = sp;
      sp += NB * mc;
      g.sn = sp;
      sp += NB * mc;
      g.g = sp;
      sp += NB * (mc + 1);
      g.cw = sp;
    }
  };
  auto restart_for = [&](const std::vector<int> &list) {
    size_t nsum = 0, ntl = 0;
    for (int s : list) {
      nsum += in[s].n;
      ntl += (in[s].n + TILE - 1) / TILE;
    }
    size_t nt4l = 0;
    for (int s : list) nt4l += (in[s].n + BT - 1) / BT;
    for (int mc = MMAX; mc > m; mc -= 8)
      if ((size_t)(mc + 1) * nsum * NB <= basis_elems && list.size() * small_of(mc) <= small_per * nsys &&
          (blk ? nt4l * (mc + 2) <= (size_t)nt4_full * (m + 2) : ntl * part_stride(mc) <= (size_t)nt_full * part_stride(m)))
        return mc;
    return m;
  };
  int mc = m;
  {
    std::vector<int> all(nsys);
    for (int s = 0; s < nsys; ++s) all[s] = s;
    layout(all, m);
  }
  ANDR_HIP(ctx, hipMemsetAsync(SM.p, 0, sizeof(double) * small_per * nsys, st));
  ANDR_HIP(ctx, hipMemsetAsync(ST8.p, 0, sizeof(double) * state_per * nsys, st));
  for (int s = 0; s < nsys; ++s) ANDR_HIP(ctx, hipMemsetAsync(in[s].X, 0, sizeof(double) * in[s].n * NB, st));
  // scale factors of the segment sums (sx = 1, sy = 0): W = A z for the Krylov and residual products
  std::vector<double> h_arn(2 * NB * (size_t)nsys);
  for (int s = 0; s < nsys; ++s)
    for (int k = 0; k < NB; ++k) {
      h_arn[s * 2 * NB + k] = 1.0;
      h_arn[s * 2 * NB + NB + k] = 0.0;
    }
  DevBuf<double> d_arn;
  if (d_arn.alloc(h_arn.size(), ctx)) return andr_fail(ctx, ANDR_ERR_ALLOC, "GMRES scales");
  ANDR_HIP(ctx, hipMemcpyAsync(d_arn.p, h_arn.data(), sizeof(double) * h_arn.size(), hipMemcpyHostToDevice, st));

  const int gs = (nsys * NB + 63) / 64;
  // ANDR_GMRES_NO_SKIP=1: converged systems keep their products (and block sweeps) until the next map
  // rebuild (the rounds after convergence are never used, so the results are the same bit for bit)
  const int noskip = getenv("ANDR_GMRES_NO_SKIP") ? 1 : 0;
  // which_x 0: residual R = B - A M^-1 X (R holds B); 1: Krylov product W = A M^-1 Q_j
  auto gemv = [&](const MapSet &ms, int which_x, int jq, const double *scales, bool prof = true) {
    if (ms.n_w == 0) return;
    double n2 = 0.0, n1 = 0.0;
    for (int s = 0; s < nsys; ++s)
      if (ms.ntiles[s]) {
        const double ne = in[s].op.ne;
        n2 += 6.0 * ne * ne;
        n1 += in[s].n;
      }
    const bool krylov = which_x == 1;
    // the residual (FP64 products of every system, then R = B - W) or the Krylov product
    const int tok = prof ? andr_prof_begin(ctx, krylov ? "kb_gemv_sym" : "gmres_residual", n2 * 4.0 + 2.0 * n1 * NB * 8.0,
                                           2.0 * 1.5 * n2 * nrhs)
                         : -1;
    // the block solver's Krylov products skip converged systems' operands too (the flag read by
    // sys_iterating is kept by the block kernels only)
    const int all = (krylov && blk) ? noskip : 1;
    hipLaunchKernelGGL(kb_prep, dim3(ms.n_et), dim3(256), 0, st, dsys.p, ms.et.p, krylov ? 0 : 1, jq, partS.p, all);
    hipLaunchKernelGGL(kb_sred, dim3(nsys), dim3(256), 0, st, dsys.p, partS.p, all);
    if (krylov && f32prod)
      hipLaunchKernelGGL(kb_gemv_sym<true>, dim3((ms.n_mw + 3) / 4), dim3(256), 0, st, dsys.p, ms.mw.p, ms.n_mw,
                         noskip);
    else
      hipLaunchKernelGGL(kb_gemv_sym<false>, dim3((ms.n_mw + 3) / 4), dim3(256), 0, st, dsys.p, ms.mw.p, ms.n_mw,
                         krylov ? noskip : 1);
    if (ms.max_seg > 1) hipLaunchKernelGGL(kb_seg_sum, dim3(ms.n_t), dim3(256), 0, st, dsys.p, ms.t.p, scales);
    if (!krylov) hipLaunchKernelGGL(kb_resid_fin, dim3(ms.n_w), dim3(256), 0, st, dsys.p, ms.w.p, partW.p);
    andr_prof_end(ctx, tok);
  };
  // Autotuner of the product's column segments: when few rows are live (the slowest bodies of a
  // batch) the columns are split so that the product still fills the chip; the split aims at a
  // number of waves. The first time a size bucket (ceil log2 of the live 32-row waves) occurs, the
  // candidates are timed on the live systems' current Krylov vector and the fastest is kept in the
  // context. The split changes the summation order of the product (segments are summed apart).
  auto krylov_target = [&](const std::vector<int> &list, int jq) -> int {
    int row_waves = 0;
    for (int s : list) row_waves += (in[s].op.ne + 31) / 32;
    if (!ctx->autotune || row_waves <= 0 || row_waves >= 4096) return MF_TARGET_WAVES;
    int key = 0;
    while ((1 << key) < row_waves) ++key;
    auto it = ctx->tune_krylov.find(key);
    if (it != ctx->tune_krylov.end()) return it->second.first;
    hipEvent_t e0, e1;
    ANDR_HIP(ctx, hipEventCreate(&e0));
    ANDR_HIP(ctx, hipEventCreate(&e1));
    int best = MF_TARGET_WAVES;
    float best_ms = 1e30f;
    for (int target : {1024, 2048, 4096, 8192}) {
      if (build(list, live, target) || use_tiles(live)) break;
      gemv(live, 1, jq, d_arn.p, false);  // warm-up
      float ms = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        (void)hipEventRecord(e0, st);
        gemv(live, 1, jq, d_arn.p, false);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms = std::min(ms, t);
      }
      if (ms < best_ms) {
        best_ms = ms;
        best = target;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ANDR_HIP(ctx, hipGetLastError());
    ctx->tune_krylov[key] = {best, best_ms};
    return best;
  };
  std::vector<int> it_count((size_t)nsys * NB, 0);
  std::vector<double> hsmall(state_per * nsys);
  auto read_small = [&]() -> int {
    ANDR_HIP(ctx, hipMemcpyAsync(hsmall.data(), ST8.p, sizeof(double) * state_per * nsys, hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));
    return ANDR_OK;
  };
  auto sm_of = [&](int s, int k) { return hsmall.data() + state_per * s + k * 8; };
  // The Krylov product skips a system once every column has converged in the cycle (kb_gemv_sym);
  // the host learns the step afterwards from the columns' step counts, and takes the skipped
  // matrix reads out of the profiled totals (a system once, when it leaves the live map).
  std::vector<char> skip_counted(nsys, 0);
  auto account_skips = [&](int s, int jlast) {
    if (noskip || skip_counted[s] || !andr_prof(ctx).on) return;
    int fz = -1;
    for (int k = 0; k < NB; ++k) {
      const double *sm = sm_of(s, k);
      if (sm[4] <= 0.0) continue;
      if (sm[5] < 0.0) return;  // a live column: no launch was skipped
      fz = std::max(fz, (int)sm[5]);
    }
    skip_counted[s] = 1;
    // fz: steps used (independent: set at step fz, whose product ran) or rounds used (block: set in
    // round fz - 1); the launches after it up to jlast skipped the system
    const int first_skipped = blk ? fz : fz + 1;
    if (fz < 0 || jlast < first_skipped) return;
    const double ne = in[s].op.ne, nskip = jlast - first_skipped + 1;
    andr_prof_adjust(ctx, "kb_gemv_sym", -nskip * (6.0 * ne * ne * 4.0 + in[s].n * NB * 8.0),
                     -nskip * 2.0 * 1.5 * 6.0 * ne * ne * nrhs);
    if (blk) {  // the block sweeps skip the same rounds, and the normalising update of round fz - 1
      double units = (fz - 1) + 3.0, fl = 0.0;
      for (int jj = fz; jj <= jlast; ++jj) {
        units += 4.0 * jj + 10.0;
        fl += 4.0 * jj + 5.0;
      }
      andr_prof_adjust(ctx, "gmres_orthogonalise", -units * in[s].n * NB * 8.0, -2.0 * 64.0 * fl * in[s].n);
    }
  };
  std::vector<int> cyc_list(nsys);
  for (int s = 0; s < nsys; ++s) cyc_list[s] = s;
  int rc = build(cyc_list, cyc);
  if (rc) return rc;
  if ((rc = use_tiles(cyc))) return rc;
  int total = 0;
  bool first = true, first_cycle = true;
  const bool debug = getenv("ANDR_DEBUG_GMRES") != nullptr;
  while (true) {
    // residual r = b - A x of every system of the last cycle (x = 0 on the first), FP64 accumulation
    if (cyc.n_t) hipLaunchKernelGGL(kb_copy_b, dim3(cyc.n_t), dim3(256), 0, st, dsys.p, cyc.t.p);
    if (first) {
      if (cyc.n_w) hipLaunchKernelGGL(kb_resid_zero, dim3(cyc.n_w), dim3(256), 0, st, dsys.p, cyc.w.p, partW.p);
    } else {
      gemv(cyc, 0, 0, d_arn.p);
    }
    ANDR_HIP(ctx, hipMemsetAsync(ddone.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(kb_restart, dim3(gs), dim3(64), 0, st, dsys.p, nsys, partW.p, cyc.w0.p, cyc.nw.p, tol,
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
    mc = first_cycle ? m : restart_for(cyc_list);
    layout(cyc_list, mc);
    first_cycle = false;
    if ((rc = use_tiles(live))) return rc;
    std::fill(skip_counted.begin(), skip_counted.end(), 0);
    if (blk) {  // Q_0 R0 = R (Cholesky QR of the residual block), g = (R0; 0)
      hipLaunchKernelGGL(kb_bdots, dim3(live.n_t4, 1), dim3(256), 0, st, dsys.p, live.t4.p, 1, 0, 1, partB.p, mc + 2, 1);
      hipLaunchKernelGGL(kb_bdots_final, dim3(nsys, 1), dim3(256), 0, st, dsys.p, partB.p, mc + 2, 1, 1);
      hipLaunchKernelGGL(kb_binit, dim3(nsys), dim3(64), 0, st, dsys.p, mc, L);
      hipLaunchKernelGGL(kb_bupdate, dim3(live.n_t4), dim3(256), 0, st, dsys.p, live.t4.p, 0, 3, 1);
    } else {
      hipLaunchKernelGGL(kb_ginit, dim3(gs), dim3(64), 0, st, dsys.p, nsys, mc);
      hipLaunchKernelGGL(kb_q0, dim3(cyc.n_t), dim3(256), 0, st, dsys.p, cyc.t.p);
    }
    {  // a cycle that starts with few rows (a small batch) takes the tuned column split at once
      const int target = krylov_target(cyc_list, 0);
      if (target < 0) return target;  // a failed timing run
      if (target != MF_TARGET_WAVES) {
        if ((rc = build(cyc_list, live, target))) return rc;
        if ((rc = use_tiles(live))) return rc;
      }
    }
    const int steps_max = std::min(mc, maxit - total);
    int steps = 0;
    for (int j = 0; j < steps_max; ++j) {
      gemv(live, 1, j, d_arn.p);  // W = A Q_j
      // one sweep of dot products and one of updates over the basis (delayed reorthogonalisation)
      double nlive = 0.0;
      for (int s = 0; s < nsys; ++s)
        if (live.ntiles[s]) nlive += in[s].n;
      const double vbytes = nlive * NB * sizeof(double);
      if (blk) {
        // two dot sweeps and two update sweeps over the j + 1 basis blocks (classical Gram-Schmidt
        // twice); the second update also normalises the new block
        const int tok_o = andr_prof_begin(ctx, "gmres_orthogonalise", (4.0 * j + 10.0) * vbytes,
                                          2.0 * 64.0 * (4.0 * j + 5.0) * nlive);
        hipLaunchKernelGGL(kb_bdots, dim3(live.n_t4, (j + 16) / 16), dim3(256), 0, st, dsys.p, live.t4.p, 0, j + 1,
                           0, partB.p, mc + 2, noskip);
        hipLaunchKernelGGL(kb_bdots_final, dim3(nsys, j + 1), dim3(256), 0, st, dsys.p, partB.p, mc + 2, 0, noskip);
        hipLaunchKernelGGL(kb_bupdate, dim3(live.n_t4), dim3(256), 0, st, dsys.p, live.t4.p, j + 1, 0, noskip);
        hipLaunchKernelGGL(kb_bdots, dim3(live.n_t4, (j + 17) / 16), dim3(256), 0, st, dsys.p, live.t4.p, 0, j + 1,
                           1, partB.p, mc + 2, noskip);
        hipLaunchKernelGGL(kb_bdots_final, dim3(nsys, j + 2), dim3(256), 0, st, dsys.p, partB.p, mc + 2, 1, noskip);
        hipLaunchKernelGGL(kb_bsmall, dim3(nsys), dim3(64), 0, st, dsys.p, j, L, tol, noskip);
        hipLaunchKernelGGL(kb_bupdate, dim3(live.n_t4), dim3(256), 0, st, dsys.p, live.t4.p, j + 1, 1, noskip);
        andr_prof_end(ctx, tok_o);
      } else {
        const int tok_o = andr_prof_begin(ctx, "gmres_orthogonalise", (2.0 * j + 6.0) * vbytes,
                                          2.0 * 4.0 * j * nlive * NB);
        hipLaunchKernelGGL(kb_adots, dim3(live.n_t, std::max(1, (j + ICH - 1) / ICH)), dim3(256), 0, st, dsys.p,
                           live.t.p, j, mc, partT.p, 0);
        hipLaunchKernelGGL(kb_adots_final, dim3(nsys, (2 * j + 3 + 7) / 8), dim3(256), 0, st, dsys.p, j, mc, partT.p);
        hipLaunchKernelGGL(kb_arnoldi, dim3(nsys * NB), dim3(64), sizeof(double) * (5 * mc + 2), st, dsys.p, j, mc,
                           tol, 0);
        hipLaunchKernelGGL(kb_aupdate, dim3(live.n_t), dim3(256), 0, st, dsys.p, live.t.p, j, mc);
        andr_prof_end(ctx, tok_o);
      }
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
          else account_skips(s, j);
        }
        if (still.empty()) break;
        int nlive = 0;
        for (int s = 0; s < nsys; ++s) nlive += live.ntiles[s] ? 1 : 0;
        if ((int)still.size() < nlive) {
          const int target = krylov_target(still, j + 1);
          if (target < 0) return target;
          if ((rc = build(still, live, target))) return rc;
          if ((rc = use_tiles(live))) return rc;
        }
      }
    }
    if (!blk && steps == steps_max && live.n_t > 0) {
      // close the cycle: the last column is final once its candidate is reorthogonalised
      hipLaunchKernelGGL(kb_adots, dim3(live.n_t, std::max(1, (steps + ICH - 1) / ICH)), dim3(256), 0, st, dsys.p,
                         live.t.p, steps, mc, partT.p, 1);
      hipLaunchKernelGGL(kb_adots_final, dim3(nsys, (2 * steps + 3 + 7) / 8), dim3(256), 0, st, dsys.p, steps, mc,
                         partT.p);
      hipLaunchKernelGGL(kb_arnoldi, dim3(nsys * NB), dim3(64), sizeof(double) * (5 * mc + 2), st, dsys.p, steps, mc,
                         tol, 1);
    }
    total += steps;
    if (blk) {
      hipLaunchKernelGGL(kb_bsolve, dim3(nsys), dim3(64), 0, st, dsys.p, steps, L);
      hipLaunchKernelGGL(kb_bupdate, dim3(cyc.n_t4), dim3(256), 0, st, dsys.p, cyc.t4.p, steps, 2, 1);
    } else {
      hipLaunchKernelGGL(kb_solve_y, dim3(nsys * NB), dim3(64), sizeof(double) * (mc + 1), st, dsys.p, steps, mc);
      hipLaunchKernelGGL(kb_xupdate, dim3(cyc.n_t), dim3(256), 0, st, dsys.p, cyc.t.p, steps, mc);
    }
    ANDR_HIP(ctx, hipGetLastError());
    if ((rc = read_small())) return rc;
    for (int s : cyc_list)
      if (live.ntiles[s]) account_skips(s, steps - 1);
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
  // hsmall holds the state of the last restart test (or of the last cycle's end): a column still
  // active stopped at the iteration limit; one marked by the floor rule was accepted above tol
  if (status)
    for (int s = 0; s < nsys; ++s)
      for (int k = 0; k < nrhs; ++k) {
        const double *sm = sm_of(s, k);
        status[s * nrhs + k] = sm[4] > 0.0 ? ANDR_GMRES_MAXIT : (sm[7] > 0.0 ? ANDR_GMRES_FLOOR : 0);
      }
  return ANDR_OK;
}
