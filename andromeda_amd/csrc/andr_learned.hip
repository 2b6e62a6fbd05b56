// Learned Gauss-order selection (no counterpart in the reference, which integrates every regular
// pair with one fixed rule, src/lib_Stokes.f90:2341-2446).
//
// A small network (6 -> 64 -> 64 -> 64 -> 8, ReLU) predicts, for a near-field (collocation point,
// element) pair, the cheapest triangle rule that reproduces the fixed-order integral: classes
// 0..6 are the Dunavant rules with 3, 4, 6, 7, 12, 13 and 19 points, class 7 the mesh's own
// fixed-order rule. Far pairs take a rule from a distance table calibrated on the same labels.
// Labels are produced here from FP64 integrals (the fixed-order rule is the reference), the network
// is trained here (forward, backward and Adam kernels), and inference runs in a classification
// pass before the assembly with BF16 weights in LDS and FP32 accumulation.
#include "andr_learned.h"
#include "andr_pairs.h"
#include "andr_prof.h"
#include "andr_singular.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

namespace {

constexpr int NIN = 6, NH = 64, NOUT = 8;
constexpr int P_W1 = 0, P_B1 = P_W1 + NIN * NH, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH, P_W3 = P_B2 + NH,
              P_B3 = P_W3 + NH * NH, P_W4 = P_B3 + NH, P_B4 = P_W4 + NH * NOUT, NPARAM = P_B4 + NOUT;
static_assert(NPARAM == ANDR_MLP_NPARAM, "parameter layout");
constexpr int NCLS_DUN = 7;   // Dunavant classes; class 7 = fixed-order rule
constexpr int DUN_ID0 = 6;    // rule id of the first Dunavant rule in the quadrature table

// ---- pair geometry and features ---------------------------------------------------------------
struct PairGeo {
  double dh;      // distance from the collocation point to the element centroid / longest edge
  double aspect;  // longest edge^2 / (2 area)
  double cosn;    // cosine of the angle between the two element normals
};

__device__ inline PairGeo pair_geo(const double s[3], const double ns[3], const double *x1, const double *x2,
                                   const double *x3, const double ne[3], double area) {
  PairGeo p;
  const double c[3] = {(x1[0] + x2[0] + x3[0]) / 3.0, (x1[1] + x2[1] + x3[1]) / 3.0, (x1[2] + x2[2] + x3[2]) / 3.0};
  const double d = sqrt((s[0] - c[0]) * (s[0] - c[0]) + (s[1] - c[1]) * (s[1] - c[1]) + (s[2] - c[2]) * (s[2] - c[2]));
  auto l2 = [](const double *a, const double *b) {
    return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
  };
  const double L2 = fmax(l2(x1, x2), fmax(l2(x2, x3), l2(x3, x1)));
  p.dh = d / sqrt(L2);
  p.aspect = L2 / (2.0 * area);
  p.cosn = ns[0] * ne[0] + ns[1] * ne[1] + ns[2] * ne[2];
  return p;
}

__device__ inline void features(const PairGeo &p, int kernel, float f[NIN]) {
  f[0] = logf((float)p.dh);  // the network takes FP32 inputs; FP32 logarithms are as good here
  f[1] = logf((float)p.aspect);
  f[2] = (float)p.cosn;
  f[3] = kernel == 0 ? 1.f : 0.f;
  f[4] = kernel == 1 ? 1.f : 0.f;
  f[5] = kernel == 2 ? 1.f : 0.f;
}

__device__ inline int far_bin(double dh, double near_ratio) {
  int b = (int)floor(2.0 * log2(dh / near_ratio));
  return b < 0 ? 0 : (b > ANDR_FAR_BINS - 1 ? ANDR_FAR_BINS - 1 : b);
}

// Laplace pair in FP64: out = {G, H1, H2, H3} (BemKernel, r = s - f)
__device__ inline void lap_pair(const double s[3], const double *x1, const double *x2, const double *x3,
                                const double n[3], double area, int qoff, int nq, double out[4]) {
  double g = 0, h1 = 0, h2 = 0, h3 = 0;
  for (int q = 0; q < nq; ++q) {
    const double l1 = c_qd[qoff + q][0], l2 = c_qd[qoff + q][1], l3 = c_qd[qoff + q][2], w = c_qd[qoff + q][3];
    const double rx = s[0] - (l1 * x1[0] + l2 * x2[0] + l3 * x3[0]);
    const double ry = s[1] - (l1 * x1[1] + l2 * x2[1] + l3 * x3[1]);
    const double rz = s[2] - (l1 * x1[2] + l2 * x2[2] + l3 * x3[2]);
    const double ira = rsqrt(rx * rx + ry * ry + rz * rz);
    const double fg = (0.25 / ANDR_PI) * ira;
    const double fh = (rx * n[0] + ry * n[1] + rz * n[2]) * ira * ira * fg;
    g += w * fg;
    h1 += w * fh * l1;
    h2 += w * fh * l2;
    h3 += w * fh * l3;
  }
  out[0] = g * area;
  out[1] = h1 * area;
  out[2] = h2 * area;
  out[3] = h3 * area;
}

// values of one kernel family with one rule: Laplace 4, Stokes-G 6, Stokes-H 18
__device__ inline int kernel_values(int kernel, const double s[3], const double *x1, const double *x2,
                                    const double *x3, const double n[3], double area, int qoff, int nq,
                                    double v[18]) {
  if (kernel == 0) {
    lap_pair(s, x1, x2, x3, n, area, qoff, nq, v);
    return 4;
  }
  double G[6], T[6][3];
  stokes_pair_regular<double>(s, x1, x2, x3, n[0], n[1], n[2], area, qoff, nq, G, T);
  if (kernel == 1) {
    for (int i = 0; i < 6; ++i) v[i] = G[i];
    return 6;
  }
  for (int k = 0; k < 6; ++k)
    for (int t = 0; t < 3; ++t) v[3 * k + t] = T[k][t];
  return 18;
}

// smallest Dunavant class whose values agree with the fixed-order values to rel_tol, else 7
__device__ inline int pair_label(int kernel, const double s[3], const double *x1, const double *x2, const double *x3,
                                 const double n[3], double area, int qoff_ref, int nq_ref, const int *dun_off,
                                 const int *dun_n, double rel_tol) {
  double ref[18], v[18];
  const int nv = kernel_values(kernel, s, x1, x2, x3, n, area, qoff_ref, nq_ref, ref);
  double scale = 0.0;
  for (int i = 0; i < nv; ++i) scale = fmax(scale, fabs(ref[i]));
  for (int c = 0; c < NCLS_DUN; ++c) {
    kernel_values(kernel, s, x1, x2, x3, n, area, dun_off[c], dun_n[c], v);
    double err = 0.0;
    for (int i = 0; i < nv; ++i) err = fmax(err, fabs(v[i] - ref[i]));
    if (err <= rel_tol * scale) return c;
  }
  return NCLS_DUN;
}

struct RuleTab {
  int off[8], n[8];
};

// deterministic 32-bit hash of a pair (subsampling of the training records)
__device__ inline uint32_t pair_hash(uint32_t r, uint32_t e) {
  uint32_t h = r * 0x9E3779B1u ^ (e + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

// one block per centroid row. mode 0: near pairs with hash <= keep -> (features, label) records,
// three per pair (one per kernel); mode 1: far pairs -> bin maxima; mode 2: count near records.
__global__ __launch_bounds__(256) void k_pair_labels(BodyGeom g, int qoff_ref, int nq_ref, RuleTab rt,
                                                     double near_ratio, double rel_tol, int mode, uint32_t keep,
                                                     int64_t max_pairs, float *__restrict__ feats,
                                                     int *__restrict__ labels,
                                                     unsigned long long *__restrict__ counter,
                                                     int *__restrict__ far_max) {
  const int r = blockIdx.x;
  const int ne = g.ne;
  const double s[3] = {g.cen[3 * r], g.cen[3 * r + 1], g.cen[3 * r + 2]};
  const double ns[3] = {g.nrm[3 * r], g.nrm[3 * r + 1], g.nrm[3 * r + 2]};
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    if (e == r) continue;
    const double *x1 = g.xyz + 3 * g.con[3 * e], *x2 = g.xyz + 3 * g.con[3 * e + 1], *x3 = g.xyz + 3 * g.con[3 * e + 2];
    const double n[3] = {g.nrm[3 * e], g.nrm[3 * e + 1], g.nrm[3 * e + 2]};
    const PairGeo pg = pair_geo(s, ns, x1, x2, x3, n, g.area[e]);
    if (mode == 2 && pg.dh < near_ratio) {
      atomicAdd(counter, 3ull);
    } else if (mode == 0 && pg.dh < near_ratio && pair_hash(r, e) <= keep) {
      for (int kernel = 0; kernel < 3; ++kernel) {
        const int lab = pair_label(kernel, s, x1, x2, x3, n, g.area[e], qoff_ref, nq_ref, rt.off, rt.n, rel_tol);
        const unsigned long long idx = atomicAdd(counter, 1ull);
        if ((int64_t)idx < max_pairs) {
          float f[NIN];
          features(pg, kernel, f);
          for (int i = 0; i < NIN; ++i) feats[idx * NIN + i] = f[i];
          labels[idx] = lab;
        }
      }
    } else if (mode == 1 && pg.dh >= near_ratio && pg.dh < near_ratio * 256.0) {
      const int b = far_bin(pg.dh, near_ratio);
      int lab = pair_label(1, s, x1, x2, x3, n, g.area[e], qoff_ref, nq_ref, rt.off, rt.n, rel_tol);
      lab = max(lab, pair_label(2, s, x1, x2, x3, n, g.area[e], qoff_ref, nq_ref, rt.off, rt.n, rel_tol));
      atomicMax(far_max + b, lab);
    }
  }
}

// ---- small FP32 GEMMs on the matrix cores (v_mfma_f32_16x16x4f32), row-major -------------------
// Lane layout of one 16x16x4 step: lane 16 g + i supplies A[i][g] and B[g][i] and, after the K
// loop, holds C[4 g + v][i], v = 0..3. Inside a 16-deep K block, step t gives lane group g the index
// k0 + 4 g + t on both operands (one permutation of K, so every k is summed once); the four steps
// go to four accumulators, added at the end. Out-of-range rows, columns and k read as 0.
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Z = op(A) op(B) (+ bias[col]); H = relu(Z) if H; Z = 0 where mask <= 0 if mask (ReLU backward,
// mask = the forward pre-activation of the same shape). One wave per 16x16 tile of Z.
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float *__restrict__ A, int lda, int ta,
                                               const float *__restrict__ B, int ldb, int tb, float *__restrict__ Z,
                                               int ldz, const float *__restrict__ bias, float *__restrict__ H,
                                               const float *__restrict__ mask) {
  const int lane = threadIdx.x & 63;
  const int tiles_n = (N + 15) >> 4;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= tiles_n * ((M + 15) >> 4)) return;
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int g = lane >> 4, i = lane & 15;
  const int row = 16 * tm + i, col = 16 * tn + i;
  const bool rok = row < M, cok = col < N;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 + 4 * g + t;
      const bool kok = k < K;
      const float a = (rok && kok) ? (ta ? A[(size_t)k * lda + row] : A[(size_t)row * lda + k]) : 0.f;
      const float b = (cok && kok) ? (tb ? B[(size_t)col * ldb + k] : B[(size_t)k * ldb + col]) : 0.f;
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
    }
  }
  if (!cok) return;
  const float bc = bias ? bias[col] : 0.f;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = 16 * tm + 4 * g + v;
    if (r >= M) break;
    const size_t o = (size_t)r * ldz + col;
    float z = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]) + bc;
    if (mask && !(mask[o] > 0.f)) z = 0.f;
    Z[o] = z;
    if (H) H[o] = fmaxf(z, 0.f);
  }
}

// Weight and bias gradients of one layer: G = A^T D (A: M x NI, D: M x NO, row-major with leading
// dimensions NI and NO; G: NI x NO) and gb = the column sums of D. One 1024-thread block per 16x16
// tile of G: its 16 waves take the samples 16 at a time in turn and their partial tiles are added
// in a fixed order through LDS. The blocks of the first row of tiles (tm = 0) also write gb.
constexpr int WG_WAVES = 16;
__global__ __launch_bounds__(1024) void k_wgrad(int M, int NI, int NO, const float *__restrict__ A,
                                                const float *__restrict__ D, float *__restrict__ G,
                                                float *__restrict__ gb) {
  __shared__ float part[WG_WAVES][5][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tiles_n = (NO + 15) >> 4;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x - tm * tiles_n;
  const int g = lane >> 4, i = lane & 15;
  const int row = 16 * tm + i, col = 16 * tn + i;
  const bool rok = row < NI, cok = col < NO;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float cs = 0.f;  // this lane's share of the column sum of D
  for (int m0 = 16 * w; m0 < M; m0 += 16 * WG_WAVES) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int m = m0 + 4 * g + t;
      const bool mok = m < M;
      const float a = (rok && mok) ? A[(size_t)m * NI + row] : 0.f;
      const float d = (cok && mok) ? D[(size_t)m * NO + col] : 0.f;
      cs += d;
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, d, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) part[w][v][lane] = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
  part[w][4][lane] = cs;
  __syncthreads();
  if (w != 0) return;
  float s[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    s[v] = 0.f;
    for (int u = 0; u < WG_WAVES; ++u) s[v] += part[u][v][lane];
  }
  // the column sum: the four lane groups hold sums over disjoint samples (all lanes take part)
  float c = s[4];
  c += __shfl_xor(c, 16, 64);
  c += __shfl_xor(c, 32, 64);
  if (!cok) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = 16 * tm + 4 * g + v;
    if (r < NI) G[(size_t)r * NO + col] = s[v];
  }
  if (gb && tm == 0 && g == 0) gb[col] = c;
}

// softmax cross-entropy over NOUT logits: P, per-sample loss, dZ = (P - onehot) * scale
__global__ void k_softmax_xent(int M, const float *__restrict__ Z, const int *__restrict__ lab, float *__restrict__ P,
                               float *__restrict__ loss, float *__restrict__ dZ, float scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const float *z = Z + (size_t)i * NOUT;
  float mx = z[0];
  for (int k = 1; k < NOUT; ++k) mx = fmaxf(mx, z[k]);
  float e[NOUT], s = 0.f;
  for (int k = 0; k < NOUT; ++k) {
    e[k] = expf(z[k] - mx);
    s += e[k];
  }
  for (int k = 0; k < NOUT; ++k) {
    const float p = e[k] / s;
    if (P) P[(size_t)i * NOUT + k] = p;
    if (dZ) dZ[(size_t)i * NOUT + k] = (p - (lab && lab[i] == k ? 1.f : 0.f)) * scale;
  }
  if (loss && lab) loss[i] = -(z[lab[i]] - mx - logf(s));
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gg = g[t];
  const float mm = b1 * m[t] + (1.f - b1) * gg;
  const float vv = b2 * v[t] + (1.f - b2) * gg * gg;
  m[t] = mm;
  v[t] = vv;
  p[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_gather(int M, const int *__restrict__ idx, const float *__restrict__ F, const int *__restrict__ L,
                         float *__restrict__ Fo, int *__restrict__ Lo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int j = idx[i];
  for (int k = 0; k < NIN; ++k) Fo[(size_t)i * NIN + k] = F[(size_t)j * NIN + k];
  Lo[i] = L[j];
}

// Buffers of one forward/backward pass on a minibatch of at most B samples.
struct MlpWork {
  DevBuf<float> Z1, H1, Z2, H2, Z3, H3, Z4, P, dZ4, dH, dZb, loss;
  int B = 0;
  int alloc(int b) {
    B = b;
    if (Z1.alloc((size_t)b * NH) || H1.alloc((size_t)b * NH) || Z2.alloc((size_t)b * NH) || H2.alloc((size_t)b * NH) ||
        Z3.alloc((size_t)b * NH) || H3.alloc((size_t)b * NH) || Z4.alloc((size_t)b * NOUT) || P.alloc((size_t)b * NOUT) ||
        dZ4.alloc((size_t)b * NOUT) || dH.alloc((size_t)b * NH) || dZb.alloc((size_t)b * NH) || loss.alloc(b))
      return 1;
    return 0;
  }
};

void sgemm(hipStream_t st, int M, int N, int K, const float *A, int lda, int ta, const float *B, int ldb, int tb,
           float *Z, int ldz, const float *bias = nullptr, float *H = nullptr, const float *mask = nullptr) {
  const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
  hipLaunchKernelGGL(k_sgemm, dim3((tiles + 3) / 4), dim3(256), 0, st, M, N, K, A, lda, ta, B, ldb, tb, Z, ldz, bias, H,
                     mask);
}

void wgrad(hipStream_t st, int M, int NI, int NO, const float *A, const float *D, float *G, float *gb) {
  const int tiles = ((NI + 15) / 16) * ((NO + 15) / 16);
  hipLaunchKernelGGL(k_wgrad, dim3(tiles), dim3(64 * WG_WAVES), 0, st, M, NI, NO, A, D, G, gb);
}

// forward pass of M samples (X: M x NIN) with parameters p: Z_l = H_(l-1) W_l + b_l, H_l = relu(Z_l)
void mlp_forward(hipStream_t st, int M, const float *X, const float *p, MlpWork &w) {
  sgemm(st, M, NH, NIN, X, NIN, 0, p + P_W1, NH, 0, w.Z1.p, NH, p + P_B1, w.H1.p);
  sgemm(st, M, NH, NH, w.H1.p, NH, 0, p + P_W2, NH, 0, w.Z2.p, NH, p + P_B2, w.H2.p);
  sgemm(st, M, NH, NH, w.H2.p, NH, 0, p + P_W3, NH, 0, w.Z3.p, NH, p + P_B3, w.H3.p);
  sgemm(st, M, NOUT, NH, w.H3.p, NH, 0, p + P_W4, NOUT, 0, w.Z4.p, NOUT, p + P_B4);
}

// backward of the mean cross-entropy: grad (NPARAM) from a forward pass already in w
void mlp_backward(hipStream_t st, int M, const float *X, const int *lab, const float *p, MlpWork &w, float *grad) {
  hipLaunchKernelGGL(k_softmax_xent, dim3((M + 255) / 256), dim3(256), 0, st, M, w.Z4.p, lab, w.P.p, w.loss.p,
                     w.dZ4.p, 1.f / (float)M);
  // layer 4, then the gradient through W4 and the ReLU of layer 3
  wgrad(st, M, NH, NOUT, w.H3.p, w.dZ4.p, grad + P_W4, grad + P_B4);
  sgemm(st, M, NH, NOUT, w.dZ4.p, NOUT, 0, p + P_W4, NOUT, 1, w.dH.p, NH, nullptr, nullptr, w.Z3.p);
  // layer 3
  wgrad(st, M, NH, NH, w.H2.p, w.dH.p, grad + P_W3, grad + P_B3);
  sgemm(st, M, NH, NH, w.dH.p, NH, 0, p + P_W3, NH, 1, w.dZb.p, NH, nullptr, nullptr, w.Z2.p);
  // layer 2
  wgrad(st, M, NH, NH, w.H1.p, w.dZb.p, grad + P_W2, grad + P_B2);
  sgemm(st, M, NH, NH, w.dZb.p, NH, 0, p + P_W2, NH, 1, w.dH.p, NH, nullptr, nullptr, w.Z1.p);
  // layer 1
  wgrad(st, M, NIN, NH, X, w.dH.p, grad + P_W1, grad + P_B1);
}

// ---- inference inside the assembly: BF16 matrix cores, FP32 accumulation --------------------------
// The network runs transposed, H^T = W^T X^T: hidden units are the M dimension and pairs the N
// dimension of v_mfma_f32_16x16x16bf16_1k. With that orientation the accumulator tile of one layer
// (lane (g, c) holds rows 4g..4g+3 of a 16-row block for pair c) is, after bias, ReLU and rounding
// to BF16, exactly the B operand (rows 4g..4g+3 of a 16-deep K block, column c) of the next layer:
// the activations never leave the registers. Each wave keeps the 40 weight tiles (A operands) in
// registers, collects the near pairs of its row in an LDS queue and evaluates them 64 at a time
// (4 column blocks of 16 pairs), for the Stokes-G and Stokes-H kernel ids.
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ inline short f2bf(float x) {  // round to nearest even (v_cvt_pk_bf16_f32)
  return __builtin_bit_cast(short, (__bf16)x);
}

constexpr int CLS_THREADS = 256, CLS_WAVES = CLS_THREADS / 64, CLS_Q = 128;

// lane (g, i): A[m = 16 mt + i][k = 16 kt + 4 g + t] = W[in = 16 kt + 4 g + t][out = 16 mt + i]
// (0 outside the layer). The 40 tiles are packed on the host when the model is set, lane by lane in
// operand order: tile index CLS_T_* + ..., then lane, then t (the A operand of one lane).
constexpr int CLS_T_W1 = 0, CLS_T_W2 = 4, CLS_T_W3 = 20, CLS_T_W4 = 36, CLS_NTILES = 40;
// the device model buffer: FP32 parameters, BF16 parameters, far table, rule sizes, then the packed
// tiles (16-byte aligned)
constexpr size_t model_tiles_off() {
  return (NPARAM * (sizeof(uint16_t) + sizeof(float)) + ANDR_FAR_BINS * sizeof(int) + 8 * sizeof(int) + 15) / 16 * 16;
}
static void pack_tile(const uint16_t *w, int nin, int nout, int mt, int kt, uint16_t *out) {
  for (int lane = 0; lane < 64; ++lane) {
    const int g = lane >> 4, i = lane & 15;
    for (int t = 0; t < 4; ++t) {
      const int in = 16 * kt + 4 * g + t, o = 16 * mt + i;
      out[lane * 4 + t] = (in < nin && o < nout) ? w[in * nout + o] : (uint16_t)0;
    }
  }
}

// bias + ReLU + BF16 of an accumulator tile -> B operand of the next layer
__device__ inline bf16x4 act_tile(f32x4 c, const float *bias, int mt, int g) {
  bf16x4 r;
#pragma unroll
  for (int v = 0; v < 4; ++v) r[v] = f2bf(fmaxf(c[v] + bias[16 * mt + 4 * g + v], 0.f));
  return r;
}

// logits of 16 pairs (column block) for one kernel id; returns in lanes g == 0 the class and the
// rule class of pair c (quantile decision, see below)
__device__ inline void mlp_block(const bf16x4 *lt /* this lane's entry of tile 0, in LDS */, const float *sb,
                                 const float *feat /* [16][3] */, int kernel, int g, int c, float min_conf, int *cls,
                                 float *conf) {
  bf16x4 x;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int f = 4 * g + t;  // feature index
    float v = 0.f;
    if (f < 3) v = feat[3 * c + f];
    else if (f == 3 + kernel) v = 1.f;
    x[t] = f2bf(v);
  }
  bf16x4 h1[4], h2[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lt[(CLS_T_W1 + mt) * 64], x, acc, 0, 0, 0);
    h1[mt] = act_tile(acc, sb + 0, mt, g);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lt[(CLS_T_W2 + 4 * mt + kt) * 64], h1[kt], acc, 0, 0, 0);
    h2[mt] = act_tile(acc, sb + NH, mt, g);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lt[(CLS_T_W3 + 4 * mt + kt) * 64], h2[kt], acc, 0, 0, 0);
    h1[mt] = act_tile(acc, sb + 2 * NH, mt, g);
  }
  f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lt[(CLS_T_W4 + kt) * 64], h1[kt], z, 0, 0, 0);
  // classes 4g + v; lanes g = 0 hold classes 0..3 of pair c, lanes g = 1 classes 4..7
  float zz[8];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float mine = z[v] + (g < 2 ? sb[3 * NH + 4 * g + v] : 0.f);
    const float other = __shfl_xor(mine, 16, 64);
    zz[v] = g == 0 ? mine : other;
    zz[4 + v] = g == 0 ? other : mine;
  }
  // The classes are ordered (a rule that is good enough stays good enough for every larger class),
  // so the decision is a quantile of the predicted label distribution: the smallest class whose
  // cumulative probability reaches min_conf. *conf is 1 when that happens below the last class.
  float mx = zz[0];
#pragma unroll
  for (int o = 1; o < NOUT; ++o) mx = fmaxf(mx, zz[o]);
  float e[NOUT], sum = 0.f;
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    e[o] = __expf(zz[o] - mx);  // the quantile decision needs no last-ulp exponentials
    sum += e[o];
  }
  const float target = min_conf * sum;
  float cum = 0.f;
  int pick = NOUT - 1;
  bool found = false;
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    cum += e[o];
    if (!found && cum >= target) {
      pick = o;
      found = true;
    }
  }
  *cls = pick;
  *conf = found ? 1.f : 0.f;
}

// one block per (job, row): rule class of every pair of the row
__global__ __launch_bounds__(CLS_THREADS, 3) void k_classify(const SysJob *__restrict__ jobs,
                                                          const int2 *__restrict__ rowmap,
                                                          const bf16x4 *__restrict__ tiles, const float *__restrict__ wf,
                                                          const int *__restrict__ far_table,
                                                          double near_ratio,
                                                          float min_conf, unsigned long long *__restrict__ stats) {
  __shared__ float sb[3 * NH + NOUT];                 // biases b1, b2, b3, b4
  __shared__ int sfar[ANDR_FAR_BINS];
  __shared__ int qe[CLS_WAVES][CLS_Q];                // queued near pairs: element
  __shared__ float qf[CLS_WAVES][64][3];              // three geometric features of a batch
  __shared__ int qres[CLS_WAVES][64][2];              // class per kernel of a processed batch
  __shared__ unsigned long long red[3];
  __shared__ bf16x4 s_tiles[CLS_NTILES * 64];  // the network's operand tiles, shared by the block's waves
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  for (int i = threadIdx.x; i < NH; i += blockDim.x) {
    sb[i] = wf[P_B1 + i];
    sb[NH + i] = wf[P_B2 + i];
    sb[2 * NH + i] = wf[P_B3 + i];
  }
  if (threadIdx.x < NOUT) sb[3 * NH + threadIdx.x] = wf[P_B4 + threadIdx.x];
  if (threadIdx.x < ANDR_FAR_BINS) sfar[threadIdx.x] = far_table[threadIdx.x];
  if (threadIdx.x < 3) red[threadIdx.x] = 0ull;
  // the operand tiles (packed lane by lane when the model was set) into LDS: the four waves use the
  // same tiles, and registers stay free for occupancy
  for (int i = threadIdx.x; i < CLS_NTILES * 64; i += blockDim.x) s_tiles[i] = tiles[i];
  const bf16x4 *lt = s_tiles + lane;
  __syncthreads();
  const int2 jr = rowmap[blockIdx.x];
  const SysJob &J = jobs[jr.x];
  const BodyGeom &g_ = J.g;
  const int r = jr.y, ne = g_.ne;
  uint8_t *crow = J.cls + (size_t)r * ne;
  const float inv_nr2 = (float)(1.0 / (near_ratio * near_ratio));
  const float4 sr = g_.cenl[r];
  unsigned pts = 0, nl = 0, nf = 0;  // per lane and row: at most 8000 x 126 points
  int qn = 0;  // queue length (wave-uniform)
  // classify the first `cnt` queued pairs (all lanes take part: the matrix cores need a full wave)
  // rule points of a class: Dunavant sizes {3,4,6,7,12,13,19} (6 bits each), the fixed rule's J.nq
  auto npts = [&](int cl) { return cl == NCLS_DUN ? J.nq : (int)((0x1334C1C6103ull >> (6 * cl)) & 63); };
  auto run_batch = [&](int cnt) {
    // the features of the batch's pairs, one per lane (the FP64 pair geometry with every lane busy);
    // unused slots get harmless features
    {
      float f0 = 0.f, f1 = 0.f, f2 = 1.f;
      if (lane < cnt) {
        const int e = qe[wid][lane];
        const double *x1 = g_.xyz + 3 * g_.con[3 * e], *x2 = g_.xyz + 3 * g_.con[3 * e + 1],
                     *x3 = g_.xyz + 3 * g_.con[3 * e + 2];
        const double n[3] = {g_.nrm[3 * e], g_.nrm[3 * e + 1], g_.nrm[3 * e + 2]};
        float f[NIN];
        // the row's centroid and normal, read here (not held in registers through the scan)
        const double s[3] = {g_.cen[3 * r], g_.cen[3 * r + 1], g_.cen[3 * r + 2]};
        const double ns[3] = {g_.nrm[3 * r], g_.nrm[3 * r + 1], g_.nrm[3 * r + 2]};
        features(pair_geo(s, ns, x1, x2, x3, n, g_.area[e]), 0, f);
        f0 = f[0];
        f1 = f[1];
        f2 = f[2];
      }
      qf[wid][lane][0] = f0;
      qf[wid][lane][1] = f1;
      qf[wid][lane][2] = f2;
      __builtin_amdgcn_wave_barrier();
    }
    for (int blk = 0; blk < 4; ++blk) {
      if (16 * blk >= cnt) break;  // uniform
      const float *feat = &qf[wid][16 * blk][0];
      int a, b;
      float ca, cb;  // whether the quantile was reached below the last class (unused here)
      mlp_block(lt, sb, feat, 1, g, c, min_conf, &a, &ca);
      mlp_block(lt, sb, feat, 2, g, c, min_conf, &b, &cb);
      if (g == 0) {
        qres[wid][16 * blk + c][0] = a;
        qres[wid][16 * blk + c][1] = b;
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < cnt) {
      const int a = qres[wid][lane][0], b = qres[wid][lane][1];
      uint8_t cl;
      if (a == NCLS_DUN || b == NCLS_DUN) {
        cl = NCLS_DUN;
        ++nf;
      } else {
        cl = (uint8_t)max(a, b);
        ++nl;
      }
      pts += npts(cl);
      crow[qe[wid][lane]] = cl;
    }
    __builtin_amdgcn_wave_barrier();
  };
  // scan the row (the wave takes every fourth group of 64 elements); near pairs are queued and
  // classified 64 at a time, the rest of the queue once the scan is done (one batch call site)
  int base = wid * 64;
  for (;;) {
    if (base < ne) {
      const int e = base + lane;
      bool near = false;
      if (e < ne) {
        // far test on squared distances from the stored centroid and longest edge (FP32): dh >=
        // near_ratio is d^2 >= near_ratio^2 L^2, and the half-octave bin floor(2 log2(dh / near_ratio))
        // is the binary exponent of d^2 / (near_ratio^2 L^2)
        const float4 ce = g_.cenl[e];
        const float dx = sr.x - ce.x, dy = sr.y - ce.y, dz = sr.z - ce.z;
        const float q = (dx * dx + dy * dy + dz * dz) * ce.w * inv_nr2;
        if (e == r) {
          crow[e] = NCLS_DUN;
        } else if (q >= 1.f) {
          const int cl = sfar[min(ilogbf(q), ANDR_FAR_BINS - 1)];
          crow[e] = (uint8_t)cl;
          pts += npts(cl);
        } else {
          near = true;  // queued: its features are formed in the batch, all lanes busy
        }
      }
      const unsigned long long m = __ballot(near);
      const int pos = qn + __popcll(m & ((1ull << lane) - 1ull));
      if (near) qe[wid][pos] = e;
      qn += __popcll(m);
      base += CLS_THREADS;
      __builtin_amdgcn_wave_barrier();
    }
    const bool scanned = base >= ne;  // uniform over the wave
    if (qn >= 64 || (scanned && qn > 0)) {
      const int cnt = min(qn, 64);
      run_batch(cnt);
      qn -= cnt;
      // move the rest (< 64) to the front
      const int te = lane < qn ? qe[wid][64 + lane] : 0;
      __builtin_amdgcn_wave_barrier();
      if (lane < qn) qe[wid][lane] = te;
      __builtin_amdgcn_wave_barrier();
    } else if (scanned) {
      break;
    }
  }
  atomicAdd(&red[0], (unsigned long long)pts);
  atomicAdd(&red[1], (unsigned long long)nl);
  atomicAdd(&red[2], (unsigned long long)nf);
  __syncthreads();
  if (threadIdx.x < 3) atomicAdd(stats + 3 * (size_t)jr.x + threadIdx.x, red[threadIdx.x]);
}

RuleTab dunavant_tab(int qoff_fixed, int nq_fixed) {
  RuleTab t;
  for (int c = 0; c < NCLS_DUN; ++c) {
    QuadRule q = andr_rule(DUN_ID0 + c);
    t.off[c] = q.offset;
    t.n[c] = q.npts;
  }
  t.off[NCLS_DUN] = qoff_fixed;
  t.n[NCLS_DUN] = nq_fixed;
  return t;
}

}  // namespace

// ---- C-ABI -----------------------------------------------------------------------------------------
static int labels_common(andr_ctx *ctx, const andr_mesh *mesh, double near_ratio, double rel_tol, int mode,
                         int64_t max_pairs, float *features_out, int32_t *labels_out, int64_t *npairs,
                         int32_t *far_table) {
  if (!ctx) return ANDR_ERR_ARG;
  if (!(near_ratio > 0.0) || !(rel_tol > 0.0)) return andr_fail(ctx, ANDR_ERR_ARG, "near_ratio and rel_tol must be > 0");
  if (mesh && mesh->nvert == 4) return andr_fail(ctx, ANDR_ERR_ARG, "quadrature training takes triangle meshes");
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  DevBody b;
  int rc = andr_body_upload(ctx, mesh, b, false);
  if (rc) return rc;
  const QuadRule ref = andr_rule(b.itri);
  RuleTab rt = dunavant_tab(ref.offset, ref.npts);
  DevBuf<float> df;
  DevBuf<int> dl, dfar;
  DevBuf<unsigned long long> cnt;
  const int64_t mp = mode == 0 ? std::max<int64_t>(max_pairs, 1) : 1;
  if (df.alloc((size_t)mp * 6) || dl.alloc(mp) || dfar.alloc(ANDR_FAR_BINS) || cnt.alloc(1))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "training pair buffers");
  uint32_t keep = 0xFFFFFFFFu;
  unsigned long long n = 0;
  if (mode == 0) {  // count first; sample the pairs uniformly when they exceed max_pairs
    ANDR_HIP(ctx, hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), ctx->stream));
    hipLaunchKernelGGL(k_pair_labels, dim3(b.g.ne), dim3(256), 0, ctx->stream, b.g, ref.offset, ref.npts, rt,
                       near_ratio, rel_tol, 2, keep, mp, df.p, dl.p, cnt.p, dfar.p);
    ANDR_HIP(ctx, hipGetLastError());
    ANDR_HIP(ctx, hipMemcpyAsync(&n, cnt.p, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
    ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((double)n > (double)max_pairs)
      keep = (uint32_t)std::min(4294967295.0, 0.97 * (double)max_pairs / (double)n * 4294967296.0);
  }
  ANDR_HIP(ctx, hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), ctx->stream));
  ANDR_HIP(ctx, hipMemsetAsync(dfar.p, 0xff, sizeof(int) * ANDR_FAR_BINS, ctx->stream));  // -1: bin not seen
  hipLaunchKernelGGL(k_pair_labels, dim3(b.g.ne), dim3(256), 0, ctx->stream, b.g, ref.offset, ref.npts, rt, near_ratio,
                     rel_tol, mode, keep, mp, df.p, dl.p, cnt.p, dfar.p);
  ANDR_HIP(ctx, hipGetLastError());
  ANDR_HIP(ctx, hipMemcpyAsync(&n, cnt.p, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (mode == 0) {
    const int64_t k = std::min<int64_t>((int64_t)n, max_pairs);
    if (k > 0) {
      ANDR_HIP(ctx, hipMemcpy(features_out, df.p, sizeof(float) * 6 * k, hipMemcpyDeviceToHost));
      ANDR_HIP(ctx, hipMemcpy(labels_out, dl.p, sizeof(int) * k, hipMemcpyDeviceToHost));
    }
    if (npairs) *npairs = k;
  } else {
    ANDR_HIP(ctx, hipMemcpy(far_table, dfar.p, sizeof(int) * ANDR_FAR_BINS, hipMemcpyDeviceToHost));
  }
  return ANDR_OK;
}

extern "C" int andr_quad_training_pairs(andr_ctx *ctx, const andr_mesh *mesh, double near_ratio, double rel_tol,
                                        int64_t max_pairs, float *features_out, int32_t *labels_out, int64_t *npairs) {
  if (!features_out || !labels_out || max_pairs <= 0) return andr_fail(ctx, ANDR_ERR_ARG, "output buffers");
  return labels_common(ctx, mesh, near_ratio, rel_tol, 0, max_pairs, features_out, labels_out, npairs, nullptr);
}

extern "C" int andr_quad_far_calibration(andr_ctx *ctx, const andr_mesh *mesh, double near_ratio, double rel_tol,
                                         int32_t *far_table) {
  if (!far_table) return andr_fail(ctx, ANDR_ERR_ARG, "far_table");
  return labels_common(ctx, mesh, near_ratio, rel_tol, 1, 0, nullptr, nullptr, nullptr, far_table);
}

extern "C" int andr_mlp_loss_grad(andr_ctx *ctx, int64_t npairs, const float *features, const int32_t *labels,
                                  const float *params, float *loss, float *grad) {
  if (!ctx || npairs <= 0 || npairs > (1 << 24) || !features || !labels || !params || !grad)
    return ctx ? andr_fail(ctx, ANDR_ERR_ARG, "andr_mlp_loss_grad arguments") : ANDR_ERR_ARG;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const int M = (int)npairs;
  for (int64_t i = 0; i < npairs; ++i)
    if (labels[i] < 0 || labels[i] >= NOUT) return andr_fail(ctx, ANDR_ERR_ARG, "label out of range");
  MlpWork w;
  DevBuf<float> X, p, gr;
  DevBuf<int> L;
  if (w.alloc(M) || X.alloc((size_t)M * NIN) || p.alloc(NPARAM) || gr.alloc(NPARAM) || L.alloc(M))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "MLP buffers");
  hipStream_t st = ctx->stream;
  ANDR_HIP(ctx, hipMemcpyAsync(X.p, features, sizeof(float) * M * NIN, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(L.p, labels, sizeof(int) * M, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(p.p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, st));
  mlp_forward(st, M, X.p, p.p, w);
  mlp_backward(st, M, X.p, L.p, p.p, w, gr.p);
  ANDR_HIP(ctx, hipGetLastError());
  std::vector<float> hl(M);
  ANDR_HIP(ctx, hipMemcpyAsync(hl.data(), w.loss.p, sizeof(float) * M, hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipMemcpyAsync(grad, gr.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipStreamSynchronize(st));
  double s = 0.0;
  for (float v : hl) s += v;
  if (loss) *loss = (float)(s / M);
  return ANDR_OK;
}

extern "C" int andr_mlp_infer(andr_ctx *ctx, int64_t npairs, const float *features, const float *params, float *probs) {
  if (!ctx || npairs <= 0 || !features || !params || !probs)
    return ctx ? andr_fail(ctx, ANDR_ERR_ARG, "andr_mlp_infer arguments") : ANDR_ERR_ARG;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  const int chunk = 1 << 16;
  MlpWork w;
  DevBuf<float> X, p;
  if (w.alloc(chunk) || X.alloc((size_t)chunk * NIN) || p.alloc(NPARAM)) return andr_fail(ctx, ANDR_ERR_ALLOC, "MLP buffers");
  hipStream_t st = ctx->stream;
  ANDR_HIP(ctx, hipMemcpyAsync(p.p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, st));
  for (int64_t i0 = 0; i0 < npairs; i0 += chunk) {
    const int M = (int)std::min<int64_t>(chunk, npairs - i0);
    ANDR_HIP(ctx, hipMemcpyAsync(X.p, features + i0 * NIN, sizeof(float) * M * NIN, hipMemcpyHostToDevice, st));
    mlp_forward(st, M, X.p, p.p, w);
    hipLaunchKernelGGL(k_softmax_xent, dim3((M + 255) / 256), dim3(256), 0, st, M, w.Z4.p, (const int *)nullptr, w.P.p,
                       (float *)nullptr, (float *)nullptr, 1.f);
    ANDR_HIP(ctx, hipGetLastError());
    ANDR_HIP(ctx, hipMemcpyAsync(probs + i0 * NOUT, w.P.p, sizeof(float) * M * NOUT, hipMemcpyDeviceToHost, st));
    ANDR_HIP(ctx, hipStreamSynchronize(st));
  }
  return ANDR_OK;
}

extern "C" int andr_mlp_train(andr_ctx *ctx, int64_t npairs, const float *features, const int32_t *labels,
                              int32_t epochs, int32_t batch, float lr, uint64_t seed, float *params, float *loss_out) {
  if (!ctx || npairs <= 0 || npairs > (1 << 26) || !features || !labels || !params || epochs <= 0 || batch <= 0)
    return ctx ? andr_fail(ctx, ANDR_ERR_ARG, "andr_mlp_train arguments") : ANDR_ERR_ARG;
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  for (int64_t i = 0; i < npairs; ++i)
    if (labels[i] < 0 || labels[i] >= NOUT) return andr_fail(ctx, ANDR_ERR_ARG, "label out of range");
  const int N = (int)npairs, B = std::min<int>(batch, N);
  MlpWork w;
  DevBuf<float> X, Xb, p, gr, m, v, lossacc;
  DevBuf<int> L, Lb, idx;
  if (w.alloc(B) || X.alloc((size_t)N * NIN) || Xb.alloc((size_t)B * NIN) || p.alloc(NPARAM) || gr.alloc(NPARAM) ||
      m.alloc(NPARAM) || v.alloc(NPARAM) || L.alloc(N) || Lb.alloc(B) || idx.alloc(N))
    return andr_fail(ctx, ANDR_ERR_ALLOC, "MLP training buffers");
  hipStream_t st = ctx->stream;
  ANDR_HIP(ctx, hipMemcpyAsync(X.p, features, sizeof(float) * (size_t)N * NIN, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(L.p, labels, sizeof(int) * N, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemcpyAsync(p.p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, st));
  ANDR_HIP(ctx, hipMemsetAsync(m.p, 0, sizeof(float) * NPARAM, st));
  ANDR_HIP(ctx, hipMemsetAsync(v.p, 0, sizeof(float) * NPARAM, st));
  std::mt19937_64 rng(seed);
  std::vector<int> perm(N);
  std::vector<float> hl(B);
  int64_t t = 0;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), rng);
    ANDR_HIP(ctx, hipMemcpyAsync(idx.p, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
    double lsum = 0.0;
    int nb = 0;
    for (int i0 = 0; i0 + B <= N; i0 += B) {
      hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256), dim3(256), 0, st, B, idx.p + i0, X.p, L.p, Xb.p, Lb.p);
      mlp_forward(st, B, Xb.p, p.p, w);
      mlp_backward(st, B, Xb.p, Lb.p, p.p, w, gr.p);
      ++t;
      const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, NPARAM, p.p, gr.p, m.p, v.p, lr, b1, b2,
                         eps, c1, c2);
      if ((nb & 15) == 0) {  // sample the loss every 16 minibatches
        ANDR_HIP(ctx, hipMemcpyAsync(hl.data(), w.loss.p, sizeof(float) * B, hipMemcpyDeviceToHost, st));
        ANDR_HIP(ctx, hipStreamSynchronize(st));
        double s = 0.0;
        for (float x : hl) s += x;
        lsum += s / B;
      }
      ++nb;
    }
    ANDR_HIP(ctx, hipGetLastError());
    if (loss_out) loss_out[ep] = (float)(lsum / std::max(1, (nb + 15) / 16));
  }
  ANDR_HIP(ctx, hipMemcpyAsync(params, p.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, st));
  ANDR_HIP(ctx, hipStreamSynchronize(st));
  return ANDR_OK;
}

extern "C" int andr_set_quadrature_model(andr_ctx *ctx, const float *params, float confidence, double near_ratio,
                                         const int32_t *far_table) {
  if (!ctx || !params || !far_table || !(near_ratio > 0.0) || !(confidence >= 0.f && confidence <= 1.f))
    return ctx ? andr_fail(ctx, ANDR_ERR_ARG, "andr_set_quadrature_model arguments") : ANDR_ERR_ARG;
  for (int b = 0; b < ANDR_FAR_BINS; ++b)
    if (far_table[b] < 0 || far_table[b] > NCLS_DUN) return andr_fail(ctx, ANDR_ERR_ARG, "far_table class range");
  ANDR_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<uint16_t> wb(NPARAM);
  for (int i = 0; i < NPARAM; ++i) {  // round to nearest even BF16
    uint32_t u;
    std::memcpy(&u, &params[i], 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    wb[i] = (uint16_t)(u >> 16);
  }
  // packed operand tiles of the classifier after the tables (16-byte aligned)
  const size_t tiles_off = model_tiles_off();
  const size_t bytes = tiles_off + (size_t)CLS_NTILES * 64 * 4 * sizeof(uint16_t);
  std::vector<uint16_t> packed((size_t)CLS_NTILES * 64 * 4);
  for (int mt = 0; mt < 4; ++mt) {
    pack_tile(wb.data() + P_W1, NIN, NH, mt, 0, packed.data() + (size_t)(CLS_T_W1 + mt) * 256);
    pack_tile(wb.data() + P_W4, NH, NOUT, 0, mt, packed.data() + (size_t)(CLS_T_W4 + mt) * 256);
    for (int kt = 0; kt < 4; ++kt) {
      pack_tile(wb.data() + P_W2, NH, NH, mt, kt, packed.data() + (size_t)(CLS_T_W2 + 4 * mt + kt) * 256);
      pack_tile(wb.data() + P_W3, NH, NH, mt, kt, packed.data() + (size_t)(CLS_T_W3 + 4 * mt + kt) * 256);
    }
  }
  // the previous model is gone from here on: a learned call made after a failure below is refused
  ctx->has_model = false;
  if (ctx->d_model) (void)hipFree(ctx->d_model);
  ctx->d_model = nullptr;
  ANDR_HIP(ctx, hipMalloc(&ctx->d_model, bytes));
  char *base = (char *)ctx->d_model;
  float *fp = (float *)base;
  uint16_t *bp = (uint16_t *)(base + NPARAM * sizeof(float));
  int *far = (int *)(base + NPARAM * (sizeof(float) + sizeof(uint16_t)));
  int *npts = far + ANDR_FAR_BINS;
  int hn[8];
  for (int c = 0; c < NCLS_DUN; ++c) hn[c] = andr_rule(DUN_ID0 + c).npts;
  hn[NCLS_DUN] = 0;
  ANDR_HIP(ctx, hipMemcpy(fp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  ANDR_HIP(ctx, hipMemcpy(bp, wb.data(), sizeof(uint16_t) * NPARAM, hipMemcpyHostToDevice));
  ANDR_HIP(ctx, hipMemcpy(far, far_table, sizeof(int) * ANDR_FAR_BINS, hipMemcpyHostToDevice));
  ANDR_HIP(ctx, hipMemcpy(npts, hn, sizeof(hn), hipMemcpyHostToDevice));
  ANDR_HIP(ctx, hipMemcpy(base + tiles_off, packed.data(), sizeof(uint16_t) * packed.size(), hipMemcpyHostToDevice));
  ctx->has_model = true;
  ctx->model_conf = confidence;
  ctx->model_near = near_ratio;
  return ANDR_OK;
}

int andr_launch_stokes_system_learned(andr_ctx *ctx, SysJob *h_jobs, const SysJob *d_jobs, const int2 *d_rowmap,
                                      int nrows, int njobs, bool summed, LearnedStats *stats) {
  if (!ctx->has_model) return andr_fail(ctx, ANDR_ERR_STATE, "no quadrature model");
  char *base = (char *)ctx->d_model;
  const float *fp = (const float *)base;
  const int *far = (const int *)(base + NPARAM * (sizeof(float) + sizeof(uint16_t)));
  const size_t tiles_off = model_tiles_off();
  const bf16x4 *tiles = (const bf16x4 *)(base + tiles_off);
  DevBuf<unsigned long long> st;
  if (st.alloc(3 * (size_t)njobs)) return andr_fail(ctx, ANDR_ERR_ALLOC, "stats");
  ANDR_HIP(ctx, hipMemsetAsync(st.p, 0, 3 * sizeof(unsigned long long) * njobs, ctx->stream));
  const int tok = andr_prof_begin(ctx, "classify_pairs", 0.0, 0.0);
  hipLaunchKernelGGL(k_classify, dim3(nrows), dim3(CLS_THREADS), 0, ctx->stream, d_jobs, d_rowmap, tiles, fp, far,
                     ctx->model_near, ctx->model_conf, st.p);
  andr_prof_end(ctx, tok);
  ANDR_HIP(ctx, hipGetLastError());
  std::vector<unsigned long long> hs(3 * (size_t)njobs);
  ANDR_HIP(ctx, hipMemcpyAsync(hs.data(), st.p, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
  ANDR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // fast-form system: the six symmetric FP32 blocks, 6 F^2 entries of 4 bytes
  double bytes = 0.0, points = 0.0, near = 0.0, cls_bytes = 0.0;
  for (int i = 0; i < njobs; ++i) {
    bytes += 6.0 * h_jobs[i].g.ne * (double)h_jobs[i].g.ne * 4.0;
    cls_bytes += h_jobs[i].g.ne * (double)h_jobs[i].g.ne;  // one class byte per pair
    near += (double)(hs[3 * (size_t)i + 1] + hs[3 * (size_t)i + 2]);
    points += (double)hs[3 * (size_t)i];
    if (stats) {
      stats[i].points = (int64_t)hs[3 * (size_t)i];
      stats[i].learned = (int64_t)hs[3 * (size_t)i + 1];
      stats[i].fallback = (int64_t)hs[3 * (size_t)i + 2];
    }
  }
  if (tok >= 0) {  // the classification's work is known once the counts are back
    // two network passes per near pair (Stokeslet and stresslet), 2 x (6*64 + 64*64 + 64*64 + 64*8)
    // flops each on the BF16 matrix cores; the class bytes written
    ProfRec &r = andr_prof(ctx).pending[tok];
    r.flops = near * 2.0 * ANDR_MLP_FLOPS_PER_PASS;
    r.bytes = cls_bytes;
  }
  return andr_launch_stokes_system(ctx, d_jobs, d_rowmap, nrows, true, true, bytes,
                                   points * (summed ? ANDR_FLOPS_PER_STOKES_EVAL_SUM : ANDR_FLOPS_PER_STOKES_EVAL),
                                   summed);
}

int andr_upload_rules_learned(andr_ctx *ctx, const double *tab, const float *tabf) {
  return andr_upload_rules_here(ctx, tab, tabf);
}
