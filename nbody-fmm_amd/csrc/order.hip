// order.hip -- learned selection of the expansion order per evaluation cell.
//
// A cell that evaluates its local expansion at its particles (L2P) gets its
// own order q in [kOrderLo, p]: its M2L pairs use the leading (q+1)^2 block of
// the operators and its L2P stops at degree q.  Cells that push their local
// to their children (L2L) keep the full order p.  The order comes from a small
// MLP over 16 features of the cell (sizes, masses and the geometry of its
// interaction list; no counterpart in the reference, which has one fixed L,
// tree.cpp:94):
//   16 inputs (normalised by the mean and deviation of the training set)
//   -> 64 ReLU -> 64 ReLU -> kNumOrders logits (orders kOrderLo..kOrderHi).
// Weights are kept in FP32 for training and used as BF16 (round to nearest
// even) in every forward pass; activations and sums are FP32.  A cell takes
// the lowest order whose predicted cumulative probability reaches
// order_conf.
//
// Training data comes from the FMM itself: for one particle set, the FMM is
// run at every order kOrderLo..kOrderHi with the same tree and lists; on
// sampled targets its accelerations are compared with the FP64 direct sum,
// and every evaluation cell gets the label "lowest order whose mean squared
// error over the cell's sampled particles is within tol^2 times the mean
// squared acceleration of all samples" (an equal share of a global relative
// L2 error tol, the measure of the accuracy gate).  Training is
// minibatch Adam on the cross entropy, with hand-written forward, backward,
// weight-gradient and update kernels.
#include "nbfmm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace nbfmm {
namespace {

constexpr int kT = 256;
constexpr int kIn = kMlpIn, kH = kMlpHidden, kOut = kNumOrders;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kOut, oMu = oB3 + kOut, oSd = oMu + kIn;
static_assert(oSd + kIn == kMlpWeights, "model layout");
static_assert(oMu == kMlpParams, "trainable parameters come first");
// rows of the per-sample training record: x, a1, a2, d1, d2, d3
constexpr int rX = 0, rA1 = rX + kIn, rA2 = rA1 + kH, rD1 = rA2 + kH, rD2 = rD1 + kH, rD3 = rD2 + kH,
              kRows = rD3 + kOut;
// LDS image of the model: BF16 W1 | W2 | W3, FP32 b1 | b2 | b3 | mu | sd
constexpr int sW1 = 0, sW2 = sW1 + kIn * kH, sW3 = sW2 + kH * kH, kSW = sW3 + kH * kOut;
constexpr int sB1 = 0, sB2 = sB1 + kH, sB3 = sB2 + kH, sMu = sB3 + kOut, sSd = sMu + kIn, kSB = sSd + kIn;

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__host__ __device__ inline uint16_t f2bf16(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u); // round to nearest, ties to even
  return (uint16_t)(u >> 16);
}

__device__ void stage_model(const uint16_t *__restrict__ wb, const float *__restrict__ w, uint16_t *sW, float *sB) {
  for (int e = threadIdx.x; e < kIn * kH; e += blockDim.x) sW[sW1 + e] = wb[oW1 + e];
  for (int e = threadIdx.x; e < kH * kH; e += blockDim.x) sW[sW2 + e] = wb[oW2 + e];
  for (int e = threadIdx.x; e < kH * kOut; e += blockDim.x) sW[sW3 + e] = wb[oW3 + e];
  for (int e = threadIdx.x; e < kH; e += blockDim.x) {
    sB[sB1 + e] = w[oB1 + e];
    sB[sB2 + e] = w[oB2 + e];
  }
  for (int e = threadIdx.x; e < kOut; e += blockDim.x) sB[sB3 + e] = w[oB3 + e];
  for (int e = threadIdx.x; e < kIn; e += blockDim.x) {
    sB[sMu + e] = w[oMu + e];
    sB[sSd + e] = w[oSd + e];
  }
}

// forward pass of one sample: x (normalised input), a1 (first ReLU layer, in
// LDS at a1[j * kT], one column per thread: the second layer walks it with a
// loop index), a2 (second ReLU layer), z (logits)
__device__ __forceinline__ void mlp_forward(const float *__restrict__ f, const uint16_t *sW, const float *sB,
                                            float x[kIn], float *a1, float a2[kH], float z[kOut]) {
#pragma unroll
  for (int i = 0; i < kIn; ++i) x[i] = (f[i] - sB[sMu + i]) / sB[sSd + i];
#pragma unroll
  for (int j = 0; j < kH; ++j) a2[j] = sB[sB1 + j];
#pragma unroll
  for (int i = 0; i < kIn; ++i)
#pragma unroll
    for (int j = 0; j < kH; ++j) a2[j] = fmaf(x[i], bf16f(sW[sW1 + i * kH + j]), a2[j]);
#pragma unroll
  for (int j = 0; j < kH; ++j) a1[j * kT] = fmaxf(a2[j], 0.0f);
#pragma unroll
  for (int j = 0; j < kH; ++j) a2[j] = sB[sB2 + j];
  for (int i = 0; i < kH; ++i) {
    const float v = a1[i * kT];
#pragma unroll
    for (int j = 0; j < kH; ++j) a2[j] = fmaf(v, bf16f(sW[sW2 + i * kH + j]), a2[j]);
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) a2[j] = fmaxf(a2[j], 0.0f);
#pragma unroll
  for (int k = 0; k < kOut; ++k) z[k] = sB[sB3 + k];
#pragma unroll
  for (int i = 0; i < kH; ++i)
#pragma unroll
    for (int k = 0; k < kOut; ++k) z[k] = fmaf(a2[i], bf16f(sW[sW3 + i * kOut + k]), z[k]);
}

__device__ __forceinline__ void softmax(const float z[kOut], float p[kOut]) {
  float zm = z[0];
#pragma unroll
  for (int k = 1; k < kOut; ++k) zm = fmaxf(zm, z[k]);
  float s = 0;
#pragma unroll
  for (int k = 0; k < kOut; ++k) {
    p[k] = expf(z[k] - zm);
    s += p[k];
  }
  const float is = 1.0f / s;
#pragma unroll
  for (int k = 0; k < kOut; ++k) p[k] *= is;
}

// order of every cell: the prediction for evaluation cells (mode_l == 1),
// omin for the others (cells that push their local get theirs from k_ord_push)
__global__ void __launch_bounds__(kT) k_mlp_select(int64_t nc, const int32_t *__restrict__ mode_l,
                                                   const int32_t *__restrict__ lslot, const float *__restrict__ feat,
                                                   const uint16_t *__restrict__ wb, const float *__restrict__ w,
                                                   float conf, int omin, int omax, int8_t *__restrict__ ord) {
  __shared__ uint16_t sW[kSW];
  __shared__ float sB[kSB];
  __shared__ float sA[kH * kT];
  stage_model(wb, w, sW, sB);
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= nc) return;
  int q = omin;
  const int32_t ls = lslot[c];
  if (mode_l[c] == 1 && ls >= 0) {
    float x[kIn], a2[kH], z[kOut], p[kOut];
    mlp_forward(feat + (size_t)ls * kIn, sW, sB, x, sA + threadIdx.x, a2, z);
    softmax(z, p);
    // lowest order whose cumulative probability reaches conf
    float cum = 0;
    bool found = false;
    q = kOrderHi;
#pragma unroll
    for (int k = 0; k < kOut; ++k) {
      cum += p[k];
      if (!found && cum >= conf) {
        q = kOrderLo + k;
        found = true;
      }
    }
    q = min(max(q, omin), omax);
  }
  ord[c] = (int8_t)q;
}

// a cell that pushes its local to its children (mode_l == 2; all its
// children hold a local) takes the highest order among them; cells of one
// level, children first
__global__ void k_ord_push(int32_t cb, int32_t ce, const int32_t *__restrict__ mode_l,
                           const int32_t *__restrict__ child8, int8_t *__restrict__ ord) {
  const int32_t c = cb + (int32_t)(blockIdx.x * kT + threadIdx.x);
  if (c >= ce || mode_l[c] != 2) return;
  int q = ord[c];
  for (int k = 0; k < 8; ++k) {
    const int32_t ch = child8[(int64_t)c * 8 + k];
    if (ch >= 0) q = max(q, (int)ord[ch]);
  }
  ord[c] = (int8_t)q;
}

__global__ void k_ord_max(int64_t nc, const int32_t *__restrict__ mode_l, const int8_t *__restrict__ ord,
                          int *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  int q = (c < nc && mode_l[c] > 0) ? (int)ord[c] : 0;
  for (int o = 32; o > 0; o >>= 1) q = max(q, __shfl_xor(q, o, 64));
  if ((threadIdx.x & 63) == 0 && q > 0) atomicMax(out, q);
}

__global__ void __launch_bounds__(kT) k_mlp_probs(int64_t nb, const float *__restrict__ feat,
                                                  const uint16_t *__restrict__ wb, const float *__restrict__ w,
                                                  float *__restrict__ prob) {
  __shared__ uint16_t sW[kSW];
  __shared__ float sB[kSB];
  __shared__ float sA[kH * kT];
  stage_model(wb, w, sW, sB);
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (b >= nb) return;
  float x[kIn], a2[kH], z[kOut], p[kOut];
  mlp_forward(feat + (size_t)b * kIn, sW, sB, x, sA + threadIdx.x, a2, z);
  softmax(z, p);
#pragma unroll
  for (int k = 0; k < kOut; ++k) prob[(size_t)b * kOut + k] = p[k];
}

// forward and backward of the cross entropy (mean over the nb samples) for
// samples rows[b] (rows == null: b): the record of every sample goes to
// act[r * nb + b], its loss to loss[b]
__global__ void __launch_bounds__(kT) k_mlp_fb(int nb, const int32_t *__restrict__ rows,
                                               const float *__restrict__ feat, const int8_t *__restrict__ label,
                                               const uint16_t *__restrict__ wb, const float *__restrict__ w,
                                               float *__restrict__ act, float *__restrict__ loss) {
  __shared__ uint16_t sW[kSW];
  __shared__ float sB[kSB];
  __shared__ float sA[kH * kT];
  stage_model(wb, w, sW, sB);
  __syncthreads();
  const int b = (int)blockIdx.x * kT + threadIdx.x;
  if (b >= nb) return;
  const int64_t row = rows ? rows[b] : b;
  const int y = label[row];
  float *a1 = sA + threadIdx.x;
  float x[kIn], a2[kH], z[kOut], p[kOut];
  mlp_forward(feat + (size_t)row * kIn, sW, sB, x, a1, a2, z);
  softmax(z, p);
  loss[b] = -logf(fmaxf(p[y], 1e-30f));
  const float ib = 1.0f / (float)nb;
  float d3[kOut];
#pragma unroll
  for (int k = 0; k < kOut; ++k) d3[k] = (p[k] - (k == y ? 1.0f : 0.0f)) * ib;
#pragma unroll
  for (int i = 0; i < kIn; ++i) act[(size_t)(rX + i) * nb + b] = x[i];
  for (int j = 0; j < kH; ++j) act[(size_t)(rA1 + j) * nb + b] = a1[j * kT];
#pragma unroll
  for (int j = 0; j < kH; ++j) act[(size_t)(rA2 + j) * nb + b] = a2[j];
#pragma unroll
  for (int k = 0; k < kOut; ++k) act[(size_t)(rD3 + k) * nb + b] = d3[k];
  // d2 = (W3 d3) * [a2 > 0], kept in a2's registers
#pragma unroll
  for (int i = 0; i < kH; ++i) {
    float s = 0;
#pragma unroll
    for (int k = 0; k < kOut; ++k) s = fmaf(bf16f(sW[sW3 + i * kOut + k]), d3[k], s);
    a2[i] = a2[i] > 0.0f ? s : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) act[(size_t)(rD2 + j) * nb + b] = a2[j];
  // d1 = (W2 d2) * [a1 > 0]
  for (int i = 0; i < kH; ++i) {
    float s = 0;
#pragma unroll
    for (int j = 0; j < kH; ++j) s = fmaf(bf16f(sW[sW2 + i * kH + j]), a2[j], s);
    act[(size_t)(rD1 + i) * nb + b] = a1[i * kT] > 0.0f ? s : 0.0f;
  }
}

// gradient of every trainable parameter: sum over the samples of the
// products of its two record rows (a bias: of its delta row), in sample order
__global__ void __launch_bounds__(kT) k_mlp_wgrad(int nb, const float *__restrict__ act, float *__restrict__ g) {
  const int t = (int)blockIdx.x * kT + threadIdx.x;
  if (t >= kMlpParams) return;
  int r1, r2;
  if (t < oB1) {
    r1 = rX + (t - oW1) / kH;
    r2 = rD1 + (t - oW1) % kH;
  } else if (t < oW2) {
    r1 = -1;
    r2 = rD1 + (t - oB1);
  } else if (t < oB2) {
    r1 = rA1 + (t - oW2) / kH;
    r2 = rD2 + (t - oW2) % kH;
  } else if (t < oW3) {
    r1 = -1;
    r2 = rD2 + (t - oB2);
  } else if (t < oB3) {
    r1 = rA2 + (t - oW3) / kOut;
    r2 = rD3 + (t - oW3) % kOut;
  } else {
    r1 = -1;
    r2 = rD3 + (t - oB3);
  }
  const float *v2 = act + (size_t)r2 * nb;
  float s = 0;
  if (r1 < 0) {
    for (int b = 0; b < nb; ++b) s += v2[b];
  } else {
    const float *v1 = act + (size_t)r1 * nb;
    for (int b = 0; b < nb; ++b) s = fmaf(v1[b], v2[b], s);
  }
  g[t] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 weights; the BF16 image follows
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       uint16_t *__restrict__ wb, float lr, float c1, float c2) {
  const int t = (int)blockIdx.x * kT + threadIdx.x;
  if (t >= kMlpParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt, vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  const float wt = w[t] - lr * (mt * c1) / (sqrtf(vt * c2) + eps);
  w[t] = wt;
  wb[t] = f2bf16(wt);
}

__global__ void k_to_bf16(const float *__restrict__ w, int n, uint16_t *__restrict__ wb) {
  const int t = (int)blockIdx.x * kT + threadIdx.x;
  if (t < n) wb[t] = f2bf16(w[t]);
}

// out[0] = sum of v[0, n) in a fixed order (one block)
__global__ void __launch_bounds__(kT) k_sum_f32(const float *__restrict__ v, int n, float *__restrict__ out) {
  __shared__ float s[kT];
  float a = 0;
  for (int i = threadIdx.x; i < n; i += kT) a += v[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// mean and deviation of every feature over the training set (block f: feature f)
__global__ void __launch_bounds__(kT) k_feat_stats(int64_t n, const float *__restrict__ feat, float *__restrict__ w) {
  __shared__ double s1[kT], s2[kT];
  const int f = blockIdx.x;
  double a = 0, b = 0;
  for (int64_t i = threadIdx.x; i < n; i += kT) {
    const double x = feat[(size_t)i * kIn + f];
    a += x;
    b += x * x;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mu = n > 0 ? s1[0] / (double)n : 0.0;
    const double var = n > 0 ? s2[0] / (double)n - mu * mu : 1.0;
    w[oMu + f] = (float)mu;
    w[oSd + f] = (float)sqrt(var > 1e-6 ? var : 1e-6);
  }
}

// ---- cell features ------------------------------------------------------------------
// mass of every sorted particle (n + 1 entries; entry n = 0)
__global__ void k_masses(int64_t n, const double4 *__restrict__ pos, double *__restrict__ mv) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i <= n) mv[i] = i < n ? pos[i].w : 0.0;
}

__device__ __forceinline__ float lg(double x) { return (float)log10(x > 1e-300 ? x : 1e-300); }

// largest distance of a particle from the cell centre, in cell sizes rho_l:
// exact for cells of at most kRmaxCells particles, the half diagonal above
constexpr int kRmaxCells = 256;
__global__ void __launch_bounds__(kT) k_rmax(int64_t nc, LevelGeo g, const uint64_t *__restrict__ ckey,
                                             const int32_t *__restrict__ cbeg, const int32_t *__restrict__ cend,
                                             const double4 *__restrict__ pos, float *__restrict__ rmax) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= nc) return;
  const int l = level_of((int32_t)c, g);
  const double ir = g.inv_rho[l];
  const int32_t b = cbeg[c], e = cend[c];
  double r2 = 0;
  if (e - b <= kRmaxCells) {
    double ctr[3];
    cell_centre(ckey[c], l, g, ctr);
    for (int32_t i = b; i < e; ++i) {
      const double4 q = pos[i];
      const double dx = q.x - ctr[0], dy = q.y - ctr[1], dz = q.z - ctr[2];
      r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
    r2 *= ir * ir;
  } else {
    const double hx = g.h[l][0] * ir, hy = g.h[l][1] * ir, hz = g.h[l][2] * ir;
    r2 = 0.25 * (hx * hx + hy * hy + hz * hz);
  }
  rmax[c] = (float)sqrt(r2);
}

// 16 features of every cell with a local, by local slot; S = exclusive prefix
// sums of the sorted masses (S[i] = mass of the first i sorted particles).
// Magnitudes are logs; features 7, 8, 9, 12 (far field of the interaction
// list and its truncation-error proxies) and 11 (near-field mass over the
// cell size squared) are made relative to the system by k_feat_centre.
__global__ void __launch_bounds__(kT) k_features(int64_t nc, int64_t n, LevelGeo g, const uint64_t *__restrict__ ckey,
                                                 const int32_t *__restrict__ cbeg, const int32_t *__restrict__ cend,
                                                 const int32_t *__restrict__ cparent, const int32_t *__restrict__ lslot,
                                                 const int32_t *__restrict__ mode_l, const int32_t *__restrict__ nbr27,
                                                 const int32_t *__restrict__ off_m2l,
                                                 const int32_t *__restrict__ cnt_m2l,
                                                 const int32_t *__restrict__ cnt_sub,
                                                 const int32_t *__restrict__ m2l_src,
                                                 const uint16_t *__restrict__ m2l_cls, const double *__restrict__ S,
                                                 const float *__restrict__ rmax, float *__restrict__ feat) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= nc) return;
  const int32_t ls = lslot[c];
  if (ls < 0) return;
  const double rt = rmax[c];
  const double mtot = S[n] > 0 ? S[n] : 1.0;
  const int l = level_of((int32_t)c, g);
  const double rho = 1.0 / g.inv_rho[l];
  const double hx = g.h[l][0] * g.inv_rho[l], hy = g.h[l][1] * g.inv_rho[l], hz = g.h[l][2] * g.inv_rho[l];
  const int32_t b = cbeg[c], e = cend[c];
  const double M = S[e] - S[b];
  // interaction list (the pairs evaluated by M2L); distances in cell sizes.
  // The M2L of a pair converges like t^(q+1), t = (r_s + r_t) / d with r the
  // largest particle distance from the cell centre: e6 and e10 weigh the
  // far field of every source with t^7 and t^11 (truncation-error proxies).
  double mil = 0, f1 = 0, f2 = 0, e6 = 0, e10 = 0, tmax = 0;
  const int32_t q0 = off_m2l[c], nq = cnt_m2l[c];
  for (int32_t q = q0; q < q0 + nq; ++q) {
    const int32_t src = m2l_src[q];
    const int cls = m2l_cls[q] % kNumClasses;
    const int ddx = cls % 7 - 3, ddy = (cls / 7) % 7 - 3, ddz = cls / 49 - 3;
    const double rx = ddx * hx, ry = ddy * hy, rz = ddz * hz;
    const double d2 = rx * rx + ry * ry + rz * rz;
    const int32_t sb = cbeg[src], se = cend[src];
    const double ms = S[se] - S[sb];
    mil += ms;
    f1 += ms / d2;
    if (max(abs(ddx), max(abs(ddy), abs(ddz))) == 2) f2 += ms / d2;
    const double t = (rt + (double)rmax[src]) / sqrt(d2);
    tmax = fmax(tmax, t);
    const double t2 = t * t, t4 = t2 * t2, t7 = t4 * t2 * t;
    e6 += ms / d2 * t7;
    e10 += ms / d2 * t7 * t4;
  }
  // near field: the 27 neighbours at this level (self included)
  double mnear = 0;
  int nnb = 0;
  for (int o = 0; o < 27; ++o) {
    const int32_t nb = nbr27[c * 27 + o];
    if (nb < 0) continue;
    ++nnb;
    mnear += S[cend[nb]] - S[cbeg[nb]];
  }
  const int32_t par = cparent[c];
  const double ir2 = 1.0 / (rho * rho);
  float *f = feat + (size_t)ls * kIn;
  f[0] = (float)(g.depth - l);
  f[1] = (float)log2((double)(e - b) + 1.0);
  f[2] = lg(M / mtot);
  f[3] = (float)rt;
  f[4] = (float)nq;
  f[5] = (float)cnt_sub[c];
  f[6] = lg(mil / mtot);
  f[7] = lg(f1 * ir2);   // far field of the list, G = 1 units
  f[8] = lg(f2 * ir2);   // its nearest (Chebyshev distance 2) part
  f[9] = lg(e6 * ir2);   // truncation-error proxy at order 6
  f[10] = (float)tmax;    // slowest convergence rate in the list
  f[11] = lg(mnear * ir2); // near-field scale
  f[12] = lg(e10 * ir2);  // truncation-error proxy at order 10
  f[13] = (float)nnb;
  f[14] = f1 > 0 ? (float)(f2 / f1) : 0.0f;
  f[15] = (par >= 0 && mode_l[par] == 2) ? 1.0f : 0.0f;
}

// particle-weighted sums over the evaluation cells of features 7 and 11 (per block)
__global__ void __launch_bounds__(kT) k_feat_part(int64_t nc, const int32_t *__restrict__ cbeg,
                                                  const int32_t *__restrict__ cend, const int32_t *__restrict__ lslot,
                                                  const int32_t *__restrict__ mode_l, const float *__restrict__ feat,
                                                  double *__restrict__ part) {
  __shared__ double s[3][kT];
  double a = 0, b = 0, w = 0;
  for (int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x; c < nc; c += (int64_t)gridDim.x * kT) {
    if (mode_l[c] != 1 || lslot[c] < 0) continue;
    const double nw = (double)(cend[c] - cbeg[c]);
    const float *f = feat + (size_t)lslot[c] * kIn;
    a += nw * f[7];
    b += nw * f[11];
    w += nw;
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  s[2][threadIdx.x] = w;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[blockIdx.x * 3 + k] = s[k][0];
}

__global__ void k_feat_means(int nb, const double *__restrict__ part, double *__restrict__ mean) {
  if (threadIdx.x != 0) return;
  double a = 0, b = 0, w = 0;
  for (int k = 0; k < nb; ++k) {
    a += part[3 * k];
    b += part[3 * k + 1];
    w += part[3 * k + 2];
  }
  mean[0] = w > 0 ? a / w : 0.0;
  mean[1] = w > 0 ? b / w : 0.0;
}

// features 7, 8, 9, 12 relative to the system's mean far field, 11 to its mean near field
__global__ void k_feat_centre(int64_t nl, const double *__restrict__ mean, float *__restrict__ feat) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= nl) return;
  float *f = feat + (size_t)r * kIn;
  const float mf = (float)mean[0], ma = (float)mean[1];
  f[7] -= mf;
  f[8] -= mf;
  f[9] -= mf;
  f[11] -= ma;
  f[12] -= mf;
}

// ---- labels ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
// sample k of a stride sampling: one particle of [k stride, (k+1) stride)
__host__ __device__ inline int64_t sample_of(int64_t k, int64_t stride, int64_t n, uint64_t seed) {
  const int64_t b = k * stride, w = (n - b) < stride ? (n - b) : stride;
  return b + (int64_t)(mix64((uint64_t)k ^ (seed * 0x632be59bd9b4e019ull)) % (uint64_t)w);
}

__global__ void k_sample_targets(int64_t ns, int64_t stride, int64_t n, uint64_t seed, int64_t *__restrict__ tgt) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k < ns) tgt[k] = sample_of(k, stride, n, seed);
}

// for every sampled particle: |a - a_dir|^2 into column qi of every
// evaluation cell on its path to the root, 1 (the sample count) into the last column
__global__ void k_label_add(int64_t n, int64_t stride, uint64_t seed, const int32_t *__restrict__ idx,
                            const int32_t *__restrict__ pcell, const int32_t *__restrict__ cparent,
                            const int32_t *__restrict__ mode_l, const int32_t *__restrict__ lslot,
                            const double *__restrict__ acc, const double *__restrict__ dir, int qi,
                            double *__restrict__ sums) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int64_t in = idx[i], k = in / stride;
  if (sample_of(k, stride, n, seed) != in) return;
  const double dx = acc[3 * in] - dir[3 * k], dy = acc[3 * in + 1] - dir[3 * k + 1], dz = acc[3 * in + 2] - dir[3 * k + 2];
  const double d2 = dx * dx + dy * dy + dz * dz;
  for (int32_t c = pcell[i]; c >= 0; c = cparent[c]) {
    if (mode_l[c] != 1) continue;
    double *r = sums + (size_t)lslot[c] * (kNumOrders + 1);
    atomicAdd(r + qi, d2);
    if (qi == 0) atomicAdd(r + kNumOrders, 1.0);
  }
}

// label of every evaluation cell with samples: the lowest order whose mean
// squared error over the cell's samples is within e2 (tol^2 times the mean
// squared direct acceleration of all samples, so that cells within it keep
// the global relative L2 error within tol); the highest when none is.
// flag = 1 for labelled cells.
__global__ void k_labels(int64_t nc, const int32_t *__restrict__ mode_l, const int32_t *__restrict__ lslot,
                         const double *__restrict__ sums, double e2, int32_t *__restrict__ flag,
                         int8_t *__restrict__ lab) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c > nc) return;
  int f = 0;
  if (c < nc && mode_l[c] == 1 && lslot[c] >= 0) {
    const double *r = sums + (size_t)lslot[c] * (kNumOrders + 1);
    const double cnt = r[kNumOrders];
    if (cnt > 0) {
      int y = kNumOrders - 1;
      for (int k = 0; k < kNumOrders; ++k)
        if (r[k] <= e2 * cnt) {
          y = k;
          break;
        }
      lab[c] = (int8_t)y;
      f = 1;
    }
  }
  flag[c] = f;
}

__global__ void k_label_scatter(int64_t nc, const int32_t *__restrict__ flag, const int32_t *__restrict__ pos,
                                const int32_t *__restrict__ lslot, const int8_t *__restrict__ lab,
                                const float *__restrict__ feat, int64_t base, float *__restrict__ ds_feat,
                                int8_t *__restrict__ ds_lab) {
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= nc || !flag[c]) return;
  const int64_t r = base + pos[c];
  ds_lab[r] = lab[c];
  for (int i = 0; i < kIn; ++i) ds_feat[(size_t)r * kIn + i] = feat[(size_t)lslot[c] * kIn + i];
}

// sum |a - a_dir|^2 and |a_dir|^2 over the sampled targets
__global__ void __launch_bounds__(kT) k_sample_err(int64_t ns, const int64_t *__restrict__ tgt,
                                                   const double *__restrict__ acc, const double *__restrict__ dir,
                                                   double *__restrict__ out) {
  __shared__ double s1[kT], s2[kT];
  double a = 0, b = 0;
  for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < ns; k += (int64_t)gridDim.x * kT) {
    const int64_t in = tgt[k];
    for (int d = 0; d < 3; ++d) {
      const double df = acc[3 * in + d] - dir[3 * k + d];
      a += df * df;
      b += dir[3 * k + d] * dir[3 * k + d];
    }
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(out, s1[0]);
    atomicAdd(out + 1, s2[0]);
  }
}

// device buffer grown with its first `used` bytes kept
void grow_keep(DBuf &b, size_t bytes, size_t used, hipStream_t s) {
  if (b.cap >= bytes && b.p) return;
  void *np = nullptr;
  const size_t want = bytes + bytes / 2 + 256;
  NB_HIP(hipMalloc(&np, want));
  if (b.p && used) NB_HIP(hipMemcpyAsync(np, b.p, used, hipMemcpyDeviceToDevice, s));
  NB_HIP(hipStreamSynchronize(s));
  if (b.p) NB_HIP(hipFree(b.p));
  b.p = np;
  b.cap = want;
}

void need_model(nbfmm_ctx *c) {
  if (!c->mlp_ok) throw Error(NBFMM_ERR_ARG, "no order model: initialise or set the model first");
}

} // namespace

// features of every cell with a local (by local slot) into c->feat
void order_features(nbfmm_ctx *c, const LevelGeo &g) {
  nbr_leaf_rows(c, nullptr); // features read the neighbour rows of every evaluation cell
  Plan &P = c->plan;
  hipStream_t s = c->stream;
  const int64_t n = P.n, nc = P.n_cells;
  double *mv = dget<double>(c->msum, (size_t)(n + 1) + 3 * 256 + 8);
  hipLaunchKernelGGL(k_masses, dim3(nblk(n + 1)), dim3(kT), 0, s, n, (const double4 *)c->pos.p, mv);
  NB_LAUNCH();
  scan_excl_f64(mv, mv, n + 1, c->scan_tmp, s);
  const int64_t nl = std::max<int64_t>(1, c->n_local);
  float *feat = dget<float>(c->feat, (size_t)nl * kIn);
  float *rmax = dget<float>(c->rmax, (size_t)nc + 1);
  hipLaunchKernelGGL(k_rmax, dim3(nblk(nc)), dim3(kT), 0, s, nc, g, (const uint64_t *)c->ckey.p,
                     (const int32_t *)c->cbeg.p, (const int32_t *)c->cend.p, (const double4 *)c->pos.p, rmax);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_features, dim3(nblk(nc)), dim3(kT), 0, s, nc, n, g, (const uint64_t *)c->ckey.p,
                     (const int32_t *)c->cbeg.p, (const int32_t *)c->cend.p, (const int32_t *)c->cparent.p,
                     (const int32_t *)c->lslot.p, (const int32_t *)c->mode_l.p, (const int32_t *)c->nbr27.p,
                     (const int32_t *)c->off_m2l.p, (const int32_t *)c->cnt_m2l.p, (const int32_t *)c->cnt_sub.p,
                     (const int32_t *)c->m2l_src.p, (const uint16_t *)c->m2l_cls.p, (const double *)mv,
                     (const float *)rmax, feat);
  NB_LAUNCH();
  const unsigned npb = std::min<unsigned>(256, nblk(nc));
  double *part = mv + (n + 1);
  double *mean = part + 3 * 256;
  hipLaunchKernelGGL(k_feat_part, dim3(npb), dim3(kT), 0, s, nc, (const int32_t *)c->cbeg.p, (const int32_t *)c->cend.p,
                     (const int32_t *)c->lslot.p, (const int32_t *)c->mode_l.p, (const float *)feat, part);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_feat_means, dim3(1), dim3(64), 0, s, (int)npb, (const double *)part, mean);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_feat_centre, dim3(nblk(c->n_local)), dim3(kT), 0, s, c->n_local, (const double *)mean, feat);
  NB_LAUNCH();
}

// per-cell orders of the current evaluation into c->ord; c->ord_max = the
// highest order of a cell with a local
const int8_t *order_select(nbfmm_ctx *c, const LevelGeo &g) {
  need_model(c);
  order_features(c, g);
  const Plan &P = c->plan;
  hipStream_t s = c->stream;
  const int64_t nc = P.n_cells;
  const int32_t *mode_l = (const int32_t *)c->mode_l.p;
  int8_t *ord = dget<int8_t>(c->ord, (size_t)nc + 1);
  hipLaunchKernelGGL(k_mlp_select, dim3(nblk(nc)), dim3(kT), 0, s, nc, mode_l, (const int32_t *)c->lslot.p,
                     (const float *)c->feat.p, (const uint16_t *)c->mlp_wb.p, (const float *)c->mlp_w.p,
                     (float)c->prm.order_conf, (int)c->prm.order_min, (int)c->prm.order, ord);
  NB_LAUNCH();
  for (int d = P.geo.depth - 1; d >= 0; --d) {
    const int32_t cb = P.lvl_start[d], ce = P.lvl_start[d + 1];
    hipLaunchKernelGGL(k_ord_push, dim3(nblk(ce - cb)), dim3(kT), 0, s, cb, ce, mode_l,
                       (const int32_t *)c->child8.p, ord);
    NB_LAUNCH();
  }
  int *mx = (int *)((double *)dget<double>(c->red, 8 * 1024 + 64) + 6 * 1024 + 60);
  NB_HIP(hipMemsetAsync(mx, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_ord_max, dim3(nblk(nc)), dim3(kT), 0, s, nc, mode_l, (const int8_t *)ord, mx);
  NB_LAUNCH();
  int h = 0;
  NB_HIP(hipMemcpyAsync(&h, mx, sizeof(int), hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
  c->ord_max = std::max(h, (int)c->prm.order_min);
  return ord;
}

void mlp_set(nbfmm_ctx *c, const float *w) {
  float *dw = dget<float>(c->mlp_w, kMlpWeights);
  uint16_t *wb = dget<uint16_t>(c->mlp_wb, kMlpWeights);
  float *m = dget<float>(c->mlp_m, kMlpParams), *v = dget<float>(c->mlp_v, kMlpParams);
  NB_HIP(hipMemcpyAsync(dw, w, sizeof(float) * kMlpWeights, hipMemcpyHostToDevice, c->stream));
  NB_HIP(hipMemsetAsync(m, 0, sizeof(float) * kMlpParams, c->stream));
  NB_HIP(hipMemsetAsync(v, 0, sizeof(float) * kMlpParams, c->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(kMlpWeights)), dim3(kT), 0, c->stream, (const float *)dw, kMlpWeights, wb);
  NB_LAUNCH();
  NB_HIP(hipStreamSynchronize(c->stream));
  c->mlp_ok = true;
  c->mlp_step = 0;
}

void mlp_get(nbfmm_ctx *c, float *w) {
  need_model(c);
  NB_HIP(hipMemcpyAsync(w, c->mlp_w.p, sizeof(float) * kMlpWeights, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
}

// He-uniform weights, zero biases, identity normalisation (set from the data
// by the first training call)
void mlp_init(nbfmm_ctx *c, uint64_t seed) {
  std::vector<float> w(kMlpWeights, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int o, int fan_in, int count) {
    const double a = std::sqrt(6.0 / fan_in);
    std::uniform_real_distribution<double> u(-a, a);
    for (int k = 0; k < count; ++k) w[o + k] = (float)u(rng);
  };
  fill(oW1, kIn, kIn * kH);
  fill(oW2, kH, kH * kH);
  fill(oW3, kH, kH * kOut);
  for (int i = 0; i < kIn; ++i) w[oSd + i] = 1.0f;
  mlp_set(c, w.data());
  c->mlp_fresh = true;
}

void mlp_predict(nbfmm_ctx *c, int64_t nb, const float *feat_h, float *prob_h) {
  need_model(c);
  hipStream_t s = c->stream;
  float *f = dget<float>(c->mlp_act, (size_t)nb * (kIn + kOut));
  float *p = f + (size_t)nb * kIn;
  NB_HIP(hipMemcpyAsync(f, feat_h, sizeof(float) * nb * kIn, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_probs, dim3(nblk(nb)), dim3(kT), 0, s, nb, (const float *)f, (const uint16_t *)c->mlp_wb.p,
                     (const float *)c->mlp_w.p, p);
  NB_LAUNCH();
  NB_HIP(hipMemcpyAsync(prob_h, p, sizeof(float) * nb * kOut, hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
}

// one forward/backward over nb samples given on the host: gradient (kMlpParams) and mean loss
void mlp_grad(nbfmm_ctx *c, int nb, const float *feat_h, const int8_t *lab_h, float *grad_h, float *loss_h) {
  need_model(c);
  hipStream_t s = c->stream;
  float *f = dget<float>(c->mlp_act, (size_t)nb * (kIn + kRows + 1) + kMlpParams + 8);
  float *act = f + (size_t)nb * kIn, *loss = act + (size_t)nb * kRows, *g = loss + nb, *tot = g + kMlpParams;
  int8_t *lab = dget<int8_t>(c->mlp_perm, (size_t)nb);
  NB_HIP(hipMemcpyAsync(f, feat_h, sizeof(float) * nb * kIn, hipMemcpyHostToDevice, s));
  NB_HIP(hipMemcpyAsync(lab, lab_h, (size_t)nb, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_fb, dim3(nblk(nb)), dim3(kT), 0, s, nb, (const int32_t *)nullptr, (const float *)f,
                     (const int8_t *)lab, (const uint16_t *)c->mlp_wb.p, (const float *)c->mlp_w.p, act, loss);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(nblk(kMlpParams)), dim3(kT), 0, s, nb, (const float *)act, g);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(kT), 0, s, (const float *)loss, nb, tot);
  NB_LAUNCH();
  NB_HIP(hipMemcpyAsync(grad_h, g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
  float t = 0;
  NB_HIP(hipMemcpyAsync(&t, tot, sizeof(float), hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
  *loss_h = t / (float)nb;
}

// minibatch Adam over the context's training set; loss_h[e] = mean loss of epoch e
void mlp_train(nbfmm_ctx *c, int epochs, int batch, double lr, uint64_t seed, float *loss_h) {
  need_model(c);
  const int64_t n = c->ds_n;
  if (n < 1) throw Error(NBFMM_ERR_ARG, "the training set is empty");
  if (n > INT32_MAX) throw Error(NBFMM_ERR_UNSUPPORTED, "training set too large");
  hipStream_t s = c->stream;
  float *w = (float *)c->mlp_w.p;
  if (c->mlp_fresh) {
    hipLaunchKernelGGL(k_feat_stats, dim3(kIn), dim3(kT), 0, s, n, (const float *)c->ds_feat.p, w);
    NB_LAUNCH();
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(kMlpWeights)), dim3(kT), 0, s, (const float *)w, kMlpWeights,
                       (uint16_t *)c->mlp_wb.p);
    NB_LAUNCH();
    c->mlp_fresh = false;
  }
  const int B = std::max(1, std::min<int>(batch, (int)n));
  const int nbat = (int)((n + B - 1) / B);
  float *act = dget<float>(c->mlp_act, (size_t)B * (kRows + 1) + kMlpParams + (size_t)nbat + 8);
  float *loss = act + (size_t)B * kRows, *g = loss + B, *lsum = g + kMlpParams;
  int32_t *perm = dget<int32_t>(c->mlp_perm, (size_t)n);
  std::vector<int32_t> hp((size_t)n);
  std::vector<float> hl((size_t)nbat);
  std::mt19937_64 rng(seed);
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = 0; i < n; ++i) hp[i] = (int32_t)i;
    std::shuffle(hp.begin(), hp.end(), rng);
    NB_HIP(hipMemcpyAsync(perm, hp.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    for (int bt = 0; bt < nbat; ++bt) {
      const int b0 = bt * B, nb = (int)std::min<int64_t>(B, n - b0);
      hipLaunchKernelGGL(k_mlp_fb, dim3(nblk(nb)), dim3(kT), 0, s, nb, (const int32_t *)perm + b0,
                         (const float *)c->ds_feat.p, (const int8_t *)c->ds_label.p, (const uint16_t *)c->mlp_wb.p,
                         (const float *)w, act, loss);
      NB_LAUNCH();
      hipLaunchKernelGGL(k_mlp_wgrad, dim3(nblk(kMlpParams)), dim3(kT), 0, s, nb, (const float *)act, g);
      NB_LAUNCH();
      ++c->mlp_step;
      const double t = (double)c->mlp_step;
      const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, t))), c2 = (float)(1.0 / (1.0 - std::pow(0.999, t)));
      hipLaunchKernelGGL(k_adam, dim3(nblk(kMlpParams)), dim3(kT), 0, s, w, (const float *)g, (float *)c->mlp_m.p,
                         (float *)c->mlp_v.p, (uint16_t *)c->mlp_wb.p, (float)lr, c1, c2);
      NB_LAUNCH();
      hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(kT), 0, s, (const float *)loss, nb, lsum + bt);
      NB_LAUNCH();
    }
    NB_HIP(hipMemcpyAsync(hl.data(), lsum, sizeof(float) * nbat, hipMemcpyDeviceToHost, s));
    NB_HIP(hipStreamSynchronize(s));
    double tot = 0;
    for (int bt = 0; bt < nbat; ++bt) tot += hl[bt];
    if (loss_h) loss_h[ep] = (float)(tot / (double)n);
  }
}

// ---- training data ----------------------------------------------------------------------
// sampled targets (one per `stride` particles) and their FP64 direct
// accelerations (compensated, no softening); clears the label sums
void label_begin(nbfmm_ctx *c, int64_t n, int64_t stride, uint64_t seed, const double *d_xyz, const double *d_m) {
  hipStream_t s = c->stream;
  const int64_t ns = (n + stride - 1) / stride;
  c->lab_ns = ns;
  c->lab_stride = stride;
  c->lab_seed = seed;
  int64_t *tgt = dget<int64_t>(c->lab_tgt, (size_t)ns);
  hipLaunchKernelGGL(k_sample_targets, dim3(nblk(ns)), dim3(kT), 0, s, ns, stride, n, seed, tgt);
  NB_LAUNCH();
  double *dir = dget<double>(c->lab_dir, (size_t)ns * 3);
  const double eps2 = c->prm.eps2;
  c->prm.eps2 = 0.0;
  try {
    run_direct(c, n, d_xyz, d_m, ns, tgt, 1, dir);
  } catch (...) {
    c->prm.eps2 = eps2;
    throw;
  }
  c->prm.eps2 = eps2;
  c->lab_nl = -1;
}

// after an evaluation at order kOrderLo + qi (accelerations d_acc, input order)
void label_add(nbfmm_ctx *c, int qi, const double *d_acc) {
  Plan &P = c->plan;
  hipStream_t s = c->stream;
  const int64_t nl = std::max<int64_t>(1, c->n_local);
  if (c->lab_nl < 0) {
    c->lab_nl = nl;
    double *sums = dget<double>(c->lab_sum, (size_t)nl * (kNumOrders + 1));
    NB_HIP(hipMemsetAsync(sums, 0, sizeof(double) * nl * (kNumOrders + 1), s));
  } else if (c->lab_nl != nl) {
    throw Error(NBFMM_ERR_HIP, "label runs built different trees");
  }
  hipLaunchKernelGGL(k_label_add, dim3(nblk(P.n)), dim3(kT), 0, s, P.n, c->lab_stride, c->lab_seed,
                     (const int32_t *)c->idx[0].p, (const int32_t *)c->pcell.p, (const int32_t *)c->cparent.p,
                     (const int32_t *)c->mode_l.p, (const int32_t *)c->lslot.p, d_acc, (const double *)c->lab_dir.p,
                     qi, (double *)c->lab_sum.p);
  NB_LAUNCH();
}

// labels of the evaluation cells of the last evaluation's tree, appended
// with their features to the training set; returns the number added
int64_t label_finish(nbfmm_ctx *c, double tol) {
  Plan &P = c->plan;
  hipStream_t s = c->stream;
  if (c->lab_nl < 0 || P.n_tgt == 0) return 0;
  // mean squared direct acceleration over all samples (host, fixed order)
  std::vector<double> dir((size_t)c->lab_ns * 3);
  NB_HIP(hipMemcpyAsync(dir.data(), c->lab_dir.p, sizeof(double) * dir.size(), hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
  double a2 = 0;
  for (double v : dir) a2 += v * v;
  a2 /= (double)std::max<int64_t>(1, c->lab_ns);
  const LevelGeo g = level_geo(P);
  order_features(c, g);
  const int64_t nc = P.n_cells;
  int32_t *flag = dget<int32_t>(c->lab_flag, (size_t)2 * (nc + 1));
  int32_t *pos = flag + nc + 1;
  int8_t *lab = dget<int8_t>(c->lab_lab, (size_t)nc + 1);
  hipLaunchKernelGGL(k_labels, dim3(nblk(nc + 1)), dim3(kT), 0, s, nc, (const int32_t *)c->mode_l.p,
                     (const int32_t *)c->lslot.p, (const double *)c->lab_sum.p, tol * tol * a2, flag, lab);
  NB_LAUNCH();
  scan_excl_i32(flag, pos, nc + 1, c->scan_tmp, s);
  int32_t add = 0;
  NB_HIP(hipMemcpyAsync(&add, pos + nc, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
  const int64_t base = c->ds_n;
  grow_keep(c->ds_feat, sizeof(float) * (size_t)(base + add) * kIn + 16, sizeof(float) * (size_t)base * kIn, s);
  grow_keep(c->ds_label, (size_t)(base + add) + 16, (size_t)base, s);
  hipLaunchKernelGGL(k_label_scatter, dim3(nblk(nc)), dim3(kT), 0, s, nc, (const int32_t *)flag,
                     (const int32_t *)pos, (const int32_t *)c->lslot.p, (const int8_t *)lab,
                     (const float *)c->feat.p, base, (float *)c->ds_feat.p, (int8_t *)c->ds_label.p);
  NB_LAUNCH();
  NB_HIP(hipStreamSynchronize(s));
  c->ds_n = base + add;
  c->lab_nl = -1;
  return add;
}

// relative L2 error of d_acc against the FP64 direct sum on ~ntgt sampled targets
double sampled_error(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, const double *d_acc,
                     int64_t ntgt, uint64_t seed) {
  hipStream_t s = c->stream;
  const int64_t stride = std::max<int64_t>(1, n / std::max<int64_t>(1, ntgt));
  label_begin(c, n, stride, seed, d_xyz, d_m);
  double *out = (double *)dget<double>(c->red, 8 * 1024 + 64) + 6 * 1024 + 56;
  NB_HIP(hipMemsetAsync(out, 0, 2 * sizeof(double), s));
  hipLaunchKernelGGL(k_sample_err, dim3(std::min<unsigned>(256, nblk(c->lab_ns))), dim3(kT), 0, s, c->lab_ns,
                     (const int64_t *)c->lab_tgt.p, d_acc, (const double *)c->lab_dir.p, out);
  NB_LAUNCH();
  double h[2];
  NB_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
  return h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
}

void dataset_get(nbfmm_ctx *c, float *feat_h, int8_t *lab_h, int64_t cap) {
  const int64_t n = std::min<int64_t>(cap, c->ds_n);
  if (n <= 0) return;
  if (feat_h)
    NB_HIP(hipMemcpyAsync(feat_h, c->ds_feat.p, sizeof(float) * n * kIn, hipMemcpyDeviceToHost, c->stream));
  if (lab_h) NB_HIP(hipMemcpyAsync(lab_h, c->ds_label.p, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
}

void dataset_put(nbfmm_ctx *c, int64_t n, const float *feat_h, const int8_t *lab_h) {
  hipStream_t s = c->stream;
  const int64_t base = c->ds_n;
  grow_keep(c->ds_feat, sizeof(float) * (size_t)(base + n) * kIn + 16, sizeof(float) * (size_t)base * kIn, s);
  grow_keep(c->ds_label, (size_t)(base + n) + 16, (size_t)base, s);
  NB_HIP(hipMemcpyAsync((float *)c->ds_feat.p + (size_t)base * kIn, feat_h, sizeof(float) * n * kIn,
                        hipMemcpyHostToDevice, s));
  NB_HIP(hipMemcpyAsync((int8_t *)c->ds_label.p + base, lab_h, (size_t)n, hipMemcpyHostToDevice, s));
  NB_HIP(hipStreamSynchronize(s));
  c->ds_n = base + n;
}

} // namespace nbfmm
