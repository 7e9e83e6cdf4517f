// m2l.hip -- the M2L translation of the gfx950 FMM on the matrix cores.
//
// With coefficients scaled by the cell scale rho_l, the M2L operator of an
// integer offset class d in [-3,3]^3 is one KP x KP matrix at every level
// (built by k_build_ops in MFMA A-fragment order).  The M2L pairs of all
// levels are sorted by class; a workgroup takes one class and a range of its
// pairs, stages the class operator (row blocks [r0, r0 + RBW)) in LDS once,
// and each wave runs 16-pair tiles: B = the 16 source multipoles,
// acc[r] += T[r] * B, every MFMA column filled.  The partial local of every
// pair goes to Y[pair] (pairs numbered in the target-major list order) and
// k_m2l_sum adds the partials of each target in list order, so the result
// does not depend on the schedule.  When the partials of all pairs would
// exceed the buffer, the class-sorted pairs are taken in passes and every
// pass adds its partials to the locals.
//
// Plain precision runs v_mfma_f64_16x16x4f64 on FP64 coefficients; mixed
// precision takes FP32 coefficients (the rho scaling keeps them O(1)) and
// forms the FP32-accurate contraction from three-term bf16 splits on
// v_mfma_f32_16x16x32_bf16 (k_m2l_bf3); the per-target sum of the partials is
// in FP64 in both.
//
// With an order per pair -- the order of its target cell (learned order
// selection), capped by the order of its offset class (class_order_tol) --
// the pairs are sorted by (order, class): a pair of order q uses the leading
// (q+1)^2 x (q+1)^2 block of the class operator (coefficients of degree <= q
// come first in the storage order), the tile width is the padded block
// KPT = pad16((q+1)^2), source coefficients past (q+1)^2 are masked and the
// sum takes the pair's first (q+1)^2 rows.
#include "nbfmm_internal.h"

#include <algorithm>
#include <vector>

namespace nbfmm {
namespace {

constexpr int kT = 256;
// workgroup size: 8 waves (FP64, 2 per SIMD at ~254 VGPRs) or 16 waves (FP32)
template <class T>
constexpr int m2l_threads() {
  return sizeof(T) == 4 ? 1024 : 512;
}
constexpr int kClsChunk = 4096;


typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <class T>
struct Mfma;
// A[i][k] comes from lane i + 16k, B[k][j] from lane j + 16k (both types).
template <>
struct Mfma<double> {
  typedef dbl4 V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // v_mfma_f64_16x16x4f64: lane l, register q holds D[l/16 + 4q][l%16]
  static __device__ __forceinline__ int row(int kg, int q) { return kg + 4 * q; }
};
template <>
struct Mfma<float> {
  typedef flt4 V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // v_mfma_f32_16x16x4f32: lane l, register q holds D[4(l/16) + q][l%16]
  static __device__ __forceinline__ int row(int kg, int q) { return 4 * kg + q; }
};

// bins (kMaxBins at most): the class, plus kNumClasses * (order - kOrderLo)
// with per-pair orders.  The pair count comes from the plan.
__global__ void __launch_bounds__(kT) k_cls_hist(const uint16_t *__restrict__ cls, const DevPlan *__restrict__ dp,
                                                 int nbins, int32_t *cnt) {
  if (plan_dead(dp)) return;
  const int64_t n = (int64_t)dp->tot[1];
  __shared__ int32_t h[kMaxBins];
  for (int k = threadIdx.x; k < nbins; k += kT) h[k] = 0;
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * kClsChunk, e = min(n, b + kClsChunk);
  for (int64_t i = b + threadIdx.x; i < e; i += kT) atomicAdd(&h[cls[i]], 1);
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += kT)
    if (h[k]) atomicAdd(&cnt[k], h[k]);
}

// cursor[k] starts at the first position of bin k; the order of the pairs
// inside a bin is free (every pair's result is computed independently).
__global__ void __launch_bounds__(kT) k_cls_scatter(const uint16_t *__restrict__ cls, const DevPlan *__restrict__ dp,
                                                    int nbins, int32_t *cursor, int32_t *__restrict__ out,
                                                    int32_t *__restrict__ inv) {
  if (plan_dead(dp)) return;
  const int64_t n = (int64_t)dp->tot[1];
  __shared__ int32_t h[kMaxBins];
  __shared__ int32_t base[kMaxBins];
  for (int k = threadIdx.x; k < nbins; k += kT) h[k] = 0;
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * kClsChunk, e = min(n, b + kClsChunk);
  for (int64_t i = b + threadIdx.x; i < e; i += kT) atomicAdd(&h[cls[i]], 1);
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += kT) {
    base[k] = h[k] ? atomicAdd(&cursor[k], h[k]) : 0;
    h[k] = 0;
  }
  __syncthreads();
  for (int64_t i = b + threadIdx.x; i < e; i += kT) {
    const int k = cls[i];
    const int32_t pos = base[k] + atomicAdd(&h[k], 1);
    out[pos] = (int32_t)i;
    inv[i] = pos;
  }
}

__global__ void k_to_f32(const double *__restrict__ in, int64_t n, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}
// the M2M/L2L operators' FP32 form, unless built by an earlier evaluation
__global__ void k_ops_to_f32(const DevPlan *__restrict__ dp, const double *__restrict__ in, int64_t n,
                             float *__restrict__ out) {
  if (dp->ops_valid) return;
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}
__global__ void k_to_f64(const float *__restrict__ in, int64_t n, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) out[i] = (double)in[i];
}

// ---- mixed-precision M2L as three-term bfloat16 products -------------------------------
// An FP32 value x is split x = x1 + x2 + x3 with x1 = bf16(x), x2 = bf16(x - x1),
// x3 = bf16(x - x1 - x2) (8 significant bits each: x1 + x2 + x3 carries 24
// bits, FP32's).  A product ab is formed as the six terms ai bj with i + j <= 4;
// the dropped terms are below 2^-24 |ab|.  Products of bf16 are exact in FP32
// and v_mfma_f32_16x16x32_bf16 accumulates in FP32, so the contraction keeps
// FP32 accuracy at six bf16 MFMAs per 16x16x32 block instead of eight FP32
// v_mfma_f32_16x16x4f32 (each about 2.7x the time of a bf16 one on gfx950).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int pad32(int k) { return (k + 31) / 32 * 32; }

// Operator fragments for v_mfma_f32_16x16x32_bf16: lane l holds
// A[l % 16][8 (l / 16) + j], j < 8, of a 16 x 32 block.  Layout
// [class][row block r < KP/16][k step < pad32(KP)/32][term < 3][lane] of 8
// bf16 (16 bytes), from the FP64 operators in FP32 MFMA order
// (frag[((class RB + r) KP/4 + k/4) 64 + i + 16 (k % 4)]); columns >= KP are 0.
__global__ void k_to_bf3(const DevPlan *__restrict__ dp, int KP, const double *__restrict__ frag,
                         uint4 *__restrict__ out) {
  if (dp->ops_valid) return; // built by an earlier evaluation (k_ops_check)
  const int RB = KP / 16, KS4 = KP / 4, KSB = pad32(KP) / 32;
  const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; // (class, r, ks, lane)
  const int64_t nlane = (int64_t)kNumClasses * RB * KSB * 64;
  if (e >= nlane) return;
  const int l = (int)(e % 64);
  const int64_t g = e / 64;
  const int ks = (int)(g % KSB);
  const int64_t cr = g / KSB; // class * RB + r
  const int i = l % 16;
  bf16x8 t[3];
  for (int j = 0; j < 8; ++j) {
    const int k = ks * 32 + 8 * (l / 16) + j;
    const double x = k < KP ? frag[((size_t)cr * KS4 + k / 4) * 64 + i + 16 * (k % 4)] : 0.0;
    const __bf16 a1 = (__bf16)(float)x;
    const double r1 = x - (double)(float)a1;
    const __bf16 a2 = (__bf16)(float)r1;
    const double r2 = r1 - (double)(float)a2;
    t[0][j] = a1;
    t[1][j] = a2;
    t[2][j] = (__bf16)(float)r2;
  }
  for (int q = 0; q < 3; ++q) out[(g * 3 + q) * 64 + l] = *reinterpret_cast<const uint4 *>(&t[q]);
}

__device__ __forceinline__ void split3(const float x[8], bf16x8 &b1, bf16x8 &b2, bf16x8 &b3) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 h = (__bf16)x[j];
    const float r1 = x[j] - (float)h; // exact: h is x rounded to 8 bits
    const __bf16 m = (__bf16)r1;
    b1[j] = h;
    b2[j] = m;
    b3[j] = (__bf16)(r1 - (float)m);
  }
}

// k_m2l_cls for FP32 coefficients with the three-term bf16 products; same
// items, partials and tile/column mapping (D row 4 (l / 16) + q, column l % 16).
template <int KP, int RBW>
__global__ void __launch_bounds__(1024) k_m2l_bf3(const DevPlan *__restrict__ dp, const int32_t *__restrict__ irng,
                                                  const int4 *__restrict__ items, int32_t ybase,
                                                  const int32_t *__restrict__ cls_pairs,
                                                  const int32_t *__restrict__ m2l_src,
                                                  const int32_t *__restrict__ src_slot, const uint4 *__restrict__ fragb,
                                                  const float *__restrict__ Mexp, float *__restrict__ Y, int KPS,
                                                  int nvalid) {
  constexpr int RB = KP / 16, KSB = pad32(KP) / 32;
  constexpr int V = KSB * 3 * 64; // 16-byte vectors per staged row block
  extern __shared__ uint4 sA[];
  if (plan_dead(dp) || (int)blockIdx.x >= irng[1]) return; // items [irng[0], + irng[1])
  const int4 it = items[irng[0] + (int)blockIdx.x];
  const int cls = it.x % kNumClasses;
  const int r0 = (int)blockIdx.y * RBW;
  const int nrb = min(RBW, RB - r0);
  {
    const int RBS = KPS / 16, KSBS = pad32(KPS) / 32; // layout of the stored (order-KPS) operators
    const uint4 *A = fragb + (size_t)cls * RBS * KSBS * 192;
    for (int e = threadIdx.x; e < nrb * V; e += 1024) {
      const int r = e / V, v = e - r * V;
      const int ks = v / 192, tl = v - ks * 192;
      sA[e] = A[((size_t)(r0 + r) * KSBS + ks) * 192 + tl];
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, kg = lane >> 4;
  for (int pb = it.y + wave * 16; pb < it.z; pb += 16 * 16) {
    const int pi = pb + col;
    const int32_t pid = pi < it.z ? cls_pairs[pi] : -1;
    const float *Mb = Mexp + (size_t)(pid >= 0 ? src_slot[m2l_src[pid]] : 0) * KPS + 8 * kg;
    flt4 acc[RBW];
#pragma unroll
    for (int r = 0; r < RBW; ++r) acc[r] = flt4{0, 0, 0, 0};
#pragma unroll 1
    for (int ks = 0; ks < KSB; ++ks) {
      const int k0 = ks * 32 + 8 * kg;
      float x[8];
      if (pid >= 0 && k0 < KPS) {
        const float4 u = *reinterpret_cast<const float4 *>(Mb + ks * 32);
        const float4 w = *reinterpret_cast<const float4 *>(Mb + ks * 32 + 4);
        x[0] = u.x, x[1] = u.y, x[2] = u.z, x[3] = u.w, x[4] = w.x, x[5] = w.y, x[6] = w.z, x[7] = w.w;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (k0 + j >= nvalid) x[j] = 0.0f;
      bf16x8 b1, b2, b3;
      split3(x, b1, b2, b3);
#pragma unroll
      for (int r = 0; r < RBW; ++r)
        if (RBW == RB || r < nrb) {
          const uint4 *a = sA + ((r * KSB + ks) * 3) * 64 + lane;
          const uint4 u1 = a[0], u2 = a[64], u3 = a[128];
          const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(&u1), a2 = *reinterpret_cast<const bf16x8 *>(&u2),
                       a3 = *reinterpret_cast<const bf16x8 *>(&u3);
          flt4 c = acc[r];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, b1, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b3, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b1, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b2, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
          acc[r] = c;
        }
    }
    if (pid >= 0) {
      // rows past the pair's (q+1)^2 up to the next multiple of 4 are stored
      // as zeros, so that the sum adds whole 4-row groups; it reads no row
      // past that, so none is stored
      const int nvalid4 = (nvalid + 3) & ~3;
      float *y = Y + (size_t)(pi - ybase) * KPS + r0 * 16;
#pragma unroll
      for (int r = 0; r < RBW; ++r)
        if (r < nrb && (r0 + r) * 16 + 4 * kg < nvalid4)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = (r0 + r) * 16 + 4 * kg + q;
            y[r * 16 + 4 * kg + q] = row < nvalid ? acc[r][q] : 0.0f;
          }
    }
  }
}

// ---- target-major mixed-precision M2L ------------------------------------------------
// Reflections.  The operator of the offset class S d, S a sign change of
// some axes, is E_S T(d) E_S, E_S diagonal +-1 over the coefficients: a
// coefficient of degree n and order m changes sign under x -> -x by (-1)^m
// (once more for an imaginary part), under y -> -y when it is an imaginary
// part, under z -> -z by (-1)^(n+m) (the reflections of the regular and
// singular harmonics).  So 56 operators, one per |d| in {0..3}^3 with
// max >= 2, serve the 316 far classes.
//
// A workgroup owns 128 M2L targets, 16 per wave as the MFMA columns, and
// keeps their locals on chip across every class: it stages each of the 56
// operators (its row blocks [r0, r0 + RBW)) in LDS once, by direct
// global-to-LDS loads, and each wave applies it to the source multipoles of
// its 16 targets under every sign change S, as E_S T (E_S M): the class is
// contracted into FP32 accumulators and added, rows negated by E_S, to the
// targets' FP64 sums (the precision of the class-major per-target sum of
// pair partials).  A table (target, class) -> multipole slot gives the
// sources; a class in none of the wave's 16 lists is skipped, and the
// tile's targets are sorted by child octant first, since siblings of one
// octant have the same far classes.  Only the table, the source multipoles
// and the locals (added to) move through memory: no per-pair partials.  The
// classes are taken in a fixed order and every target is owned by one
// workgroup, so the result does not depend on the schedule.
//
// It is the tile option m2l_mode 2, not the default: an order-11 operator
// fills 138 KB of the 160 KB LDS, so a CU holds one workgroup, two waves per
// SIMD, too few to hide the dependent gathers (table, then multipoles)
// behind a class's MFMAs; the class-major contraction is faster on gfx950.
constexpr int kTgtWaves = 8;
constexpr int kTgtTile = 16 * kTgtWaves; // targets per workgroup

// bit 0: the sign change of storage coefficient k under x -> -x, bit 1 under
// y -> -y, bit 2 under z -> -z (storage: n^2 -> Re X_n^0, n^2+2m-1 ->
// Re X_n^m, n^2+2m -> Im X_n^m)
__device__ __forceinline__ uint32_t coef_sign_code(int k) {
  int n = (int)sqrtf((float)k);
  while (n * n > k) --n;
  while ((n + 1) * (n + 1) <= k) ++n;
  const int rr = k - n * n, m = (rr + 1) >> 1, im = (rr > 0 && (rr & 1) == 0) ? 1 : 0;
  return (uint32_t)(((m + im) & 1) | (im << 1) | (((n + m) & 1) << 2));
}

__device__ __forceinline__ uint64_t sign_pick(int s, uint64_t px, uint64_t py, uint64_t pz) {
  return ((s & 1) ? px : 0ull) ^ ((s & 2) ? py : 0ull) ^ ((s & 4) ? pz : 0ull);
}

__device__ __forceinline__ float neg_if(float x, uint64_t bits, int i) {
  return __uint_as_float(__float_as_uint(x) ^ ((uint32_t)((bits >> i) & 1ull) << 31));
}

// class -> multipole slot of each M2L target's sources: tab[t * kNumClasses +
// class] (cleared to -1 before), one wave per target, a fixed grid striding
// over the targets; each workgroup's algorithmic flops (2 n^2 per pair,
// n = (q+1)^2) and order sum go to part[2 * block] (k_m2l_tab_sum adds them).
constexpr int kTabBlocks = 1024;
__global__ void __launch_bounds__(kT) k_m2l_tab(const DevPlan *__restrict__ dp, int p, int binned,
                                                const int32_t *__restrict__ tgt_cells,
                                                const int32_t *__restrict__ off_m2l,
                                                const int32_t *__restrict__ cnt_m2l,
                                                const int32_t *__restrict__ m2l_src,
                                                const uint16_t *__restrict__ m2l_cls,
                                                const int32_t *__restrict__ mslot, int32_t *__restrict__ tab,
                                                double *__restrict__ part) {
  __shared__ double red[kT / 64][2];
  if (plan_dead(dp)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t nt = dp->nst[1];
  double fl = 0, qs = 0;
  for (int64_t t = (int64_t)blockIdx.x * (kT / 64) + wave; t < nt; t += (int64_t)gridDim.x * (kT / 64)) {
    const int32_t c = tgt_cells[t];
    const int32_t b = off_m2l[c], e = b + cnt_m2l[c];
    for (int32_t i = b + lane; i < e; i += 64) {
      const int cls = (int)m2l_cls[i] % kNumClasses;
      tab[t * kNumClasses + cls] = mslot[m2l_src[i]];
      const int q = binned ? (int)dp->qc.q[cls] : p;
      const double nq = (double)((q + 1) * (q + 1));
      fl += 2.0 * nq * nq;
      qs += q;
    }
  }
  for (int o = 32; o >= 1; o >>= 1) {
    fl += __shfl_xor(fl, o);
    qs += __shfl_xor(qs, o);
  }
  if (lane == 0) {
    red[wave][0] = fl;
    red[wave][1] = qs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int w = 0; w < kT / 64; ++w) {
      a += red[w][0];
      b += red[w][1];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// the sums of k_m2l_tab's kTabBlocks partials into dp->flop_m2l, dp->order_sum
__global__ void __launch_bounds__(kT) k_m2l_tab_sum(DevPlan *__restrict__ dp, const double *__restrict__ part) {
  __shared__ double red[kT / 64][2];
  if (plan_dead(dp)) return;
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < kTabBlocks; i += kT) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  for (int o = 32; o >= 1; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave][0] = a;
    red[wave][1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0;
    for (int w = 0; w < kT / 64; ++w) {
      x += red[w][0];
      y += red[w][1];
    }
    dp->flop_m2l = x;
    dp->order_sum = y;
  }
}

// RBW row blocks per workgroup (slice blockIdx.y), KSBM k steps of the widest
// operator (the expansion's KP).  Class orders: binned, dp->qc (one order per
// |d|: the rule depends on d through |d_i| only); else p for every class.
template <int RBW, int KSBM>
__global__ void __launch_bounds__(64 * kTgtWaves) k_m2l_tgt(const DevPlan *__restrict__ dp, int p, int binned,
                                                            const int32_t *__restrict__ tgt_cells,
                                                            const int32_t *__restrict__ lslot,
                                                            const uint64_t *__restrict__ ckey,
                                                            const int32_t *__restrict__ tab,
                                                            const uint4 *__restrict__ fragb,
                                                            const float *__restrict__ Mexp, float *__restrict__ Lexp,
                                                            int KP) {
  extern __shared__ uint4 sA[];
  __shared__ int32_t perm[kTgtTile];
  __shared__ int32_t boff[9];
  if (plan_dead(dp)) return;
  const int32_t nt = dp->nst[1];
  const int32_t t0 = (int32_t)blockIdx.x * kTgtTile;
  if (t0 >= nt) return;
  const int r0 = (int)blockIdx.y * RBW;
  const int RBS = KP / 16, KSBS = pad32(KP) / 32; // layout of the stored operators
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kg = lane >> 4;
  // the tile's targets sorted by child octant (a counting sort in LDS)
  if (tid < 9) boff[tid] = 0;
  if (tid < kTgtTile) perm[tid] = -1;
  __syncthreads();
  int oct = 8, rank = 0;
  if (tid < kTgtTile && t0 + tid < nt) {
    oct = (int)(ckey[tgt_cells[t0 + tid]] & 7u);
    rank = atomicAdd(&boff[oct + 1], 1);
  }
  __syncthreads();
  if (tid == 0)
    for (int b = 1; b <= 8; ++b) boff[b] += boff[b - 1];
  __syncthreads();
  if (oct < 8) perm[boff[oct] + rank] = t0 + tid;
  __syncthreads();
  const int32_t ti = perm[wave * 16 + col];
  // sign patterns of this lane's accumulator rows and source coefficients
  uint64_t rx = 0, ry = 0, rz = 0, kx = 0, ky = 0, kz = 0;
#pragma unroll
  for (int r = 0; r < RBW; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint64_t cd = coef_sign_code((r0 + r) * 16 + 4 * kg + q);
      const int i = r * 4 + q;
      rx |= (cd & 1) << i;
      ry |= ((cd >> 1) & 1) << i;
      rz |= ((cd >> 2) & 1) << i;
    }
#pragma unroll
  for (int ks = 0; ks < KSBM; ++ks)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint64_t cd = coef_sign_code(ks * 32 + 8 * kg + j);
      const int i = ks * 8 + j;
      kx |= (cd & 1) << i;
      ky |= ((cd >> 1) & 1) << i;
      kz |= ((cd >> 2) & 1) << i;
    }
  // the targets' locals in FP64 (as the class-major sum of FP32 pair
  // partials): each class is contracted into FP32 accumulators from zero and
  // added with the signs of its reflection
  double tot[RBW][4];
#pragma unroll
  for (int r = 0; r < RBW; ++r)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) tot[r][qq] = 0.0;
  bool any = false; // the wave added a class
  for (int ub = 0; ub < 64; ++ub) {
    const int ax = ub & 3, ay = (ub >> 2) & 3, az = ub >> 4;
    if (max(ax, max(ay, az)) < 2) continue;
    const int cb = (az + 3) * 49 + (ay + 3) * 7 + (ax + 3);
    const int q = binned ? (int)dp->qc.q[cb] : p;
    const int nvalid = (q + 1) * (q + 1);
    const int KPT = (nvalid + 15) / 16 * 16, KSB = pad32(KPT) / 32;
    const int nrb = min(RBW, KPT / 16 - r0);
    if (nrb <= 0) continue;
    // the multipole slots of the eight sign changes (in flight during the staging)
    int32_t sl[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bool ok = !(((s & 1) && !ax) || ((s & 2) && !ay) || ((s & 4) && !az));
      const int dx = (s & 1) ? -ax : ax, dy = (s & 2) ? -ay : ay, dz = (s & 4) ? -az : az;
      sl[s] = ok && ti >= 0 ? tab[(size_t)ti * kNumClasses + (dz + 3) * 49 + (dy + 3) * 7 + (dx + 3)] : -1;
    }
    __syncthreads(); // every wave is done with the previous operator
    {
      // the operator's row blocks straight into LDS: chunks of 64 vectors
      // (one k step's term of one row block), one per lane
      const uint4 *A = fragb + (size_t)cb * RBS * KSBS * 192;
      const int nch = nrb * KSB * 3;
      for (int ch = wave; ch < nch; ch += kTgtWaves) {
        const int r = ch / (KSB * 3), rem = ch - r * KSB * 3;
        const int ks = rem / 3, term = rem - ks * 3;
        __builtin_amdgcn_global_load_lds(A + ((size_t)(r0 + r) * KSBS + ks) * 192 + term * 64 + lane,
                                         sA + ((r * KSBM + ks) * 3 + term) * 64, 16, 0, 0);
      }
      __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();
    uint32_t smask = 0; // the wave's classes among the eight (wave-uniform)
#pragma unroll
    for (int s = 0; s < 8; ++s)
      if (__ballot(sl[s] >= 0) != 0) smask |= 1u << s;
    if (smask == 0) continue;
    any = true;
    auto pick = [&](int s) {
      int32_t v = sl[0];
#pragma unroll
      for (int j = 1; j < 8; ++j) v = s == j ? sl[j] : v;
      return v;
    };
    // k step ks of a class's source coefficients for this lane
    auto load = [&](int32_t slot, int ks, float v[8]) {
      const float *Mb = Mexp + (size_t)(slot >= 0 ? slot : 0) * KP + 8 * kg + ks * 32;
      if (slot >= 0 && ks * 32 + 8 * kg < KP) {
        const float4 u = *reinterpret_cast<const float4 *>(Mb);
        const float4 w = *reinterpret_cast<const float4 *>(Mb + 4);
        v[0] = u.x, v[1] = u.y, v[2] = u.z, v[3] = u.w, v[4] = w.x, v[5] = w.y, v[6] = w.z, v[7] = w.w;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = 0.0f;
      }
    };
    // the steps (class s, k step ks) in one sequence, the next step's
    // coefficients in flight while the current one is contracted
    int s = __builtin_ctz(smask), ks = 0;
    float xn[8];
    load(pick(s), 0, xn);
    flt4 acc[RBW];
#pragma unroll
    for (int r = 0; r < RBW; ++r) acc[r] = flt4{0, 0, 0, 0};
    while (true) {
      const uint64_t pk = sign_pick(s, kx, ky, kz);
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        x[j] = ks * 32 + 8 * kg + j < nvalid ? neg_if(xn[j], pk, ks * 8 + j) : 0.0f;
      int s2 = s, ks2 = ks + 1;
      if (ks2 == KSB) {
        const uint32_t later = smask & ~((2u << s) - 1u);
        s2 = later ? __builtin_ctz(later) : -1;
        ks2 = 0;
      }
      if (s2 >= 0) load(pick(s2), ks2, xn);
      bf16x8 b1, b2, b3;
      split3(x, b1, b2, b3);
#pragma unroll
      for (int r = 0; r < RBW; ++r)
        if (r < nrb) {
          const uint4 *a = sA + ((r * KSBM + ks) * 3) * 64 + lane;
          const uint4 u1 = a[0], u2 = a[64], u3 = a[128];
          const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(&u1), a2 = *reinterpret_cast<const bf16x8 *>(&u2),
                       a3 = *reinterpret_cast<const bf16x8 *>(&u3);
          flt4 c = acc[r];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, b1, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b3, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b1, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b2, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
          acc[r] = c;
        }
      if (ks == KSB - 1) {
        // the class is complete: its rows up to the order, with the signs of
        // its reflection, into the FP64 sums
        const uint64_t pm = sign_pick(s, rx, ry, rz);
#pragma unroll
        for (int r = 0; r < RBW; ++r)
          if (r < nrb)
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
              if ((r0 + r) * 16 + 4 * kg + qq < nvalid) tot[r][qq] += (double)neg_if(acc[r][qq], pm, r * 4 + qq);
              acc[r][qq] = 0.0f;
            }
      }
      if (s2 < 0) break;
      s = s2;
      ks = ks2;
    }
  }
  if (!any || ti < 0) return;
  float *L = Lexp + (size_t)lslot[tgt_cells[ti]] * KP + r0 * 16 + 4 * kg;
#pragma unroll
  for (int r = 0; r < RBW; ++r)
    if (r0 + r < RBS) {
      float4 v = *reinterpret_cast<float4 *>(L + r * 16);
      v.x = (float)((double)v.x + tot[r][0]);
      v.y = (float)((double)v.y + tot[r][1]);
      v.z = (float)((double)v.z + tot[r][2]);
      v.w = (float)((double)v.w + tot[r][3]);
      *reinterpret_cast<float4 *>(L + r * 16) = v;
    }
}

// bin of every M2L pair with per-pair orders: class + kNumClasses * (order -
// kOrderLo), order = min(target order (ord, or p when null), class order)
__global__ void k_order_bins(const DevPlan *__restrict__ dp, const int32_t *__restrict__ tgt_cells,
                             const int32_t *__restrict__ off_m2l, const int32_t *__restrict__ cnt_m2l,
                             const int8_t *__restrict__ ord, int p, uint16_t *__restrict__ cls) {
  if (plan_dead(dp)) return;
  __shared__ ClsOrd qc;
  for (int k = threadIdx.x; k < kNumClasses; k += kT) qc.q[k] = dp->qc.q[k];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (t >= dp->nst[1]) return;
  const int32_t c = tgt_cells[t];
  const int ot = ord ? (int)ord[c] : p;
  for (int32_t q = off_m2l[c] + (threadIdx.x & 63), e = off_m2l[c] + cnt_m2l[c]; q < e; q += 64) {
    const int k = cls[q];
    cls[q] = (uint16_t)(k + kNumClasses * (min(ot, (int)qc.q[k]) - kOrderLo));
  }
}

// KP: tile width (padded block of the pair's order); KPS: storage width of
// the operators, multipoles and partials; nvalid: source coefficients used
// ((q+1)^2 for order q; KPS when every pair has the storage order).
template <int KP, int RBW, class T>
__global__ void __launch_bounds__(m2l_threads<T>()) k_m2l_cls(const DevPlan *__restrict__ dp,
                                                         const int32_t *__restrict__ irng,
                                                         const int4 *__restrict__ items, int32_t ybase,
                                                         const int32_t *__restrict__ cls_pairs,
                                                         const int32_t *__restrict__ m2l_src,
                                                         const int32_t *__restrict__ src_slot, const T *__restrict__ frag,
                                                         const T *__restrict__ Mexp, T *__restrict__ Y, int KPS,
                                                         int nvalid) {
  constexpr int RB = KP / 16, KS = KP / 4;
  extern __shared__ double4 s_raw[];
  T *sA = reinterpret_cast<T *>(s_raw);
  if (plan_dead(dp) || (int)blockIdx.x >= irng[1]) return; // items [irng[0], + irng[1])
  const int4 it = items[irng[0] + (int)blockIdx.x];
  const int cls = it.x % kNumClasses;
  const int r0 = (int)blockIdx.y * RBW;
  const int nrb = min(RBW, RB - r0);
  {
    // row blocks r0.. of the class operator, k steps [0, KS) of the KPS / 4 stored
    constexpr int V = (int)(KS * 64 * sizeof(T) / 32); // 32-byte vectors per row block
    const size_t rowv = (size_t)(KPS / 4) * 64 * sizeof(T) / 32;
    const double4 *A = reinterpret_cast<const double4 *>(frag + (size_t)cls * KPS * KPS) + (size_t)r0 * rowv;
    for (int e = threadIdx.x; e < nrb * V; e += m2l_threads<T>()) {
      const int r = e / V, v = e - r * V;
      s_raw[e] = A[(size_t)r * rowv + v];
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, kg = lane >> 4;
  for (int pb = it.y + wave * 16; pb < it.z; pb += (m2l_threads<T>() / 64) * 16) {
    const int pi = pb + col;
    const int32_t pid = pi < it.z ? cls_pairs[pi] : -1;
    const T *Mb = Mexp + (size_t)(pid >= 0 ? src_slot[m2l_src[pid]] : 0) * KPS + kg;
    typename Mfma<T>::V acc[RBW];
#pragma unroll
    for (int r = 0; r < RBW; ++r) acc[r] = typename Mfma<T>::V{0, 0, 0, 0};
    if constexpr (sizeof(T) == 8) {
      T bv[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) bv[ks] = (pid >= 0 && ks * 4 + kg < nvalid) ? Mb[ks * 4] : (T)0;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int r = 0; r < RBW; ++r)
          if (r < nrb) acc[r] = Mfma<T>::op(sA[(r * KS + ks) * 64 + lane], bv[ks], acc[r]);
    } else {
      // k in chunks of 4 with the next chunk's multipole values loaded ahead
      // (a fully unrolled k loop exhausts the registers for KP = 96..128)
      T bc[4], bn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) bc[u] = (pid >= 0 && u * 4 + kg < nvalid) ? Mb[u * 4] : (T)0;
#pragma unroll 1
      for (int k0 = 0; k0 < KS; k0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          bn[u] = (pid >= 0 && k0 + 4 < KS && (k0 + 4 + u) * 4 + kg < nvalid) ? Mb[(k0 + 4 + u) * 4] : (T)0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int r = 0; r < RBW; ++r)
            if (RBW == RB || r < nrb) acc[r] = Mfma<T>::op(sA[(r * KS + k0 + u) * 64 + lane], bc[u], acc[r]);
#pragma unroll
        for (int u = 0; u < 4; ++u) bc[u] = bn[u];
      }
    }
    if (pid >= 0) {
      T *y = Y + (size_t)(pi - ybase) * KPS + r0 * 16;
#pragma unroll
      for (int r = 0; r < RBW; ++r)
        if (r < nrb)
#pragma unroll
          for (int q = 0; q < 4; ++q) y[r * 16 + Mfma<T>::row(kg, q)] = acc[r][q];
    }
  }
}

// L of every M2L target += sum of the partials of its pairs in the pass
// [cb, ce) of class-sorted positions (pos[q] of list entry q), in list order,
// in FP64.  With per-pair orders (bins != null: the pair's bin, order
// kOrderLo + bin / kNumClasses) only the (q+1)^2 coefficients of a pair's
// order q are added.
__device__ __forceinline__ int pair_rows(const uint16_t *bins, int32_t q, int KP) {
  if (!bins) return KP;
  const int o = kOrderLo + (int)bins[q] / kNumClasses;
  return (o + 1) * (o + 1);
}

// Pairs whose partial loads k_m2l_sum keeps in flight (measured at 4M, the
// sum alone: FP32 2 and 4 best, 8 and 16 slower; FP64 4 and 8 best)
#ifndef NBFMM_M2L_SUM_UNROLL_F32
#define NBFMM_M2L_SUM_UNROLL_F32 2
#endif
#ifndef NBFMM_M2L_SUM_UNROLL_F64
#define NBFMM_M2L_SUM_UNROLL_F64 4
#endif
template <class T>
__global__ void __launch_bounds__(kT) k_m2l_sum(int KP, const DevPlan *__restrict__ dp,
                                                const int32_t *__restrict__ tgt_cells,
                                                const int32_t *__restrict__ off_m2l, const int32_t *__restrict__ cnt_m2l,
                                                const int32_t *__restrict__ lslot, const int32_t *__restrict__ pos,
                                                int32_t cb, int32_t ce, const T *__restrict__ Y,
                                                T *__restrict__ Lexp, const uint16_t *__restrict__ bins) {
  if (plan_dead(dp)) return;
  const int64_t t = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= dp->nst[1]) return;
  const int32_t c = tgt_cells[t];
  const int32_t b = off_m2l[c], e = b + cnt_m2l[c];
  T *L = Lexp + (size_t)lslot[c] * KP;
  // rows lane, lane + 64, ... (KP <= 256)
  double s[4] = {0, 0, 0, 0};
  int nmax = 0; // rows any pair added
  for (int32_t q0 = b; q0 < e; q0 += 64) {
    // the next 64 entries' partial positions and row counts, one per lane,
    // so that the partial loads below depend on no other load: each pair's
    // position is read out of its lane (v_readlane, wavefront-uniform)
    const int32_t qi = q0 + lane;
    int32_t prl = cb, nvl = 0, nva = 0;
    if (qi < e) {
      const int32_t pr = pos[qi];
      nva = pair_rows(bins, qi, KP);
      if (pr >= cb && pr < ce) {
        prl = pr;
        nvl = nva;
      }
    }
    for (int o = 32; o >= 1; o >>= 1) nva = max(nva, __shfl_xor(nva, o));
    nmax = max(nmax, nva);
    const int m = min(64, e - q0);
    int j = 0;
    // kU pairs' loads in flight, then their sums in list order
    constexpr int kU = sizeof(T) == 4 ? NBFMM_M2L_SUM_UNROLL_F32 : NBFMM_M2L_SUM_UNROLL_F64;
    for (; j + kU <= m; j += kU) {
      T v[kU][4];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int32_t pr = __builtin_amdgcn_readlane(prl, j + u);
        const int nv = __builtin_amdgcn_readlane(nvl, j + u);
        const T *y = Y + (size_t)(pr - cb) * KP;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[u][r] = lane + 64 * r < nv ? y[lane + 64 * r] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) s[r] += (double)v[u][r];
    }
    for (; j < m; ++j) {
      const int32_t pr = __builtin_amdgcn_readlane(prl, j);
      const int nv = __builtin_amdgcn_readlane(nvl, j);
      const T *y = Y + (size_t)(pr - cb) * KP;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (lane + 64 * r < nv) s[r] += (double)y[lane + 64 * r];
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (lane + 64 * r < nmax) L[lane + 64 * r] = (T)((double)L[lane + 64 * r] + s[r]);
}

// FP32 partials with one 16-byte load per pair: lane l (< KP / 4) sums rows
// 4l .. 4l + 3 in FP64 (k_m2l_sum<float> needs KP / 64 load instructions per
// pair).  Pairs in flight per wavefront, measured at 4M with P2P alongside:
// 1 best (M2L alone 5.23 -> 4.72 ms, the overlapped evaluation 17.95 ->
// 17.82 ms); 2, 4 and 8 were faster alone but slower overlapped (18.0-18.8 ms:
// more loads in flight take memory bandwidth from the P2P beside them).
#ifndef NBFMM_M2L_SUM4_UNROLL
#define NBFMM_M2L_SUM4_UNROLL 1
#endif
__global__ void __launch_bounds__(kT) k_m2l_sum4(int KP, const DevPlan *__restrict__ dp,
                                                 const int32_t *__restrict__ tgt_cells,
                                                 const int32_t *__restrict__ off_m2l,
                                                 const int32_t *__restrict__ cnt_m2l, const int32_t *__restrict__ lslot,
                                                 const int32_t *__restrict__ pos, int32_t cb, int32_t ce,
                                                 const float *__restrict__ Y, float *__restrict__ Lexp,
                                                 const uint16_t *__restrict__ bins) {
  if (plan_dead(dp)) return;
  const int64_t t = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= dp->nst[1]) return;
  const int32_t c = tgt_cells[t];
  const int32_t b = off_m2l[c], e = b + cnt_m2l[c];
  float *L = Lexp + (size_t)lslot[c] * KP;
  const int r0 = 4 * lane;
  const bool ld = r0 < KP;
  double s[4] = {0, 0, 0, 0};
  int nmax = 0; // rows any pair added
  for (int32_t q0 = b; q0 < e; q0 += 64) {
    // the next 64 entries' partial positions and row counts, one per lane (as in k_m2l_sum)
    const int32_t qi = q0 + lane;
    int32_t prl = cb, nvl = 0, nva = 0;
    if (qi < e) {
      const int32_t pr = pos[qi];
      nva = pair_rows(bins, qi, KP);
      if (pr >= cb && pr < ce) {
        prl = pr;
        nvl = nva;
      }
    }
    for (int o = 32; o >= 1; o >>= 1) nva = max(nva, __shfl_xor(nva, o));
    nmax = max(nmax, nva);
    const int m = min(64, e - q0);
    int j = 0;
    constexpr int kU = NBFMM_M2L_SUM4_UNROLL;
    for (; j + kU <= m; j += kU) {
      float4 v[kU];
      int nv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int32_t pr = __builtin_amdgcn_readlane(prl, j + u);
        nv[u] = __builtin_amdgcn_readlane(nvl, j + u);
        v[u] = ld && r0 < nv[u] ? *reinterpret_cast<const float4 *>(Y + (size_t)(pr - cb) * KP + r0)
                                : make_float4(0, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        // a loaded group has r0 < nv; its rows from nv on are zeros (k_m2l_bf3)
        s[0] += (double)v[u].x;
        s[1] += (double)v[u].y;
        s[2] += (double)v[u].z;
        s[3] += (double)v[u].w;
      }
    }
    for (; j < m; ++j) {
      const int32_t pr = __builtin_amdgcn_readlane(prl, j);
      const int nv = __builtin_amdgcn_readlane(nvl, j);
      if (ld && r0 < nv) {
        const float4 v = *reinterpret_cast<const float4 *>(Y + (size_t)(pr - cb) * KP + r0);
        s[0] += (double)v.x;
        s[1] += r0 + 1 < nv ? (double)v.y : 0.0;
        s[2] += r0 + 2 < nv ? (double)v.z : 0.0;
        s[3] += r0 + 3 < nv ? (double)v.w : 0.0;
      }
    }
  }
  if (ld)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (r0 + i < nmax) L[r0 + i] = (float)((double)L[r0 + i] + s[i]);
}

// Row blocks per workgroup: the fewest slices of the KP / 16 row blocks that
// fit the LDS, split evenly (at KP = 144 in FP64: 5 + 4, not 8 + 1, so that
// no slice reloads the multipoles for a single row block).
template <class T>
constexpr int rbw_for(int KP, int lds, int more = 0) {
  const int fit = lds / (16 * KP * (int)sizeof(T)), rb = KP / 16;
  const int slices = (fit >= rb ? 1 : (rb + fit - 1) / fit) + more; // more: slices beyond the fewest
  return (rb + slices - 1) / slices;
}

template <int KP, int RBW, class T>
void launch_rbw(nbfmm_ctx *c, unsigned nitems, const int32_t *irng, const int4 *items, int32_t ybase, const int32_t *cls_pairs,
                const T *frag, int KPS, int nvalid) {
  constexpr int RB = KP / 16;
  const size_t lds = (size_t)RBW * 16 * KP * sizeof(T);
  static bool attr_set[64] = {};
  if (lds > 65536 && !attr_set[c->device & 63]) {
    NB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_cls<KP, RBW, T>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set[c->device & 63] = true;
  }
  hipLaunchKernelGGL((k_m2l_cls<KP, RBW, T>), dim3(nitems, (RB + RBW - 1) / RBW), dim3(m2l_threads<T>()), lds, c->stream,
                     (const DevPlan *)dplan(c), irng, items, ybase, cls_pairs, (const int32_t *)c->m2l_src.p, (const int32_t *)c->mslot.p, frag,
                     (const T *)c->mexp.p, (T *)c->ypair.p, KPS, nvalid);
  NB_LAUNCH();
}

// gfx950 gives a workgroup up to 160 KB of LDS.
constexpr int kLds = 163840;

template <int KP, class T>
void launch_kp(nbfmm_ctx *c, unsigned nitems, const int32_t *irng, const int4 *items, int32_t ybase, const int32_t *cls_pairs,
               const T *frag, int KPS, int nvalid) {
  if (c->lds_max < kLds) throw Error(NBFMM_ERR_UNSUPPORTED, "M2L needs 160 KB of LDS per workgroup (gfx950)");
  if (c->shapes.m2l_slices == 2)
    launch_rbw<KP, rbw_for<T>(KP, kLds, 1), T>(c, nitems, irng, items, ybase, cls_pairs, frag, KPS, nvalid);
  else
    launch_rbw<KP, rbw_for<T>(KP, kLds), T>(c, nitems, irng, items, ybase, cls_pairs, frag, KPS, nvalid);
}

// as rbw_for, for the three-term bf16 operator blocks
constexpr int rbw_bf3(int KP, int more = 0) {
  const int fit = kLds / (pad32(KP) / 32 * 3072), rb = KP / 16;
  const int slices = (fit >= rb ? 1 : (rb + fit - 1) / fit) + more; // more: slices beyond the fewest
  return (rb + slices - 1) / slices;
}

template <int KP, int RBW>
void launch_bf3_rbw(nbfmm_ctx *c, unsigned nitems, const int32_t *irng, const int4 *items, int32_t ybase, const int32_t *cls_pairs,
                    const uint4 *fragb, int KPS, int nvalid) {
  constexpr int RB = KP / 16;
  if (c->lds_max < kLds) throw Error(NBFMM_ERR_UNSUPPORTED, "M2L needs 160 KB of LDS per workgroup (gfx950)");
  const size_t lds = (size_t)RBW * (pad32(KP) / 32) * 3072;
  static bool attr_set[64] = {};
  if (lds > 65536 && !attr_set[c->device & 63]) {
    NB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_bf3<KP, RBW>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set[c->device & 63] = true;
  }
  hipLaunchKernelGGL((k_m2l_bf3<KP, RBW>), dim3(nitems, (RB + RBW - 1) / RBW), dim3(1024), lds, c->stream,
                     (const DevPlan *)dplan(c), irng, items, ybase, cls_pairs, (const int32_t *)c->m2l_src.p, (const int32_t *)c->mslot.p, fragb,
                     (const float *)c->mexp.p, (float *)c->ypair.p, KPS, nvalid);
  NB_LAUNCH();
}

template <int KP>
void launch_bf3_kp(nbfmm_ctx *c, unsigned nitems, const int32_t *irng, const int4 *items, int32_t ybase, const int32_t *cls_pairs,
                   const uint4 *fragb, int KPS, int nvalid) {
  if (c->shapes.m2l_slices == 2)
    launch_bf3_rbw<KP, rbw_bf3(KP, 1)>(c, nitems, irng, items, ybase, cls_pairs, fragb, KPS, nvalid);
  else
    launch_bf3_rbw<KP, rbw_bf3(KP)>(c, nitems, irng, items, ybase, cls_pairs, fragb, KPS, nvalid);
}

void launch_bf3(nbfmm_ctx *c, int KP, unsigned nitems, const int32_t *irng, const int4 *items, int32_t ybase, const int32_t *cls_pairs,
                const uint4 *fragb, int KPS, int nvalid) {
  switch (KP) {
#define NB_BF3_CASE(k)                                                                \
  case k:                                                                             \
    launch_bf3_kp<k>(c, nitems, irng, items, ybase, cls_pairs, fragb, KPS, nvalid);         \
    break;
    NB_BF3_CASE(16)
    NB_BF3_CASE(32)
    NB_BF3_CASE(48)
    NB_BF3_CASE(64)
    NB_BF3_CASE(96)
    NB_BF3_CASE(112)
    NB_BF3_CASE(128)
    NB_BF3_CASE(144)
    NB_BF3_CASE(176)
    NB_BF3_CASE(208)
    NB_BF3_CASE(240)
#undef NB_BF3_CASE
    default:
      throw Error(NBFMM_ERR_UNSUPPORTED, "unsupported expansion order");
  }
}

// row blocks per workgroup of the target-major contraction: the fewest
// slices whose operator rows fit the LDS beside the tile's target table
constexpr int rbw_tgt(int KP, int more = 0) {
  const int fit = (kLds - 1024) / (pad32(KP) / 32 * 3072), rb = KP / 16;
  const int slices = (fit >= rb ? 1 : (rb + fit - 1) / fit) + more;
  return (rb + slices - 1) / slices;
}

template <int RBW, int KSBM>
void launch_tgt_rbw(nbfmm_ctx *c, int KP, unsigned ntile, int p, bool binned, const int32_t *tab, const uint4 *fragb) {
  if (c->lds_max < kLds) throw Error(NBFMM_ERR_UNSUPPORTED, "M2L needs 160 KB of LDS per workgroup (gfx950)");
  const size_t lds = (size_t)RBW * KSBM * 3072;
  static bool attr_set[64] = {};
  if (lds > 65536 && !attr_set[c->device & 63]) {
    NB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_tgt<RBW, KSBM>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set[c->device & 63] = true;
  }
  const int RB = KP / 16;
  hipLaunchKernelGGL((k_m2l_tgt<RBW, KSBM>), dim3(ntile, (RB + RBW - 1) / RBW), dim3(64 * kTgtWaves), lds, c->stream,
                     (const DevPlan *)dplan(c), p, binned ? 1 : 0, (const int32_t *)c->tgt_cells.p,
                     (const int32_t *)c->lslot.p, (const uint64_t *)c->ckey.p, tab, fragb, (const float *)c->mexp.p,
                     (float *)c->lexp.p, KP);
  NB_LAUNCH();
}

template <int KP>
void launch_tgt_kp(nbfmm_ctx *c, unsigned ntile, int p, bool binned, const int32_t *tab, const uint4 *fragb) {
  if (c->shapes.m2l_slices == 2)
    launch_tgt_rbw<rbw_tgt(KP, 1), pad32(KP) / 32>(c, KP, ntile, p, binned, tab, fragb);
  else
    launch_tgt_rbw<rbw_tgt(KP), pad32(KP) / 32>(c, KP, ntile, p, binned, tab, fragb);
}

void launch_tgt(nbfmm_ctx *c, int KP, unsigned ntile, int p, bool binned, const int32_t *tab, const uint4 *fragb) {
  switch (KP) {
#define NB_TGT_CASE(k)                                      \
  case k:                                                   \
    launch_tgt_kp<k>(c, ntile, p, binned, tab, fragb);      \
    break;
    NB_TGT_CASE(16)
    NB_TGT_CASE(32)
    NB_TGT_CASE(48)
    NB_TGT_CASE(64)
    NB_TGT_CASE(96)
    NB_TGT_CASE(112)
    NB_TGT_CASE(128)
    NB_TGT_CASE(144)
    NB_TGT_CASE(176)
    NB_TGT_CASE(208)
    NB_TGT_CASE(240)
#undef NB_TGT_CASE
    default:
      throw Error(NBFMM_ERR_UNSUPPORTED, "unsupported expansion order");
  }
}

template <class T>
void launch(nbfmm_ctx *c, int KP, unsigned nitems, const int32_t *irng, const int4 *items, int32_t ybase, const int32_t *cls_pairs,
            const T *frag, int KPS, int nvalid) {
  switch (KP) {
#define NB_M2L_CASE(k)                                                      \
  case k:                                                                   \
    launch_kp<k, T>(c, nitems, irng, items, ybase, cls_pairs, frag, KPS, nvalid); \
    break;
    NB_M2L_CASE(16)
    NB_M2L_CASE(32)
    NB_M2L_CASE(48)
    NB_M2L_CASE(64)
    NB_M2L_CASE(96)
    NB_M2L_CASE(112)
    NB_M2L_CASE(128)
    NB_M2L_CASE(144)
    NB_M2L_CASE(176)
    NB_M2L_CASE(208)
    NB_M2L_CASE(240)
#undef NB_M2L_CASE
    default:
      throw Error(NBFMM_ERR_UNSUPPORTED, "unsupported expansion order");
  }
}

// ---- M2M and L2L on MFMA ---------------------------------------------------------------
// Parent-stationary: a workgroup takes 16 parents of one level (the MFMA
// columns) and its 4 waves split the 8 octant operators (8 KP x KP matrices,
// small enough to stay in the caches, read in A-fragment order): wave w runs
// octants w and w + 4.  Splitting the octants keeps each wave's chain of
// dependent MFMAs and operator loads short, which sets the time of the small
// levels near the root.  M2M sums the children's multipoles into the parents'
// (columns of a missing child are zero; the waves' partial sums meet in LDS,
// added in a fixed order); L2L adds the translated parent local to every
// child's local.
// The parents cells[rng[0], rng[1]) (a device range: one level's).
template <int KP, class T>
__global__ void __launch_bounds__(256) k_m2m(const DevPlan *__restrict__ dp, const int32_t *__restrict__ rng,
                                             const int32_t *__restrict__ cells,
                                             const int32_t *__restrict__ child8, const int32_t *__restrict__ mslot,
                                             const T *__restrict__ frag8, T *__restrict__ Mexp) {
  if (plan_dead(dp)) return;
  constexpr int RB = KP / 16, KS = KP / 4;
  __shared__ T part[2][RB * 4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, kg = lane >> 4;
  const int pi = (int)blockIdx.x * 16 + col;
  const int32_t ncell = rng[1] - rng[0];
  const int32_t cell = pi < ncell ? cells[rng[0] + pi] : -1;
  typename Mfma<T>::V acc[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) acc[r] = typename Mfma<T>::V{0, 0, 0, 0};
  for (int o = wave; o < 8; o += 4) {
    const int32_t ch = cell >= 0 ? child8[(int64_t)cell * 8 + o] : -1;
    if (!__any(ch >= 0)) continue;
    const T *Mb = Mexp + (size_t)(ch >= 0 ? mslot[ch] : 0) * KP + kg;
    const T *A = frag8 + (size_t)o * KP * KP + lane;
#pragma unroll 1
    for (int k0 = 0; k0 < KS; k0 += 4) {
      T b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = ch >= 0 ? Mb[(k0 + u) * 4] : (T)0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] = Mfma<T>::op(A[(size_t)(r * KS + k0 + u) * 64], b[u], acc[r]);
    }
  }
  // (w0 + w2) + (w1 + w3)
  if (wave >= 2) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) part[wave - 2][r * 4 + q][lane] = acc[r][q];
  }
  __syncthreads();
  if (wave < 2) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[r][q] += part[wave][r * 4 + q][lane];
  }
  __syncthreads();
  if (wave == 1) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) part[0][r * 4 + q][lane] = acc[r][q];
  }
  __syncthreads();
  if (wave > 0 || cell < 0) return;
  T *M = Mexp + (size_t)mslot[cell] * KP;
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) M[r * 16 + Mfma<T>::row(kg, q)] = acc[r][q] + part[0][r * 4 + q][lane];
}

template <int KP, class T>
__global__ void __launch_bounds__(256) k_l2l(const DevPlan *__restrict__ dp, const int32_t *__restrict__ rng,
                                             const int32_t *__restrict__ cells,
                                             const int32_t *__restrict__ child8, const int32_t *__restrict__ lslot,
                                             const T *__restrict__ frag8, T *__restrict__ Lexp) {
  if (plan_dead(dp)) return;
  constexpr int RB = KP / 16, KS = KP / 4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, kg = lane >> 4;
  const int pi = (int)blockIdx.x * 16 + col;
  const int32_t ncell = rng[1] - rng[0];
  const int32_t cell = pi < ncell ? cells[rng[0] + pi] : -1;
  const T *Lb = Lexp + (size_t)(cell >= 0 ? lslot[cell] : 0) * KP + kg;
  for (int o = wave; o < 8; o += 4) {
    const int32_t ch = cell >= 0 ? child8[(int64_t)cell * 8 + o] : -1;
    if (!__any(ch >= 0)) continue;
    typename Mfma<T>::V acc[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = typename Mfma<T>::V{0, 0, 0, 0};
    const T *A = frag8 + (size_t)o * KP * KP + lane;
#pragma unroll 1
    for (int k0 = 0; k0 < KS; k0 += 4) {
      T b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = cell >= 0 ? Lb[(k0 + u) * 4] : (T)0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] = Mfma<T>::op(A[(size_t)(r * KS + k0 + u) * 64], b[u], acc[r]);
    }
    if (ch >= 0) {
      T *L = Lexp + (size_t)lslot[ch] * KP;
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) L[r * 16 + Mfma<T>::row(kg, q)] += acc[r][q];
    }
  }
}

// the parents cells[rng[0], rng[1]) of one level, at most `grid_cells` of them
template <class T>
void launch_trans(nbfmm_ctx *c, bool l2l, int KP, const int32_t *cells, const int32_t *rng, int64_t grid_cells,
                  const int32_t *slot, const T *frag8, T *X) {
  if (grid_cells <= 0) return;
  const unsigned nb = (unsigned)((grid_cells + 15) / 16);
  const int32_t *child8 = (const int32_t *)c->child8.p;
  const DevPlan *dp = dplan(c);
  switch (KP) {
#define NB_TR_CASE(k)                                                                                          \
  case k:                                                                                                      \
    if (l2l)                                                                                                   \
      hipLaunchKernelGGL((k_l2l<k, T>), dim3(nb), dim3(256), 0, c->stream, dp, rng, cells, child8, slot, frag8, X); \
    else                                                                                                       \
      hipLaunchKernelGGL((k_m2m<k, T>), dim3(nb), dim3(256), 0, c->stream, dp, rng, cells, child8, slot, frag8, X); \
    break;
    NB_TR_CASE(16)
    NB_TR_CASE(32)
    NB_TR_CASE(48)
    NB_TR_CASE(64)
    NB_TR_CASE(96)
    NB_TR_CASE(112)
    NB_TR_CASE(128)
    NB_TR_CASE(144)
    NB_TR_CASE(176)
    NB_TR_CASE(208)
    NB_TR_CASE(240)
#undef NB_TR_CASE
    default:
      throw Error(NBFMM_ERR_UNSUPPORTED, "unsupported expansion order");
  }
  NB_LAUNCH();
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

// Distance-dependent orders.  The M2L of class d truncated after degree q
// has the worst-case bound rate_d^(q+1) / |d|^2 (relative to the field of
// the pair's source at unit distance), rate_d = (cell diagonal) / |d|; for
// particles spread over their cells the error falls per degree by about
// kOrderAlpha * rate_d instead (measured for cubes: 0.50 for the face class
// (2,0,0) of rate 0.87, 0.15 for (3,3,3) of rate 0.33).  A class gets the
// lowest q in [kOrderLo, p] with
//   (a rate_d)^(q+1) / |d|^2 <= tol (a rate_0)^(p+1) / |d_0|^2,
// class 0 being the slowest M2L class (rates and |d| the same for every
// level: cells keep the box's aspect ratio); classes evaluated directly or
// not admissible, and class 0, keep p (every class keeps p when tol = 0).
// The oracle computes the same table with the same operations (orc_exact.c).
// One workgroup of 512 threads, from the plan's box into dp->qc.
__global__ void __launch_bounds__(512) k_class_orders(DevPlan *dp, int p, double tol) {
#pragma clang fp contract(off)
  __shared__ double r2s[kNumClasses], d2s[kNumClasses];
  __shared__ double r0s, d0s;
  const int t = threadIdx.x;
  const double hx = dp->hi[0] - dp->lo[0], hy = dp->hi[1] - dp->lo[1], hz = dp->hi[2] - dp->lo[2];
  const double diag2 = hx * hx + hy * hy + hz * hz;
  if (t < kNumClasses) {
    const int dx = t % 7 - 3, dy = (t / 7) % 7 - 3, dz = t / 49 - 3;
    double r2 = -1, d2 = 0;
    if (max(abs(dx), max(abs(dy), abs(dz))) >= 2) {
      d2 = dx * dx * hx * hx + dy * dy * hy * hy + dz * dz * hz * hz;
      if (!(diag2 > kMacRatio * kMacRatio * d2)) r2 = diag2 / d2;
    }
    r2s[t] = r2;
    d2s[t] = d2;
  }
  __syncthreads();
  if (t == 0) {
    double r0 = 0, d0 = 0; // the largest squared rate of an admissible class (the first), its |d|^2
    for (int k = 0; k < kNumClasses; ++k)
      if (r2s[k] > r0) {
        r0 = r2s[k];
        d0 = d2s[k];
      }
    r0s = r0;
    d0s = d0;
  }
  __syncthreads();
  if (t < kNumClasses) {
    int q = p;
    const double r2 = r2s[t], r0 = r0s, d0 = d0s;
    if (tol > 0 && r2 > 0 && r2 < r0) {
      const double la = log(kOrderAlpha);
      const double need = (log(tol) + (p + 1) * (la + 0.5 * log(r0)) + log(d2s[t] / d0)) / (la + 0.5 * log(r2));
      q = (int)ceil(need) - 1;
      q = min(p, max(kOrderLo, q));
    }
    dp->qc.q[t] = (uint8_t)q;
  }
}

// The M2L schedule of one pass on the device: every bin's first position
// (the scatter's cursors), items of at most `per` consecutive pairs of one
// bin in bin order, each order's item range (bins are order-major), and the
// algorithmic flops (2 n^2 per pair, n = (q+1)^2) and order sum of the pairs.
// Items are written only when they fit `items_cap` (a planned evaluation
// flags the overflow instead).  One workgroup; bins in runs of kSchedPer per
// thread.
constexpr int kSchedThreads = 1024;
constexpr int kSchedPer = (kMaxBins + kSchedThreads - 1) / kSchedThreads;
struct OrdCaps {
  int32_t v[kNumOrders];
};
__global__ void __launch_bounds__(kSchedThreads) k_m2l_sched(DevPlan *dp, int nbins, int per, int binned, int p,
                                                            int check, int64_t items_cap, OrdCaps cap_ord,
                                                            int4 *__restrict__ items) {
  if (plan_dead(dp)) return;
  __shared__ int32_t wsum[kSchedThreads / 64][2];
  __shared__ double wfl[kSchedThreads / 64][2];
  __shared__ int32_t ostart[kNumOrders + 1];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int k0 = t * kSchedPer;
  int32_t cnt[kSchedPer], ni[kSchedPer];
  int32_t sc = 0, si = 0;
  double fl = 0, qs = 0;
#pragma unroll
  for (int j = 0; j < kSchedPer; ++j) {
    const int k = k0 + j;
    cnt[j] = k < nbins ? dp->bin_cnt[k] : 0;
    ni[j] = (cnt[j] + per - 1) / per;
    sc += cnt[j];
    si += ni[j];
    const int q = binned ? kOrderLo + k / kNumClasses : p;
    const double nq = (double)(q + 1) * (q + 1);
    fl += 2.0 * nq * nq * cnt[j];
    qs += (double)q * cnt[j];
  }
  int32_t xc = sc, xi = si;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t yc = __shfl_up(xc, o, 64), yi = __shfl_up(xi, o, 64);
    if (lane >= o) {
      xc += yc;
      xi += yi;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    fl += __shfl_xor(fl, o, 64);
    qs += __shfl_xor(qs, o, 64);
  }
  if (lane == 63) {
    wsum[wv][0] = xc;
    wsum[wv][1] = xi;
  }
  if (lane == 0) {
    wfl[wv][0] = fl;
    wfl[wv][1] = qs;
  }
  __syncthreads();
  int32_t pc = 0, pi = 0, ti = 0;
  for (int q = 0; q < kSchedThreads / 64; ++q) {
    if (q < wv) {
      pc += wsum[q][0];
      pi += wsum[q][1];
    }
    ti += wsum[q][1];
  }
  const bool fits = ti <= items_cap;
  int32_t bc = pc + xc - sc, bi = pi + xi - si; // this thread's first bin: first pair, first item
#pragma unroll
  for (int j = 0; j < kSchedPer; ++j) {
    const int k = k0 + j;
    if (k >= nbins) break;
    dp->bin_cur[k] = bc;
    if (binned && k % kNumClasses == 0) ostart[k / kNumClasses] = bi;
    if (fits)
      for (int m = 0; m < ni[j]; ++m)
        items[bi + m] = make_int4(k, bc + m * per, min(bc + (m + 1) * per, bc + cnt[j]), 0);
    bc += cnt[j];
    bi += ni[j];
  }
  __syncthreads();
  if (t != 0) return;
  const int nord = binned ? nbins / kNumClasses : 1;
  if (!binned) ostart[0] = 0;
  ostart[nord] = ti;
  int ovf = check && !fits ? 2 : 0;
  for (int oi = 0; oi < kNumOrders; ++oi) {
    const int32_t f = oi < nord ? ostart[oi] : ti, e = oi < nord ? ostart[oi + 1] : ti;
    dp->ord_items[oi][0] = f;
    dp->ord_items[oi][1] = e - f;
    if (check && e - f > cap_ord.v[oi]) ovf = 2;
  }
  dp->n_items = ti;
  double a = 0, b = 0;
  for (int q = 0; q < kSchedThreads / 64; ++q) {
    a += wfl[q][0];
    b += wfl[q][1];
  }
  dp->flop_m2l = a;
  dp->order_sum = b;
  if (ovf) dp->ovf |= ovf;
}

} // namespace

// Sorts the pairs by bin (counts -> schedule -> scatter), runs the MFMA
// contraction (between events 7 and 8) and the per-target sum.  Expects the
// FP64 operators in c->tmat and the multipoles in c->mexp (FP32 when `f32`),
// both of width KP.  `ord` (device, per cell) gives the order of every
// target, or is null when all have the order of KP.  One pass (the partials
// of every pair in one buffer) is scheduled on the device; several passes
// (partials beyond kPassBytes, or m2l_pass_pairs) are scheduled on the host
// (direct evaluations only).
void run_m2l(nbfmm_ctx *c, int KP, bool f32, const int8_t *ord, int p) {
  Plan &P = c->plan;
  const Caps &K = c->cap;
  DevPlan *dp = dplan(c);
  hipStream_t s = c->stream;
  const int64_t nt = K.on ? K.n_tgt : P.n_tgt, npair = K.on ? K.n_m2l : P.n_m2l; // exact, or the capacities
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  // per-pair orders: the targets' (learned) and/or the classes'
  const bool binned = ord || c->prm.class_order_tol > 0;
  const int nbins = binned ? kMaxBins : kNumClasses;
  if (f32 && !ord && c->shapes.m2l_mode != 1) {
    // ---- target-major contraction (fixed or class orders): no pair partials ----
    if (binned) {
      hipLaunchKernelGGL(k_class_orders, dim3(1), dim3(512), 0, s, dp, p, c->prm.class_order_tol);
      NB_LAUNCH();
    }
    const int64_t ntc = (std::max<int64_t>(1, nt) + 1) / 2 * 2; // even: the partials after the table are 8-byte aligned
    int32_t *tab = dget<int32_t>(c->m2l_tab, (size_t)ntc * kNumClasses + 4 * kTabBlocks);
    double *part = reinterpret_cast<double *>(tab + (size_t)ntc * kNumClasses);
    fill_i32(s, tab, ntc * kNumClasses, -1);
    hipLaunchKernelGGL(k_m2l_tab, dim3(kTabBlocks), dim3(kT), 0, s, (const DevPlan *)dp, p, binned ? 1 : 0,
                       (const int32_t *)c->tgt_cells.p, (const int32_t *)c->off_m2l.p, (const int32_t *)c->cnt_m2l.p,
                       (const int32_t *)c->m2l_src.p, (const uint16_t *)c->m2l_cls.p, (const int32_t *)c->mslot.p,
                       tab, part);
    NB_LAUNCH();
    hipLaunchKernelGGL(k_m2l_tab_sum, dim3(1), dim3(kT), 0, s, dp, (const double *)part);
    NB_LAUNCH();
    if (!K.on) // the counts for the statistics, read when they are asked for
      NB_HIP(hipMemcpyAsync(&c->hplan->flop_m2l, &dp->flop_m2l, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    // the operators' three-term bf16 fragments
    const int64_t nl = (int64_t)kNumClasses * (KP / 16) * (pad32(KP) / 32) * 64;
    uint4 *fb = dget<uint4>(c->tmat32, (size_t)nl * 3);
    hipLaunchKernelGGL(k_to_bf3, dim3(nblk(nl)), dim3(kT), 0, s, (const DevPlan *)dp, KP, (const double *)c->tmat.p, fb);
    NB_LAUNCH();
    const unsigned ntile = (unsigned)((ntc + kTgtTile - 1) / kTgtTile);
    P.n_m2l_tiles = ntile;
    P.m2l_stats_dev = !K.on;
    mark(c, 7);
    launch_tgt(c, KP, ntile, p, binned, tab, fb);
    mark(c, 8);
    launch_p2p(c, 1); // the rest of P2P beside the rest of the chain
    c->p2p_started = true;
    return;
  }
  if (binned) {
    hipLaunchKernelGGL(k_class_orders, dim3(1), dim3(512), 0, s, dp, p, c->prm.class_order_tol);
    NB_LAUNCH();
    hipLaunchKernelGGL(k_order_bins, dim3((unsigned)std::max<int64_t>(1, (nt + 3) / 4)), dim3(kT), 0, s,
                       (const DevPlan *)dp, (const int32_t *)c->tgt_cells.p, (const int32_t *)c->off_m2l.p,
                       (const int32_t *)c->cnt_m2l.p, ord, p, (uint16_t *)c->m2l_cls.p);
    NB_LAUNCH();
  }
  fill_i32(s, dp->bin_cnt, nbins, 0);
  const unsigned nchunk = (unsigned)std::max<int64_t>(1, (npair + kClsChunk - 1) / kClsChunk);
  hipLaunchKernelGGL(k_cls_hist, dim3(nchunk), dim3(kT), 0, s, (const uint16_t *)c->m2l_cls.p, (const DevPlan *)dp,
                     nbins, dp->bin_cnt);
  NB_LAUNCH();
  // passes: consecutive ranges of the sorted pairs whose partials fit the
  // buffer (at most kPassBytes of the 288 GB, or the caller's pairs-per-pass limit)
  constexpr int64_t kPassBytes = (int64_t)16 << 30;
  int64_t pass = c->prm.m2l_pass_pairs > 0 ? c->prm.m2l_pass_pairs : kPassBytes / (KP * (int64_t)esz);
  pass = std::max<int64_t>(1, std::min<int64_t>(pass, std::max<int64_t>(1, npair)));
  const bool one = npair <= pass;
  if (K.on && !one) throw Error(NBFMM_ERR_UNSUPPORTED, "planned evaluations need the M2L partials in one pass");
  int64_t per = npair / 2048;
  per = std::max<int64_t>(128, std::min<int64_t>(c->shapes.m2l_item_pairs, per)) / 128 * 128;
  int32_t *cls_pairs = dget<int32_t>(c->cls_pairs, (size_t)npair * 2 + 2);
  int32_t *pos = cls_pairs + npair + 1; // sorted position of every list entry
  void *Lexp = c->lexp.p; // sized and cleared by the caller (one slot per cell that holds a local)
  auto fragments = [&]() -> const void * {
    if (!f32) return c->tmat.p;
    // three-term bf16 fragments of the operators
    const int64_t nl = (int64_t)kNumClasses * (KP / 16) * (pad32(KP) / 32) * 64;
    uint4 *fb = dget<uint4>(c->tmat32, (size_t)nl * 3);
    hipLaunchKernelGGL(k_to_bf3, dim3(nblk(nl)), dim3(kT), 0, s, (const DevPlan *)dp, KP, (const double *)c->tmat.p, fb);
    NB_LAUNCH();
    return fb;
  };
  auto contract = [&](int oi, unsigned nitems, const int32_t *irng, const int4 *items, int32_t pb, const void *frag) {
    const int q = binned ? kOrderLo + oi : p;
    const int KPT = binned ? padded_k(q) : KP, nvalid = binned ? (q + 1) * (q + 1) : KP;
    if (f32)
      launch_bf3(c, KPT, nitems, irng, items, pb, cls_pairs, (const uint4 *)frag, KP, nvalid);
    else
      launch<double>(c, KPT, nitems, irng, items, pb, cls_pairs, (const double *)frag, KP, nvalid);
  };
  auto sum = [&](int32_t pb, int32_t pe) {
    const unsigned nb = (unsigned)std::max<int64_t>(1, (nt + 3) / 4);
    const uint16_t *bins = binned ? (const uint16_t *)c->m2l_cls.p : nullptr;
    if (f32 && KP % 4 == 0 && KP <= 4 * 64)
      hipLaunchKernelGGL(k_m2l_sum4, dim3(nb), dim3(kT), 0, s, KP, (const DevPlan *)dp, (const int32_t *)c->tgt_cells.p,
                         (const int32_t *)c->off_m2l.p, (const int32_t *)c->cnt_m2l.p, (const int32_t *)c->lslot.p,
                         (const int32_t *)pos, pb, pe, (const float *)c->ypair.p, (float *)Lexp, bins);
    else if (f32)
      hipLaunchKernelGGL(k_m2l_sum<float>, dim3(nb), dim3(kT), 0, s, KP, (const DevPlan *)dp,
                         (const int32_t *)c->tgt_cells.p, (const int32_t *)c->off_m2l.p, (const int32_t *)c->cnt_m2l.p,
                         (const int32_t *)c->lslot.p, (const int32_t *)pos, pb, pe, (const float *)c->ypair.p,
                         (float *)Lexp, bins);
    else
      hipLaunchKernelGGL(k_m2l_sum<double>, dim3(nb), dim3(kT), 0, s, KP, (const DevPlan *)dp,
                         (const int32_t *)c->tgt_cells.p, (const int32_t *)c->off_m2l.p, (const int32_t *)c->cnt_m2l.p,
                         (const int32_t *)c->lslot.p, (const int32_t *)pos, pb, pe, (const double *)c->ypair.p,
                         (double *)Lexp, bins);
    NB_LAUNCH();
  };
  const int nord = binned ? (ord ? kNumOrders : p - kOrderLo + 1) : 1; // order indices that may hold pairs

  if (one) {
    const int64_t icap = K.on ? K.items : npair / per + nbins + 1;
    int4 *ditems = dget<int4>(c->m2l_items, (size_t)icap + 1);
    OrdCaps oc{};
    for (int oi = 0; oi < kNumOrders; ++oi) oc.v[oi] = K.items_ord[oi];
    hipLaunchKernelGGL(k_m2l_sched, dim3(1), dim3(kSchedThreads), 0, s, dp, nbins, (int)per, binned ? 1 : 0, p,
                       K.on ? 1 : 0, icap, oc, ditems);
    NB_LAUNCH();
    int32_t cnt_ord[kNumOrders] = {};
    if (!K.on) {
      read_plan(c, &c->hplan->ord_items[0][0],
                offsetof(DevPlan, order_sum) + sizeof(double) - offsetof(DevPlan, ord_items), s);
      const DevPlan *hp = c->hplan;
      P.flop_m2l = hp->flop_m2l;
      P.mean_order = npair > 0 ? hp->order_sum / (double)npair : (double)p;
      P.n_m2l_tiles = hp->n_items;
      c->last.items = hp->n_items;
      for (int oi = 0; oi < kNumOrders; ++oi) {
        cnt_ord[oi] = hp->ord_items[oi][1];
        c->last.items_ord[oi] = cnt_ord[oi];
      }
    } else {
      for (int oi = 0; oi < kNumOrders; ++oi) cnt_ord[oi] = K.items_ord[oi];
    }
    hipLaunchKernelGGL(k_cls_scatter, dim3(nchunk), dim3(kT), 0, s, (const uint16_t *)c->m2l_cls.p,
                       (const DevPlan *)dp, nbins, dp->bin_cur, cls_pairs, pos);
    NB_LAUNCH();
    dbuf_ensure(c->ypair, (size_t)std::max<int64_t>(1, npair) * KP * esz + 16);
    const void *frag = fragments();
    mark(c, 7);
    for (int oi = 0; oi < nord; ++oi)
      if (cnt_ord[oi] > 0) contract(oi, (unsigned)cnt_ord[oi], &dp->ord_items[oi][0], ditems, 0, frag);
    mark(c, 8);
    launch_p2p(c, 1); // the rest of P2P beside the rest of the chain
    c->p2p_started = true;
    sum(0, INT32_MAX);
    return;
  }

  // ---- several passes (direct evaluations): the schedule on the host ----
  int32_t *hcnt = c->hplan->bin_cnt;
  read_plan(c, hcnt, sizeof(int32_t) * nbins, s);
  std::vector<int32_t> &hoff = c->h_cls_off;
  hoff.assign(nbins + 1, 0);
  for (int k = 0; k < nbins; ++k) hoff[k + 1] = hoff[k] + hcnt[k];
  if (hoff[nbins] != npair) throw Error(NBFMM_ERR_HIP, "M2L class histogram does not match the pair count");
  // algorithmic flops: 2 n^2 per pair, n = (q+1)^2 coefficients of the pair's order q
  {
    double fl = 0, qs = 0;
    for (int k = 0; k < nbins; ++k) {
      const int q = binned ? kOrderLo + k / kNumClasses : p;
      const double nq = (double)(q + 1) * (q + 1);
      fl += 2.0 * nq * nq * hcnt[k];
      qs += (double)q * hcnt[k];
    }
    P.flop_m2l = fl;
    P.mean_order = npair > 0 ? qs / (double)npair : (double)p;
  }
  std::vector<int4> &hitems = c->h_m2l_items;
  // launches: (pass, first item, item count, order index) -- items of one pass
  // and one order form one launch
  std::vector<int4> runs;
  hitems.clear();
  for (int64_t pb = 0; pb < npair; pb += pass) {
    const int64_t pe = std::min<int64_t>(npair, pb + pass);
    for (int k = 0; k < nbins; ++k) {
      const int64_t kb = std::max<int64_t>(hoff[k], pb), ke = std::min<int64_t>(hoff[k + 1], pe);
      if (kb >= ke) continue;
      const int oi = k / kNumClasses;
      const int ps = (int)(pb / pass);
      if (runs.empty() || runs.back().x != ps || runs.back().w != oi)
        runs.push_back(make_int4(ps, (int)hitems.size(), 0, oi));
      for (int64_t b = kb; b < ke; b += per) {
        hitems.push_back(make_int4(k, (int)b, (int)std::min<int64_t>(b + per, ke), 0));
        ++runs.back().z;
      }
    }
  }
  NB_HIP(hipMemcpyAsync(dp->bin_cur, hoff.data(), sizeof(int32_t) * nbins, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_cls_scatter, dim3(nchunk), dim3(kT), 0, s, (const uint16_t *)c->m2l_cls.p, (const DevPlan *)dp,
                     nbins, dp->bin_cur, cls_pairs, pos);
  NB_LAUNCH();
  const unsigned nitems = (unsigned)hitems.size();
  int4 *ditems = dget<int4>(c->m2l_items, (size_t)nitems + 1);
  NB_HIP(hipMemcpyAsync(ditems, hitems.data(), sizeof(int4) * nitems, hipMemcpyHostToDevice, s));
  // each launch's item range (first, count) for the kernels
  std::vector<int32_t> hr(2 * runs.size());
  for (size_t r = 0; r < runs.size(); ++r) {
    hr[2 * r] = runs[r].y;
    hr[2 * r + 1] = runs[r].z;
  }
  int32_t *drng = dget<int32_t>(c->cls_cnt, hr.size() + 2);
  NB_HIP(hipMemcpyAsync(drng, hr.data(), sizeof(int32_t) * hr.size(), hipMemcpyHostToDevice, s));
  P.n_m2l_tiles = nitems;
  dbuf_ensure(c->ypair, (size_t)pass * KP * esz + 16);
  const void *frag = fragments();
  // the host vectors above are read by the copies: wait before they change
  NB_HIP(hipStreamSynchronize(s));
  mark(c, 7);
  const int npass = (int)((npair + pass - 1) / pass);
  size_t r = 0;
  for (int ps = 0; ps < npass; ++ps) {
    const int32_t pb = (int32_t)((int64_t)ps * pass), pe = (int32_t)std::min<int64_t>(npair, (int64_t)pb + pass);
    for (; r < runs.size() && runs[r].x == ps; ++r)
      contract(runs[r].w, (unsigned)runs[r].z, drng + 2 * r, ditems, pb, frag);
    if (ps == 0) {
      mark(c, 8);
      launch_p2p(c, 1); // the rest of P2P beside the rest of the chain
      c->p2p_started = true;
    }
    sum(pb, pe);
  }
}

void run_trans(nbfmm_ctx *c, bool l2l, int KP, bool f32, const int32_t *cells, const int32_t *rng,
               int64_t grid_cells) {
  const int32_t *slot = (const int32_t *)(l2l ? c->lslot.p : c->mslot.p);
  void *X = l2l ? c->lexp.p : c->mexp.p;
  const size_t off = l2l ? (size_t)8 * KP * KP : 0;
  if (f32)
    launch_trans<float>(c, l2l, KP, cells, rng, grid_cells, slot, (const float *)c->trans32.p + off, (float *)X);
  else
    launch_trans<double>(c, l2l, KP, cells, rng, grid_cells, slot, (const double *)c->trans.p + off, (double *)X);
}

void run_trans64(nbfmm_ctx *c, bool l2l, int KP, const int32_t *cells, const int32_t *rng, int64_t grid_cells,
                 double *X) {
  const int32_t *slot = (const int32_t *)(l2l ? c->lslot.p : c->mslot.p);
  const size_t off = l2l ? (size_t)8 * KP * KP : 0;
  launch_trans<double>(c, l2l, KP, cells, rng, grid_cells, slot, (const double *)c->trans.p + off, X);
}

void widen(nbfmm_ctx *c, const float *in, int64_t count, double *out) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_to_f64, dim3(nblk(count)), dim3(kT), 0, c->stream, in, count, out);
  NB_LAUNCH();
}

void narrow(nbfmm_ctx *c, const double *in, int64_t count, float *out) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_to_f32, dim3(nblk(count)), dim3(kT), 0, c->stream, in, count, out);
  NB_LAUNCH();
}

void trans_to_f32(nbfmm_ctx *c, int KP) {
  const int64_t n = (int64_t)16 * KP * KP;
  float *t32 = dget<float>(c->trans32, (size_t)n);
  hipLaunchKernelGGL(k_ops_to_f32, dim3(nblk(n)), dim3(kT), 0, c->stream, (const DevPlan *)dplan(c),
                     (const double *)c->trans.p, n, t32);
  NB_LAUNCH();
}

const void *ops_f32_buffers(nbfmm_ctx *c, int KP, const void **trans32) {
  const int64_t nl = (int64_t)kNumClasses * (KP / 16) * (pad32(KP) / 32) * 64;
  *trans32 = dget<float>(c->trans32, (size_t)16 * KP * KP);
  return dget<uint4>(c->tmat32, (size_t)nl * 3);
}

} // namespace nbfmm
