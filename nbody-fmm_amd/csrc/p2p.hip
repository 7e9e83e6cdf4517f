// p2p.hip -- particle-particle interactions of the gfx950 FMM and the direct sum.
//
// P2P follows Leaf::compute_acceleration (leaf.cpp:105-143): a particle
// receives grad -= m_j (x_i - x_j) / |x_i - x_j|^3 from the particles of its
// own leaf (j != i) and of its neighbour leaves, here also from the
// interaction-list cells whose pair (t, s) is evaluated directly.
//
// Work is organised as sub-tasks (P2PTask): up to kTaskTargets targets of
// one cell against a range of the concatenated particles of a list of
// source cells (a 27-neighbour row, or a cell's directly evaluated
// interaction-list cells).  One wavefront runs a task: it prefix-sums the
// list's cell sizes into LDS and stages the task's targets there; each lane
// then walks its own sources of the concatenation (its cell found by binary
// search once, then moved forward), keeping one sum per target in
// registers.  Tasks of blocks with more than kNarrowMax targets (which this
// kernel would walk in several passes) go to k_p2p_wide instead: four
// targets per lane, one pass.  Tasks of one cell and of cells at different
// levels add into the same targets, so the partial sums are added with FP64
// atomics into the near-field gradient.
// P2P runs on its own stream, concurrently with the expansion passes (vector
// ALU work beside MFMA work); k_finish writes acc = G * (far + near) in input
// order (particle.cpp:16-21).
#include "nbfmm_internal.h"

#include <algorithm>
#include <initializer_list>
#include <mutex>
#include <string>

namespace nbfmm {
namespace {

constexpr int kT = 256;
constexpr int kW = 64; // one wavefront per task
// Targets per pass over the sources: each takes 3 (4 with potentials) sums in
// registers, and the register count sets how many wavefronts share a SIMD.
#ifndef NBFMM_P2P_GROUP_F32
#define NBFMM_P2P_GROUP_F32 16
#endif
#ifndef NBFMM_P2P_GROUP_F64
#define NBFMM_P2P_GROUP_F64 8
#endif
// Targets a wavefront-uniform branch covers in the pair loops of
// k_p2p_tasks (a block's last group computes up to this many minus one
// unused slots): measured at 4M, 1, 2 and 4 are within 0.2 ms of each other;
// 2 for FP32 (9.49 ms against 9.65 with 4), 1 for FP64 (14.06 against 14.51)
#ifndef NBFMM_P2P_TGT_STEP_F32
#define NBFMM_P2P_TGT_STEP_F32 2
#endif
#ifndef NBFMM_P2P_TGT_STEP_F64
#define NBFMM_P2P_TGT_STEP_F64 1
#endif
static_assert(NBFMM_P2P_GROUP_F32 % 4 == 0 && NBFMM_P2P_GROUP_F64 % 4 == 0, "targets go in blocks of four");

// list cell of concatenated source index g: the last with lp <= g (lp[0] = 0)
__device__ __forceinline__ int first_cell(const int32_t *lp, int ne, int32_t g) {
  int lo = 0, hi = ne;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (lp[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
static_assert(kTaskTargets == 4 * kW, "a P2P task holds at most four targets per lane");

// 1/sqrt(r2) for positive finite r2: v_rsq_f64 estimate y refined by the
// third-order step y + y e (1/2 + 3e/8), e = 1 - r2 y^2 (the Taylor series of
// (1 - e)^(-1/2)), whose error is O(e^3).
__device__ __forceinline__ double rsqrt_pos(double r2) {
  const double y = __builtin_amdgcn_rsq(r2);
  const double e = fma(-r2 * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// FP64: positions as stored, FP64 sums.  A particle meets itself (and
// coincident particles meet each other) with r2 = 0: those pairs are
// dropped by a select.
struct PairF64 {
  typedef double4 S;
  typedef double A;
  static constexpr bool kAnchor = false;
  static constexpr int kGroup = NBFMM_P2P_GROUP_F64; // targets per pass (3 or 4 FP64 sums each in registers)
  static constexpr int kStep = NBFMM_P2P_TGT_STEP_F64; // targets per branch of the pair loops
  static __device__ __forceinline__ S make(const double4 &p, const double4 &) { return p; }
  static __device__ __forceinline__ S tgt(const double4 &v) { return v; }
  static __device__ __forceinline__ void massless(S &v) { v.w = 0.0; }
  template <bool POT>
  static __device__ __forceinline__ void pair(const S &xi, const S &xj, A &ax, A &ay, A &az, A &ap) {
    const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double inv = r2 > 0.0 ? rsqrt_pos(r2) : 0.0;
    const double mi = xj.w * inv;
    const double w = mi * inv * inv;
    ax -= dx * w;
    ay -= dy * w;
    az -= dz * w;
    if (POT) ap += mi;
  }
  static __device__ __forceinline__ void add(A &s, A &, A v) { s += v; }
};

// FP64 for sources in cells that do not touch the target cell: r > 0 for
// every pair, so no select.
struct PairF64Far : PairF64 {
  template <bool POT>
  static __device__ __forceinline__ void pair(const S &xi, const S &xj, A &ax, A &ay, A &az, A &ap) {
    const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double inv = rsqrt_pos(r2);
    const double mi = xj.w * inv;
    const double w = mi * inv * inv;
    ax -= dx * w;
    ay -= dy * w;
    az -= dz * w;
    if (POT) ap += mi;
  }
};

// FP32 (mixed precision): offsets from an anchor (the task's first target,
// so every target lies within its cell's diagonal of it) formed in FP64 and kept as FP32 pairs hi + lo; the separation
// (xi_hi - xj_hi) + (xi_lo - xj_lo) is accurate to a few FP32 ulps of the
// separation itself however large the offsets are, which the close pairs of
// the target's own and adjacent cells need.  v_rsq_f32.  A chunk's
// contributions (at most 64 sources) are summed in FP32 and the chunk sums
// added with Kahan compensation into an FP32 (sum, correction) pair.
struct f2pos {
  float4 hi; // offset (x, y, z) and the mass
  float4 lo; // remainder of the offset
};
struct PairF32 {
  typedef f2pos S;
  typedef float A;
  static constexpr bool kAnchor = true;
  static constexpr int kGroup = NBFMM_P2P_GROUP_F32; // targets per pass
  static constexpr int kStep = NBFMM_P2P_TGT_STEP_F32; // targets per branch of the pair loops
  static __device__ __forceinline__ S make(const double4 &p, const double4 &a) {
    const double dx = p.x - a.x, dy = p.y - a.y, dz = p.z - a.z;
    const float hx = (float)dx, hy = (float)dy, hz = (float)dz;
    return {make_float4(hx, hy, hz, (float)p.w),
            make_float4((float)(dx - (double)hx), (float)(dy - (double)hy), (float)(dz - (double)hz), 0.0f)};
  }
  static __device__ __forceinline__ S tgt(const f2pos &v) { return v; }
  static __device__ __forceinline__ void massless(S &v) { v.hi.w = 0.0f; }
  template <bool POT>
  static __device__ __forceinline__ void pair(const S &xi, const S &xj, A &ax, A &ay, A &az, A &ap) {
    const float dx = (xi.hi.x - xj.hi.x) + (xi.lo.x - xj.lo.x);
    const float dy = (xi.hi.y - xj.hi.y) + (xi.lo.y - xj.lo.y);
    const float dz = (xi.hi.z - xj.hi.z) + (xi.lo.z - xj.lo.z);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float inv = r2 > 0.0f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    const float mi = xj.hi.w * inv;
    const float w = mi * inv * inv;
    ax -= dx * w;
    ay -= dy * w;
    az -= dz * w;
    if (POT) ap += mi;
  }
  static __device__ __forceinline__ void add(A &s, A &c, A v) { // Kahan
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// FP32 for sources in cells that do not touch the target cell (separated
// from every target by at least the target cell's edge h, while every target
// lies within sqrt(3) h of the anchor): single FP32 offsets, whose rounding
// moves a separation r >= h by at most 2^-24 (|x_s - a| + sqrt(3) h) <=
// 2^-24 (r + 2 sqrt(3) h), a relative 2.7e-7 at worst; and r > 0, so no
// select.  9 vector operations per pair fewer than PairF32.
struct PairF32Far {
  typedef float4 S;
  typedef float A;
  static __device__ __forceinline__ S make(const double4 &p, const double4 &a) {
    return make_float4((float)(p.x - a.x), (float)(p.y - a.y), (float)(p.z - a.z), (float)p.w);
  }
  static __device__ __forceinline__ S tgt(const f2pos &v) { return v.hi; }
  static __device__ __forceinline__ void massless(S &v) { v.w = 0.0f; }
  template <bool POT>
  static __device__ __forceinline__ void pair(const S &xi, const S &xj, A &ax, A &ay, A &az, A &ap) {
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float mi = xj.w * inv;
    const float w = mi * inv * inv;
    ax -= dx * w;
    ay -= dy * w;
    az -= dz * w;
    if (POT) ap += mi;
  }
};

// The task's list cells in evaluation order: lb[k] first particle, lp[k]
// first concatenated index of cell k; returns the number of concatenated
// sources, and in near_end the first concatenated index of the sources in
// cells that do not touch the target cell (those run with the far pair).
//  - a substituted interaction list (list >= 0): no cell touches the target;
//  - a leaf's own 27-neighbour row: every cell touches it;
//  - the 27-neighbour row of the parent P of a closed cell A: the eight
//    cells touching A (per axis, offset 0 or towards A's octant) go first,
//    the nineteen others (at least one edge of A away) after them.  (The
//    eight touching cells replaced by their children, of which only A's
//    27-neighbourhood touches A, gave more far pairs but lists of up to 83
//    small cells, whose walk cost more: P2P 9.4 -> 10.5 ms at 4M.)
// tmp: 2 x 27 LDS words.  One wavefront; ends with a barrier.
__device__ __forceinline__ int32_t stage_list(const P2PTask &tk, const int32_t *__restrict__ nbr27,
                                              const int32_t *__restrict__ sub_src, const int32_t *__restrict__ cbeg,
                                              const int32_t *__restrict__ cend, const uint64_t *__restrict__ ckey,
                                              int32_t *lb, int32_t *lp, int32_t *tmp, int lane, int32_t &near_end) {
  const int ne = tk.nlist;
  if (tk.list < 0 && -tk.list - 1 != tk.tcell) {
    // parent row of a closed cell: near cells first
    const int32_t *row = nbr27 + (int64_t)(-tk.list - 1) * 27;
    const uint32_t oct = (uint32_t)(ckey[tk.tcell] & 7u); // A's octant: bit d set = upper half along axis d
    const int32_t q = lane < 27 ? row[lane] : -1;
    const int32_t sb = q >= 0 ? cbeg[q] : 0;
    const int32_t cnt = q >= 0 ? cend[q] - sb : 0;
    bool near = true;
    for (int d = 0; d < 3; ++d) {
      const int od = (d == 0 ? lane % 3 : d == 1 ? (lane / 3) % 3 : lane / 9) - 1;
      near = near && (od == 0 || od == (int)((oct >> d) & 1u) * 2 - 1);
    }
    const uint64_t nm = __ballot(lane < 27 && near), fm = __ballot(lane < 27 && !near);
    const uint64_t below = (1ull << lane) - 1ull;
    const int nn = __popcll(nm);
    const int slot = near ? __popcll(nm & below) : nn + __popcll(fm & below);
    if (lane < 27) {
      tmp[slot] = sb;
      tmp[27 + slot] = cnt;
    }
    __syncthreads();
    const int32_t c2 = lane < 27 ? tmp[27 + lane] : 0;
    int32_t x = c2;
    for (int off = 1; off < kW; off <<= 1) {
      const int32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane < 27) {
      lb[lane] = tmp[lane];
      lp[lane] = x - c2;
    }
    near_end = __shfl(x, nn - 1);
    const int32_t total = __shfl(x, kW - 1);
    __syncthreads();
    return total;
  }
  const int32_t *list = tk.list >= 0 ? sub_src + tk.list : nbr27 + (int64_t)(-tk.list - 1) * 27;
  int32_t carry = 0;
  for (int base = 0; base < ne; base += kW) {
    const int e = base + lane;
    const int32_t q = e < ne ? list[e] : -1;
    const int32_t sb = q >= 0 ? cbeg[q] : 0;
    const int32_t cnt = q >= 0 ? cend[q] - sb : 0;
    int32_t x = cnt;
    for (int off = 1; off < kW; off <<= 1) {
      const int32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (e < ne) {
      lb[e] = sb;
      lp[e] = carry + x - cnt;
    }
    carry += __shfl(x, kW - 1);
  }
  near_end = tk.list >= 0 ? 0 : carry;
  __syncthreads();
  return carry;
}

// Sum over the 64 lanes of each of G values (G a power of two <= 64) in a
// fixed order, transposed: at every halving step a lane keeps one half of
// its values and adds the other half of its partner's (offset 32, 16, ...),
// then the remaining lanes of each value are added by plain xor-shuffles.
// G - 1 + log2(64 / G) shuffles instead of 6 G; lane l ends with the total
// of value l >> (6 - log2 G).  The first KLO steps add in the values' own
// type T (FP32 in the mixed kernels: each step adds one rounding to lane sums
// of tens of FP32 terms, and those steps hold most of the shuffles), the
// rest in FP64.
template <int G, int KLO, class T>
__device__ __forceinline__ double tr_reduce(T (&v)[G], int lane) {
  static_assert(G >= 1 && G <= kW && (G & (G - 1)) == 0, "G is a power of two up to the wavefront");
  static_assert((G >> KLO) >= 1, "at most log2 G steps in T");
#pragma unroll
  for (int l = 0; l < KLO; ++l) {
    const int S = G >> l, o = (kW / 2) >> l;
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int k = 0; k < S / 2; ++k) {
      const T a = v[k], c = v[k + S / 2];
      if (sizeof(T) == 4 && (o == 32 || o == 16)) {
        // gfx950 lane swaps: the half (o = 32) or the odd rows (o = 16) of a
        // trade places with the other half / even rows of c, so r[0] + r[1]
        // holds a's pair sums where lane & o is clear and c's where it is set
        const unsigned ua = __builtin_bit_cast(unsigned, (float)a), uc = __builtin_bit_cast(unsigned, (float)c);
        const auto r = o == 32 ? __builtin_amdgcn_permlane32_swap(ua, uc, false, false)
                               : __builtin_amdgcn_permlane16_swap(ua, uc, false, false);
        v[k] = (T)(__builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]));
      } else {
        const T send = up ? a : c, keep = up ? c : a;
        v[k] = keep + __shfl_xor(send, o);
      }
    }
  }
  constexpr int G2 = G >> KLO;
  double d[G2];
#pragma unroll
  for (int k = 0; k < G2; ++k) d[k] = (double)v[k];
#pragma unroll
  for (int S = G2, o = (kW / 2) >> KLO; S > 1; S >>= 1, o >>= 1) {
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int k = 0; k < S / 2; ++k) {
      const double a = d[k], c = d[k + S / 2];
      const double send = up ? a : c, keep = up ? c : a;
      d[k] = keep + __shfl_xor(send, o);
    }
  }
  double x = d[0];
#pragma unroll
  for (int o = kW / G / 2; o >= 1; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Lanes take sources, targets are unrolled: the task's targets are staged
// in LDS and, kGroup at a time, every lane keeps one sum per target in
// registers while it walks its sources (concatenated index lane, lane + 64,
// ...); each target read is an LDS broadcast.  So every lane works whatever
// the number of targets (blocks hold 1 to a few dozen particles against
// thousands of sources), and a target's partial sums meet once, in an FP64
// reduction across the wavefront.  A lane's sum runs over at most
// kTaskPairs / 64 sources.  The sources in cells touching the target cell
// run with the pair Tr, the others with TrFar, in two loops one after the
// other; both keep the kernel at 128 registers (four wavefronts per SIMD).
// A source's particle comes from an LDS table of the list cell of every
// concatenated index in a window of kSrcWin indices (one byte each) and the
// cell's offset particle - index: two LDS reads, where walking the cell
// boundaries cost a loop of dependent reads per source.  One wavefront per
// task (tasks blockIdx.x, blockIdx.x + gridDim.x, ... when the grid is
// smaller than the task list).  POT also sums m_j / r_ij into pot.
// Wavefronts per SIMD the narrow kernel is held to: four (128 registers) for
// the group sizes measured to fit without spills in their loops, otherwise
// left to the compiler
template <class Tr, int KG, bool POT>
constexpr int p2p_waves() {
  return (!POT && KG <= (Tr::kAnchor ? 16 : 8)) ? 4 : 1;
}
constexpr int pow2_at_least(int x) { return x <= 1 ? 1 : 2 * pow2_at_least((x + 1) / 2); }

constexpr int kSrcWin = 4096; // concatenated source indices per table window (list cells < 256)
static_assert(kMaxList <= 256, "list cells are stored in one byte");

// tab[g - w0] = list cell of concatenated index g, for g in [w0, w1)
// (lp: first index of each of the ne cells, total: their sum).  Lane e
// writes the entries of cell e (e + 64, ... for longer lists) itself when
// they are at most kFillRun, one byte per step; longer runs are written by
// the whole wavefront afterwards.  Ends with a barrier.
constexpr int kFillRun = 32;
__device__ __forceinline__ void fill_cells(uint8_t *tab, const int32_t *lp, int ne, int32_t total, int32_t w0,
                                           int32_t w1, int lane) {
  for (int e0 = 0; e0 < ne; e0 += kW) {
    const int e = e0 + lane;
    int32_t a = w1, b = w1;
    if (e < ne) {
      a = max(lp[e], w0);
      b = min(e + 1 < ne ? lp[e + 1] : total, w1);
    }
    const bool own = b - a <= kFillRun;
    if (own)
      for (int32_t g = a; g < b; ++g) tab[g - w0] = (uint8_t)e;
    for (uint64_t m = __ballot(!own); m; m &= m - 1) {
      const int k = __ffsll((unsigned long long)m) - 1;
      const int32_t ka = __shfl(a, k), kb = __shfl(b, k);
      for (int32_t g = ka + lane; g < kb; g += kW) tab[g - w0] = (uint8_t)(e0 + k);
    }
  }
  __syncthreads();
}

// The pairs of targets st[t0, t0 + ng) with the sources [g0, g1) of the
// table window at w0, two per lane and iteration (g and g + 64): the two
// sources are converted, the next two loads are issued into the same
// registers, and the pairs are evaluated while they are in flight.  The
// loads are unconditional (past the end they repeat a source and are not
// used) and a scheduling barrier keeps them after the conversions: hoisted
// above them, both pairs of positions are live at once and the loop copies
// them at every iteration (16 64-bit moves) or spills.
template <class P, class SS, int KG, int STEP, bool POT>
__device__ __forceinline__ void src_pairs(int32_t g0, int32_t g1, int32_t w0, const uint8_t *tab, const int32_t *dl,
                                          const double4 *__restrict__ pos, const double4 &anchor, const SS *st,
                                          int t0, int ng, typename P::A (&ax)[KG], typename P::A (&ay)[KG],
                                          typename P::A (&az)[KG], typename P::A (&ap)[KG], int lane) {
  typedef typename P::S S;
  // particle of source g < g1; a missing second source repeats the first
  // (its mass is zeroed), so every pair stays finite
  auto load = [&](int32_t g, int32_t g_first) -> double4 {
    const int32_t gg = g < g1 ? g : g_first;
    return pos[dl[tab[gg - w0]] + gg];
  };
  auto pairs2 = [&](const S &x1, const S &x2) {
#pragma unroll
    for (int t4 = 0; t4 < KG; t4 += STEP)
      if (t4 < ng) // STEP targets a branch (sums of slots past ng are never used)
#pragma unroll
        for (int t = t4; t < t4 + STEP; ++t) {
          const auto xt = P::tgt(st[t0 + t]);
          P::template pair<POT>(xt, x1, ax[t], ay[t], az[t], ap[t]);
          P::template pair<POT>(xt, x2, ax[t], ay[t], az[t], ap[t]);
        }
  };
  if (g0 >= g1) return;
  int32_t g = g0 + lane, base = g0; // the lane's source, the wavefront's first
  // a lane without a source (fewer than 64) loads the first one and leaves
  // before the tail
  const int32_t gv = g < g1 ? g : g0;
  double4 q1 = load(gv, gv), q2 = load(g + kW, gv);
  // full iterations: every lane has both sources
  for (; base + 2 * kW <= g1; base += 2 * kW, g += 2 * kW) {
    const S x1 = P::make(q1, anchor), x2 = P::make(q2, anchor);
    __builtin_amdgcn_sched_barrier(0);
    {
      const int32_t n1 = g + 2 * kW < g1 ? g + 2 * kW : g;
      q1 = load(n1, n1);
      q2 = load(g + 3 * kW, n1);
    }
    pairs2(x1, x2);
  }
  if (base >= g1 || g >= g1) return; // no tail, or none for this lane
  // the tail (at most 127 sources): two per lane while some lane has a second
  // one (a missing second one is massless), else one
  const S x1 = P::make(q1, anchor);
  if (base + kW < g1) {
    S x2 = P::make(q2, anchor);
    if (!(g + kW < g1)) P::massless(x2);
    pairs2(x1, x2);
  } else {
#pragma unroll
    for (int t4 = 0; t4 < KG; t4 += STEP)
      if (t4 < ng)
#pragma unroll
        for (int t = t4; t < t4 + STEP; ++t)
          P::template pair<POT>(P::tgt(st[t0 + t]), x1, ax[t], ay[t], az[t], ap[t]);
  }
}

// One narrow task (the body of k_p2p_tasks and of the queue kernel's narrow
// branch); st, lb, lp, tmp, tab are the wavefront's LDS.
template <class Tr, class TrFar, int KG, bool POT>
__device__ __forceinline__ void narrow_task(const P2PTask &tk, const int32_t *__restrict__ nbr27,
                                            const int32_t *__restrict__ sub_src, const int32_t *__restrict__ cbeg,
                                            const int32_t *__restrict__ cend, const uint64_t *__restrict__ ckey,
                                            const double4 *__restrict__ pos, double *__restrict__ grad,
                                            double *__restrict__ pot, typename Tr::S *st, int32_t *lb, int32_t *lp,
                                            int32_t *tmp, uint8_t *tab, int lane) {
  typedef typename Tr::A A;
  const int32_t b = cbeg[tk.tcell] + tk.toff, nc = min(cend[tk.tcell] - b, kTaskTargets);
  // the offsets' origin: the task's first target (every target lies in its
  // cell, within the cell's diagonal of it)
  double4 anchor = make_double4(0, 0, 0, 0);
  if (Tr::kAnchor) anchor = pos[b];
  for (int t = lane; t < nc; t += kW) st[t] = Tr::make(pos[b + t], anchor);
  int32_t near_end;
  const int32_t total = stage_list(tk, nbr27, sub_src, cbeg, cend, ckey, lb, lp, tmp, lane, near_end);
  const int ne = tk.nlist;
  const int32_t f0 = tk.f0, f1 = min(tk.f1, total);
  const int32_t fm = max(f0, min(f1, near_end)); // [f0, fm) near cells, [fm, f1) the others
  // lb becomes the offset particle - concatenated index of every list cell
  for (int e = lane; e < ne; e += kW) lb[e] -= lp[e];
  const bool one_win = f1 - f0 <= kSrcWin;
  if (one_win) fill_cells(tab, lp, ne, total, f0, f1, lane); // ends with a barrier
  else __syncthreads();
  constexpr int kGroup = KG;
  for (int t0 = 0; t0 < nc; t0 += kGroup) {
    const int ng = min(kGroup, nc - t0);
    A ax[kGroup], ay[kGroup], az[kGroup], ap[kGroup];
#pragma unroll
    for (int t = 0; t < kGroup; ++t) ax[t] = ay[t] = az[t] = ap[t] = 0;
    for (int32_t w0 = f0; w0 < f1; w0 += kSrcWin) {
      const int32_t w1 = min(f1, w0 + kSrcWin);
      if (!one_win) {
        __syncthreads(); // the previous window's table reads are done
        fill_cells(tab, lp, ne, total, w0, w1, lane);
      }
      src_pairs<Tr, typename Tr::S, KG, Tr::kStep, POT>(max(f0, w0), min(fm, w1), w0, tab, lb, pos, anchor, st, t0, ng, ax, ay,
                                             az, ap, lane);
      // the far loop's set-up stays after the near loop: moved ahead of it,
      // its registers would be held through it
      asm volatile("" ::: "memory");
      src_pairs<TrFar, typename Tr::S, KG, Tr::kStep, POT>(max(fm, w0), w1, w0, tab, lb, pos, anchor, st, t0, ng, ax, ay, az,
                                                ap, lane);
    }

    // each component's kGroup sums (zero-padded to a power of two RG) meet
    // in one transposed FP64 reduction: afterwards lane l holds the total
    // of target l >> (6 - log2 RG)
    constexpr int RG = pow2_at_least(kGroup);
    constexpr int kSpan = kW / RG; // lanes sharing one target's total
    const int tl = lane / kSpan;
    const bool writer = (lane % kSpan) == 0 && tl < ng;
    const int64_t i = b + t0 + tl;
    constexpr int kLo = sizeof(A) == 4 ? 2 : 0; // steps of the reduction in FP32
    A v[RG];
#pragma unroll
    for (int t = 0; t < RG; ++t) v[t] = t < kGroup ? ax[t] : A(0);
    const double gx = tr_reduce<RG, kLo>(v, lane);
#pragma unroll
    for (int t = 0; t < RG; ++t) v[t] = t < kGroup ? ay[t] : A(0);
    const double gy = tr_reduce<RG, kLo>(v, lane);
#pragma unroll
    for (int t = 0; t < RG; ++t) v[t] = t < kGroup ? az[t] : A(0);
    const double gz = tr_reduce<RG, kLo>(v, lane);
    double gp = 0.0;
    if (POT) {
#pragma unroll
      for (int t = 0; t < RG; ++t) v[t] = t < kGroup ? ap[t] : A(0);
      gp = tr_reduce<RG, kLo>(v, lane);
    }
    if (writer) {
      atomicAdd(&grad[3 * i], gx);
      atomicAdd(&grad[3 * i + 1], gy);
      atomicAdd(&grad[3 * i + 2], gz);
      if (POT) atomicAdd(&pot[i], gp);
    }
  }
}

template <class Tr, class TrFar, int KG, bool POT>
__global__ void __launch_bounds__(kW) __attribute__((amdgpu_waves_per_eu(p2p_waves<Tr, KG, POT>()))) k_p2p_tasks(const DevPlan *__restrict__ dp, const int32_t *__restrict__ rng, const P2PTask *__restrict__ tasks,
                                                  const int32_t *__restrict__ nbr27,
                                                  const int32_t *__restrict__ sub_src,
                                                  const int32_t *__restrict__ cbeg, const int32_t *__restrict__ cend,
                                                  const uint64_t *__restrict__ ckey,
                                                  const double4 *__restrict__ pos, double *__restrict__ grad,
                                                  double *__restrict__ pot) {
  // narrow tasks hold at most kNarrowCap targets; the last pass reads up to KG - 1 slots past nc
  __shared__ typename Tr::S st[kNarrowCap + KG];
  __shared__ int32_t lb[kMaxList];
  __shared__ int32_t lp[kMaxList];
  __shared__ int32_t tmp[54];
  __shared__ uint8_t tab[kSrcWin];
  const int lane = threadIdx.x;
  if (plan_dead(dp)) return;
  const int64_t t0 = rng[0], ntask = rng[1] - t0; // the tasks [rng[0], rng[1])
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    __syncthreads(); // the previous task's reads of lb, lp, st, tab are done
    narrow_task<Tr, TrFar, KG, POT>(tasks[t0 + task], nbr27, sub_src, cbeg, cend, ckey, pos, grad, pot, st, lb, lp, tmp,
                                    tab, lane);
  }
}

// One stretch of k_p2p_wide over the sources [g0, g1): staged 64 at a time in
// LDS (each lane finds its source's cell by binary search) and read once per
// 4 pairs; a chunk's contributions are summed per lane and added with the
// precision's add (Kahan in FP32).
template <class Pr, class Ad, bool POT, class TS>
__device__ __forceinline__ void wide_sweep(int32_t g0, int32_t g1, const int32_t *lb, const int32_t *lp, int ne,
                                           const double4 *__restrict__ pos, const double4 &anchor,
                                           typename Pr::S *sh, const TS (&xi)[4], int si, int sub, int lane,
                                           typename Pr::A (&sx)[4], typename Pr::A (&sy)[4], typename Pr::A (&sz)[4],
                                           typename Pr::A (&sp)[4], typename Pr::A (&cx)[4],
                                           typename Pr::A (&cy)[4], typename Pr::A (&cz)[4],
                                           typename Pr::A (&cp)[4]) {
  typedef typename Pr::S S;
  typedef typename Pr::A A;
  for (int32_t fb = g0; fb < g1; fb += kW) {
    const int32_t g = fb + lane;
    __syncthreads();
    if (g < g1) {
      int lo = 0, hi = ne; // last list cell with lp <= g (lp[0] = 0 <= g)
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lp[mid] <= g)
          lo = mid;
        else
          hi = mid;
      }
      sh[lane] = Pr::make(pos[lb[lo] + (g - lp[lo])], anchor);
    }
    __syncthreads();
    const int cnt = min(kW, g1 - fb);
    A ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0}, az[4] = {0, 0, 0, 0}, ap[4] = {0, 0, 0, 0};
    for (int k = si; k < cnt; k += sub) {
      const S xj = sh[k];
#pragma unroll
      for (int q = 0; q < 4; ++q) Pr::template pair<POT>(Pr::tgt(xi[q]), xj, ax[q], ay[q], az[q], ap[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Ad::add(sx[q], cx[q], ax[q]);
      Ad::add(sy[q], cy[q], ay[q]);
      Ad::add(sz[q], cz[q], az[q]);
      if (POT) Ad::add(sp[q], cp[q], ap[q]);
    }
  }
}

// The kernel for blocks of many targets (more than kNarrowMax: one pass over
// the sources instead of several): lanes are tl target lanes (power of two,
// 4 targets each: t = ti + tl q) times 64 / tl source slots; the sources are
// walked by wide_sweep, the cells touching the target cell with the pair Tr
// and the others with TrFar, and the slots are combined with xor-shuffles at
// the end.  One wavefront per task.
// One wide task (the body of k_p2p_wide and of the queue kernel's wide
// branch); sh, shf, lb, lp, tmp are the wavefront's LDS.
template <class Tr, class TrFar, bool POT>
__device__ __forceinline__ void wide_task(const P2PTask &tk, const int32_t *__restrict__ nbr27,
                                          const int32_t *__restrict__ sub_src, const int32_t *__restrict__ cbeg,
                                          const int32_t *__restrict__ cend, const uint64_t *__restrict__ ckey,
                                          const double4 *__restrict__ pos, double *__restrict__ grad,
                                          double *__restrict__ pot, typename Tr::S *sh, typename TrFar::S *shf,
                                          int32_t *lb, int32_t *lp, int32_t *tmp, int lane) {
  typedef typename Tr::S S;
  typedef typename Tr::A A;
  const int32_t b = cbeg[tk.tcell] + tk.toff, nc = min(cend[tk.tcell] - b, kTaskTargets);
  // the offsets' origin: the task's first target (every target lies in its
  // cell, within the cell's diagonal of it)
  double4 anchor = make_double4(0, 0, 0, 0);
  if (Tr::kAnchor) anchor = pos[b];
  int32_t near_end;
  const int32_t total = stage_list(tk, nbr27, sub_src, cbeg, cend, ckey, lb, lp, tmp, lane, near_end);
  const int ne = tk.nlist;
  const int32_t f0 = tk.f0, f1 = min(tk.f1, total);
  const int32_t fm = max(f0, min(f1, near_end)); // [f0, fm) near cells, [fm, f1) the others
  int tl = 1;
  while (tl * 4 < nc) tl <<= 1;
  const int sub = kW / tl, ti = lane & (tl - 1), si = lane / tl;
  S xi[4];
  int32_t ii[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int32_t t = ti + tl * q;
    ii[q] = t < nc ? b + t : -1;
    xi[q] = Tr::make(pos[t < nc ? b + t : b], anchor);
  }
  A sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, sz[4] = {0, 0, 0, 0}, sp[4] = {0, 0, 0, 0};
  A cx[4] = {0, 0, 0, 0}, cy[4] = {0, 0, 0, 0}, cz[4] = {0, 0, 0, 0}, cp[4] = {0, 0, 0, 0};
  wide_sweep<Tr, Tr, POT>(f0, fm, lb, lp, ne, pos, anchor, sh, xi, si, sub, lane, sx, sy, sz, sp, cx, cy, cz, cp);
  wide_sweep<TrFar, Tr, POT>(fm, f1, lb, lp, ne, pos, anchor, shf, xi, si, sub, lane, sx, sy, sz, sp, cx, cy, cz,
                             cp);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double gx = (double)sx[q] + (double)cx[q], gy = (double)sy[q] + (double)cy[q],
           gz = (double)sz[q] + (double)cz[q];
    for (int o = tl; o < kW; o <<= 1) {
      gx += __shfl_xor(gx, o);
      gy += __shfl_xor(gy, o);
      gz += __shfl_xor(gz, o);
    }
    double gp = 0;
    if (POT) {
      gp = (double)sp[q] + (double)cp[q];
      for (int o = tl; o < kW; o <<= 1) gp += __shfl_xor(gp, o);
    }
    if (si == 0 && ii[q] >= 0) {
      atomicAdd(&grad[3 * (int64_t)ii[q]], gx);
      atomicAdd(&grad[3 * (int64_t)ii[q] + 1], gy);
      atomicAdd(&grad[3 * (int64_t)ii[q] + 2], gz);
      if (POT) atomicAdd(&pot[ii[q]], gp);
    }
  }
}

template <class Tr, class TrFar, bool POT>
__global__ void __launch_bounds__(kW) k_p2p_wide(const DevPlan *__restrict__ dp, const int32_t *__restrict__ rng,
                                                  const P2PTask *__restrict__ tasks,
                                                  const int32_t *__restrict__ nbr27,
                                                  const int32_t *__restrict__ sub_src,
                                                  const int32_t *__restrict__ cbeg, const int32_t *__restrict__ cend,
                                                  const uint64_t *__restrict__ ckey,
                                                  const double4 *__restrict__ pos, double *__restrict__ grad,
                                                  double *__restrict__ pot) {
  __shared__ typename Tr::S sh[kW];
  __shared__ typename TrFar::S shf[kW];
  __shared__ int32_t lb[kMaxList];
  __shared__ int32_t lp[kMaxList];
  __shared__ int32_t tmp[54];
  const int lane = threadIdx.x;
  if (plan_dead(dp)) return;
  const int64_t t0 = rng[0], ntask = rng[1] - t0; // the tasks [rng[0], rng[1])
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    __syncthreads(); // the previous task's reads of lb, lp, sh are done
    wide_task<Tr, TrFar, POT>(tasks[t0 + task], nbr27, sub_src, cbeg, cend, ckey, pos, grad, pot, sh, shf, lb, lp, tmp, lane);
  }
}

// Persistent P2P over a work queue: a fixed grid of one-wavefront workgroups
// takes tasks in queue order from an atomic counter until the queue is
// drained, so one launch covers every task of the share (every system's, in
// a batch) however uneven their sizes.  The share's wide tasks are
// [rw[0], rw[1]), its narrow ones [rn[0], rn[1]) (device ranges); queue index
// q < nwide is the wide task wide0 + q (the larger ones first), the others
// the narrow tasks narrow0 + (q - nwide).  A wavefront takes `chunk` consecutive tasks per
// atomic: one task per grab serialises some 10^5 atomics on one address.
// *ctr is zero at the launch.
template <class Tr, class TrFar, int KG, bool POT>
__global__ void __launch_bounds__(kW) __attribute__((amdgpu_waves_per_eu(p2p_waves<Tr, KG, POT>()))) k_p2p_queue(
    const DevPlan *__restrict__ dp, const int32_t *__restrict__ rw, const int32_t *__restrict__ rn, int chunk,
    int32_t *__restrict__ ctr,
    const P2PTask *__restrict__ tasks, const int32_t *__restrict__ nbr27, const int32_t *__restrict__ sub_src,
    const int32_t *__restrict__ cbeg, const int32_t *__restrict__ cend, const uint64_t *__restrict__ ckey,
    const double4 *__restrict__ pos, double *__restrict__ grad, double *__restrict__ pot) {
  typedef typename Tr::S S;
  typedef typename TrFar::S FS;
  // the narrow branch's targets and the wide branch's source chunks share one buffer
  constexpr size_t kNarrowBytes = sizeof(S) * (kNarrowCap + KG);
  constexpr size_t kWideBytes = sizeof(S) * kW + sizeof(FS) * kW;
  __shared__ __align__(16) char buf[kNarrowBytes > kWideBytes ? kNarrowBytes : kWideBytes];
  __shared__ int32_t lb[kMaxList];
  __shared__ int32_t lp[kMaxList];
  __shared__ int32_t tmp[54];
  __shared__ uint8_t tab[kSrcWin];
  __shared__ int32_t next;
  const int lane = threadIdx.x;
  if (plan_dead(dp)) return;
  const int64_t wide0 = rw[0], nwide = rw[1] - rw[0], narrow0 = rn[0];
  const int64_t total = nwide + (rn[1] - rn[0]);
  for (;;) {
    __syncthreads(); // the previous task's LDS reads are done
    if (lane == 0) next = atomicAdd(ctr, chunk);
    __syncthreads();
    const int64_t q0 = next;
    if (q0 >= total) break; // uniform: every wavefront leaves once the queue is drained
    const int64_t q1 = min(q0 + chunk, total);
    for (int64_t q = q0; q < q1; ++q) {
      if (q > q0) __syncthreads();
      if (q < nwide)
        wide_task<Tr, TrFar, POT>(tasks[wide0 + q], nbr27, sub_src, cbeg, cend, ckey, pos, grad, pot,
                                  reinterpret_cast<S *>(buf), reinterpret_cast<FS *>(buf + sizeof(S) * kW), lb, lp,
                                  tmp, lane);
      else
        narrow_task<Tr, TrFar, KG, POT>(tasks[narrow0 + (q - nwide)], nbr27, sub_src, cbeg, cend, ckey, pos, grad,
                                        pot, reinterpret_cast<S *>(buf), lb, lp, tmp, tab, lane);
    }
  }
}

__global__ void k_zero_i32(int32_t *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// acc = G * grad; with potentials, phi = -G * sum_j m_j / r_ij
__global__ void k_finish(int64_t n, const double *__restrict__ far, const double *__restrict__ near,
                         const double *__restrict__ pfar, const double *__restrict__ pnear,
                         const int32_t *__restrict__ idx, double G, double *__restrict__ acc_out,
                         double *__restrict__ phi_out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int64_t out = idx[i];
  acc_out[3 * out] = G * (far[3 * i] + near[3 * i]);
  acc_out[3 * out + 1] = G * (far[3 * i + 1] + near[3 * i + 1]);
  acc_out[3 * out + 2] = G * (far[3 * i + 2] + near[3 * i + 2]);
  if (phi_out) phi_out[out] = -G * (pfar[i] + pnear[i]);
}

// ---- direct sums ----------------------------------------------------------------------
// mode 0: NaiveUpdate (naive_update.cpp:20-30) with the reference's rounding:
//   s = m_j / ((r^2 + eps2) * r); grad -= d * s; acc = grad * G, sources in
//   input order, every operation individually rounded (no contraction).
__global__ void __launch_bounds__(kT) k_direct_naive(int64_t n, const double *__restrict__ xyz,
                                                     const double *__restrict__ m, int64_t ntgt,
                                                     const int64_t *__restrict__ tgt, double G, double eps2,
                                                     double *__restrict__ acc) {
#pragma clang fp contract(off)
  __shared__ double4 sh[kT];
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool valid = t < ntgt;
  const int64_t i = valid ? (tgt ? tgt[t] : t) : 0;
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int64_t base = 0; base < n; base += kT) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    sh[threadIdx.x] = j < n ? make_double4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], m[j]) : make_double4(0, 0, 0, 0);
    __syncthreads();
    const int cnt = (int)((n - base) < kT ? (n - base) : kT);
    for (int k = 0; k < cnt; ++k) {
      const double4 q = sh[k];
      if (base + k == i) continue;
      const double d0 = xi - q.x, d1 = yi - q.y, d2 = zi - q.z;
      double r2 = d0 * d0;
      r2 = r2 + d1 * d1;
      r2 = r2 + d2 * d2;
      if (r2 == 0.0) continue;
      const double s = q.w / ((r2 + eps2) * __dsqrt_rn(r2));
      s0 = s0 - d0 * s;
      s1 = s1 - d1 * s;
      s2 = s2 - d2 * s;
    }
  }
  if (!valid) return;
  acc[3 * t] = s0 * G;
  acc[3 * t + 1] = s1 * G;
  acc[3 * t + 2] = s2 * G;
}

// mode 1: grad += m_j (x_j - x_i) / (r^2 + eps2)^{3/2} with Neumaier-compensated
// sums per component (the FP64 oracle of the accuracy metric).  Block (x, y)
// sums source chunk y for 256 targets; k_direct_combine adds the chunk
// partials in chunk order, again compensated.
__device__ __forceinline__ void neumaier(double &s, double &c, double v) {
  const double u = s + v;
  c += fabs(s) >= fabs(v) ? (s - u) + v : (v - u) + s;
  s = u;
}

__global__ void __launch_bounds__(kT) k_direct_part(int64_t n, const double *__restrict__ xyz,
                                                    const double *__restrict__ m, int64_t ntgt,
                                                    const int64_t *__restrict__ tgt, double eps2, int64_t chunk,
                                                    double *__restrict__ part) {
  __shared__ double4 sh[kT];
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool valid = t < ntgt;
  const int64_t i = valid ? (tgt ? tgt[t] : t) : 0;
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  double s[3] = {0, 0, 0}, cc[3] = {0, 0, 0};
  const int64_t j0 = (int64_t)blockIdx.y * chunk, j1 = min(n, j0 + chunk);
  for (int64_t base = j0; base < j1; base += kT) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    sh[threadIdx.x] = j < j1 ? make_double4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], m[j]) : make_double4(0, 0, 0, 0);
    __syncthreads();
    const int cnt = (int)((j1 - base) < kT ? (j1 - base) : kT);
    for (int k = 0; k < cnt; ++k) {
      const double4 q = sh[k];
      if (base + k == i) continue;
      const double d0 = q.x - xi, d1 = q.y - yi, d2 = q.z - zi;
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2 + eps2;
      if (r2 == 0.0) continue;
      const double w = q.w / (r2 * sqrt(r2));
      neumaier(s[0], cc[0], d0 * w);
      neumaier(s[1], cc[1], d1 * w);
      neumaier(s[2], cc[2], d2 * w);
    }
  }
  if (!valid) return;
  double *o = part + ((size_t)blockIdx.y * ntgt + t) * 6;
  for (int d = 0; d < 3; ++d) {
    o[2 * d] = s[d];
    o[2 * d + 1] = cc[d];
  }
}

__global__ void k_direct_combine(int64_t ntgt, int nchunk, const double *__restrict__ part, double G,
                                 double *__restrict__ acc) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= ntgt) return;
  for (int d = 0; d < 3; ++d) {
    double s = 0, c = 0;
    for (int k = 0; k < nchunk; ++k) {
      const double *o = part + ((size_t)k * ntgt + t) * 6;
      neumaier(s, c, o[2 * d]);
      neumaier(s, c, o[2 * d + 1]);
    }
    acc[3 * t + d] = G * (s + c);
  }
}

typedef void (*P2PKernel)(const DevPlan *, const int32_t *, const P2PTask *, const int32_t *, const int32_t *,
                          const int32_t *, const int32_t *, const uint64_t *, const double4 *, double *, double *);

// the narrow kernel's instance for a target group size (the sizes the tiles allow)
P2PKernel narrow_kernel(bool f32, bool pot, int group) {
#define NB_NARROW(TR, FAR, G) \
  case G:                     \
    return pot ? k_p2p_tasks<TR, FAR, G, true> : k_p2p_tasks<TR, FAR, G, false>;
  if (f32) {
    switch (group) {
      NB_NARROW(PairF32, PairF32Far, 8)
      NB_NARROW(PairF32, PairF32Far, 12)
      NB_NARROW(PairF32, PairF32Far, 16)
      NB_NARROW(PairF32, PairF32Far, 24)
    }
  } else {
    switch (group) {
      NB_NARROW(PairF64, PairF64Far, 4)
      NB_NARROW(PairF64, PairF64Far, 8)
      NB_NARROW(PairF64, PairF64Far, 12)
    }
  }
#undef NB_NARROW
  throw Error(NBFMM_ERR_ARG, "no P2P kernel for target group " + std::to_string(group));
}

typedef void (*P2PQueue)(const DevPlan *, const int32_t *, const int32_t *, int, int32_t *, const P2PTask *,
                         const int32_t *, const int32_t *, const int32_t *, const int32_t *, const uint64_t *,
                         const double4 *, double *, double *);

// the queue kernel's instance for a target group size
P2PQueue queue_kernel(bool f32, bool pot, int group) {
#define NB_QUEUE(TR, FAR, G) \
  case G:                    \
    return pot ? k_p2p_queue<TR, FAR, G, true> : k_p2p_queue<TR, FAR, G, false>;
  if (f32) {
    switch (group) {
      NB_QUEUE(PairF32, PairF32Far, 8)
      NB_QUEUE(PairF32, PairF32Far, 12)
      NB_QUEUE(PairF32, PairF32Far, 16)
      NB_QUEUE(PairF32, PairF32Far, 24)
    }
  } else {
    switch (group) {
      NB_QUEUE(PairF64, PairF64Far, 4)
      NB_QUEUE(PairF64, PairF64Far, 8)
      NB_QUEUE(PairF64, PairF64Far, 12)
    }
  }
#undef NB_QUEUE
  throw Error(NBFMM_ERR_ARG, "no P2P kernel for target group " + std::to_string(group));
}

} // namespace

// Share of the P2P tasks run in part 0 (mixed precision, FP64).
#ifndef NBFMM_P2P_EARLY
#define NBFMM_P2P_EARLY 0.5
#endif
#ifndef NBFMM_P2P_EARLY_F64
#define NBFMM_P2P_EARLY_F64 0.2
#endif
// Part 0's narrow tasks go out in this many launches (mixed precision, FP64).
// While one P2P launch still has workgroups to place, the kernels of the
// main stream get almost no CU slots (measured: the up-pass planning stalled
// until part 0 was placed); between launches they are placed.  Each launch
// waits for the previous one to drain, which costs more with the longer FP64
// tasks: measured best at 16 launches for mixed precision, 1 for FP64.
#ifndef NBFMM_P2P_CHUNKS0
#define NBFMM_P2P_CHUNKS0 16
#endif
#ifndef NBFMM_P2P_CHUNKS0_F64
#define NBFMM_P2P_CHUNKS0_F64 1
#endif

nbfmm_tiles default_tiles() {
  nbfmm_tiles t{};
  t.p2p_group_f32 = PairF32::kGroup;
  t.p2p_group_f64 = PairF64::kGroup;
  t.p2p_narrow_max_f32 = NBFMM_P2P_NARROW_F32;
  t.p2p_narrow_max_f64 = NBFMM_P2P_NARROW_F64;
  t.p2p_early_share_f32 = NBFMM_P2P_EARLY;
  t.p2p_early_share_f64 = NBFMM_P2P_EARLY_F64;
  t.p2p_early_launches_f32 = NBFMM_P2P_CHUNKS0;
  t.p2p_early_launches_f64 = NBFMM_P2P_CHUNKS0_F64;
  t.m2l_slices = 1;
  t.m2l_item_pairs = 1024;
  t.p2p_rest_launches = 4;
  t.m2l_mode = 1;
  t.p2p_queue = 0;
  t.p2p_queue_waves0 = 8;
  t.p2p_queue_chunk = 4;
  return t;
}

bool tiles_ok(const nbfmm_tiles &t, std::string &why) {
  auto in = [](int v, std::initializer_list<int> set) {
    for (int x : set)
      if (v == x) return true;
    return false;
  };
  if (!in(t.p2p_group_f32, {8, 12, 16, 24})) why = "p2p_group_f32 must be 8, 12, 16 or 24";
  else if (!in(t.p2p_group_f64, {4, 8, 12})) why = "p2p_group_f64 must be 4, 8 or 12";
  else if (t.p2p_narrow_max_f32 < 1 || t.p2p_narrow_max_f32 > kNarrowCap) why = "p2p_narrow_max_f32 must be in [1, 64]";
  else if (t.p2p_narrow_max_f64 < 1 || t.p2p_narrow_max_f64 > kNarrowCap) why = "p2p_narrow_max_f64 must be in [1, 64]";
  else if (!(t.p2p_early_share_f32 >= 0.0 && t.p2p_early_share_f32 <= 1.0)) why = "p2p_early_share_f32 must be in [0, 1]";
  else if (!(t.p2p_early_share_f64 >= 0.0 && t.p2p_early_share_f64 <= 1.0)) why = "p2p_early_share_f64 must be in [0, 1]";
  else if (t.p2p_early_launches_f32 < 1 || t.p2p_early_launches_f32 > 64) why = "p2p_early_launches_f32 must be in [1, 64]";
  else if (t.p2p_early_launches_f64 < 1 || t.p2p_early_launches_f64 > 64) why = "p2p_early_launches_f64 must be in [1, 64]";
  else if (t.m2l_slices != 1 && t.m2l_slices != 2) why = "m2l_slices must be 1 or 2";
  else if (t.m2l_item_pairs < 128 || t.m2l_item_pairs > 16384 || t.m2l_item_pairs % 128)
    why = "m2l_item_pairs must be a multiple of 128 in [128, 16384]";
  else if (t.p2p_rest_launches < 1 || t.p2p_rest_launches > kMaxChunks1)
    why = "p2p_rest_launches must be in [1, 8]";
  else if (t.m2l_mode != 1 && t.m2l_mode != 2) why = "m2l_mode must be 1 or 2";
  else if (t.p2p_queue < 0 || t.p2p_queue > 2) why = "p2p_queue must be 0, 1 or 2";
  else if (t.p2p_queue_waves0 < 1 || t.p2p_queue_waves0 > 16) why = "p2p_queue_waves0 must be in [1, 16]";
  else if (t.p2p_queue_chunk < 1 || t.p2p_queue_chunk > 64) why = "p2p_queue_chunk must be in [1, 64]";
  else return true;
  return false;
}

// The task ranges of the P2P launches (DevPlan::p2p_rng) from the task
// counts: the same share `early` of the narrow tasks [0, nn) and of the wide
// ones [nn, ntask) goes first, its narrow tasks in nch launches.
__global__ void k_p2p_ranges(DevPlan *dp, double early, int nch, int nch1) {
  if (threadIdx.x != 0) return;
  const int64_t nn = dp->nnarrow, nt = (int64_t)dp->ntask, nw = nt - nn;
  const int64_t sn = (int64_t)((double)nn * early), sw = nn + (int64_t)((double)nw * early);
  int32_t *r = dp->p2p_rng;
  for (int ch = 0; ch < nch; ++ch) {
    r[2 * ch] = (int32_t)(sn * ch / nch);
    r[2 * ch + 1] = (int32_t)(sn * (ch + 1) / nch);
  }
  for (int ch = 0; ch < nch1; ++ch) {
    r[kRngRest0 + 2 * ch] = (int32_t)(sn + (nn - sn) * ch / nch1);
    r[kRngRest0 + 2 * ch + 1] = (int32_t)(sn + (nn - sn) * (ch + 1) / nch1);
  }
  const int32_t v[12] = {0, (int32_t)sn, (int32_t)nn, (int32_t)sw, (int32_t)sn, (int32_t)nn,
                         (int32_t)sw, (int32_t)nt, 0, (int32_t)nn, (int32_t)nn, (int32_t)nt};
  for (int k = 0; k < 12; ++k) r[kRngNarrow0 + k] = v[k];
}

// The shares as k_p2p_ranges computes them, for the launch sizes: exact (the
// task counts of a direct evaluation) or from the capacities (planned; the
// kernels stride over their tasks, so any grid covers them)
struct P2PSplit {
  int nch, nch1;
  double early;
  int64_t rng[kRngCount];
};
static P2PSplit p2p_split(nbfmm_ctx *c) {
  const Plan &P = c->plan;
  const bool f32 = c->prm.precision == NBFMM_MIXED;
  const nbfmm_tiles &tl = c->shapes;
  P2PSplit sp{};
  sp.early = f32 ? tl.p2p_early_share_f32 : tl.p2p_early_share_f64;
  sp.nch = std::min(kMaxChunks0, f32 ? tl.p2p_early_launches_f32 : tl.p2p_early_launches_f64);
  const int64_t nn = c->cap.on ? c->cap.nnarrow : P.n_tasks_narrow;
  const int64_t nt = c->cap.on ? c->cap.ntask : P.n_tasks, nw = std::max<int64_t>(0, nt - nn);
  const int64_t sn = (int64_t)((double)nn * sp.early), sw = nn + (int64_t)((double)nw * sp.early);
  for (int ch = 0; ch < sp.nch; ++ch) {
    sp.rng[2 * ch] = sn * ch / sp.nch;
    sp.rng[2 * ch + 1] = sn * (ch + 1) / sp.nch;
  }
  sp.nch1 = std::min(kMaxChunks1, tl.p2p_rest_launches);
  for (int ch = 0; ch < sp.nch1; ++ch) {
    sp.rng[kRngRest0 + 2 * ch] = sn + (nn - sn) * ch / sp.nch1;
    sp.rng[kRngRest0 + 2 * ch + 1] = sn + (nn - sn) * (ch + 1) / sp.nch1;
  }
  const int64_t v[12] = {0, sn, nn, sw, sn, nn, sw, nt, 0, nn, nn, nt};
  for (int k = 0; k < 12; ++k) sp.rng[kRngNarrow0 + k] = v[k];
  return sp;
}

void fork_p2p(nbfmm_ctx *c) {
  const P2PSplit sp = p2p_split(c);
  hipLaunchKernelGGL(k_p2p_ranges, dim3(1), dim3(64), 0, c->stream, dplan(c), sp.early, sp.nch, sp.nch1);
  NB_LAUNCH();
  NB_HIP(hipEventRecord(c->ev_fork, c->stream));
}

void launch_p2p(nbfmm_ctx *c, int part) {
  // isolated timing: parts 0 and 1 are left out and run_finish runs every
  // task (part 2) once the main stream is idle
  if (c->isolated && part != 2) return;
  Plan &P = c->plan;
  hipStream_t s2 = c->stream2;
  const bool pot = c->phi_out != nullptr;
  double *gn = dget<double>(c->grad_near, (size_t)P.n * (pot ? 4 : 3));
  double *pn = pot ? gn + 3 * (size_t)P.n : nullptr;
  const bool f32 = c->prm.precision == NBFMM_MIXED;
  const nbfmm_tiles &tl = c->shapes;
  const DevPlan *dp = dplan(c);
  const int32_t *rg = dp->p2p_rng;
  const P2PSplit sp = p2p_split(c);
  if (part == 2) { // every task, one launch per kernel
    NB_HIP(hipEventRecord(c->ev_fork2, c->stream));
    NB_HIP(hipStreamWaitEvent(s2, c->ev_fork2, 0));
    fill_zero(s2, gn, sizeof(double) * (pot ? 4 : 3) * (size_t)P.n);
    mark2(c, 9);
    mark2(c, 10);
  }
  if (part == 1) {
    // after the point of the main stream where part 1 is enqueued (the first M2L contraction)
    NB_HIP(hipEventRecord(c->ev_fork2, c->stream));
    NB_HIP(hipStreamWaitEvent(s2, c->ev_fork2, 0));
  }
  if (part == 0) {
    NB_HIP(hipStreamWaitEvent(s2, c->ev_fork, 0));
    // near-field gradient (3 per particle) and, with potentials, near-field sums of m / r
    fill_zero(s2, gn, sizeof(double) * (pot ? 4 : 3) * (size_t)P.n);
  }
  if (part != 2) mark2(c, part == 0 ? 9 : 11);
  if (part == 2) mark2(c, 11);
  // this part's narrow and wide ranges (indices into p2p_rng)
  const int in = part == 0 ? kRngNarrow0 : part == 1 ? kRngNarrow1 : kRngNarrowAll;
  const int iw = part == 0 ? kRngWide0 : part == 1 ? kRngWide1 : kRngWideAll;
  auto len = [&](int k) { return sp.rng[k + 1] - sp.rng[k]; };
  auto grid_of = [&](int64_t cnt) { // a planned evaluation's counts are capacities: a quarter more
    return std::max<int64_t>(1, std::min<int64_t>(cnt, ((int64_t)1 << 31) - 1));
  };
  if (tl.p2p_queue == 2 || (tl.p2p_queue == 1 && part >= 1)) {
    // one persistent launch over this share's wide then narrow tasks
    const int64_t tot = len(iw) + len(in);
    if (tot > 0 || c->cap.on) {
      int32_t *ctr = dget<int32_t>(c->p2p_ctr, 4) + part;
      hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s2, ctr, 1);
      NB_LAUNCH();
      // a CU holds 16 of these wavefronts (128 VGPRs each): 16 per CU fill the device
      const int per_cu = part == 0 ? tl.p2p_queue_waves0 : 16;
      const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(tot, 1), (int64_t)c->n_cu * per_cu));
      const int chunk = tl.p2p_queue_chunk;
      if (tot + grid * chunk >= INT32_MAX) throw Error(NBFMM_ERR_ARG, "too many P2P tasks for the work queue");
      hipLaunchKernelGGL(queue_kernel(f32, pot, f32 ? tl.p2p_group_f32 : tl.p2p_group_f64), dim3((unsigned)grid),
                         dim3(kW), 0, s2, dp, rg + iw, rg + in, chunk, ctr, (const P2PTask *)c->tasks.p,
                         (const int32_t *)c->nbr27.p, (const int32_t *)c->sub_src.p, (const int32_t *)c->cbeg.p,
                         (const int32_t *)c->cend.p, (const uint64_t *)c->ckey.p, (const double4 *)c->pos.p, gn,
                         pn);
      NB_LAUNCH();
    }
    mark2(c, part == 0 ? 10 : 12);
    if (part >= 1) NB_HIP(hipEventRecord(c->ev_join, s2));
    return;
  }
  // part 1 in mixed precision starts with the wide tasks, so that its last
  // workgroups are short narrow tasks (measured 0.2 ms sooner at 4M; FP64,
  // whose narrow tasks are longer, measured slower that way)
  for (int wi = 0; wi < 2; ++wi) {
    const int w = (part == 1 && f32) ? 1 - wi : wi;
    const int nch = w == 0 && part == 0 ? sp.nch : w == 0 && part == 1 ? sp.nch1 : 1;
    for (int ch = 0; ch < nch; ++ch) {
      // the range's index: part 0's or part 1's narrow chunks, or the part's whole range
      const int k = w == 0 && part == 0 ? 2 * ch : w == 0 && part == 1 ? kRngRest0 + 2 * ch : (w == 0 ? in : iw);
      const int64_t cnt = len(k);
      if (cnt <= 0 && !c->cap.on) continue;
      auto kern = w == 0 ? narrow_kernel(f32, pot, f32 ? tl.p2p_group_f32 : tl.p2p_group_f64)
                         : (f32 ? (pot ? k_p2p_wide<PairF32, PairF32Far, true> : k_p2p_wide<PairF32, PairF32Far, false>)
                                : (pot ? k_p2p_wide<PairF64, PairF64Far, true> : k_p2p_wide<PairF64, PairF64Far, false>));
      hipLaunchKernelGGL(kern, dim3((unsigned)grid_of(cnt)), dim3(kW), 0, s2, dp, rg + k, (const P2PTask *)c->tasks.p,
                         (const int32_t *)c->nbr27.p, (const int32_t *)c->sub_src.p, (const int32_t *)c->cbeg.p,
                         (const int32_t *)c->cend.p, (const uint64_t *)c->ckey.p, (const double4 *)c->pos.p, gn, pn);
      NB_LAUNCH();
    }
  }
  mark2(c, part == 0 ? 10 : 12);
  if (part >= 1) NB_HIP(hipEventRecord(c->ev_join, s2));
}

void run_finish(nbfmm_ctx *c, double *d_acc_out) {
  Plan &P = c->plan;
  if (c->isolated) launch_p2p(c, 2);
  NB_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  const bool pot = c->phi_out != nullptr;
  const double *gn = (const double *)c->grad_near.p;
  hipLaunchKernelGGL(k_finish, dim3((unsigned)((P.n + kT - 1) / kT)), dim3(kT), 0, c->stream, P.n,
                     (const double *)c->grad.p, gn, pot ? (const double *)c->grad.p + 3 * (size_t)P.n : nullptr,
                     pot ? gn + 3 * (size_t)P.n : nullptr, (const int32_t *)c->idx[0].p, c->prm.G, d_acc_out,
                     c->phi_out);
  NB_LAUNCH();
}

namespace {
// The partial sums of the compensated direct sum for callers that cannot
// allocate (nbfmm_direct_accel_noalloc_dev): a fixed device array, one copy
// per GPU, kept until the process ends.  Host code never uses the array's
// name as an address: scratch_ptr(device) is its device address, looked up
// once per device by scratch_init.
constexpr size_t kScratchDoubles = (size_t)NBFMM_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// source chunks of the compensated direct sum for ntgt targets over n sources
int64_t direct_chunks(int64_t n, int64_t ntgt) {
  const int64_t gx = (ntgt + kT - 1) / kT;
  // enough chunks to give every CU several blocks
  int64_t nchunk = (2048 + gx - 1) / gx;
  if (nchunk > 256) nchunk = 256;
  if (nchunk > (n + kT - 1) / kT) nchunk = (n + kT - 1) / kT;
  if (nchunk < 1) nchunk = 1;
  return nchunk;
}
} // namespace

void scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(NBFMM_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return;
  // the caller has made `device` current; the symbol's address is that GPU's copy
  void *ptr = nullptr;
  NB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch[device] = (double *)ptr;
}

double *scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  double *p = device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
  if (!p) throw Error(NBFMM_ERR_HIP, "the fixed scratch of this device was not set up");
  return p;
}

void run_direct(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, int64_t ntgt, const int64_t *d_tgt,
                int mode, double *d_acc, double *part_fixed) {
  const unsigned gx = (unsigned)((ntgt + kT - 1) / kT);
  if (mode == 0) {
    hipLaunchKernelGGL(k_direct_naive, dim3(gx), dim3(kT), 0, c->stream, n, d_xyz, d_m, ntgt, d_tgt, c->prm.G,
                       c->prm.eps2, d_acc);
    NB_LAUNCH();
    return;
  }
  const int64_t nchunk = direct_chunks(n, ntgt);
  const int64_t chunk = (n + nchunk - 1) / nchunk;
  const size_t need = (size_t)nchunk * (size_t)ntgt * 6;
  double *part;
  if (part_fixed) {
    // every entry the combine reads is written by k_direct_part first: nothing to clear
    if (ntgt > NBFMM_NOALLOC_MAX_TARGETS)
      throw Error(NBFMM_ERR_ARG, "the allocation-free direct sum takes at most " +
                                     std::to_string(NBFMM_NOALLOC_MAX_TARGETS) + " targets in mode 1, got " +
                                     std::to_string(ntgt));
    // (up to NBFMM_NOALLOC_MAX_TARGETS targets the partial sums fit: nchunk <= 2^19 / ntgt + 1 chunks)
    if (need > kScratchDoubles)
      throw Error(NBFMM_ERR_ARG, "the direct sum of " + std::to_string(ntgt) + " targets needs " +
                                     std::to_string(need * sizeof(double)) + " bytes of partial sums, more than the " +
                                     std::to_string((size_t)NBFMM_NOALLOC_SCRATCH_BYTES) + " of the fixed scratch");
    part = part_fixed;
  } else {
    part = dget<double>(c->dpart, need);
  }
  hipLaunchKernelGGL(k_direct_part, dim3(gx, (unsigned)nchunk), dim3(kT), 0, c->stream, n, d_xyz, d_m, ntgt, d_tgt,
                     c->prm.eps2, chunk, part);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_direct_combine, dim3(gx), dim3(kT), 0, c->stream, ntgt, (int)nchunk, (const double *)part,
                     c->prm.G, d_acc);
  NB_LAUNCH();
}

} // namespace nbfmm
