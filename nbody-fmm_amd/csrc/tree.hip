// tree.hip -- the reference's uniform-depth sparse octree on the GPU.
//
// Follows src/system_updaters/tree_components/tree.cpp:15-90:
//   bbox (tree.cpp:20-30) -> depth loop (tree.cpp:32-69: quantise, Morton key,
//   max occupancy, repeat) -> leaves holding their particles in input order
//   (tree.cpp:71-74) -> ancestors (tree.cpp:76-84) -> neighbour and
//   interaction lists (node_i.cpp:23-99).
// Keys are 64-bit (the reference's 32-bit truncation, F6, only matters for
// depth >= 10); the key layout is Morton::encode's (morton_code.cpp:9-22):
// bit k of coordinate i -> bit 3k+i.
//
// The sort is an LSD radix sort on (key, input index) with 8-bit digits:
// per pass a block histogram in LDS, a device-wide exclusive scan of the
// digit-major histogram, and a stable scatter whose in-block ranks come from
// wavefront ballots (match of equal digits) plus LDS prefix counts.  LSD
// passes are stable, and the initial order is the input order, so equal keys
// keep input order exactly as the reference's leaves do.
#include "nbfmm_internal.h"

#include <algorithm>
#include <cstddef>
#include <cstring>

namespace nbfmm {
namespace {

constexpr int kT = 256;           // threads per block
constexpr int kSortItems = 16;    // items per thread per radix tile
constexpr int kSortTile = kT * kSortItems;

__device__ __forceinline__ uint64_t spread3(uint64_t x) {
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
__device__ __forceinline__ uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ull;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x1fffffull;
  return (uint32_t)x;
}

// ---- bbox + finiteness --------------------------------------------------------
__global__ void __launch_bounds__(kT) k_bbox(const double *__restrict__ xyz, const double *__restrict__ m,
                                             int64_t n, double *part, int *bad) {
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = xyz[d];
    hi[d] = xyz[d];
  }
  int b = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    for (int d = 0; d < 3; ++d) {
      double v = xyz[3 * i + d];
      if (!isfinite(v)) b = 1;
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
    if (!isfinite(m[i])) b = 1;
  }
  if (b) atomicOr(bad, 1);
  __shared__ double s[6][kT];
  for (int d = 0; d < 3; ++d) {
    s[d][threadIdx.x] = lo[d];
    s[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + o]);
        s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// The box (tree.cpp:20-30) into the device plan; with `cubic` the bounding
// cube centred on it, widened outward where rounding would cut the box.
__global__ void __launch_bounds__(kT) k_bbox_final(const double *part, int nb, int cubic, DevPlan *dp) {
#pragma clang fp contract(off)
  __shared__ double s[6][kT];
  double v[6];
  for (int d = 0; d < 6; ++d) v[d] = part[d];
  for (int b = threadIdx.x; b < nb; b += kT)
    for (int d = 0; d < 6; ++d) v[d] = d < 3 ? fmin(v[d], part[b * 6 + d]) : fmax(v[d], part[b * 6 + d]);
  for (int d = 0; d < 6; ++d) s[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int d = 0; d < 6; ++d)
        s[d][threadIdx.x] = d < 3 ? fmin(s[d][threadIdx.x], s[d][threadIdx.x + o]) : fmax(s[d][threadIdx.x], s[d][threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = s[d][0];
    hi[d] = s[3 + d][0];
  }
  if (cubic) {
    double ext = 0;
    for (int d = 0; d < 3; ++d) ext = fmax(ext, hi[d] - lo[d]);
    for (int d = 0; d < 3; ++d) {
      const double cc = 0.5 * (lo[d] + hi[d]);
      double a = cc - 0.5 * ext, b = cc + 0.5 * ext;
      if (a > lo[d]) a = lo[d];
      if (b < hi[d]) b = hi[d];
      lo[d] = a;
      hi[d] = b;
    }
  }
  for (int d = 0; d < 3; ++d) {
    dp->lo[d] = lo[d];
    dp->hi[d] = hi[d];
  }
}

// ---- quantisation + Morton key (tree.cpp:55-61) -------------------------------
// q = (x - a) / leaf_size (IEEE division), truncated, clamped to 2^depth - 1.
// A zero extent gives NaN, mapped to 0 (what the reference's x86-64 build does).
__device__ __forceinline__ uint32_t quantize(double x, double a, double ls, uint32_t nld) {
  double q = __ddiv_rn(x - a, ls);
  uint32_t c;
  if (!(q == q) || q < 0.0)
    c = 0;
  else if (q >= 4294967295.0)
    c = 0xffffffffu;
  else
    c = (uint32_t)q;
  return c >= nld ? nld - 1 : c;
}

// leaf size per dimension (hi - lo) / 2^depth (IEEE division) from the plan's box
__global__ void __launch_bounds__(kT) k_keys(const double *__restrict__ xyz, int64_t n, const DevPlan *__restrict__ dp,
                                             uint32_t nld, uint64_t *__restrict__ key, int32_t *__restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double a0 = dp->lo[0], a1 = dp->lo[1], a2 = dp->lo[2];
  const double l0 = __ddiv_rn(dp->hi[0] - a0, (double)nld), l1 = __ddiv_rn(dp->hi[1] - a1, (double)nld),
               l2 = __ddiv_rn(dp->hi[2] - a2, (double)nld);
  uint32_t cx = quantize(xyz[3 * i], a0, l0, nld);
  uint32_t cy = quantize(xyz[3 * i + 1], a1, l1, nld);
  uint32_t cz = quantize(xyz[3 * i + 2], a2, l2, nld);
  key[i] = spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
  idx[i] = (int32_t)i;
}

// ---- radix sort ------------------------------------------------------------------
__global__ void __launch_bounds__(kT) k_rs_hist(const uint64_t *__restrict__ key, int64_t n, int shift,
                                                int nblocks, int32_t *__restrict__ hist) {
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    int64_t i = base + (int64_t)k * kT + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kT) k_rs_scatter(const uint64_t *__restrict__ kin,
                                                   const int32_t *__restrict__ vin, int64_t n, int shift,
                                                   int nblocks, const int32_t *__restrict__ off,
                                                   uint64_t *__restrict__ kout, int32_t *__restrict__ vout) {
  __shared__ int32_t run[256];      // running count per digit in this tile
  __shared__ int32_t wcnt[4][256];  // per-wave digit counts of the current round
  __shared__ int32_t goff[256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  run[threadIdx.x] = 0;
  goff[threadIdx.x] = off[(int64_t)threadIdx.x * nblocks + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    int64_t i = base + (int64_t)k * kT + threadIdx.x;
    bool valid = i < n;
    uint64_t kv = valid ? kin[i] : 0ull;
    int32_t vv = valid ? vin[i] : 0;
    uint32_t dg = valid ? (uint32_t)((kv >> shift) & 255u) : 0u;
    // lanes of this wave holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = __ballot((dg >> b) & 1u);
      peers &= ((dg >> b) & 1u) ? bb : ~bb;
    }
    for (int j = threadIdx.x; j < 4 * 256; j += kT) (&wcnt[0][0])[j] = 0;
    __syncthreads();
    int rank_w = __popcll(peers & lt);
    bool leader = valid && (peers & lt) == 0ull;
    if (leader) wcnt[wave][dg] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int r = run[dg] + rank_w;
      for (int w = 0; w < wave; ++w) r += wcnt[w][dg];
      int64_t dst = (int64_t)goff[dg] + r;
      kout[dst] = kv;
      vout[dst] = vv;
    }
    __syncthreads();
    run[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
    __syncthreads();
  }
}

__global__ void k_gather_pos(const double *__restrict__ xyz, const double *__restrict__ m,
                             const int32_t *__restrict__ idx, int64_t n, double4 *__restrict__ pos) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  int64_t j = idx[i];
  pos[i] = make_double4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], m[j]);
}

// ---- levels ------------------------------------------------------------------------
// For consecutive sorted keys, the coarsest level at which they differ is
// D - floor(msb/3); every level from there to D starts a new cell at i.
__device__ inline int start_level(const uint64_t *__restrict__ key, int64_t i, int64_t n, int D) {
  if (i >= n) return D + 1;
  if (i == 0) return 0;
  const uint64_t x = key[i] ^ key[i - 1];
  return x == 0 ? D + 1 : D - (63 - __clzll((long long)x)) / 3;
}

// Cells that start in each block of kT particles, per level: bcnt[d * nb + b];
// with them the block's first and last cell start (-1: none) and the longest
// cell that starts and ends inside the block (bfirst, blast, bgap, same layout).
__global__ void __launch_bounds__(kT) k_level_blocks(const uint64_t *__restrict__ key, int64_t n, int D,
                                                     int32_t *__restrict__ bcnt, int32_t *__restrict__ bfirst,
                                                     int32_t *__restrict__ blast, int32_t *__restrict__ bgap) {
  __shared__ int32_t w[kT / 64][kMaxDepth + 1];
  __shared__ int32_t wf[kT / 64][kMaxDepth + 1], wl[kT / 64][kMaxDepth + 1], wg[kT / 64][kMaxDepth + 1];
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int d0 = start_level(key, i, n, D);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int d = 0; d <= D; ++d) {
    const uint64_t m = __ballot(d0 <= d);
    // a start's distance to the next start of its wavefront (0: none)
    const uint64_t above = lane == 63 ? 0ull : m >> (lane + 1);
    int g = (d0 <= d && above) ? __builtin_ctzll(above) + 1 : 0;
    if (m & (m - 1)) { // two starts or more: the longest such distance
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) g = max(g, __shfl_xor(g, o, 64));
    }
    if (lane == 0) {
      w[wv][d] = (int32_t)__popcll(m);
      wf[wv][d] = m ? __builtin_ctzll(m) : -1;
      wl[wv][d] = m ? 63 - __builtin_clzll(m) : -1;
      wg[wv][d] = g;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x <= D) {
    const int d = threadIdx.x;
    const int32_t base = (int32_t)blockIdx.x * kT;
    int32_t t = 0, F = -1, S = -1, G = 0;
    for (int q = 0; q < kT / 64; ++q) {
      t += w[q][d];
      if (wf[q][d] >= 0) {
        const int32_t pf = base + 64 * q + wf[q][d];
        if (S >= 0) G = max(G, pf - S);
        if (F < 0) F = pf;
        S = base + 64 * q + wl[q][d];
        G = max(G, wg[q][d]);
      }
    }
    const int64_t o = (int64_t)d * gridDim.x + blockIdx.x;
    bcnt[o] = t;
    bfirst[o] = F;
    blast[o] = S;
    bgap[o] = G;
  }
}

// One block per level: exclusive scan of that level's block counts in place,
// the level's cell count (total[d]) and its fullest cell (total[kMaxDepth + 1 + d]):
// the longest of the cells inside one block and of the cells from a block's
// first start back to the last start before it (a max-scan of blast), the
// last cell ending at n.
constexpr int kOffThreads = 1024, kOffItems = 8;
__global__ void __launch_bounds__(kOffThreads) k_level_offsets(int32_t *bcnt, const int32_t *__restrict__ bfirst,
                                                               const int32_t *__restrict__ blast,
                                                               const int32_t *__restrict__ bgap, int nb, int64_t n,
                                                               uint32_t *total) {
  __shared__ int32_t ws[kOffThreads / 64], wm[kOffThreads / 64], wg[kOffThreads / 64];
  const int64_t ro = (int64_t)blockIdx.x * nb;
  int32_t *row = bcnt + ro;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t carry = 0, carryS = -1, gmax = 0;
  for (int base = 0; base < nb; base += kOffThreads * kOffItems) {
    const int j0 = base + (int)threadIdx.x * kOffItems;
    int32_t v[kOffItems], f[kOffItems], l[kOffItems], loc = 0, sl = -1;
#pragma unroll
    for (int k = 0; k < kOffItems; ++k) {
      const bool in = j0 + k < nb;
      v[k] = in ? row[j0 + k] : 0;
      f[k] = in ? bfirst[ro + j0 + k] : -1;
      l[k] = in ? blast[ro + j0 + k] : -1;
      if (in) gmax = max(gmax, bgap[ro + j0 + k]);
      loc += v[k];
      sl = max(sl, l[k]);
    }
    int32_t x = loc, y = sl; // inclusive sum and max scans over the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t xs = __shfl_up(x, o, 64), ys = __shfl_up(y, o, 64);
      if (lane >= o) {
        x += xs;
        y = max(y, ys);
      }
    }
    int32_t yprev = __shfl_up(y, 1, 64);
    if (lane == 0) yprev = -1;
    if (lane == 63) {
      ws[wv] = x;
      wm[wv] = y;
    }
    __syncthreads();
    int32_t pre = 0, all = 0, preS = carryS, allS = carryS;
    for (int q = 0; q < kOffThreads / 64; ++q) {
      const int32_t t = ws[q], u = wm[q];
      if (q < wv) {
        pre += t;
        preS = max(preS, u);
      }
      all += t;
      allS = max(allS, u);
    }
    int32_t run = carry + pre + x - loc;
    int32_t last = max(preS, yprev); // the last cell start before this thread's blocks
#pragma unroll
    for (int k = 0; k < kOffItems; ++k) {
      if (j0 + k < nb) row[j0 + k] = run;
      run += v[k];
      if (f[k] >= 0) {
        if (last >= 0) gmax = max(gmax, f[k] - last);
        last = l[k];
      }
    }
    carry += all;
    carryS = allS;
    __syncthreads();
  }
  // the fullest cell over the block's threads
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) gmax = max(gmax, __shfl_xor(gmax, o, 64));
  if (lane == 0) wg[wv] = gmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t g = carryS >= 0 ? (int32_t)(n - carryS) : 0;
    for (int q = 0; q < kOffThreads / 64; ++q) g = max(g, wg[q]);
    total[blockIdx.x] = (uint32_t)carry;
    total[kMaxDepth + 1 + blockIdx.x] = (uint32_t)g;
  }
}

// The plan of the tree sorted at depth D, from the last level count: level
// starts, cell count, per-level geometry, the operator geometry and the mask
// of offset classes evaluated directly (the box's aspect ratio, lists.hip).
// A planned evaluation (check != 0) also replays the reference's depth loop
// (tree.cpp:32-69) on the levels' fullest cells and flags a depth other than
// D, more cells than the capacities, and a non-finite input.  One workgroup
// of 512 threads.
__global__ void __launch_bounds__(512) k_tree_plan(DevPlan *dp, int D, int64_t n, uint32_t thr, int list_level,
                                                   int check, int64_t cap_nc, CapLevels cap_lev) {
#pragma clang fp contract(off)
  __shared__ uint32_t maskw[(kNumClasses + 31) / 32];
  __shared__ int32_t ls[kMaxDepth + 2];
  const int t = threadIdx.x;
  const uint32_t *cnt = dp->lcount;
  const double lo[3] = {dp->lo[0], dp->lo[1], dp->lo[2]}, hi[3] = {dp->hi[0], dp->hi[1], dp->hi[2]};
  if (t == 0) {
    int32_t s = 0;
    ls[0] = 0;
    for (int d = 0; d <= D; ++d) {
      s += (int32_t)cnt[d];
      ls[d + 1] = s;
    }
    for (int d = D + 2; d < kMaxDepth + 2; ++d) ls[d] = s;
    dp->nc = s;
    dp->one_nc[0] = 1;
    dp->one_nc[1] = s;
    dp->depth = D;
    if (check) {
      int ovf = 0;
      // the loop as the host runs it from the depth the keys were sorted at
      auto more = [&](int depth, uint32_t maxcnt) {
        return depth < kMaxDepth && (maxcnt > thr || depth < list_level);
      };
      int depth = 0, it = 0;
      uint32_t maxcnt = (uint32_t)n;
      for (;;) {
        const uint64_t nl = (uint64_t)((maxcnt - 1u) / thr + 1u);
        const int bw = nl ? 64 - __clzll((long long)nl) : 0;
        const int nd = min(kMaxDepth, depth + bw / 3 + 1);
        if (nd > D) break;
        depth = nd;
        maxcnt = cnt[kMaxDepth + 1 + depth];
        ++it;
        if (!more(depth, maxcnt)) break;
      }
      if (more(depth, maxcnt) || depth != D) ovf |= 1;
      if ((int64_t)s > cap_nc) ovf |= 2;
      for (int d = 0; d <= D; ++d)
        if ((int32_t)cnt[d] > cap_lev.v[d]) ovf |= 2;
      if (dp->bad) ovf |= 4;
      dp->iterations = it;
      if (ovf) dp->ovf |= ovf;
    }
  }
  // per-level geometry: h = (hi - lo) / 2^l, rho the largest edge (1 for a point)
  LevelGeo &g = dp->g;
  if (t < 3) g.lo[t] = lo[t];
  if (t <= D) {
    double r = 0;
    for (int d = 0; d < 3; ++d) {
      const double h = (hi[d] - lo[d]) / (double)(1ull << t);
      g.h[t][d] = h;
      if (d == 0 || h > r) r = h;
    }
    g.inv_rho[t] = 1.0 / (r > 0 ? r : 1.0);
    if (t == D) {
      const double rho = r > 0 ? r : 1.0;
      for (int d = 0; d < 3; ++d) dp->hs[d] = (hi[d] - lo[d]) / (double)(1ull << t) / rho;
    }
  }
  if (t == 0) g.depth = D;
  __syncthreads(); // ls written
  if (t < kMaxDepth + 2) {
    dp->lvl_start[t] = ls[t];
    g.lvl_start[t] = ls[t];
  }
  // offset classes (lists.hip: class 0 is the slowest admissible one)
  if (t < (kNumClasses + 31) / 32) maskw[t] = 0;
  const double hx = hi[0] - lo[0], hy = hi[1] - lo[1], hz = hi[2] - lo[2];
  const double diag2 = hx * hx + hy * hy + hz * hz;
  __syncthreads();
  if (t < kNumClasses) {
    const int dx = t % 7 - 3, dy = (t / 7) % 7 - 3, dz = t / 49 - 3;
    if (max(abs(dx), max(abs(dy), abs(dz))) >= 2) {
      const double d2 = dx * dx * hx * hx + dy * dy * hy * hy + dz * dz * hz * hz;
      if (diag2 > kMacRatio * kMacRatio * d2) atomicOr(&maskw[t >> 5], 1u << (t & 31));
    }
  }
  __syncthreads();
  if (t < (kNumClasses + 31) / 32) dp->bad_cls.w[t] = maskw[t];
}

// Every level's cells at once: particle i finds the index of its cell at each
// level from the block offsets and the ranks within its block; where it starts
// a cell it writes the cell's key, first particle and parent, and ends the
// level's previous cell.  Levels are concatenated from the root; keys of level
// d are the leaf keys >> 3(D - d).
__global__ void __launch_bounds__(kT) k_cells_fill(const uint64_t *__restrict__ key, int64_t n, int D,
                                                   const int32_t *__restrict__ boff, const DevPlan *__restrict__ dp,
                                                   uint64_t *__restrict__ ckey, int32_t *__restrict__ cbeg,
                                                   int32_t *__restrict__ cend, int32_t *__restrict__ cpar,
                                                   int32_t *__restrict__ pcell) {
  if (plan_dead(dp)) return;
  const int32_t *__restrict__ Ls = dp->lvl_start;
  __shared__ int32_t w[kT / 64][kMaxDepth + 1];
  __shared__ int32_t bo[kMaxDepth + 1];
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int d0 = start_level(key, i, n, D);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t upto = ~0ull >> (63 - lane); // lanes 0..lane
  if ((int)threadIdx.x <= D) bo[threadIdx.x] = boff[(int64_t)threadIdx.x * gridDim.x + blockIdx.x];
  for (int d = 0; d <= D; ++d) {
    const uint64_t m = __ballot(d0 <= d);
    if (lane == 0) w[wv][d] = (int32_t)__popcll(m);
  }
  __syncthreads();
  const uint64_t k = i < n ? key[i] : 0;
  int32_t up = -1; // the cell one level up (-1 above the root)
  for (int d = 0; d <= D; ++d) {
    const uint64_t m = __ballot(d0 <= d);
    int32_t pre = 0;
    for (int q = 0; q < wv; ++q) pre += w[q][d];
    const int32_t cell = Ls[d] + bo[d] + pre + (int32_t)__popcll(m & upto) - 1;
    if (d0 <= d) {
      ckey[cell] = k >> (3 * (D - d));
      cbeg[cell] = (int32_t)i;
      cpar[cell] = up;
      if (cell > Ls[d]) cend[cell - 1] = (int32_t)i;
    }
    if (i == n - 1) cend[cell] = (int32_t)n;
    up = cell;
  }
  if (i < n) pcell[i] = up;
}

// cells [rng[0], rng[1]) (device range) and the plan's guard
__global__ void k_child8(const uint64_t *__restrict__ ckey, const int32_t *__restrict__ cparent,
                         const int32_t *__restrict__ rng, const DevPlan *__restrict__ dp, int32_t *child8) {
  if (plan_dead(dp)) return;
  int32_t c = rng[0] + (int32_t)(blockIdx.x * kT + threadIdx.x);
  if (c >= rng[1]) return;
  child8[(int64_t)cparent[c] * 8 + (int)(ckey[c] & 7u)] = c;
}

// 27-neighbourhood (offsets x fastest, self at 13) from the parent's table;
// one thread per cell: the cell's coordinates are decoded once and, per axis,
// the three offsets fall in two of the parent's neighbours (the parent itself
// and the one on the side of the cell's octant), so 8 parent entries and 27
// child entries are read.  Rows go through LDS (stride 27, odd: no bank
// conflicts) and leave as contiguous stores.  With `skip`, rows of cells with
// skip[c] != 0 are left unwritten.
constexpr int kNbrCells = 256; // cells per workgroup
__global__ void __launch_bounds__(kNbrCells) k_nbr27(const uint64_t *__restrict__ ckey,
                                                     const int32_t *__restrict__ cparent,
                                                     const int32_t *__restrict__ child8,
                                                     const DevPlan *__restrict__ dp, int lev,
                                                     const int32_t *__restrict__ skip, int32_t *nbr27) {
  __shared__ int32_t rows[kNbrCells * 27];
  if (plan_dead(dp)) return;
  const int32_t c0 = dp->lvl_start[lev] + (int32_t)blockIdx.x * kNbrCells;
  const int32_t nb = min(kNbrCells, dp->lvl_start[lev + 1] - c0);
  if (nb <= 0) return;
  const int t = threadIdx.x;
  const int32_t c = c0 + t;
  if (t < nb && !(skip && skip[c])) {
    const uint64_t k = ckey[c];
    const int x = (int)compact3(k), y = (int)compact3(k >> 1), z = (int)compact3(k >> 2);
    const int32_t *prow = nbr27 + (int64_t)cparent[c] * 27;
    // Along an axis with coordinate u and bit b = u & 1, the offset slots
    // o (offset o - 1) are taken as k = 0: the octant's side (o = 2b), in
    // the parent's neighbour on that side, child bit b ^ 1; k = 1: o = 1, in
    // the parent, bit b; k = 2: o = 2 - 2b, in the parent, bit b ^ 1.  A
    // neighbour outside the domain lies across the parent's edge, where the
    // parent's row holds -1 (the root's row is -1 but for itself).
    const int bx = x & 1, by = y & 1, bz = z & 1;
    int32_t pn[8]; // parent neighbours: bit d set = the octant's side along axis d
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int px = (j & 1) ? 2 * bx : 1, py = (j & 2) ? 2 * by : 1, pz = (j & 4) ? 2 * bz : 1;
      pn[j] = prow[pz * 9 + py * 3 + px];
    }
    int32_t *row = rows + t * 27;
#pragma unroll
    for (int kz = 0; kz < 3; ++kz)
#pragma unroll
      for (int ky = 0; ky < 3; ++ky)
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
          const int ox = kx == 0 ? 2 * bx : kx == 1 ? 1 : 2 - 2 * bx;
          const int oy = ky == 0 ? 2 * by : ky == 1 ? 1 : 2 - 2 * by;
          const int oz = kz == 0 ? 2 * bz : kz == 1 ? 1 : 2 - 2 * bz;
          const int32_t P = pn[(kx == 0 ? 1 : 0) | (ky == 0 ? 2 : 0) | (kz == 0 ? 4 : 0)];
          const int cb = (kx == 1 ? bx : bx ^ 1) | ((ky == 1 ? by : by ^ 1) << 1) | ((kz == 1 ? bz : bz ^ 1) << 2);
          row[oz * 9 + oy * 3 + ox] = P >= 0 ? child8[(int64_t)P * 8 + cb] : -1;
        }
  }
  __syncthreads();
  int32_t *out = nbr27 + (int64_t)c0 * 27;
  for (int i = t; i < nb * 27; i += kNbrCells) {
    if (skip && skip[c0 + i / 27]) continue;
    out[i] = rows[i];
  }
}

inline unsigned nbr_blocks(int64_t cells) {
  return (unsigned)std::max<int64_t>(1, (cells + kNbrCells - 1) / kNbrCells);
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

// Sorts (key, idx) by the low `bits` bits of key.  Returns the buffer index
// (0 or 1) holding the result.
int radix_sort(nbfmm_ctx *c, int64_t n, int bits, int cur) {
  const int nb = (int)((n + kSortTile - 1) / kSortTile);
  int32_t *hist = dget<int32_t>(c->sort_hist, (size_t)256 * nb);
  for (int shift = 0; shift < bits; shift += 8) {
    uint64_t *kin = (uint64_t *)c->key[cur].p, *kout = (uint64_t *)c->key[cur ^ 1].p;
    int32_t *vin = (int32_t *)c->idx[cur].p, *vout = (int32_t *)c->idx[cur ^ 1].p;
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kT), 0, c->stream, kin, n, shift, nb, hist);
    NB_LAUNCH();
    scan_excl_i32(hist, hist, (int64_t)256 * nb, c->scan_tmp, c->stream);
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kT), 0, c->stream, kin, vin, n, shift, nb, hist, kout, vout);
    NB_LAUNCH();
    cur ^= 1;
  }
  return cur;
}

int bit_width_u64(uint64_t v) {
  int w = 0;
  while (v) {
    ++w;
    v >>= 1;
  }
  return w;
}

} // namespace

// Cells per level of the keys sorted at depth D (buffer cur), and every
// level's fullest cell, into the device plan.  A direct evaluation reads them
// (P.lvl_start, lm: one transfer, one sync); a planned one does not.  The
// per-block offsets stay in c->flags for fill_levels.
static void count_levels(nbfmm_ctx *c, int D, int cur, uint32_t *lm) {
  Plan &P = c->plan;
  const int64_t n = P.n;
  hipStream_t s = c->stream;
  DevPlan *dp = dplan(c);
  const uint64_t *key = (const uint64_t *)c->key[cur].p;
  const int64_t nbl = nblk(n);
  const size_t per = (size_t)(D + 1) * (size_t)nbl;
  int32_t *bcnt = dget<int32_t>(c->flags, std::max<size_t>((size_t)n + 1, per));
  int32_t *aux = dget<int32_t>(c->fscan, std::max<size_t>((size_t)n + 1, 3 * per));
  hipLaunchKernelGGL(k_level_blocks, dim3((unsigned)nbl), dim3(kT), 0, s, key, n, D, bcnt, aux, aux + per,
                     aux + 2 * per);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_level_offsets, dim3(D + 1), dim3(kOffThreads), 0, s, bcnt, (const int32_t *)aux,
                     (const int32_t *)(aux + per), (const int32_t *)(aux + 2 * per), (int)nbl, n, dp->lcount);
  NB_LAUNCH();
  if (c->cap.on) return;
  read_plan_field(c, &DevPlan::lcount, s);
  const uint32_t *counts = c->hplan->lcount;
  P.lvl_start.assign(D + 2, 0);
  for (int d = 0; d <= D; ++d) {
    P.lvl_start[d + 1] = P.lvl_start[d] + (int32_t)counts[d];
    lm[d] = counts[kMaxDepth + 1 + d];
  }
}

// Every level's cells, the particles in key order and the neighbour tables,
// from the keys sorted at `depth` (buffer `cur`) and count_levels' offsets.
static void fill_levels(nbfmm_ctx *c, const double *d_xyz, const double *d_m, int depth, int cur) {
  Plan &P = c->plan;
  const int64_t n = P.n;
  hipStream_t s = c->stream;
  DevPlan *dp = dplan(c);
  P.geo.depth = depth;
  if (cur != 0) { // keep the sorted keys/indices in buffer 0
    std::swap(c->key[0], c->key[1]);
    std::swap(c->idx[0], c->idx[1]);
  }
  const uint64_t *key = (const uint64_t *)c->key[0].p;
  const int32_t *idx = (const int32_t *)c->idx[0].p;
  double4 *pos = dget<double4>(c->pos, (size_t)n);
  // the particle gather runs on the second stream, beside the cells, the
  // neighbour tables and the lists; P2P follows it on that stream, and the
  // main stream waits for it at the end of build_lists
  NB_HIP(hipEventRecord(c->ev_pos, s));
  NB_HIP(hipStreamWaitEvent(c->stream2, c->ev_pos, 0));
  hipLaunchKernelGGL(k_gather_pos, dim3(nblk(n)), dim3(kT), 0, c->stream2, d_xyz, d_m, idx, n, pos);
  NB_LAUNCH();
  NB_HIP(hipEventRecord(c->ev_pos, c->stream2));

  // ---- cells of every level ----
  const int D = depth;
  const int64_t nbl = nblk(n);
  const int32_t *bcnt = (const int32_t *)c->flags.p;
  if (!c->cap.on) {
    P.n_leaves = P.lvl_start[D + 1] - P.lvl_start[D];
    P.n_cells = P.lvl_start[D + 1];
  }
  const int64_t nc = c->cap.on ? c->cap.nc : P.n_cells; // cells, or their capacity
  uint64_t *ckey = dget<uint64_t>(c->ckey, (size_t)nc);
  int32_t *cbeg = dget<int32_t>(c->cbeg, (size_t)nc);
  int32_t *cend = dget<int32_t>(c->cend, (size_t)nc);
  int32_t *cpar = dget<int32_t>(c->cparent, (size_t)nc);
  int32_t *pcell = dget<int32_t>(c->pcell, (size_t)n);
  hipLaunchKernelGGL(k_cells_fill, dim3((unsigned)nbl), dim3(kT), 0, s, key, n, D, bcnt, (const DevPlan *)dp, ckey,
                     cbeg, cend, cpar, pcell);
  NB_LAUNCH();
  // neighbour tables
  int32_t *child8 = dget<int32_t>(c->child8, (size_t)nc * 8);
  int32_t *nbr27 = dget<int32_t>(c->nbr27, (size_t)nc * 27);
  fill_i32(s, child8, nc * 8, -1);
  hipLaunchKernelGGL(k_child8, dim3(nblk(nc - 1)), dim3(kT), 0, s, ckey, cpar, (const int32_t *)dp->one_nc,
                     (const DevPlan *)dp, child8);
  NB_LAUNCH();
  fill_i32(s, nbr27, 27, -1);
  fill_i32(s, nbr27 + 13, 1, 0);
  // the leaf level's rows follow in build_lists, for the leaves that are not
  // closed (the only ones the evaluation reads); nbr_leaf_rows completes them
  for (int d = 1; d < D; ++d) {
    hipLaunchKernelGGL(k_nbr27, dim3(nbr_blocks(level_cells(c, d))), dim3(kNbrCells), 0, s, ckey, cpar, child8,
                       (const DevPlan *)dp, d, nullptr, nbr27);
    NB_LAUNCH();
  }
  c->nbr_leaf = D == 0 ? 2 : 0;
}

// ---------------------------------------------------------------------------------
void build_tree(nbfmm_ctx *c, const double *d_xyz, const double *d_m) {
  Plan &P = c->plan;
  const int64_t n = P.n;
  hipStream_t s = c->stream;
  const Caps &K = c->cap;
  DevPlan *dp = dplan(c);
  // ---- bbox (tree.cpp:20-30) ----
  int nbb = (int)std::min<int64_t>(1024, (n + kT - 1) / kT);
  double *red = dget<double>(c->red, 6 * 1024 + 64);
  fill_i32(s, &dp->bad, 1, 0);
  hipLaunchKernelGGL(k_bbox, dim3(nbb), dim3(kT), 0, s, d_xyz, d_m, n, red, &dp->bad);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kT), 0, s, (const double *)red, nbb, (int)c->prm.cubic_box, dp);
  NB_LAUNCH();
  if (!K.on) {
    // the box and the finiteness flag (adjacent in the plan)
    read_plan(c, &c->hplan->lo[0], offsetof(DevPlan, bad) + sizeof(int32_t), s);
    if (c->hplan->bad) throw Error(NBFMM_ERR_DEGENERATE, "non-finite coordinate or mass");
    for (int d = 0; d < 3; ++d) {
      P.geo.lo[d] = c->hplan->lo[d];
      P.geo.hi[d] = c->hplan->hi[d];
    }
  }
  // ---- depth loop (tree.cpp:32-69) ----
  const uint32_t thr = c->prm.leaf_capacity;
  dget<uint64_t>(c->key[0], (size_t)n);
  dget<uint64_t>(c->key[1], (size_t)n);
  dget<int32_t>(c->idx[0], (size_t)n);
  dget<int32_t>(c->idx[1], (size_t)n);
  uint32_t lm[kMaxDepth + 1]; // fullest cell of every level of the last count
  int cur = 0;
  P.iterations = 0;
  P.depth_sorts = 0;
  // Morton keys at `depth`, sorted (buffer `cur`), and their cells per level
  auto sort_at = [&](int depth) {
    hipLaunchKernelGGL(k_keys, dim3(nblk(n)), dim3(kT), 0, s, d_xyz, n, (const DevPlan *)dp, (uint32_t)(1u << depth),
                       (uint64_t *)c->key[0].p, (int32_t *)c->idx[0].p);
    NB_LAUNCH();
    cur = radix_sort(c, n, 3 * depth, 0);
    P.depth_sorts++;
    count_levels(c, depth, cur, lm);
  };
  // the reference's step: the depth after `depth` when its fullest cell holds maxcnt
  auto next_depth = [&](int depth, uint32_t maxcnt) {
    const uint64_t nl = (uint64_t)((maxcnt - 1u) / thr + 1u);
    return std::min(kMaxDepth, depth + bit_width_u64(nl) / 3 + 1);
  };
  // the loop's condition; a batch's lists start at level list_level (below
  // its system cells), so its tree reaches that level even when every system
  // fits in one leaf; for one system list_level is 1 and the loop is the reference's
  auto more = [&](int depth, uint32_t maxcnt) {
    return depth < kMaxDepth && (maxcnt > thr || depth < c->list_level);
  };
  int depth = 0;
  if (K.on) {
    // planned: the keys sorted once at the planned depth; k_tree_plan replays
    // the loop on the levels' fullest cells and flags another depth
    depth = K.depth;
    sort_at(depth);
  } else {
    uint32_t maxcnt = (uint32_t)n;
    bool done = false;
    const int guess = c->depth_guess;
    if (guess >= 1 && guess <= kMaxDepth) {
      // Sort once at the depth the previous evaluation ended at.  A level-d
      // cell is a run of the level-d key, the sorted key's top 3d bits, so the
      // cell counts give the fullest cell of every level <= guess, and the loop
      // is replayed on them.  It ends at `guess` (the usual case: particles move
      // little between evaluations), before it (sorted again at that depth) or
      // beyond it (the loop goes on from its last depth <= guess); its result
      // is the loop's in every case.
      sort_at(guess);
      for (;;) {
        const int nd = next_depth(depth, maxcnt);
        if (nd > guess) break;
        depth = nd;
        maxcnt = lm[depth];
        P.iterations++;
        if (!more(depth, maxcnt)) break;
      }
      if (!more(depth, maxcnt)) {
        done = true;
        if (depth != guess) sort_at(depth);
      }
    }
    while (!done) {
      depth = next_depth(depth, maxcnt);
      sort_at(depth);
      maxcnt = lm[depth];
      P.iterations++;
      done = !more(depth, maxcnt);
    }
  }
  CapLevels cl{};
  for (int d = 0; d <= kMaxDepth; ++d) cl.v[d] = K.lev[d];
  hipLaunchKernelGGL(k_tree_plan, dim3(1), dim3(512), 0, s, dp, depth, n, thr, c->list_level, K.on ? 1 : 0, K.nc, cl);
  NB_LAUNCH();
  fill_levels(c, d_xyz, d_m, depth, cur);
  c->depth_guess = depth;
}

// rows of the leaf level: of the leaves with closed[c] == 0 (closed != null),
// or of every leaf
void nbr_leaf_rows(nbfmm_ctx *c, const int32_t *closed) {
  Plan &P = c->plan;
  const int D = P.geo.depth;
  if (c->nbr_leaf == 2 || D == 0 || (closed && c->nbr_leaf == 1)) return;
  hipLaunchKernelGGL(k_nbr27, dim3(nbr_blocks(level_cells(c, D))), dim3(kNbrCells), 0, c->stream, (const uint64_t *)c->ckey.p,
                     (const int32_t *)c->cparent.p, (const int32_t *)c->child8.p, (const DevPlan *)dplan(c), D,
                     closed, (int32_t *)c->nbr27.p);
  NB_LAUNCH();
  c->nbr_leaf = closed ? 1 : 2;
}

} // namespace nbfmm
