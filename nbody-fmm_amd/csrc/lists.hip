// lists.hip -- interaction lists, their M2L / direct classification and the
// P2P work list of the gfx950 FMM.
//
// Interaction lists follow NodeI::compute_interaction_list (node_i.cpp:23-99):
// for a cell c the reference scans the 6x6x6 window of its parent's
// neighbours' children, offsets [-2-(c%2), 3-(c%2)] per axis with x fastest;
// occupied cells at Chebyshev distance >= 2 form the interaction list, in that
// order.  The window is the same for the eight children of one parent (child
// coordinates 2P-2 .. 2P+3 per axis), so one wavefront per parent stages the
// 216 candidate cells in LDS once, in window order, and classifies them for
// every child; ballot prefix counts keep the reference order when the entries
// are written.  A listed pair (t, s) is evaluated by M2L, or directly (P2P)
// when n_t * n_s < p2p_cut or when its offset class is not admissible for
// the cell shape (see class_mask).
#include "nbfmm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstddef>

namespace nbfmm {

// The reference's cells are boxes of the bounding box's aspect ratio
// (tree.cpp:20-30, not cubed).  The M2L of an offset class d converges at the
// rate (R_s + R_t) / |d|, R the half diagonal of a cell: below 0.87 for cubes
// (the nearest classes, |d| = 2 h), but at or above 1 for the short axis of a
// box whose longest side exceeds sqrt(2) times its shortest, where the series
// no longer converges for particles near the cell corners.  Pairs of classes
// with a rate above kMacRatio are evaluated directly, like the pairs below
// p2p_cut (exact; for bounding boxes within 1.15:1 no class is affected),
// as long as n_t * n_s < kBadDirect: larger pairs stay M2L, since evaluating
// them directly grows as N^2 (the cubic_box option avoids such classes).
// The mask of these classes is built on the device with the tree
// (k_tree_plan, tree.hip) and read from the plan here.

namespace {

constexpr int kT = 256;
constexpr int kW = 64;
constexpr int kWin = 216;
// workgroups of the list passes (four parents per round)
constexpr int64_t kListBlocks = 4096;
// direct-evaluation work per P2P sub-task (particle pairs)
constexpr int64_t kTaskPairs = 1 << 17;

// One wavefront per parent, four parents per block.  A child c with
// n_c * (largest particle count among its parent's 27 neighbours) < cut has
// every interaction-list pair of its subtree evaluated directly (the target
// of such a pair lies inside c and its source inside one of those
// neighbours), so c and all its descendants are closed.  When the caller
// does not keep the lists and c may be merged into a closed-cell task (cell
// index >= skip_from), its lists are never read: they are skipped and its
// counts stay 0.  skip_from < 0 classifies every cell.
// One thread per parent: does it have a child to classify (the skip rule of
// k_lists)?  Those parents are appended to active[] (order free; one counter
// update per wavefront, the count at *nactive).
// The range of parents and the skip level come from the plan: parents of
// levels [lc - 1, D) (cells [pbeg, nparents)), cells of level >= lc may be
// skipped (skip_from = their first cell, or -1 when every cell is classified).
struct ParentRange {
  int32_t pbeg, nparents, skip_from;
};
__device__ __forceinline__ ParentRange parent_range(const DevPlan *dp, int lc, int D, int keep) {
  ParentRange r;
  r.pbeg = dp->lvl_start[min(lc - 1, D)];
  r.nparents = dp->lvl_start[D];
  r.skip_from = keep ? -1 : dp->lvl_start[lc];
  return r;
}

__global__ void __launch_bounds__(kT) k_list_filter(const DevPlan *__restrict__ dp, int lc, int D, int keep,
                                                    const int32_t *__restrict__ nbr27,
                                                    const int32_t *__restrict__ child8,
                                                    const int32_t *__restrict__ cbeg,
                                                    const int32_t *__restrict__ cend, int64_t cut, int32_t *active,
                                                    int32_t *nactive) {
  if (plan_dead(dp)) return;
  const ParentRange pr = parent_range(dp, lc, D, keep);
  const int32_t nparents = pr.nparents, skip_from = pr.skip_from;
  const int32_t P = pr.pbeg + (int32_t)(blockIdx.x * kT + threadIdx.x);
  bool todo = false;
  if (P < nparents) {
    int64_t maxn = 0;
    if (skip_from >= 0) {
      int32_t q[27];
#pragma unroll
      for (int k = 0; k < 27; ++k) q[k] = nbr27[(int64_t)P * 27 + k];
#pragma unroll
      for (int k = 0; k < 27; ++k)
        if (q[k] >= 0) maxn = max(maxn, (int64_t)(cend[q[k]] - cbeg[q[k]]));
    }
    int32_t ch[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) ch[o] = child8[(int64_t)P * 8 + o];
#pragma unroll
    for (int o = 0; o < 8; ++o)
      todo |= ch[o] >= 0 &&
              !(skip_from >= 0 && ch[o] >= skip_from && (int64_t)(cend[ch[o]] - cbeg[ch[o]]) * maxn < cut);
  }
  const uint64_t m = __ballot(todo);
  const int lane = threadIdx.x & 63;
  int32_t base = 0;
  if (lane == 0 && m) base = atomicAdd(nactive, (int32_t)__popcll(m));
  base = __shfl(base, 0);
  if (todo) active[base + __popcll(m & (lane == 0 ? 0ull : (~0ull >> (64 - lane))))] = P;
}

template <bool FILL>
__global__ void __launch_bounds__(kT) k_lists(const DevPlan *__restrict__ dp, int lc, int D, int keep,
                                              const int32_t *__restrict__ nbr27,
                                              const int32_t *__restrict__ child8, const int32_t *__restrict__ cbeg,
                                              const int32_t *__restrict__ cend, int64_t cut,
                                              int32_t *cnt_il,
                                              int32_t *cnt_m2l, int32_t *cnt_sub, int64_t *sub_pairs,
                                              const int32_t *off_il, const int32_t *off_m2l, const int32_t *off_sub,
                                              int32_t *il, int32_t *m2l_src, uint16_t *m2l_cls, int32_t *sub_src,
                                              int32_t *is_src, const int32_t *active) {
  if (plan_dead(dp)) return;
  __shared__ int32_t cand_s[kT / kW][kWin];
  __shared__ int32_t csz_s[kT / kW][kWin];
  __shared__ ClsMask bad;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x < (kNumClasses + 31) / 32) bad.w[threadIdx.x] = dp->bad_cls.w[threadIdx.x];
  const ParentRange pr = parent_range(dp, lc, D, keep);
  const int32_t nparents = pr.nparents, skip_from = pr.skip_from;
  // both passes take the parents k_list_filter listed (their count at
  // dp->nactive), four per workgroup round: the grid strides over them
  const int32_t nact = dp->nactive;
  int32_t *cand = cand_s[wave], *csz = csz_s[wave];
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int32_t b = (int32_t)blockIdx.x; b < (nact + kT / kW - 1) / (kT / kW); b += (int32_t)gridDim.x) {
    const int32_t pw = b * (kT / kW) + wave;
    const int32_t P = pw < nact ? active[pw] : nparents;
    const bool live = P < nparents;
    int64_t maxn = 0;
    // children to classify (bit o): all when skipping is off
    uint32_t todo = 0;
    if (live) {
      if (skip_from >= 0) {
        const int32_t q = lane < 27 ? nbr27[(int64_t)P * 27 + lane] : -1;
        maxn = q >= 0 ? cend[q] - cbeg[q] : 0;
        for (int off = 32; off >= 1; off >>= 1) maxn = max(maxn, __shfl_xor(maxn, off));
      }
      const int32_t ch = lane < 8 ? child8[(int64_t)P * 8 + lane] : -1;
      const bool need =
          ch >= 0 && !(skip_from >= 0 && ch >= skip_from && (int64_t)(cend[ch] - cbeg[ch]) * maxn < cut);
      todo = (uint32_t)__ballot(need);
      if (todo) {
        // the four window entries of a lane: each load level issued for all
        // four before the next (three dependent gathers, not twelve)
        constexpr int kR = (kWin + kW - 1) / kW;
        int32_t pn[kR], q[kR], nq[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const int w = lane + r * kW;
          const int wx = w % 6, wy = (w / 6) % 6, wz = w / 36;
          pn[r] = w < kWin ? nbr27[(int64_t)P * 27 + (wz >> 1) * 9 + (wy >> 1) * 3 + (wx >> 1)] : -1;
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const int w = lane + r * kW;
          const int wx = w % 6, wy = (w / 6) % 6, wz = w / 36;
          q[r] = pn[r] >= 0 ? child8[(int64_t)pn[r] * 8 + ((wx & 1) | ((wy & 1) << 1) | ((wz & 1) << 2))] : -1;
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) nq[r] = q[r] >= 0 ? cend[q[r]] - cbeg[q[r]] : 0;
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const int w = lane + r * kW;
          if (w < kWin) {
            cand[w] = q[r];
            csz[w] = nq[r];
          }
        }
      }
    }
    __syncthreads();
    for (int o = 0; o < 8; ++o) {
      if (!((todo >> o) & 1u)) continue;
      const int32_t s = child8[(int64_t)P * 8 + o];
      const int bx = o & 1, by = (o >> 1) & 1, bz = o >> 2;
      const int64_t nt = cend[s] - cbeg[s];
      int32_t a_il = 0, a_m = 0, a_s = 0;
      if (FILL) {
        a_il = il ? off_il[s] : 0;
        a_m = off_m2l[s];
        a_s = off_sub[s];
      }
      int32_t c_il = 0, c_m = 0, c_s = 0;
      int64_t sp = 0;
      for (int base = 0; base < kWin; base += kW) {
        const int w = base + lane;
        const bool valid = w < kWin;
        const int32_t q = valid ? cand[w] : -1;
        const int64_t ns = valid ? csz[w] : 0;
        const int dx = w % 6 - 2 - bx, dy = (w / 6) % 6 - 2 - by, dz = w / 36 - 2 - bz;
        const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
        const bool far = q >= 0 && !near;
        const int cls = (dz + 3) * 49 + (dy + 3) * 7 + (dx + 3);
        const bool sub =
            far && (nt * ns < cut || (((bad.w[cls >> 5] >> (cls & 31)) & 1u) && nt * ns < kBadDirect));
        const bool m2l = far && !sub;
        const uint64_t bf = __ballot(far), bs = __ballot(sub), bm = __ballot(m2l);
        if (FILL) {
          if (il && far) il[a_il + c_il + __popcll(bf & lt)] = q;
          if (sub) sub_src[a_s + c_s + __popcll(bs & lt)] = q;
          if (m2l) {
            const int32_t k = a_m + c_m + __popcll(bm & lt);
            m2l_src[k] = q;
            m2l_cls[k] = (uint16_t)cls;
            is_src[q] = 1;
          }
        } else if (sub) {
          sp += nt * ns;
        }
        c_il += __popcll(bf);
        c_s += __popcll(bs);
        c_m += __popcll(bm);
      }
      if (!FILL) {
        for (int off = 32; off >= 1; off >>= 1) sp += __shfl_xor(sp, off);
        if (lane == 0) {
          cnt_il[s] = c_il;
          cnt_m2l[s] = c_m;
          cnt_sub[s] = c_s;
          sub_pairs[s] = sp;
        }
      }
    }
    __syncthreads(); // the window is restaged in the next round
  }
}

// Two compactions in one scan: cells with a[i] > 0 (flags [0, n)) and cells
// with b[i] > 0 (flags [n, 2n)).
__global__ void k_flag_pos2(const int32_t *a, const int32_t *b, int64_t n, int32_t *flags) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    flags[i] = a[i] > 0 ? 1 : 0;
    flags[n + i] = b[i] > 0 ? 1 : 0;
  }
}
// slot (or -1) and list of each half; tot[0], tot[1]: the two list lengths
__global__ void k_compact2(const int32_t *flags, const int32_t *scan, int64_t n, int32_t *slot_a, int32_t *list_a,
                           int32_t *slot_b, int32_t *list_b, int32_t *tot) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int32_t na = scan[n];
  if (flags[i]) {
    slot_a[i] = scan[i];
    list_a[scan[i]] = (int32_t)i;
  } else {
    slot_a[i] = -1;
  }
  const int32_t k = scan[n + i] - na;
  if (flags[n + i]) {
    slot_b[i] = k;
    list_b[k] = (int32_t)i;
  } else {
    slot_b[i] = -1;
  }
  if (i == n - 1) {
    tot[0] = na;
    tot[1] = scan[2 * n - 1] + flags[2 * n - 1] - na;
  }
}
__global__ void k_sum_i32(const int32_t *v, int64_t n, unsigned long long *out) {
  __shared__ unsigned long long s[kT];
  unsigned long long a = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) a += (unsigned)v[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, s[0]);
}
// the sums of three arrays of n entries into out[0..2]
__global__ void k_sum3_i32(const int32_t *a, const int32_t *b, const int32_t *c, int64_t n, unsigned long long *out) {
  __shared__ unsigned long long s[3][kT];
  unsigned long long x = 0, y = 0, z = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    x += (unsigned)a[i];
    y += (unsigned)b[i];
    z += (unsigned)c[i];
  }
  s[0][threadIdx.x] = x;
  s[1][threadIdx.x] = y;
  s[2][threadIdx.x] = z;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicAdd(out + threadIdx.x, s[threadIdx.x][0]);
}
__global__ void k_sum_i64(const int64_t *v, int64_t n, unsigned long long *out) {
  __shared__ unsigned long long s[kT];
  unsigned long long a = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) a += (unsigned long long)v[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

// ---- P2P sub-tasks ------------------------------------------------------------
// A cell is closed when its own and all its descendants' interaction-list
// pairs are evaluated directly (no M2L at or below it).  For the particles
// of a closed cell A the directly evaluated pairs of the leaves below it
// (near field plus every substituted interaction list down the tree) are
// exactly the particles of the 27 neighbours of A's parent, so a maximal
// closed cell (parent not closed) becomes one dense block of work.  Other
// leaves get their near field (27-neighbour row) and other cells their
// substituted list.  Targets are cut into groups of kTaskTargets, the
// concatenated sources into whole 64-particle chunks of about kTaskPairs
// pair interactions per task.
// closed[c] = 1 for the cells of levels >= lc, 0 below (k_open then clears
// the cells whose subtree holds an M2L list)
__global__ void k_closed_init(const DevPlan *__restrict__ dp, int lc, int64_t S, int32_t *closed) {
  if (plan_dead(dp)) return;
  const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (c >= S) return;
  closed[c] = c >= dp->lvl_start[lc] && c < dp->nc ? 1 : 0;
}
// A cell is closed when no cell of its subtree (itself included) has an M2L
// list: every cell with one clears itself and its ancestors, stopping at the
// first one found cleared (whoever cleared it goes on upwards, so every
// ancestor of a cleared cell ends cleared; two walks may clear the same
// cells, with the same value).  The reference's recursion (closed = no list
// and every child closed) gives the same flags.
__global__ void k_open(const DevPlan *__restrict__ dp, const int32_t *__restrict__ cnt_m2l,
                       const int32_t *__restrict__ cparent, int32_t *closed) {
  if (plan_dead(dp)) return;
  const int32_t c = (int32_t)(blockIdx.x * kT + threadIdx.x);
  if (c >= dp->nc || cnt_m2l[c] == 0) return;
  volatile int32_t *cl = closed;
  for (int32_t x = c; x >= 0; x = cparent[x]) {
    if (x != c && cl[x] == 0) break;
    cl[x] = 0;
  }
}

__device__ __forceinline__ int64_t row_particles(const int32_t *row, const int32_t *cbeg, const int32_t *cend) {
  int32_t q[27]; // the row's loads first, then the counts' (two rounds of gathers)
#pragma unroll
  for (int o = 0; o < 27; ++o) q[o] = row[o];
  int64_t t = 0;
#pragma unroll
  for (int o = 0; o < 27; ++o)
    if (q[o] >= 0) t += cend[q[o]] - cbeg[q[o]];
  return t;
}

// source groups of cell c: up to two (list, nlist, particles)
struct SrcGroups {
  int n;
  int32_t list[2], nlist[2];
  int64_t np[2];
};

__device__ __forceinline__ SrcGroups src_groups(int32_t c, int32_t leaf_begin, const int32_t *cparent,
                                                const int32_t *nbr27, const int32_t *closed, const int32_t *cbeg,
                                                const int32_t *cend, const int32_t *cnt_sub, const int32_t *off_sub,
                                                const int64_t *sub_pairs) {
  SrcGroups g;
  g.n = 0;
  if (closed[c]) {
    const int32_t P = cparent[c];
    if (!closed[P]) {
      g.list[0] = -(P + 1);
      g.nlist[0] = 27;
      g.np[0] = row_particles(nbr27 + (int64_t)P * 27, cbeg, cend);
      g.n = 1;
    }
    return g;
  }
  if (c >= leaf_begin) {
    g.list[g.n] = -(c + 1);
    g.nlist[g.n] = 27;
    g.np[g.n] = row_particles(nbr27 + (int64_t)c * 27, cbeg, cend);
    ++g.n;
  }
  if (cnt_sub[c] > 0) {
    g.list[g.n] = off_sub[c];
    g.nlist[g.n] = cnt_sub[c];
    g.np[g.n] = sub_pairs[c] / (cend[c] - cbeg[c]);
    ++g.n;
  }
  return g;
}

__device__ __forceinline__ int32_t seg_len(int64_t ng, int64_t np) {
  int64_t k = (ng * np + kTaskPairs - 1) / kTaskPairs;
  if (k < 1) k = 1;
  int64_t sl = (np + k - 1) / k;
  sl = (sl + 63) & ~(int64_t)63;
  return (int32_t)sl;
}

// Also the particle pairs each cell's tasks evaluate (a 27-neighbour row
// contains the target cell itself, whose particles skip themselves).  The
// counts of cells of at most narrow_max particles go to cnt[c], the others
// to cnt[S + c] (S >= nc + 1; the entries past the cells stay zero), so one
// scan over both halves numbers the narrow tasks first and the wide ones
// after them.
__global__ void k_task_count(const DevPlan *__restrict__ dp, int D, int64_t S, const int32_t *cparent,
                             const int32_t *nbr27, const int32_t *closed, const int32_t *cbeg, const int32_t *cend,
                             const int32_t *cnt_sub, const int32_t *off_sub, const int64_t *sub_pairs, int narrow_max,
                             int32_t *cnt, int64_t *pairs) {
  if (plan_dead(dp)) return;
  const int32_t nc = dp->nc, leaf_begin = dp->lvl_start[D];
  const int32_t c = 1 + (int32_t)(blockIdx.x * kT + threadIdx.x);
  if (c >= nc) return;
  const SrcGroups g = src_groups(c, leaf_begin, cparent, nbr27, closed, cbeg, cend, cnt_sub, off_sub, sub_pairs);
  const int32_t nt = cend[c] - cbeg[c];
  const int32_t ntg = (nt + kTaskTargets - 1) / kTaskTargets;
  int32_t k = 0;
  int64_t np = 0;
  for (int q = 0; q < g.n; ++q) {
    if (g.np[q] <= 0) continue;
    const int32_t sl = seg_len(min(nt, kTaskTargets), g.np[q]);
    k += (int32_t)((g.np[q] + sl - 1) / sl);
    np += (int64_t)nt * (g.np[q] - (g.list[q] < 0 ? 1 : 0));
  }
  const bool wide = nt > narrow_max;
  cnt[c] = wide ? 0 : k * ntg;
  cnt[S + c] = wide ? k * ntg : 0;
  pairs[c] = np;
}

__global__ void k_task_fill(const DevPlan *__restrict__ dp, int D, int64_t S, const int32_t *cparent,
                            const int32_t *nbr27, const int32_t *closed, const int32_t *cbeg, const int32_t *cend,
                            const int32_t *cnt_sub, const int32_t *off_sub, const int64_t *sub_pairs, int narrow_max,
                            const int32_t *toff, P2PTask *tasks) {
  if (plan_dead(dp)) return;
  const int32_t nc = dp->nc, leaf_begin = dp->lvl_start[D];
  const int32_t c = 1 + (int32_t)(blockIdx.x * kT + threadIdx.x);
  if (c >= nc) return;
  const SrcGroups g = src_groups(c, leaf_begin, cparent, nbr27, closed, cbeg, cend, cnt_sub, off_sub, sub_pairs);
  const int32_t nt = cend[c] - cbeg[c];
  const int32_t ntg = (nt + kTaskTargets - 1) / kTaskTargets;
  int32_t w = nt > narrow_max ? toff[S + c] : toff[c];
  for (int32_t tg = 0; tg < ntg; ++tg)
    for (int q = 0; q < g.n; ++q) {
      if (g.np[q] <= 0) continue;
      const int64_t sl = seg_len(min(nt, kTaskTargets), g.np[q]);
      for (int64_t f = 0; f < g.np[q]; f += sl) {
        P2PTask t;
        t.tcell = c;
        t.toff = tg * kTaskTargets;
        t.list = g.list[q];
        t.nlist = g.nlist[q];
        t.f0 = (int32_t)f;
        t.f1 = (int32_t)min(g.np[q], f + sl);
        tasks[w++] = t;
      }
    }
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

} // namespace

// Planned evaluations: sizes beyond the capacities (k_list_check after the
// list counts, k_task_totals after the task counts) set the plan's ovf.
__global__ void k_list_check(DevPlan *dp, int64_t cap_m2l, int64_t cap_sub) {
  if (threadIdx.x == 0 && (dp->tot[1] > (unsigned long long)cap_m2l || dp->tot[2] > (unsigned long long)cap_sub))
    dp->ovf |= 2;
}
// the narrow task count (the scan at the start of the wide half) into the
// plan; with check, the task and target capacities
__global__ void k_task_totals(DevPlan *dp, const int32_t *toff, int64_t S, int check, int64_t cap_task,
                              int64_t cap_tgt) {
  if (threadIdx.x != 0) return;
  dp->nnarrow = toff[S];
  if (check && (dp->ntask > (unsigned long long)cap_task || dp->nst[1] > cap_tgt)) dp->ovf |= 2;
}

void build_lists(nbfmm_ctx *c) {
  Plan &P = c->plan;
  hipStream_t s = c->stream;
  const Caps &K = c->cap;
  DevPlan *dp = dplan(c);
  const int D = P.geo.depth;
  // cell arrays hold nc cells (direct) or the capacity (planned); S >= nc + 1
  // is the stride of the arrays of two or five cell-indexed halves
  const int64_t nc = K.on ? K.nc : P.n_cells;
  const int64_t S = nc + 1;
  const int32_t *cbeg = (const int32_t *)c->cbeg.p, *cend = (const int32_t *)c->cend.p;
  const int32_t *child8 = (const int32_t *)c->child8.p, *nbr27 = (const int32_t *)c->nbr27.p;
  const int64_t cut = c->prm.p2p_cut;
  int32_t *cnt_il = dget<int32_t>(c->cnt_il, (size_t)S);
  int32_t *cnt_m2l = dget<int32_t>(c->cnt_m2l, (size_t)S);
  int32_t *cnt_sub = dget<int32_t>(c->cnt_sub, (size_t)S);
  int32_t *off_il = dget<int32_t>(c->off_il, (size_t)S);
  int32_t *off_m2l = dget<int32_t>(c->off_m2l, (size_t)S);
  int32_t *off_sub = dget<int32_t>(c->off_sub, (size_t)S);
  int64_t *sub_pairs = dget<int64_t>(c->sub_pairs, (size_t)S);
  // Closed cells (below) merge the near field of their whole subtree into one
  // P2P task; cells of level >= lc may be merged.  Cells of levels below lc
  // have no lists (lc = 1 for one system; below the system cells for a batch).
  const int lc = std::max(1, c->list_level);
  if (lc > D)
    throw Error(NBFMM_ERR_UNSUPPORTED, "batched systems need a tree deeper than the system cells (more particles per "
                                       "system than leaf_capacity)");
  if (K.on && c->keep_lists) throw Error(NBFMM_ERR_UNSUPPORTED, "planned evaluations keep no lists");
  const int keep = c->keep_lists ? 1 : 0;
  // cells that are not classified (the root, skipped children) have empty lists
  fill_i32(s, cnt_il, S, 0);
  fill_i32(s, cnt_m2l, S, 0);
  fill_i32(s, cnt_sub, S, 0);
  fill_zero(s, sub_pairs, sizeof(int64_t) * (size_t)S);
  // every non-leaf cell is a parent: the grids cover all of them, the kernels
  // take those of levels [lc - 1, D) from the plan
  int64_t npar = 0;
  if (K.on)
    npar = K.parents;
  else
    npar = P.lvl_start[D] - P.lvl_start[std::min(lc - 1, D)];
  // the list passes stride over the active parents (few of all at the usual
  // cuts): a grid of at most kListBlocks workgroups
  const unsigned nlb = (unsigned)std::clamp<int64_t>((npar + kT / kW - 1) / (kT / kW), 1, kListBlocks);
  // parents with a child to classify, their count in the plan
  int32_t *active = dget<int32_t>(c->act_par, (size_t)npar + 2);
  fill_i32(s, &dp->nactive, 1, 0);
  hipLaunchKernelGGL(k_list_filter, dim3(nblk(npar)), dim3(kT), 0, s, (const DevPlan *)dp, lc, D, keep, nbr27, child8,
                     cbeg, cend, cut, active, &dp->nactive);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_lists<false>, dim3(nlb), dim3(kT), 0, s, (const DevPlan *)dp, lc, D, keep, nbr27, child8, cbeg,
                     cend, cut, cnt_il, cnt_m2l, cnt_sub, sub_pairs, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr, nullptr, (const int32_t *)active);
  NB_LAUNCH();
  if (keep) scan_excl_i32(cnt_il, off_il, S, c->scan_tmp, s); // the full lists' offsets only when they are kept
  scan_excl_i32(cnt_m2l, off_m2l, S, c->scan_tmp, s);
  scan_excl_i32(cnt_sub, off_sub, S, c->scan_tmp, s);
  // list lengths in 64 bits: the offsets above are 32-bit
  unsigned long long *tot = dp->tot;
  fill_zero(s, tot, sizeof(unsigned long long) * 3);
  hipLaunchKernelGGL(k_sum3_i32, dim3(std::min<unsigned>(512, nblk(S))), dim3(kT), 0, s, (const int32_t *)cnt_il,
                     (const int32_t *)cnt_m2l, (const int32_t *)cnt_sub, S, tot);
  NB_LAUNCH();
  const unsigned long long lim = 0x7fffffffull - 1;
  // the host reads the counts (direct) or the plan checks them (planned)
  const DevPlan *hp = c->hplan;
  auto read_counts = [&]() {
    read_plan(c, &c->hplan->nactive, offsetof(DevPlan, npairs) + sizeof(unsigned long long) - offsetof(DevPlan, nactive),
              s);
  };
  int32_t nactive = 0;
  if (K.on) {
    hipLaunchKernelGGL(k_list_check, dim3(1), dim3(64), 0, s, dp, K.n_m2l, K.n_sub);
    NB_LAUNCH();
  } else {
    read_counts();
    nactive = hp->nactive;
    if (hp->tot[1] > lim || hp->tot[2] > lim || (c->keep_lists && hp->tot[0] > lim))
      throw Error(NBFMM_ERR_UNSUPPORTED, "interaction lists exceed 2^31 entries (raise leaf_capacity or p2p_cut)");
    P.n_il = (int64_t)hp->tot[0];
    P.n_m2l = (int64_t)hp->tot[1];
    P.n_sub = (int64_t)hp->tot[2];
  }
  const int64_t n_m2l = K.on ? K.n_m2l : P.n_m2l, n_sub = K.on ? K.n_sub : P.n_sub;
  int32_t *il = c->keep_lists ? dget<int32_t>(c->il, (size_t)P.n_il + 1) : nullptr;
  int32_t *m2l_src = dget<int32_t>(c->m2l_src, (size_t)n_m2l + 1);
  uint16_t *m2l_cls = dget<uint16_t>(c->m2l_cls, (size_t)n_m2l + 1);
  int32_t *sub_src = dget<int32_t>(c->sub_src, (size_t)n_sub + 1);
  int32_t *is_src = dget<int32_t>(c->is_src, (size_t)S);
  fill_i32(s, is_src, S, 0);
  const int64_t nfill = K.on ? npar : nactive;
  if (nfill > 0) {
    const unsigned nfb = (unsigned)std::min<int64_t>((nfill + kT / kW - 1) / (kT / kW), kListBlocks);
    hipLaunchKernelGGL(k_lists<true>, dim3(nfb), dim3(kT), 0, s, (const DevPlan *)dp, lc, D, keep, nbr27, child8, cbeg,
                       cend, cut, nullptr, nullptr, nullptr, nullptr, il ? off_il : nullptr, off_m2l, off_sub, il,
                       m2l_src, m2l_cls, sub_src, is_src, (const int32_t *)active);
    NB_LAUNCH();
  }

  // ---- M2L source / target compaction ----
  const size_t fl = (size_t)std::max<int64_t>(2 * S, P.n) + 1;
  int32_t *flags = dget<int32_t>(c->flags, fl);
  int32_t *fscan = dget<int32_t>(c->fscan, fl);
  int32_t *src_slot = dget<int32_t>(c->src_slot, (size_t)S);
  int32_t *tgt_slot = dget<int32_t>(c->tgt_slot, (size_t)S);
  int32_t *src_cells = dget<int32_t>(c->src_cells, (size_t)S);
  int32_t *tgt_cells = dget<int32_t>(c->tgt_cells, (size_t)S);
  // sources: cells in some M2L list; targets: cells with an M2L list (the
  // entries past the cells are zero).  Their counts reach the plan.
  hipLaunchKernelGGL(k_flag_pos2, dim3(nblk(S)), dim3(kT), 0, s, (const int32_t *)is_src, (const int32_t *)cnt_m2l, S,
                     flags);
  NB_LAUNCH();
  scan_excl_i32(flags, fscan, 2 * S, c->scan_tmp, s);
  hipLaunchKernelGGL(k_compact2, dim3(nblk(S)), dim3(kT), 0, s, flags, fscan, S, src_slot, src_cells, tgt_slot,
                     tgt_cells, dp->nst);
  NB_LAUNCH();

  // ---- P2P sub-tasks ----
  int32_t *closed = dget<int32_t>(c->closed, (size_t)S);
  // cells of levels < lc are never task cells
  hipLaunchKernelGGL(k_closed_init, dim3(nblk(S)), dim3(kT), 0, s, (const DevPlan *)dp, lc, S, closed);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_open, dim3(nblk(nc)), dim3(kT), 0, s, (const DevPlan *)dp, (const int32_t *)cnt_m2l,
                     (const int32_t *)c->cparent.p, closed);
  NB_LAUNCH();
  nbr_leaf_rows(c, closed); // the P2P work list reads the rows of the leaves that are not closed
  const int32_t *cpar = (const int32_t *)c->cparent.p;
  // task counts of narrow cells, then of wide cells (two halves of S entries)
  const int narrow_max = c->prm.precision == NBFMM_MIXED ? c->shapes.p2p_narrow_max_f32 : c->shapes.p2p_narrow_max_f64;
  int32_t *tcnt = dget<int32_t>(c->task_cnt, (size_t)2 * S);
  int32_t *toff = dget<int32_t>(c->task_off, (size_t)2 * S);
  fill_i32(s, tcnt, 2 * S, 0);
  int64_t *cell_pairs = dget<int64_t>(c->near_pairs, (size_t)S);
  fill_zero(s, cell_pairs, sizeof(int64_t) * (size_t)S);
  fill_zero(s, &dp->npairs, sizeof(unsigned long long));
  hipLaunchKernelGGL(k_task_count, dim3(nblk(nc)), dim3(kT), 0, s, (const DevPlan *)dp, D, S, cpar, nbr27,
                     (const int32_t *)closed, cbeg, cend, (const int32_t *)cnt_sub, (const int32_t *)off_sub,
                     (const int64_t *)sub_pairs, narrow_max, tcnt, cell_pairs);
  NB_LAUNCH();
  scan_excl_i32(tcnt, toff, 2 * S, c->scan_tmp, s);
  hipLaunchKernelGGL(k_sum_i64, dim3(std::min<unsigned>(512, nblk(S))), dim3(kT), 0, s, (const int64_t *)cell_pairs,
                     S, &dp->npairs);
  NB_LAUNCH();
  fill_zero(s, &dp->ntask, sizeof(unsigned long long));
  hipLaunchKernelGGL(k_sum_i32, dim3(std::min<unsigned>(512, nblk(2 * S))), dim3(kT), 0, s, (const int32_t *)tcnt,
                     2 * S, &dp->ntask);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_task_totals, dim3(1), dim3(64), 0, s, dp, (const int32_t *)toff, S, K.on ? 1 : 0, K.ntask,
                     K.n_tgt);
  NB_LAUNCH();
  if (!K.on) {
    read_counts();
    if (hp->ntask > lim) throw Error(NBFMM_ERR_UNSUPPORTED, "P2P work list exceeds 2^31 tasks");
    P.n_src = hp->nst[0];
    P.n_tgt = hp->nst[1];
    P.n_tasks = (int64_t)hp->ntask;
    P.n_tasks_narrow = hp->nnarrow;
    P.n_p2p_pairs = (int64_t)hp->npairs;
  }
  const int64_t ntask = K.on ? K.ntask : P.n_tasks;
  P2PTask *tasks = dget<P2PTask>(c->tasks, (size_t)ntask + 1);
  hipLaunchKernelGGL(k_task_fill, dim3(nblk(nc)), dim3(kT), 0, s, (const DevPlan *)dp, D, S, cpar, nbr27,
                     (const int32_t *)closed, cbeg, cend, (const int32_t *)cnt_sub, (const int32_t *)off_sub,
                     (const int64_t *)sub_pairs, narrow_max, (const int32_t *)toff, tasks);
  NB_LAUNCH();
  NB_HIP(hipStreamWaitEvent(s, c->ev_pos, 0)); // the sorted particles (build_tree)
}

} // namespace nbfmm
