// batch.hip -- many independent systems in one evaluation (a forest).
//
// System k (particles [off[k], off[k+1]) of the concatenated arrays) is
// scaled by a power of two s_k (exact) and translated into its own cell of a
// cubic grid of W = 2^L0 cells per axis: cell (2a+1, 2b+1, 2c+1), a, b, c < g
// = ceil(cbrt(nsys)), so every system cell is surrounded by empty cells.  Two
// massless particles at (0,0,0) and (W,W,W) make the bounding box that grid.
// Interaction lists are built only for cells below level L0 (a cell of level
// L0 + 1 or deeper lies inside one system cell; its neighbours and the
// children of its parent's neighbours lie in that cell or in empty ones), so
// no pair of particles of different systems is ever evaluated, by M2L or
// directly.  Every phase runs once for all systems, in the same launches.
// Accelerations are scaled back by s_k^2 (G m / r^2 with r scaled by s_k).
#include "nbfmm_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nbfmm {
namespace {

constexpr int kT = 256;
inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

__device__ __forceinline__ int sys_of(int64_t i, const int64_t *__restrict__ off, int nsys) {
  int lo = 0, hi = nsys; // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// per-system bounding box (block k: system k), out[6k..6k+5] = lo xyz, hi xyz
__global__ void __launch_bounds__(kT) k_sys_bbox(const int64_t *__restrict__ off, const double *__restrict__ x,
                                                 double *__restrict__ out) {
  __shared__ double s[6][kT];
  const int k = blockIdx.x;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = off[k] + threadIdx.x; i < off[k + 1]; i += kT)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], x[3 * i + d]);
      hi[d] = fmax(hi[d], x[3 * i + d]);
    }
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
  if (threadIdx.x < 6) out[6 * k + threadIdx.x] = s[threadIdx.x][0];
}

// placed copy: x' = s_k x + t_k, the two grid corners appended (mass 0)
__global__ void k_place(int64_t n, int nsys, const int64_t *__restrict__ off, const double4 *__restrict__ st,
                        const double *__restrict__ x, const double *__restrict__ m, double W, double *__restrict__ xb,
                        double *__restrict__ mb) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n + 2) return;
  if (i >= n) {
    const double v = i == n ? 0.0 : W;
    xb[3 * i] = v;
    xb[3 * i + 1] = v;
    xb[3 * i + 2] = v;
    mb[i] = 0.0;
    return;
  }
  const double4 q = st[sys_of(i, off, nsys)]; // (t_x, t_y, t_z, s)
  xb[3 * i] = fma(x[3 * i], q.w, q.x);
  xb[3 * i + 1] = fma(x[3 * i + 1], q.w, q.y);
  xb[3 * i + 2] = fma(x[3 * i + 2], q.w, q.z);
  mb[i] = m[i];
}

__global__ void k_unscale(int64_t n, int nsys, const int64_t *__restrict__ off, const double4 *__restrict__ st,
                          const double *__restrict__ ab, double *__restrict__ acc) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double s = st[sys_of(i, off, nsys)].w, s2 = s * s;
  acc[3 * i] = ab[3 * i] * s2;
  acc[3 * i + 1] = ab[3 * i + 1] * s2;
  acc[3 * i + 2] = ab[3 * i + 2] * s2;
}

// Placement of every system: s = 2^e with s * extent <= 0.98 (e = floor(log2(0.98 /
// extent)), from the exponent of 0.98 / extent: exact), so the system stays
// inside its cell; t moves the box centre to the cell centre (2a+1.5, 2b+1.5,
// 2c+1.5), a, b, c the system's grid coordinates.
__global__ void k_place_plan(int nsys, int g, const double *__restrict__ bb, double4 *__restrict__ st) {
#pragma clang fp contract(off)
  const int k = (int)(blockIdx.x * kT + threadIdx.x);
  if (k >= nsys) return;
  const double *b = bb + 6 * k;
  const double ext = fmax(b[3] - b[0], fmax(b[4] - b[1], b[5] - b[2]));
  int e = 0;
  if (ext > 0) {
    int x2 = 0;
    (void)frexp(0.98 / ext, &x2); // 0.98 / ext = f 2^x2, f in [0.5, 1)
    e = x2 - 1;
  }
  const double sc = ldexp(1.0, e);
  const int a = k % g, bq = (k / g) % g, cq = k / (g * g);
  const double org[3] = {2.0 * a + 1.5, 2.0 * bq + 1.5, 2.0 * cq + 1.5}; // cell centres
  double t[3];
  for (int d = 0; d < 3; ++d) t[d] = org[d] - sc * 0.5 * (b[d] + b[3 + d]);
  st[k] = make_double4(t[0], t[1], t[2], sc);
}

} // namespace

void batch_eval(nbfmm_ctx *c, int nsys, const int64_t *d_off, int64_t n, const double *d_xyz, const double *d_m,
                double *d_acc, const std::function<void(int64_t, const double *, const double *, double *)> &eval) {
  hipStream_t s = c->stream;
  double *bb = dget<double>(c->bat_bb, (size_t)6 * nsys);
  hipLaunchKernelGGL(k_sys_bbox, dim3(nsys), dim3(kT), 0, s, d_off, d_xyz, bb);
  NB_LAUNCH();
  // grid: g systems per axis at odd cell coordinates, W = 2^L0 >= 2g + 2 cells
  int g = 1;
  while ((int64_t)g * g * g < nsys) ++g;
  int L0 = 0;
  while ((1 << L0) < 2 * g + 2) ++L0;
  if (L0 + 1 > kMaxDepth) throw Error(NBFMM_ERR_UNSUPPORTED, "too many systems for one batch");
  const double W = (double)(1 << L0);
  double4 *dst = dget<double4>(c->bat_st, (size_t)nsys);
  hipLaunchKernelGGL(k_place_plan, dim3(nblk(nsys)), dim3(kT), 0, s, nsys, g, (const double *)bb, dst);
  NB_LAUNCH();
  double *xb = dget<double>(c->bat_x, (size_t)3 * (n + 2));
  double *mb = dget<double>(c->bat_m, (size_t)(n + 2));
  double *ab = dget<double>(c->bat_a, (size_t)3 * (n + 2));
  hipLaunchKernelGGL(k_place, dim3(nblk(n + 2)), dim3(kT), 0, s, n, nsys, d_off, (const double4 *)dst, d_xyz, d_m, W,
                     xb, mb);
  NB_LAUNCH();
  struct Level {
    nbfmm_ctx *c;
    int old;
    ~Level() { c->list_level = old; }
  } restore{c, c->list_level};
  c->list_level = L0 + 1;
  eval(n + 2, xb, mb, ab);
  hipLaunchKernelGGL(k_unscale, dim3(nblk(n)), dim3(kT), 0, s, n, nsys, d_off, (const double4 *)dst,
                     (const double *)ab, d_acc);
  NB_LAUNCH();
}

int64_t *upload_offsets(nbfmm_ctx *c, int nsys, const int64_t *h_off) {
  int64_t *doff = dget<int64_t>(c->bat_off, (size_t)nsys + 1);
  NB_HIP(hipMemcpy(doff, h_off, sizeof(int64_t) * (nsys + 1), hipMemcpyHostToDevice));
  return doff;
}

} // namespace nbfmm
