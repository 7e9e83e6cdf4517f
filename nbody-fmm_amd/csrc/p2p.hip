// p2p.hip -- particle-particle interactions of the gfx950 FMM and the direct sum.
//
// P2P follows Leaf::compute_acceleration (leaf.cpp:105-143): every particle
// of a target leaf receives grad -= m_j (x_i - x_j) / |x_i - x_j|^3 from the
// particles of its own leaf (j != i) and of its neighbour leaves; here also
// from the interaction-list cells whose pair is evaluated directly.  The
// result is added to the far field of the L2P pass and written as
// acc = G * grad in input order (particle.cpp:16-21).
//
// One wavefront per target leaf.  Leaves of the reference tree hold few
// particles (often 1-4), so the 64 lanes are split as (target, source slot):
// `sub` = 64 / next_pow2(n_t) lanes share one target and stride over the
// source particles; their partial sums are combined with xor-shuffles.
#include "nbfmm_internal.h"

namespace nbfmm {
namespace {

constexpr int kT = 256;

__global__ void __launch_bounds__(kT) k_p2p(int64_t nl, int32_t lb, const int32_t *__restrict__ cbeg,
                                            const int32_t *__restrict__ cend, const int32_t *__restrict__ p2p_off,
                                            const int32_t *__restrict__ p2p_src, const double4 *__restrict__ pos,
                                            const double *__restrict__ gfar, const int32_t *__restrict__ idx, double G,
                                            double *__restrict__ acc_out) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t leaf = (int64_t)blockIdx.x * (kT / 64) + wave;
  if (leaf >= nl) return;
  const int32_t c = lb + (int32_t)leaf;
  const int32_t b = cbeg[c], nt = cend[c] - b;
  int np2 = 1;
  while (np2 < nt && np2 < 64) np2 <<= 1;
  const int sub = 64 / np2;
  const int per = 64 / sub;
  const int32_t e0 = p2p_off[leaf], e1 = p2p_off[leaf + 1];
  for (int g0 = 0; g0 < nt; g0 += per) {
    const int ti = lane / sub, si = lane % sub;
    const bool valid = g0 + ti < nt;
    const int32_t i = b + g0 + (valid ? ti : 0);
    const double4 xi = pos[i];
    double ax = 0, ay = 0, az = 0;
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t s = p2p_src[e];
      const int32_t je = cend[s];
      for (int32_t j = cbeg[s] + si; j < je; j += sub) {
        const double4 xj = pos[j];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (j != i && r2 > 0.0) {
          const double inv = rsqrt(r2);
          const double w = xj.w * inv * inv * inv;
          ax -= dx * w;
          ay -= dy * w;
          az -= dz * w;
        }
      }
    }
    for (int o = sub >> 1; o >= 1; o >>= 1) {
      ax += __shfl_xor(ax, o);
      ay += __shfl_xor(ay, o);
      az += __shfl_xor(az, o);
    }
    if (valid && si == 0) {
      const int64_t out = idx[i];
      acc_out[3 * out] = G * (gfar[3 * (int64_t)i] + ax);
      acc_out[3 * out + 1] = G * (gfar[3 * (int64_t)i + 1] + ay);
      acc_out[3 * out + 2] = G * (gfar[3 * (int64_t)i + 2] + az);
    }
  }
}

// Direct sum over all sources for the listed targets (or all, tgt == null).
// mode 0: NaiveUpdate (naive_update.cpp:20-30) with the reference's rounding:
//   s = m_j / ((r^2 + eps2) * r); grad -= d * s; acc = grad * G, sources in
//   input order, every operation individually rounded (no contraction).
// mode 1: grad += m_j (x_j - x_i) / (r^2 + eps2)^{3/2} with Neumaier-compensated
//   sums per component (the FP64 oracle of the accuracy metric).
__global__ void __launch_bounds__(kT) k_direct(int64_t n, const double *__restrict__ xyz, const double *__restrict__ m,
                                               int64_t ntgt, const int64_t *__restrict__ tgt, int mode, double G,
                                               double eps2, double *__restrict__ acc) {
  __shared__ double4 sh[kT];
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool valid = t < ntgt;
  const int64_t i = valid ? (tgt ? tgt[t] : t) : 0;
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  double s0 = 0, s1 = 0, s2 = 0, c0 = 0, c1 = 0, c2 = 0;
  for (int64_t base = 0; base < n; base += kT) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    sh[threadIdx.x] = j < n ? make_double4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], m[j]) : make_double4(0, 0, 0, 0);
    __syncthreads();
    const int cnt = (int)((n - base) < kT ? (n - base) : kT);
    if (mode == 0) {
      for (int k = 0; k < cnt; ++k) {
        const double4 q = sh[k];
        if (base + k == i) continue;
        const double d0 = __dsub_rn(xi, q.x), d1 = __dsub_rn(yi, q.y), d2 = __dsub_rn(zi, q.z);
        double r2 = __dmul_rn(d0, d0);
        r2 = __dadd_rn(r2, __dmul_rn(d1, d1));
        r2 = __dadd_rn(r2, __dmul_rn(d2, d2));
        if (r2 == 0.0) continue;
        const double s = __ddiv_rn(q.w, __dmul_rn(__dadd_rn(r2, eps2), __dsqrt_rn(r2)));
        s0 = __dsub_rn(s0, __dmul_rn(d0, s));
        s1 = __dsub_rn(s1, __dmul_rn(d1, s));
        s2 = __dsub_rn(s2, __dmul_rn(d2, s));
      }
    } else {
      for (int k = 0; k < cnt; ++k) {
        const double4 q = sh[k];
        if (base + k == i) continue;
        const double d0 = q.x - xi, d1 = q.y - yi, d2 = q.z - zi;
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2 + eps2;
        if (r2 == 0.0) continue;
        const double inv = 1.0 / (r2 * sqrt(r2));
        const double w = q.w * inv;
        double v, u;
        v = d0 * w;
        u = s0 + v;
        c0 += fabs(s0) >= fabs(v) ? (s0 - u) + v : (v - u) + s0;
        s0 = u;
        v = d1 * w;
        u = s1 + v;
        c1 += fabs(s1) >= fabs(v) ? (s1 - u) + v : (v - u) + s1;
        s1 = u;
        v = d2 * w;
        u = s2 + v;
        c2 += fabs(s2) >= fabs(v) ? (s2 - u) + v : (v - u) + s2;
        s2 = u;
      }
    }
  }
  if (!valid) return;
  if (mode == 0) {
    acc[3 * t] = __dmul_rn(s0, G);
    acc[3 * t + 1] = __dmul_rn(s1, G);
    acc[3 * t + 2] = __dmul_rn(s2, G);
  } else {
    acc[3 * t] = G * (s0 + c0);
    acc[3 * t + 1] = G * (s1 + c1);
    acc[3 * t + 2] = G * (s2 + c2);
  }
}

} // namespace

void run_p2p(nbfmm_ctx *c, double *d_acc_out) {
  Plan &P = c->plan;
  const int D = P.geo.depth;
  const int32_t lb = P.lvl_start[D];
  const int64_t nl = P.lvl_start[D + 1] - lb;
  mark(c, 9);
  hipLaunchKernelGGL(k_p2p, dim3((unsigned)((nl + 3) / 4)), dim3(kT), 0, c->stream, nl, lb,
                     (const int32_t *)c->cbeg.p, (const int32_t *)c->cend.p, (const int32_t *)c->p2p_off.p,
                     (const int32_t *)c->p2p_src.p, (const double4 *)c->pos.p, (const double *)c->grad.p,
                     (const int32_t *)c->idx[0].p, c->prm.G, d_acc_out);
  NB_LAUNCH();
  mark(c, 10);
}

void run_direct(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, int64_t ntgt, const int64_t *d_tgt,
                int mode, double *d_acc) {
  hipLaunchKernelGGL(k_direct, dim3((unsigned)((ntgt + kT - 1) / kT)), dim3(kT), 0, c->stream, n, d_xyz, d_m, ntgt,
                     d_tgt, mode, c->prm.G, c->prm.eps2, d_acc);
  NB_LAUNCH();
}

} // namespace nbfmm
