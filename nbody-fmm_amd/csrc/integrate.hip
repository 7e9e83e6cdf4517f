// integrate.hip -- device-side integrator updates (FP64, each operation
// individually rounded so that the steps match the reference's
// SyplecticEuler / RungeKutta4 arithmetic operation for operation).
//
// SyplecticEuler::integrate (symplectic_euler.cpp:8-15):  v += dt*a; x += dt*v
// RungeKutta4::integrate (runge_kutta_4.cpp:10-93): stage positions
//   x_k = x + h*v_{k-1}, stage velocities v_k = v + h*a_vel, sums
//   asum = a1 + 2a2 + 2a3 + a4, vsum = v1 + 2v2 + 2v3 + v4, then
//   v += dt/6*asum, x += dt/6*vsum.  (Tensor operator*(double) multiplies
//   each component by the scalar: tensor.cpp:226-232.)
#include "nbfmm_internal.h"

#include <algorithm>

// HIP contracts a*b+c into FMA by default (and __dmul_rn is a plain multiply);
// the reference's x86-64 build rounds every operation, so contraction is off.
#pragma clang fp contract(off)

namespace nbfmm {
namespace {
constexpr int kT = 256;

// The updates of the state (x, v, a) leave it unchanged when the planned
// evaluations of the step overflowed (the step is then run again directly).
__global__ void k_se(int64_t m, double *x, double *v, const double *a, double dt, const DevPlan *dp) {
  if (plan_dead(dp)) return;
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= m) return;
  double vi = __dadd_rn(v[i], __dmul_rn(a[i], dt));
  v[i] = vi;
  x[i] = __dadd_rn(x[i], __dmul_rn(vi, dt));
}

__global__ void k_rk4_stage(int64_t m, const double *x, const double *v, const double *vkp, const double *avel,
                            double h, double *xk, double *vk) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= m) return;
  double xn = __dadd_rn(x[i], __dmul_rn(vkp[i], h));
  double vn = __dadd_rn(v[i], __dmul_rn(avel[i], h));
  xk[i] = xn;
  vk[i] = vn;
}

__global__ void k_rk4_accum(int64_t m, double *asum, double *vsum, const double *a, const double *vk, double w,
                            int first) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= m) return;
  if (first) {
    asum[i] = a[i];
    vsum[i] = vk[i];
  } else if (w == 1.0) {
    asum[i] = __dadd_rn(asum[i], a[i]);
    vsum[i] = __dadd_rn(vsum[i], vk[i]);
  } else {
    asum[i] = __dadd_rn(asum[i], __dmul_rn(a[i], w));
    vsum[i] = __dadd_rn(vsum[i], __dmul_rn(vk[i], w));
  }
}

__global__ void k_rk4_final(int64_t m, double *x, double *v, const double *asum, const double *vsum, double h6,
                            const DevPlan *dp) {
  if (plan_dead(dp)) return;
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= m) return;
  v[i] = __dadd_rn(v[i], __dmul_rn(asum[i], h6));
  x[i] = __dadd_rn(x[i], __dmul_rn(vsum[i], h6));
}

__global__ void k_copy(int64_t m, double *dst, const double *src, const DevPlan *dp) {
  if (plan_dead(dp)) return;
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < m) dst[i] = src[i];
}

inline unsigned nb(int64_t m) { return (unsigned)std::max<int64_t>(1, (m + kT - 1) / kT); }
// ---- diagnostics --------------------------------------------------------------------------
// Sums of a particle state in FP64, in a fixed order (strided per thread, tree
// per block, blocks in order): 0 kinetic energy 1/2 m v^2, 1 potential energy
// 1/2 m phi, 3..5 momentum m v, 6..8 angular momentum m x cross v, 9 mass.
constexpr int kDiagBlocks = 256;
__global__ void __launch_bounds__(kT) k_diag_part(int64_t n, const double *__restrict__ x, const double *__restrict__ v,
                                                  const double *__restrict__ m, const double *__restrict__ phi,
                                                  double *__restrict__ part) {
  __shared__ double sh[10][kT];
  double q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double mi = m[i];
    const double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
    const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
    q[0] += 0.5 * mi * (vx * vx + vy * vy + vz * vz);
    if (phi) q[1] += 0.5 * mi * phi[i];
    q[3] += mi * vx;
    q[4] += mi * vy;
    q[5] += mi * vz;
    q[6] += mi * (py * vz - pz * vy);
    q[7] += mi * (pz * vx - px * vz);
    q[8] += mi * (px * vy - py * vx);
    q[9] += mi;
  }
  for (int k = 0; k < 10; ++k) sh[k][threadIdx.x] = q[k];
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 10; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 10) part[(size_t)blockIdx.x * 10 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_diag_final(int nb, const double *__restrict__ part, double *__restrict__ out) {
  const int k = threadIdx.x;
  if (k >= 10) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * 10 + k];
  out[k] = s;
}

} // namespace

void se_update(nbfmm_ctx *c, int64_t n, double *x, double *v, const double *a, double dt) {
  hipLaunchKernelGGL(k_se, dim3(nb(3 * n)), dim3(kT), 0, c->stream, 3 * n, x, v, a, dt, (const DevPlan *)dplan(c));
  NB_LAUNCH();
}
void rk4_stage(nbfmm_ctx *c, int64_t n, const double *x, const double *v, const double *vk_prev,
               const double *a_vel, double h, double *xk, double *vk_out) {
  hipLaunchKernelGGL(k_rk4_stage, dim3(nb(3 * n)), dim3(kT), 0, c->stream, 3 * n, x, v, vk_prev, a_vel, h, xk, vk_out);
  NB_LAUNCH();
}
void rk4_accum(nbfmm_ctx *c, int64_t n, double *asum, double *vsum, const double *a, const double *vk, double w,
               int first) {
  hipLaunchKernelGGL(k_rk4_accum, dim3(nb(3 * n)), dim3(kT), 0, c->stream, 3 * n, asum, vsum, a, vk, w, first);
  NB_LAUNCH();
}
void rk4_final(nbfmm_ctx *c, int64_t n, double *x, double *v, const double *asum, const double *vsum, double h6) {
  hipLaunchKernelGGL(k_rk4_final, dim3(nb(3 * n)), dim3(kT), 0, c->stream, 3 * n, x, v, asum, vsum, h6,
                     (const DevPlan *)dplan(c));
  NB_LAUNCH();
}
void copy_d(nbfmm_ctx *c, double *dst, const double *src, int64_t count) {
  hipLaunchKernelGGL(k_copy, dim3(nb(count)), dim3(kT), 0, c->stream, count, dst, src, (const DevPlan *)dplan(c));
  NB_LAUNCH();
}

void diagnostics(nbfmm_ctx *c, int64_t n, const double *x, const double *v, const double *m, const double *phi,
                 double *h_out) {
  double *part = dget<double>(c->dpart, (size_t)kDiagBlocks * 10 + 16);
  double *d_out = part + (size_t)kDiagBlocks * 10;
  hipLaunchKernelGGL(k_diag_part, dim3(kDiagBlocks), dim3(kT), 0, c->stream, n, x, v, m, phi, part);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c->stream, kDiagBlocks, (const double *)part, d_out);
  NB_LAUNCH();
  NB_HIP(hipMemcpyAsync(h_out, d_out, sizeof(double) * 10, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  h_out[2] = h_out[0] + h_out[1];
}

} // namespace nbfmm
