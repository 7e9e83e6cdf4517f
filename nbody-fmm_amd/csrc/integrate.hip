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

// HIP contracts a*b+c into FMA by default (and __dmul_rn is a plain multiply);
// the reference's x86-64 build rounds every operation, so contraction is off.
#pragma clang fp contract(off)

namespace nbfmm {
namespace {
constexpr int kT = 256;

__global__ void k_se(int64_t m, double *x, double *v, const double *a, double dt) {
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

__global__ void k_rk4_final(int64_t m, double *x, double *v, const double *asum, const double *vsum, double h6) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= m) return;
  v[i] = __dadd_rn(v[i], __dmul_rn(asum[i], h6));
  x[i] = __dadd_rn(x[i], __dmul_rn(vsum[i], h6));
}

inline unsigned nb(int64_t m) { return (unsigned)((m + kT - 1) / kT); }
} // namespace

void se_update(nbfmm_ctx *c, int64_t n, double *x, double *v, const double *a, double dt) {
  hipLaunchKernelGGL(k_se, dim3(nb(3 * n)), dim3(kT), 0, c->stream, 3 * n, x, v, a, dt);
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
  hipLaunchKernelGGL(k_rk4_final, dim3(nb(3 * n)), dim3(kT), 0, c->stream, 3 * n, x, v, asum, vsum, h6);
  NB_LAUNCH();
}
void copy_d(nbfmm_ctx *c, double *dst, const double *src, int64_t count) {
  NB_HIP(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)count, hipMemcpyDeviceToDevice, c->stream));
}

} // namespace nbfmm
