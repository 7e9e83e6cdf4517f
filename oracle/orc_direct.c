/*
 * orc_direct.c -- direct sums and integrators of the oracle.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* NaiveUpdate::update (naive_update.cpp:9-35), 3D branch, with the reference's
 * operation order; eps2 is the reference's hardcoded 1e-4 made a parameter.
 * Coincident distinct particles (r^2 == 0) are skipped. */
void orc_direct_naive(int64_t n, const double *xyz, const double *m, double G, double eps2,
                      double *acc) {
  for (int64_t i = 0; i < n; ++i) {
    double g[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double d[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2]};
      double r2 = 0.0;
      r2 += d[0] * d[0];
      r2 += d[1] * d[1];
      r2 += d[2] * d[2];
      if (r2 == 0.0) continue;
      double s = m[j] / ((r2 + eps2) * sqrt(r2));
      for (int k = 0; k < 3; ++k) g[k] -= d[k] * s;
    }
    for (int k = 0; k < 3; ++k) acc[3 * i + k] = g[k] * G;
  }
}

/* Neumaier-compensated direct sum for selected targets (the metric oracle). */
void orc_direct_targets(int64_t n, const double *xyz, const double *m, double G, double eps2,
                        int64_t ntgt, const int64_t *tgt, double *acc) {
  for (int64_t t = 0; t < ntgt; ++t) {
    int64_t i = tgt[t];
    double s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      double d[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1], xyz[3 * j + 2] - xyz[3 * i + 2]};
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2;
      if (r2 == 0.0) continue;
      double w = m[j] / (r2 * sqrt(r2));
      for (int k = 0; k < 3; ++k) {
        double v = d[k] * w;
        double u = s[k] + v;
        if (fabs(s[k]) >= fabs(v))
          c[k] += (s[k] - u) + v;
        else
          c[k] += (v - u) + s[k];
        s[k] = u;
      }
    }
    for (int k = 0; k < 3; ++k) acc[3 * t + k] = G * (s[k] + c[k]);
  }
}

static int run_update(int64_t n, const double *xyz, const double *m, const orc_upd_cfg *cfg,
                      double *acc) {
  switch (cfg->updater) {
    case ORC_UPD_NAIVE:
      orc_direct_naive(n, xyz, m, cfg->G, cfg->eps2, acc);
      return 0;
    case ORC_UPD_FMM_REF:
      return orc_fmm_refcompat(n, xyz, m, cfg->order, cfg->thr, cfg->G, acc);
    case ORC_UPD_FMM_EXACT:
      return orc_fmm_exact(n, xyz, m, cfg->order, cfg->thr, cfg->G, cfg->center_mode, cfg->p2p_cut, acc);
  }
  return -1;
}

/* SyplecticEuler::integrate (symplectic_euler.cpp:8-15):
 * update; v += dt*a; x += dt*v. */
int orc_step_se(int64_t n, double *xyz, double *vel, double *acc, const double *m,
                const orc_upd_cfg *cfg, double dt) {
  int rc = run_update(n, xyz, m, cfg, acc);
  if (rc) return rc;
  for (int64_t i = 0; i < 3 * n; ++i) {
    vel[i] += acc[i] * dt;
    xyz[i] += vel[i] * dt;
  }
  return 0;
}

/* RungeKutta4::integrate (runge_kutta_4.cpp:10-93).
 * mode 0 restates the reference's observable semantics: each stage copies the
 * input particles (positions, velocities and the carried acceleration a_in,
 * lines 41, 57, 73), so stage velocities are v + h*a_in (lines 46-48, 62-64,
 * 77-79) and the carried acceleration is left unchanged.
 * mode 1 is classic RK4; the carried acceleration becomes a(x_n). */
int orc_step_rk4(int64_t n, double *xyz, double *vel, double *acc, const double *m,
                 const orc_upd_cfg *cfg, double dt, int mode) {
  size_t sz = sizeof(double) * 3 * (size_t)n;
  double *xk = (double *)malloc(sz), *vk = (double *)malloc(sz), *ak = (double *)malloc(sz);
  double *asum = (double *)malloc(sz), *vsum = (double *)malloc(sz);
  double *a1 = NULL;
  double h2 = dt / 2, h6 = dt / 6;
  int rc;
  /* stage 1 */
  memcpy(vk, vel, sz);
  rc = run_update(n, xyz, m, cfg, ak);
  if (rc) goto done;
  memcpy(asum, ak, sz);
  memcpy(vsum, vk, sz);
  if (mode == 1) {
    a1 = (double *)malloc(sz);
    memcpy(a1, ak, sz);
  }
  for (int stage = 2; stage <= 4; ++stage) {
    double h = stage == 4 ? dt : h2;
    double wgt = stage == 4 ? 1.0 : 2.0;
    /* carried acceleration used for the stage velocity */
    const double *ac = mode == 0 ? acc : ak;
    for (int64_t i = 0; i < 3 * n; ++i) xk[i] = xyz[i] + vk[i] * h;
    for (int64_t i = 0; i < 3 * n; ++i) vk[i] = vel[i] + ac[i] * h;
    rc = run_update(n, xk, m, cfg, ak);
    if (rc) goto done;
    for (int64_t i = 0; i < 3 * n; ++i) {
      if (stage == 4) {
        asum[i] += ak[i];
        vsum[i] += vk[i];
      } else {
        asum[i] += ak[i] * wgt;
        vsum[i] += vk[i] * wgt;
      }
    }
  }
  for (int64_t i = 0; i < 3 * n; ++i) {
    vel[i] += asum[i] * h6;
    xyz[i] += vsum[i] * h6;
  }
  if (mode == 1) memcpy(acc, a1, sz);
done:
  free(a1);
  free(xk);
  free(vk);
  free(ak);
  free(asum);
  free(vsum);
  return rc;
}
