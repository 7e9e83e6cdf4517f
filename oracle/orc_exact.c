/*
 * orc_exact.c -- correct solid-harmonic operators and the exact FMM on the
 * reference tree.  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Normalisation (Yoshida; Epton & Dembart):
 *   R_n^m(r) = r^n P_n^m(cos t) e^{i m phi} / (n+m)!
 *   S_n^m(r) = (n-m)! P_n^m(cos t) e^{i m phi} / r^{n+1}
 * with the Condon-Shortley phase in P_n^m and X_n^{-m} = (-1)^m conj(X_n^m).
 * Then 1/|x-y| = sum_{n,m} conj(R_n^m(y)) S_n^m(x) for |y| < |x|, and
 *   P2M  M_n^m  = sum_i q_i conj(R_n^m(x_i - c))
 *   M2M  M'_n^m = sum_{j,k} conj(R_j^k(c - c')) M_{n-j}^{m-k}
 *   M2L  L_j^k  = (-1)^{j+k} sum_{n,m} M_n^m S_{n+j}^{m-k}(c_L - c_M)
 *   L2L  L'_j^k = sum_{n>=j,m} L_n^m R_{n-j}^{m-k}(c' - c)
 *   L2P  phi = sum L_n^m R_n^m(x - c); d/dz R_n^m = R_{n-1}^m,
 *        (d/dx + i d/dy) R_n^m = R_{n-1}^{m+1}, (d/dx - i d/dy) R_n^m = -R_{n-1}^{m-1}
 *   M2P  phi = sum M_n^m S_n^m(x - c); d/dz S_n^m = -S_{n+1}^m,
 *        (d/dx + i d/dy) S_n^m = S_{n+1}^{m+1}, (d/dx - i d/dy) S_n^m = -S_{n+1}^{m-1}
 * The FMM driver follows the reference's pass structure (tree.cpp:93-111):
 * upward pass on every level, M2L over each cell's interaction list
 * (node_i.cpp:105-120), top-down L2L (node.cpp:33-43), leaf L2P plus P2P over
 * the own leaf and the neighbour leaves (leaf.cpp:105-143).  The expansion
 * centres are geometric cell centres (centre_mode 0) or mass centres (1).
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int tri(int n, int m) { return n * (n + 1) / 2 + m; }
static int tsize(int p) { return (p + 1) * (p + 2) / 2; }

static cplx tget(const double *X, int n, int m) {
  if (m >= 0) {
    int i = tri(n, m);
    return cmk(X[2 * i], X[2 * i + 1]);
  }
  int i = tri(n, -m);
  cplx v = cmk(X[2 * i], -X[2 * i + 1]);
  if ((-m) & 1) v = cmk(-v.re, -v.im);
  return v;
}
static void tset(double *X, int n, int m, cplx v) {
  int i = tri(n, m);
  X[2 * i] = v.re;
  X[2 * i + 1] = v.im;
}

void orc_exact_regular(uint32_t pu, const double r[3], double *R) {
  int p = (int)pu;
  double x = r[0], y = r[1], z = r[2], r2 = x * x + y * y + z * z;
  memset(R, 0, sizeof(double) * 2 * (size_t)tsize(p));
  tset(R, 0, 0, cmk(1.0, 0.0));
  for (int m = 0; m <= p; ++m) {
    if (m > 0) {
      cplx prev = tget(R, m - 1, m - 1);
      cplx v = cmul(cmk(-x, -y), prev);
      tset(R, m, m, cscale(1.0 / (2.0 * m), v));
    }
    if (m + 1 <= p) tset(R, m + 1, m, cscale(z, tget(R, m, m)));
    for (int n = m + 2; n <= p; ++n) {
      cplx a = cscale((2.0 * n - 1.0) * z, tget(R, n - 1, m));
      cplx b = cscale(r2, tget(R, n - 2, m));
      tset(R, n, m, cscale(1.0 / ((double)(n + m) * (double)(n - m)), csub(a, b)));
    }
  }
}

void orc_exact_singular(uint32_t pu, const double r[3], double *S) {
  int p = (int)pu;
  double x = r[0], y = r[1], z = r[2], r2 = x * x + y * y + z * z;
  double ir2 = 1.0 / r2;
  memset(S, 0, sizeof(double) * 2 * (size_t)tsize(p));
  tset(S, 0, 0, cmk(1.0 / sqrt(r2), 0.0));
  for (int m = 0; m <= p; ++m) {
    if (m > 0) {
      cplx prev = tget(S, m - 1, m - 1);
      cplx v = cmul(cmk(-x, -y), prev);
      tset(S, m, m, cscale((2.0 * m - 1.0) * ir2, v));
    }
    if (m + 1 <= p) tset(S, m + 1, m, cscale((2.0 * m + 1.0) * z * ir2, tget(S, m, m)));
    for (int n = m + 2; n <= p; ++n) {
      cplx a = cscale((2.0 * n - 1.0) * z, tget(S, n - 1, m));
      cplx b = cscale((double)((n - 1) * (n - 1) - m * m), tget(S, n - 2, m));
      tset(S, n, m, cscale(ir2, csub(a, b)));
    }
  }
}

void orc_exact_p2m_add(uint32_t pu, const double r[3], double q, double *M) {
  int p = (int)pu;
  double *R = (double *)malloc(sizeof(double) * 2 * (size_t)tsize(p));
  orc_exact_regular(pu, r, R);
  for (int i = 0; i < tsize(p); ++i) {
    M[2 * i] += q * R[2 * i];
    M[2 * i + 1] -= q * R[2 * i + 1];
  }
  free(R);
}

void orc_exact_m2m(uint32_t pu, const double *M, const double d[3], double *out) {
  int p = (int)pu;
  double *R = (double *)malloc(sizeof(double) * 2 * (size_t)tsize(p));
  orc_exact_regular(pu, d, R);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      cplx acc = cmk(0, 0);
      for (int j = 0; j <= n; ++j)
        for (int k = -j; k <= j; ++k) {
          int mm = m - k;
          if (mm < -(n - j) || mm > n - j) continue;
          acc = cadd(acc, cmul(cconj(tget(R, j, k)), tget(M, n - j, mm)));
        }
      tset(out, n, m, acc);
    }
  free(R);
}

void orc_exact_m2l(uint32_t pu, const double *M, const double D[3], double *out) {
  int p = (int)pu;
  double *S = (double *)malloc(sizeof(double) * 2 * (size_t)tsize(2 * p));
  orc_exact_singular((uint32_t)(2 * p), D, S);
  for (int j = 0; j <= p; ++j)
    for (int k = 0; k <= j; ++k) {
      cplx acc = cmk(0, 0);
      for (int n = 0; n <= p; ++n)
        for (int m = -n; m <= n; ++m) acc = cadd(acc, cmul(tget(M, n, m), tget(S, n + j, m - k)));
      if ((j + k) & 1) acc = cmk(-acc.re, -acc.im);
      tset(out, j, k, acc);
    }
  free(S);
}

void orc_exact_l2l(uint32_t pu, const double *Lc, const double d[3], double *out) {
  int p = (int)pu;
  double *R = (double *)malloc(sizeof(double) * 2 * (size_t)tsize(p));
  orc_exact_regular(pu, d, R);
  for (int j = 0; j <= p; ++j)
    for (int k = 0; k <= j; ++k) {
      cplx acc = cmk(0, 0);
      for (int n = j; n <= p; ++n)
        for (int m = -n; m <= n; ++m) {
          int mk = m - k;
          if (mk < -(n - j) || mk > n - j) continue;
          acc = cadd(acc, cmul(tget(Lc, n, m), tget(R, n - j, mk)));
        }
      tset(out, j, k, acc);
    }
  free(R);
}

/* sum_m X_n^m Y_n^m over m=-n..n for two coefficient families that both obey
 * X^{-m} = (-1)^m conj(X^m); the result is real. */
void orc_exact_l2p(uint32_t pu, const double *Lc, const double t[3], double *phi, double g[3]) {
  int p = (int)pu;
  double *R = (double *)malloc(sizeof(double) * 2 * (size_t)tsize(p));
  orc_exact_regular(pu, t, R);
  double ph = 0, gx = 0, gy = 0, gz = 0;
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      double w = m == 0 ? 1.0 : 2.0;
      cplx L = tget(Lc, n, m);
      ph += w * cmul(L, tget(R, n, m)).re;
      if (n == 0) continue;
      cplx up = (m + 1 <= n - 1) ? tget(R, n - 1, m + 1) : cmk(0, 0);
      cplx dn = (m - 1 >= -(n - 1)) ? tget(R, n - 1, m - 1) : cmk(0, 0);
      cplx zz = (m <= n - 1) ? tget(R, n - 1, m) : cmk(0, 0);
      cplx dx = cscale(0.5, csub(up, dn));
      cplx dy = cmul(cmk(0, -0.5), cadd(up, dn));
      gx += w * cmul(L, dx).re;
      gy += w * cmul(L, dy).re;
      gz += w * cmul(L, zz).re;
    }
  if (phi) *phi = ph;
  g[0] = gx;
  g[1] = gy;
  g[2] = gz;
  free(R);
}

void orc_exact_m2p(uint32_t pu, const double *M, const double t[3], double *phi, double g[3]) {
  int p = (int)pu;
  double *S = (double *)malloc(sizeof(double) * 2 * (size_t)tsize(p + 1));
  orc_exact_singular((uint32_t)(p + 1), t, S);
  double ph = 0, gx = 0, gy = 0, gz = 0;
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      double w = m == 0 ? 1.0 : 2.0;
      cplx Mv = tget(M, n, m);
      ph += w * cmul(Mv, tget(S, n, m)).re;
      cplx up = tget(S, n + 1, m + 1);
      cplx dn = tget(S, n + 1, m - 1);
      cplx dx = cscale(0.5, csub(up, dn));
      cplx dy = cmul(cmk(0, -0.5), cadd(up, dn));
      cplx dz = cscale(-1.0, tget(S, n + 1, m));
      gx += w * cmul(Mv, dx).re;
      gy += w * cmul(Mv, dy).re;
      gz += w * cmul(Mv, dz).re;
    }
  if (phi) *phi = ph;
  g[0] = gx;
  g[1] = gy;
  g[2] = gz;
  free(S);
}

/* ------------------------------------------------------------------------ */

static void p2p_cells(const orc_tree *t, const double *xyz, const double *mass, int64_t tc,
                      int64_t sc, double *grad) {
  for (int64_t k = t->pbeg[tc]; k < t->pend[tc]; ++k) {
    int64_t i = t->perm[k];
    for (int64_t s = t->pbeg[sc]; s < t->pend[sc]; ++s) {
      int64_t j = t->perm[s];
      if (j == i) continue;
      double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 == 0.0) continue;
      double inv = mass[j] / (r2 * sqrt(r2));
      grad[3 * i] -= dx * inv;
      grad[3 * i + 1] -= dy * inv;
      grad[3 * i + 2] -= dz * inv;
    }
  }
}

int orc_fmm_exact(int64_t n, const double *xyz, const double *mass, uint32_t pu, uint32_t thr,
                  double G, int center_mode, int64_t p2p_cut, double *acc) {
  return orc_fmm_exact_box(n, xyz, mass, pu, thr, G, center_mode, p2p_cut, 0, acc);
}

int orc_fmm_exact_box(int64_t n, const double *xyz, const double *mass, uint32_t pu, uint32_t thr,
                      double G, int center_mode, int64_t p2p_cut, int cubic, double *acc) {
  return orc_fmm_exact_opt(n, xyz, mass, pu, thr, G, center_mode, p2p_cut, cubic, 0.0, acc);
}

/* Distance-dependent M2L orders (the library's class_order_tol; not in the
 * reference).  For the offset classes d (integers in [-3,3]^3, Chebyshev
 * distance >= 2) that are admissible (diag^2 <= 0.81 |d|^2), rate_d^2 =
 * diag^2 / |d|^2 in units of the box extents; class 0 has the largest rate.
 * A class with rate below it gets the lowest q in [4, p] with
 * (0.55 rate_d)^(q+1) / |d|^2 <= tol (0.55 rate_0)^(p+1) / |d_0|^2; every
 * other class keeps p.  The same operations as the library's table, so both
 * give the same integers. */
void orc_class_orders(const double ext[3], int p, double tol, int *q) {
  const double ex = ext[0], ey = ext[1], ez = ext[2];
  const double diag2 = ex * ex + ey * ey + ez * ez;
  double r2[343], dd[343], r0 = 0, d0 = 0;
  for (int k = 0; k < 343; ++k) {
    const double ox = k % 7 - 3, oy = (k / 7) % 7 - 3, oz = k / 49 - 3;
    r2[k] = -1;
    if (fabs(ox) < 2 && fabs(oy) < 2 && fabs(oz) < 2) continue;
    const double d2 = ox * ox * ex * ex + oy * oy * ey * ey + oz * oz * ez * ez;
    if (diag2 > 0.9 * 0.9 * d2) continue;
    r2[k] = diag2 / d2;
    dd[k] = d2;
    if (r2[k] > r0) {
      r0 = r2[k];
      d0 = d2;
    }
  }
  const double la = log(0.55);
  for (int k = 0; k < 343; ++k) {
    int qk = p;
    if (tol > 0 && r2[k] > 0 && r2[k] < r0) {
      const double need = (log(tol) + (p + 1) * (la + 0.5 * log(r0)) + log(dd[k] / d0)) / (la + 0.5 * log(r2[k]));
      qk = (int)ceil(need) - 1;
      if (qk < 4) qk = 4;
      if (qk > p) qk = p;
    }
    q[k] = qk;
  }
}

int orc_fmm_exact_opt(int64_t n, const double *xyz, const double *mass, uint32_t pu, uint32_t thr, double G,
                      int center_mode, int64_t p2p_cut, int cubic, double class_order_tol, double *acc) {
  if (n < 1 || pu < 1) return -1;
  orc_tree *t = orc_tree_build_box(n, xyz, thr, cubic);
  if (!t) return -1;
  int p = (int)pu, D = t->depth;
  int64_t nc = t->ncells;
  size_t cs = 2 * (size_t)tsize(p);
  double *ctr = (double *)calloc((size_t)nc * 3, sizeof(double));
  double *Mc = (double *)calloc((size_t)nc * cs, sizeof(double));
  double *Lc = (double *)calloc((size_t)nc * cs, sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * cs);
  double *mtot = (double *)calloc((size_t)nc, sizeof(double));
  double *grad = (double *)calloc((size_t)n * 3, sizeof(double));

  /* expansion centres */
  for (int lev = 0; lev <= D; ++lev) {
    double h[3];
    for (int d = 0; d < 3; ++d) h[d] = (t->hi[d] - t->lo[d]) / (double)((uint64_t)1 << lev);
    for (int64_t c = t->level_start[lev]; c < t->level_start[lev + 1]; ++c) {
      uint32_t q[3];
      orc_morton_decode3(t->key[c], &q[0], &q[1], &q[2]);
      double mt = 0, s[3] = {0, 0, 0};
      for (int64_t k = t->pbeg[c]; k < t->pend[c]; ++k) {
        int64_t i = t->perm[k];
        mt += mass[i];
        for (int d = 0; d < 3; ++d) s[d] += mass[i] * xyz[3 * i + d];
      }
      mtot[c] = mt;
      for (int d = 0; d < 3; ++d)
        ctr[3 * c + d] = center_mode == 1 ? s[d] / mt : t->lo[d] + ((double)q[d] + 0.5) * h[d];
    }
  }
  /* P2M at leaves, M2M upward */
  for (int64_t c = t->level_start[D]; c < t->level_start[D + 1]; ++c)
    for (int64_t k = t->pbeg[c]; k < t->pend[c]; ++k) {
      int64_t i = t->perm[k];
      double r[3] = {xyz[3 * i] - ctr[3 * c], xyz[3 * i + 1] - ctr[3 * c + 1], xyz[3 * i + 2] - ctr[3 * c + 2]};
      orc_exact_p2m_add(pu, r, mass[i], Mc + cs * c);
    }
  for (int lev = D; lev >= 1; --lev)
    for (int64_t c = t->level_start[lev]; c < t->level_start[lev + 1]; ++c) {
      int64_t par = t->parent[c];
      double d[3] = {ctr[3 * c] - ctr[3 * par], ctr[3 * c + 1] - ctr[3 * par + 1], ctr[3 * c + 2] - ctr[3 * par + 2]};
      orc_exact_m2m(pu, Mc + cs * c, d, tmp);
      for (size_t q = 0; q < cs; ++q) Mc[cs * par + q] += tmp[q];
    }
  /* interaction lists: M2L, or direct P2P between the two cells when
   * n_t * n_s < p2p_cut, or when n_t * n_s < 2^20 and the offset is not
   * admissible for the cell shape: (R_s + R_t) / |d| > 0.9 with R the half
   * diagonal of a cell (the library's rule for boxes of the bounding box's
   * aspect ratio) */
  const double ex = t->hi[0] - t->lo[0], ey = t->hi[1] - t->lo[1], ez = t->hi[2] - t->lo[2];
  const double diag2 = ex * ex + ey * ey + ez * ez;
  int qcls[343];
  {
    const double ext[3] = {ex, ey, ez};
    orc_class_orders(ext, p, class_order_tol, qcls);
  }
  for (int64_t c = 0; c < nc; ++c)
    for (int64_t q = t->il_off[c]; q < t->il_off[c + 1]; ++q) {
      int64_t s = t->il[q];
      uint32_t qc[3], qs[3];
      orc_morton_decode3(t->key[c], &qc[0], &qc[1], &qc[2]);
      orc_morton_decode3(t->key[s], &qs[0], &qs[1], &qs[2]);
      const double ox = (double)qs[0] - (double)qc[0], oy = (double)qs[1] - (double)qc[1],
                   oz = (double)qs[2] - (double)qc[2];
      const double d2 = ox * ox * ex * ex + oy * oy * ey * ey + oz * oz * ez * ez;
      const int bad = diag2 > 0.9 * 0.9 * d2;
      const int64_t np = (t->pend[c] - t->pbeg[c]) * (t->pend[s] - t->pbeg[s]);
      if ((bad && np < ((int64_t)1 << 20)) || np < p2p_cut) {
        p2p_cells(t, xyz, mass, c, s, grad);
        continue;
      }
      double Dv[3] = {ctr[3 * c] - ctr[3 * s], ctr[3 * c + 1] - ctr[3 * s + 1], ctr[3 * c + 2] - ctr[3 * s + 2]};
      /* the pair's order: its class's (p unless class_order_tol > 0);
       * degrees above it stay zero */
      const int cls = ((int)oz + 3) * 49 + ((int)oy + 3) * 7 + ((int)ox + 3);
      for (size_t k = 0; k < cs; ++k) tmp[k] = 0;
      orc_exact_m2l((uint32_t)qcls[cls], Mc + cs * s, Dv, tmp);
      for (size_t k = 0; k < cs; ++k) Lc[cs * c + k] += tmp[k];
    }
  /* L2L downward */
  for (int lev = 1; lev <= D; ++lev)
    for (int64_t c = t->level_start[lev]; c < t->level_start[lev + 1]; ++c) {
      int64_t par = t->parent[c];
      double d[3] = {ctr[3 * c] - ctr[3 * par], ctr[3 * c + 1] - ctr[3 * par + 1], ctr[3 * c + 2] - ctr[3 * par + 2]};
      orc_exact_l2l(pu, Lc + cs * par, d, tmp);
      for (size_t k = 0; k < cs; ++k) Lc[cs * c + k] += tmp[k];
    }
  /* leaves: L2P + P2P (own leaf and neighbour leaves) */
  for (int64_t c = t->level_start[D]; c < t->level_start[D + 1]; ++c)
    for (int64_t k = t->pbeg[c]; k < t->pend[c]; ++k) {
      int64_t i = t->perm[k];
      double tt[3] = {xyz[3 * i] - ctr[3 * c], xyz[3 * i + 1] - ctr[3 * c + 1], xyz[3 * i + 2] - ctr[3 * c + 2]};
      double g[3];
      orc_exact_l2p(pu, Lc + cs * c, tt, NULL, g);
      for (int d = 0; d < 3; ++d) grad[3 * i + d] += g[d];
    }
  for (int64_t c = t->level_start[D]; c < t->level_start[D + 1]; ++c) {
    p2p_cells(t, xyz, mass, c, c, grad);
    for (int64_t q = t->nbr_off[c]; q < t->nbr_off[c + 1]; ++q) p2p_cells(t, xyz, mass, c, t->nbr[q], grad);
  }
  for (int64_t i = 0; i < 3 * n; ++i) acc[i] = G * grad[i];
  free(ctr);
  free(Mc);
  free(Lc);
  free(tmp);
  free(mtot);
  free(grad);
  orc_tree_free(t);
  return 0;
}
