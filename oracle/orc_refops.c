/*
 * orc_refops.c -- the reference's own 3D expansion operators and FMM driver,
 * restated in C with the reference's operation order.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Storage layouts are the reference's:
 *   PowerSet<3> / Irregular : (L+1)^2 doubles, idx(n,0)=n^2,
 *                             idx(n,m>0)=n^2+2m-1 (re), +1 (im)
 *                             (power_set3D.cpp:49-63, irregular.cpp:43-57)
 *   MultipoleSet<3>         : 2L^2+4L+1 doubles, idx(0,0)=0 (re only),
 *                             idx(n>=1,m)=2n^2+2n+2m-1 (re), +1 (im), m=-n..n
 *                             (multipole_set3D.cpp:18-29)
 *   LocalSet<3>             : 2(L+1)^2 doubles, idx(n,m)=2n^2+2n+2m (re), +1
 *                             (local_set3D.cpp:17-31)
 * These operators carry the reference's conventions unchanged (the far field
 * they produce does not converge to the direct sum); they exist so that the
 * reference's FMM::update output can be reproduced and checked.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const cplx IPOW[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};

/* PowerSet<3>::operator() / Irregular::operator() (power_set3D.cpp:49-63) */
static cplx ps_get(const double *e, int l, int m) {
  if (m == 0) return cmk(e[l * l], 0.0);
  if (m > 0) return cmk(e[l * l + 2 * m - 1], e[l * l + 2 * m]);
  if (m % 2 == 0) return cmk(e[l * l - 2 * m - 1], -e[l * l - 2 * m]);
  return cmk(-e[l * l - 2 * m - 1], e[l * l - 2 * m]);
}

/* MultipoleSet<3>::operator() (multipole_set3D.cpp:18-23) */
static cplx ms_get(const double *e, int l, int m) {
  if (l == 0) return cmk(e[0], 0.0);
  unsigned idx = (unsigned)(2 * l * l + 2 * l + 2 * m - 1);
  return cmk(e[idx], e[idx + 1]);
}

/* LocalSet<3>::operator() (local_set3D.cpp:25-29) */
static cplx ls_get(const double *e, int l, int m) {
  unsigned idx = (unsigned)(2 * l * l + 2 * l + 2 * m);
  return cmk(e[idx], e[idx + 1]);
}
static void ls_set(double *e, int n, int m, cplx v) {
  unsigned idx = (unsigned)(2 * n * n + 2 * n + 2 * m);
  e[idx] = v.re;
  e[idx + 1] = v.im;
}

/* power_set3D.cpp:10-47 */
void orc_ref_powerset(uint32_t Lu, const double r[3], double *e) {
  int L = (int)Lu;
  int k = 0;
  e[k++] = 1;
  if (L >= 1) {
    e[k++] = r[2];
    e[k++] = -1 / ORC_SQRT2 * r[0];
    e[k++] = -1 / ORC_SQRT2 * r[1];
  }
  double rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  for (int l = 2; l <= L; l++) {
    e[k++] = (r[2] * (2 * l - 1) * ps_get(e, l - 1, 0).re -
              (l - 1) * rsq * ps_get(e, l - 2, 0).re) /
             (double)l;
    for (int m = 1; m < l - 1; m++) {
      cplx a = cscale((2 * l - 1) * r[2], ps_get(e, l - 1, m));
      cplx b = cscale(sqrt((double)((l - 1) * (l - 1) - m * m)) * rsq, ps_get(e, l - 2, m));
      cplx el = csub(a, b);
      double den = sqrt((double)(l * l - m * m));
      e[k++] = el.re / den;
      e[k++] = el.im / den;
    }
    cplx el = cscale(r[2] * sqrt((double)(2 * l - 1)), ps_get(e, l - 1, l - 1));
    e[k++] = el.re;
    e[k++] = el.im;
    double s = -sqrt((double)(2 * l - 1) / (2 * l));
    el = cmul(cmk(r[0] * s, r[1] * s), ps_get(e, l - 1, l - 1));
    e[k++] = el.re;
    e[k++] = el.im;
  }
}

/* irregular.cpp:10-41 */
void orc_ref_irregular(uint32_t Lu, const double r[3], double *e) {
  int L = (int)Lu;
  double rsq = 0.0;
  rsq += r[0] * r[0];
  rsq += r[1] * r[1];
  rsq += r[2] * r[2];
  int k = 0;
  e[k++] = 1 / sqrt(rsq);
  if (L >= 1) {
    e[k++] = r[2] * e[0] / rsq;
    e[k++] = -1 / ORC_SQRT2 * r[0] * e[0] / rsq;
    e[k++] = -1 / ORC_SQRT2 * r[1] * e[0] / rsq;
  }
  for (int l = 2; l <= L; l++) {
    e[k++] = (r[2] * (2 * l - 1) * ps_get(e, l - 1, 0).re - (l - 1) * ps_get(e, l - 2, 0).re) /
             (l * rsq);
    for (int m = 1; m < l - 1; m++) {
      cplx a = cscale((2 * l - 1) * r[2], ps_get(e, l - 1, m));
      cplx b = cscale(sqrt((double)((l - 1) * (l - 1) - m * m)), ps_get(e, l - 2, m));
      cplx el = csub(a, b);
      double den = sqrt((double)(l * l - m * m)) * rsq;
      e[k++] = el.re / den;
      e[k++] = el.im / den;
    }
    cplx el = cscale(r[2] * sqrt((double)(2 * l - 1)), ps_get(e, l - 1, l - 1));
    e[k++] = el.re / rsq;
    e[k++] = el.im / rsq;
    double s = -sqrt((double)(2 * l - 1) / (2 * l));
    el = cmul(cmk(r[0] * s, r[1] * s), ps_get(e, l - 1, l - 1));
    e[k++] = el.re / rsq;
    e[k++] = el.im / rsq;
  }
}

static int ms_size(int L) { return 2 * L * L + 4 * L + 1; }
static int ls_size(int L) { return 2 * (L + 1) * (L + 1); }

/* leaf.cpp:15-35 with PowerSetI::operator*= (power_set.cpp:12-18) and
 * MultipoleSet<3>::operator+=(PowerSetI) (multipole_set3D.cpp:31-43) */
void orc_ref_p2m_add(uint32_t Lu, const double r[3], double q, double *M) {
  int L = (int)Lu;
  int np = (L + 1) * (L + 1);
  if (np < 4) np = 4;
  double *z = (double *)malloc(sizeof(double) * (size_t)np);
  orc_ref_powerset(Lu, r, z);
  for (int i = 0; i < np; ++i) z[i] *= q;
  M[0] += ps_get(z, 0, 0).re;
  int idx = 1;
  for (int n = 1; n <= L; ++n)
    for (int m = -n; m <= n; ++m) {
      cplx zc = ps_get(z, n, -m);
      M[idx] += zc.re;
      M[idx + 1] += zc.im;
      idx += 2;
    }
  free(z);
}

static int ipow_idx(int v) { return (4 + v % 4) % 4; }

/* MultipoleSet<3>::weigh_children_with_distance (multipole_set3D.cpp:45-86) */
void orc_ref_m2m(uint32_t Lu, const double *M, const double d[3], double *out) {
  int L = (int)Lu;
  double *R = (double *)malloc(sizeof(double) * (size_t)((L + 1) * (L + 1) + 4));
  orc_ref_powerset(Lu, d, R);
  int o = 0;
  out[o++] = M[0];
  for (int j = 1; j <= L; ++j) {
    for (int k = -j; k <= j; ++k) {
      double A_row = 1;
      cplx mc = cmul(cscale(A_row, ms_get(M, 0, 0)), ps_get(R, j, -k));
      int m_row = k;
      for (int n = j - 1; n >= 0; --n) {
        if (m_row < 0) {
          ++m_row;
          A_row *= sqrt((double)((2 * n - j + k + 1) * (2 * n - j + k + 2)) /
                        ((2 * j - 2 * n) * (2 * j - 2 * n - 1)));
        } else if (m_row > 0) {
          --m_row;
          A_row *= sqrt((double)((2 * n - j + k + 1) * (2 * n - j + k + 2)) /
                        ((2 * j - 2 * n) * (2 * j - 2 * n - 1)));
        } else {
          A_row *= (n + 1) / sqrt((double)((j - k - n) * (j + k - n)));
        }
        double A = A_row;
        {
          int ip = ipow_idx(abs(k) - abs(m_row) - abs(k - m_row));
          cplx t = cmul(cscale(A, cmul(ms_get(M, j - n, k - m_row), IPOW[ip])),
                        ps_get(R, n, -m_row));
          mc = cadd(mc, t);
        }
        for (int m = m_row - 1; m >= -n && m >= k + n - j; --m) {
          A *= sqrt((double)((n + m + 1) * (j - k - n + m + 1)) / ((n - m) * (j + k - n - m)));
          int ip = ipow_idx(abs(k) - abs(m) - abs(k - m));
          cplx t = cmul(cscale(A, cmul(ms_get(M, j - n, k - m), IPOW[ip])), ps_get(R, n, -m));
          mc = cadd(mc, t);
        }
        A = A_row;
        for (int m = m_row + 1; m <= n && m <= k + j - n; ++m) {
          A *= sqrt((double)((n - m + 1) * (j + k - n - m + 1)) / ((n + m) * (j - k - n + m)));
          int ip = ipow_idx(abs(k) - abs(m) - abs(k - m));
          cplx t = cmul(cscale(A, cmul(ms_get(M, j - n, k - m), IPOW[ip])), ps_get(R, n, -m));
          mc = cadd(mc, t);
        }
      }
      out[o++] = mc.re;
      out[o++] = mc.im;
    }
  }
  free(R);
}

/* MultipoleSet<3>::to_local (multipole_set3D.cpp:88-118).  The reference
 * mixes unsigned and int arithmetic; every product it forms is non-negative,
 * so plain int arithmetic reproduces it. */
void orc_ref_m2l(uint32_t Lu, const double *M, const double d[3], double *out) {
  int L = (int)Lu;
  int L2 = 2 * L;
  double *I = (double *)malloc(sizeof(double) * (size_t)((L2 + 1) * (L2 + 1) + 4));
  orc_ref_irregular((uint32_t)L2, d, I);
  memset(out, 0, sizeof(double) * (size_t)ls_size(L));
  for (int j = 0; j <= L; ++j) {
    for (int k = -j; k <= j; ++k) {
      double A_row = 1;
      int sign = 1;
      cplx lc = cmul(cscale(A_row, ms_get(M, 0, 0)), ps_get(I, j, -k));
      for (int n = 1; n <= L; ++n) {
        A_row = sqrt((double)((unsigned)((n + j + k) * (n + j - k)))) / n;
        sign = -sign;
        double A = A_row;
        lc = cadd(lc, cmul(cscale(A, cscale((double)sign, ms_get(M, n, 0))), ps_get(I, j + n, -k)));
        for (int m = -1; m >= -n; --m) {
          A *= sqrt(((double)((n - m + j + k) * (n + m + 1)) / ((n + m + j - k + 1) * (n - m))));
          int ip = ipow_idx(abs(k - m) - abs(k) - abs(m));
          cplx t = cmul(cscale(A, cmul(cscale((double)sign, ms_get(M, n, m)), IPOW[ip])),
                        ps_get(I, j + n, m - k));
          lc = cadd(lc, t);
        }
        A = A_row;
        for (int m = 1; m <= n; ++m) {
          A *= sqrt(((double)((n + m + j - k) * (n - m + 1)) / ((n - m + j + k + 1) * (n + m))));
          int ip = ipow_idx(abs(k - m) - abs(k) - abs(m));
          cplx t = cmul(cscale(A, cmul(cscale((double)sign, ms_get(M, n, m)), IPOW[ip])),
                        ps_get(I, j + n, m - k));
          lc = cadd(lc, t);
        }
      }
      ls_set(out, j, k, lc);
    }
  }
  free(I);
}

/* LocalSet<3>::distribute_parent_with_distance (local_set3D.cpp:51-81) */
void orc_ref_l2l(uint32_t Lu, const double *Lc, const double d[3], double *out) {
  int L = (int)Lu;
  int L2 = 2 * L;
  double *R = (double *)malloc(sizeof(double) * (size_t)((L2 + 1) * (L2 + 1) + 4));
  orc_ref_powerset((uint32_t)L2, d, R);
  memset(out, 0, sizeof(double) * (size_t)ls_size(L));
  for (int j = 0; j <= L; ++j) {
    for (int k = -j; k <= j; ++k) {
      double A_row = 1;
      int sign = 1;
      cplx lc = cmul(ls_get(Lc, j, k), ps_get(R, 0, 0));
      for (int n = j + 1; n <= L; ++n) {
        A_row *= sqrt((double)((unsigned)((n - k) * (n + k)))) / (n - j);
        sign = (n + j) % 2 ? -1 : 1;
        double A = A_row;
        lc = cadd(lc, cmul(cscale(A, cscale((double)sign, ls_get(Lc, n, k))), ps_get(R, n - j, 0)));
        for (int m = k - 1; m >= k - n + j; --m) {
          A *= sqrt((double)((n - m) * (n + m - j - k + 1)) / ((n + m + 1) * (n - m - j + k)));
          int ip = ipow_idx(abs(m) - abs(m - k) - abs(k));
          cplx t = cmul(cscale(A, cmul(cscale((double)sign, ls_get(Lc, n, m)), IPOW[ip])),
                        ps_get(R, n - j, m - k));
          lc = cadd(lc, t);
        }
        A = A_row;
        for (int m = k + 1; m <= k + n - j; ++m) {
          A *= sqrt((double)((n + m) * (n - m - j + k + 1)) / ((n - m + 1) * (n + m - j - k)));
          int ip = ipow_idx(abs(m) - abs(m - k) - abs(k));
          cplx t = cmul(cscale(A, cmul(cscale((double)sign, ls_get(Lc, n, m)), IPOW[ip])),
                        ps_get(R, n - j, m - k));
          lc = cadd(lc, t);
        }
      }
      ls_set(out, j, k, lc);
    }
  }
  free(R);
}

/* LocalSet<3>::get_gradient (local_set3D.cpp:83-210) */
void orc_ref_l2p(uint32_t Lu, const double *Lc, const double d[3], double grad[3]) {
  int L = (int)Lu;
  double *R = (double *)malloc(sizeof(double) * (size_t)((L + 1) * (L + 1) + 4));
  orc_ref_powerset(Lu, d, R);
  grad[0] = grad[1] = grad[2] = 0.0;
  const cplx MI = cmk(0.0, -0.5);
  for (int j = 1; j <= L; ++j) {
    for (int k = -j; k <= j; ++k) {
      cplx Ljk = ls_get(Lc, j, k);
      cplx dx = cmk(0.0, 0.0), dy = cmk(0.0, 0.0), dz = cmk(0.0, 0.0);
      if (k >= 0) {
        if (abs(k - 1) <= j - 1)
          dx = csub(dx, cscale(0.5 * sqrt((double)((j + k) * (j + k - 1))), ps_get(R, j - 1, k - 1)));
        if (abs(k + 1) <= j - 1)
          dx = cadd(dx, cscale(0.5 * sqrt((double)((j - k) * (j - k - 1))), ps_get(R, j - 1, k + 1)));
      } else {
        int m = -k;
        cplx p = cmk(0.0, 0.0);
        if (abs(m - 1) <= j - 1)
          p = csub(p, cscale(0.5 * sqrt((double)((j + m) * (j + m - 1))), ps_get(R, j - 1, m - 1)));
        if (abs(m + 1) <= j - 1)
          p = cadd(p, cscale(0.5 * sqrt((double)((j - m) * (j - m - 1))), ps_get(R, j - 1, m + 1)));
        dx = cscale((m % 2) ? -1.0 : 1.0, cconj(p));
      }
      if (k >= 0) {
        if (abs(k - 1) <= j - 1)
          dy = cadd(dy, cmul(cscale(sqrt((double)((j + k) * (j + k - 1))), MI), ps_get(R, j - 1, k - 1)));
        if (abs(k + 1) <= j - 1)
          dy = cadd(dy, cmul(cscale(sqrt((double)((j - k) * (j - k - 1))), MI), ps_get(R, j - 1, k + 1)));
      } else {
        int m = -k;
        cplx p = cmk(0.0, 0.0);
        if (abs(m - 1) <= j - 1)
          p = cadd(p, cmul(cscale(sqrt((double)((j + m) * (j + m - 1))), MI), ps_get(R, j - 1, m - 1)));
        if (abs(m + 1) <= j - 1)
          p = cadd(p, cmul(cscale(sqrt((double)((j - m) * (j - m - 1))), MI), ps_get(R, j - 1, m + 1)));
        dy = cscale((m % 2) ? -1.0 : 1.0, cconj(p));
      }
      if (abs(k) <= j - 1)
        dz = cscale(sqrt((double)((j - abs(k)) * (j + abs(k)))), ps_get(R, j - 1, k));
      grad[0] += cmul(Ljk, dx).re;
      grad[1] += cmul(Ljk, dy).re;
      grad[2] -= cmul(Ljk, dz).re;
    }
  }
  free(R);
}

/* ------------------------------------------------------------------------ */
/* Reference FMM driver (tree.cpp:93-111 and the NodeI/Node/Leaf methods).   */

typedef struct {
  const orc_tree *t;
  const double *xyz, *m;
  int L;
  double *mc;   /* 3 per cell: mass centres (leaf.cpp:37-54, node.cpp:70-86) */
  double *M;    /* ms_size per cell */
  double *Lo;   /* ls_size per cell */
  char *has_local;
} refctx;

static void ref_upward(refctx *c) {
  const orc_tree *t = c->t;
  int L = c->L, msz = ms_size(L);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)msz);
  for (int lev = t->depth; lev >= 0; --lev) {
    for (int64_t cell = t->level_start[lev]; cell < t->level_start[lev + 1]; ++cell) {
      double *mc = c->mc + 3 * cell, *M = c->M + (size_t)msz * cell;
      memset(M, 0, sizeof(double) * (size_t)msz);
      if (lev == t->depth) {
        /* Leaf::calculateMC (leaf.cpp:37-54) + Leaf::compute_multipoles */
        double tot = 0;
        double acc[3] = {0, 0, 0};
        for (int64_t s = t->pbeg[cell]; s < t->pend[cell]; ++s) {
          int64_t i = t->perm[s];
          tot += c->m[i];
          for (int d = 0; d < 3; ++d) acc[d] += c->xyz[3 * i + d] * c->m[i];
        }
        for (int d = 0; d < 3; ++d) mc[d] = acc[d] / tot;
        for (int64_t s = t->pbeg[cell]; s < t->pend[cell]; ++s) {
          int64_t i = t->perm[s];
          double r[3] = {c->xyz[3 * i] - mc[0], c->xyz[3 * i + 1] - mc[1], c->xyz[3 * i + 2] - mc[2]};
          orc_ref_p2m_add((uint32_t)L, r, c->m[i], M);
        }
      } else {
        /* children of `cell` at lev+1 in key order (node.cpp:50-68) */
        int64_t ce = t->level_start[lev + 2];
        int64_t ch0 = orc_tree_first_child(t, lev, cell);
        /* Node::calculateMC (node.cpp:70-86) */
        double tot = 0;
        double acc[3] = {0, 0, 0};
        for (int64_t ch = ch0; ch < ce && (t->key[ch] >> 3) == t->key[cell]; ++ch) {
          double cm = c->M[(size_t)msz * ch];
          tot += cm;
          for (int d = 0; d < 3; ++d) acc[d] += c->mc[3 * ch + d] * cm;
        }
        for (int d = 0; d < 3; ++d) mc[d] = acc[d] / tot;
        /* Node::compute_multipoles (node.cpp:14-24) */
        for (int64_t ch = ch0; ch < ce && (t->key[ch] >> 3) == t->key[cell]; ++ch) {
          double dd[3] = {c->mc[3 * ch] - mc[0], c->mc[3 * ch + 1] - mc[1], c->mc[3 * ch + 2] - mc[2]};
          orc_ref_m2m((uint32_t)L, c->M + (size_t)msz * ch, dd, tmp);
          for (int q = 0; q < msz; ++q) M[q] += tmp[q];
        }
      }
    }
  }
  free(tmp);
}

/* NodeI::collect_multipoles_to_locals (node_i.cpp:105-120) */
static void ref_m2l_all(refctx *c) {
  const orc_tree *t = c->t;
  int L = c->L, msz = ms_size(L), lsz = ls_size(L);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)lsz);
  for (int64_t cell = 0; cell < t->ncells; ++cell) {
    double *Lo = c->Lo + (size_t)lsz * cell;
    c->has_local[cell] = 0;
    for (int64_t q = t->il_off[cell]; q < t->il_off[cell + 1]; ++q) {
      int64_t s = t->il[q];
      double d[3] = {c->mc[3 * cell] - c->mc[3 * s], c->mc[3 * cell + 1] - c->mc[3 * s + 1],
                     c->mc[3 * cell + 2] - c->mc[3 * s + 2]};
      if (!c->has_local[cell]) {
        orc_ref_m2l((uint32_t)L, c->M + (size_t)msz * s, d, Lo);
        c->has_local[cell] = 1;
      } else {
        orc_ref_m2l((uint32_t)L, c->M + (size_t)msz * s, d, tmp);
        for (int i = 0; i < lsz; ++i) Lo[i] += tmp[i];
      }
    }
  }
  free(tmp);
}

/* Node::propagate_locals (node.cpp:33-43) + NodeI::propagate_locals
 * (node_i.cpp:122-130), top-down; children in key order. */
static void ref_downward(refctx *c) {
  const orc_tree *t = c->t;
  int L = c->L, lsz = ls_size(L);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)lsz);
  for (int lev = 1; lev <= t->depth; ++lev) {
    for (int64_t cell = t->level_start[lev]; cell < t->level_start[lev + 1]; ++cell) {
      int64_t par = t->parent[cell];
      if (!c->has_local[par]) continue;
      double d[3] = {c->mc[3 * cell] - c->mc[3 * par], c->mc[3 * cell + 1] - c->mc[3 * par + 1],
                     c->mc[3 * cell + 2] - c->mc[3 * par + 2]};
      orc_ref_l2l((uint32_t)L, c->Lo + (size_t)lsz * par, d, tmp);
      double *Lo = c->Lo + (size_t)lsz * cell;
      if (c->has_local[cell]) {
        for (int i = 0; i < lsz; ++i) Lo[i] += tmp[i];
      } else {
        memcpy(Lo, tmp, sizeof(double) * (size_t)lsz);
        c->has_local[cell] = 1;
      }
    }
  }
  free(tmp);
}

/* P2P term of Leaf::compute_acceleration (leaf.cpp:119-142) */
static void ref_p2p_leaf(const refctx *c, int64_t i, int64_t src_cell, double *g) {
  const orc_tree *t = c->t;
  for (int64_t s = t->pbeg[src_cell]; s < t->pend[src_cell]; ++s) {
    int64_t j = t->perm[s];
    if (j == i) continue;
    double d[3] = {c->xyz[3 * i] - c->xyz[3 * j], c->xyz[3 * i + 1] - c->xyz[3 * j + 1],
                   c->xyz[3 * i + 2] - c->xyz[3 * j + 2]};
    double r2 = 0.0;
    r2 += d[0] * d[0];
    r2 += d[1] * d[1];
    r2 += d[2] * d[2];
    if (r2 == 0.0) continue; /* coincident pair: defined as no contribution */
    double pw = pow(r2, 1.5);
    for (int k = 0; k < 3; ++k) g[k] -= (d[k] * c->m[j]) / pw;
  }
}

int orc_fmm_refcompat(int64_t n, const double *xyz, const double *m, uint32_t L, uint32_t thr,
                      double G, double *acc) {
  if (n < 1 || L < 1) return -1;
  orc_tree *t = orc_tree_build(n, xyz, thr);
  if (!t) return -1;
  refctx c;
  c.t = t;
  c.xyz = xyz;
  c.m = m;
  c.L = (int)L;
  int64_t nc = t->ncells;
  c.mc = (double *)calloc((size_t)nc * 3, sizeof(double));
  c.M = (double *)calloc((size_t)nc * (size_t)ms_size((int)L), sizeof(double));
  c.Lo = (double *)calloc((size_t)nc * (size_t)ls_size((int)L), sizeof(double));
  c.has_local = (char *)calloc((size_t)nc, 1);
  ref_upward(&c);
  ref_m2l_all(&c);
  ref_downward(&c);
  int lsz = ls_size((int)L);
  int D = t->depth;
  for (int64_t cell = t->level_start[D]; cell < t->level_start[D + 1]; ++cell) {
    for (int64_t s = t->pbeg[cell]; s < t->pend[cell]; ++s) {
      int64_t i = t->perm[s];
      double g[3] = {0.0, 0.0, 0.0};
      if (c.has_local[cell]) {
        double d[3] = {xyz[3 * i] - c.mc[3 * cell], xyz[3 * i + 1] - c.mc[3 * cell + 1],
                       xyz[3 * i + 2] - c.mc[3 * cell + 2]};
        double gl[3];
        orc_ref_l2p(L, c.Lo + (size_t)lsz * cell, d, gl);
        for (int k = 0; k < 3; ++k) g[k] += gl[k];
      }
      ref_p2p_leaf(&c, i, cell, g);
      for (int64_t q = t->nbr_off[cell]; q < t->nbr_off[cell + 1]; ++q)
        ref_p2p_leaf(&c, i, t->nbr[q], g);
      for (int k = 0; k < 3; ++k) acc[3 * i + k] = g[k] * G;
    }
  }
  free(c.mc);
  free(c.M);
  free(c.Lo);
  free(c.has_local);
  orc_tree_free(t);
  return 0;
}
