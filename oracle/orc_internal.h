/* orc_internal.h -- private definitions shared by the oracle's C files.
 * TEST INFRASTRUCTURE ONLY (see orc.h). */
#ifndef NBFMM_ORC_INTERNAL_H
#define NBFMM_ORC_INTERNAL_H
#include "orc.h"

struct orc_tree {
  int64_t n;
  uint32_t thr;
  int depth;
  double lo[3], hi[3], leaf_size[3];
  int niter;
  int trace_depth[32];
  uint32_t trace_maxcnt[32];
  int64_t *perm;        /* sorted position -> input index */
  uint64_t *pkey;       /* input index -> leaf key */
  int64_t *level_start; /* depth+2 */
  int64_t ncells;
  uint64_t *key;
  int64_t *pbeg, *pend, *parent;
  int64_t *nbr_off, *nbr, *il_off, *il;
};

/* the value of M_SQRT2 in <math.h> (used by power_set3D.cpp:14 and irregular.cpp:15) */
#define ORC_SQRT2 1.41421356237309504880

int64_t orc_tree_find(const orc_tree *t, int lev, uint64_t k);
int64_t orc_tree_first_child(const orc_tree *t, int lev, int64_t cell);

/* complex helpers with the reference's operation order (std::complex<double>
 * multiplication (ac-bd) + i(ad+bc) for finite operands) */
typedef struct {
  double re, im;
} cplx;
static inline cplx cmk(double re, double im) {
  cplx c = {re, im};
  return c;
}
static inline cplx cmul(cplx a, cplx b) {
  return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cplx cscale(double s, cplx a) { return cmk(s * a.re, s * a.im); }
static inline cplx cconj(cplx a) { return cmk(a.re, -a.im); }

#endif
