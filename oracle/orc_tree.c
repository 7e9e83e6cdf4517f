/*
 * orc_tree.c -- oracle restatement of the reference tree rule.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Follows src/system_updaters/tree_components/tree.cpp:15-90 (bbox, depth
 * loop, leaves, ancestors), morton_code.cpp:9-36 (keys) and
 * node_i.cpp:23-99 (neighbour / interaction lists), with 64-bit keys where
 * the reference truncates to 32 bits (identical for leaf depth <= 9).
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* morton_code.cpp:9-22: bit k of coordinate i goes to bit 3k+i (x lowest). */
uint64_t orc_morton_encode3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t code = 0;
  uint32_t c[3] = {x, y, z};
  for (unsigned i = 0; i < 3; ++i) {
    unsigned digit = i;
    uint32_t v = c[i];
    while (v > 0) {
      code += (uint64_t)(v % 2u) << digit;
      v /= 2u;
      digit += 3;
    }
  }
  return code;
}

/* morton_code.cpp:24-36 */
void orc_morton_decode3(uint64_t key, uint32_t *x, uint32_t *y, uint32_t *z) {
  uint32_t c[3] = {0, 0, 0};
  unsigned digit = 0;
  while (key > 0) {
    for (unsigned i = 0; i < 3; ++i) {
      c[i] += (uint32_t)(key % 2u) << digit;
      key /= 2u;
    }
    ++digit;
  }
  *x = c[0];
  *y = c[1];
  *z = c[2];
}

static int bit_width_u64(uint64_t v) {
  int w = 0;
  while (v) {
    ++w;
    v >>= 1;
  }
  return w;
}

/* Stable LSD radix sort of (key, idx) by key, 8 bits per pass over `bits`. */
static void radix_sort_pairs(int64_t n, uint64_t *key, int64_t *idx, int bits) {
  uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  int64_t *i2 = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  for (int shift = 0; shift < bits; shift += 8) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; ++i) cnt[((key[i] >> shift) & 255u) + 1]++;
    for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
    for (int64_t i = 0; i < n; ++i) {
      int64_t dst = cnt[(key[i] >> shift) & 255u]++;
      k2[dst] = key[i];
      i2[dst] = idx[i];
    }
    memcpy(key, k2, sizeof(uint64_t) * (size_t)n);
    memcpy(idx, i2, sizeof(int64_t) * (size_t)n);
  }
  free(k2);
  free(i2);
}

/* tree.cpp:55-61 quantisation.  (x - a) / leaf_size truncated to unsigned and
 * clamped to 2^depth - 1.  A zero extent makes the quotient NaN, which the
 * reference converts with cvttsd2si to 0 on x86-64; the oracle defines it as 0. */
static uint32_t quantize(double x, double a, double ls, uint64_t nld) {
  double q = (x - a) / ls;
  uint32_t c;
  if (!(q == q) || q < 0.0)
    c = 0;
  else if (q >= 4294967295.0)
    c = 0xffffffffu;
  else
    c = (uint32_t)q;
  if ((uint64_t)c >= nld) c = (uint32_t)(nld - 1);
  return c;
}

static void compute_keys(int64_t n, const double *xyz, const double *a, const double *ls,
                         uint64_t nld, uint64_t *keys) {
  for (int64_t i = 0; i < n; ++i) {
    uint32_t c[3];
    for (int d = 0; d < 3; ++d) c[d] = quantize(xyz[3 * i + d], a[d], ls[d], nld);
    keys[i] = orc_morton_encode3(c[0], c[1], c[2]);
  }
}

static int64_t lower_bound_u64(const uint64_t *arr, int64_t n, uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (arr[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* Index of the cell with key `k` at level `lev`, or -1. */
int64_t orc_tree_find(const orc_tree *t, int lev, uint64_t k) {
  int64_t b = t->level_start[lev], e = t->level_start[lev + 1];
  int64_t pos = lower_bound_u64(t->key + b, e - b, k);
  if (pos < e - b && t->key[b + pos] == k) return b + pos;
  return -1;
}

/* First child (smallest key >= key<<3) of `cell` at level lev+1; children of
 * one parent are contiguous because every level is sorted by key. */
int64_t orc_tree_first_child(const orc_tree *t, int lev, int64_t cell) {
  int64_t b = t->level_start[lev + 1], e = t->level_start[lev + 2];
  return b + lower_bound_u64(t->key + b, e - b, t->key[cell] << 3);
}

/* node_i.cpp:23-99: for cell c at level lev, scan the 6x6x6 window of the
 * parent's neighbours' children, x fastest; Chebyshev distance <= 1 (self
 * excluded) -> neighbour list, otherwise -> interaction list. */
static void cell_lists(const orc_tree *t, int lev, int64_t c, int64_t *nbr, int64_t *nn,
                       int64_t *il, int64_t *ni) {
  uint32_t cc[3];
  orc_morton_decode3(t->key[c], &cc[0], &cc[1], &cc[2]);
  int st[3], en[3], dir[3];
  for (int i = 0; i < 3; ++i) {
    st[i] = -(int)(cc[i] % 2u) - 2;
    en[i] = 3 - (int)(cc[i] % 2u);
    dir[i] = st[i];
  }
  const int64_t side = (int64_t)1 << lev;
  *nn = 0;
  *ni = 0;
  for (;;) {
    int64_t q[3];
    int out = 0;
    for (int i = 0; i < 3; ++i) {
      q[i] = (int64_t)cc[i] + dir[i];
      if (q[i] < 0 || q[i] >= side) out = 1;
    }
    if (!out) {
      uint64_t k = orc_morton_encode3((uint32_t)q[0], (uint32_t)q[1], (uint32_t)q[2]);
      int is_nbr = 1;
      for (int i = 0; i < 3; ++i)
        if (dir[i] < -1 || dir[i] > 1) is_nbr = 0;
      int64_t f = orc_tree_find(t, lev, k);
      if (f >= 0) {
        if (is_nbr) {
          if (f != c) nbr[(*nn)++] = f;
        } else {
          il[(*ni)++] = f;
        }
      }
    }
    int ended = 1;
    for (int i = 0; i < 3 && ended; ++i) {
      if (dir[i] == en[i]) {
        dir[i] = st[i];
      } else {
        ++dir[i];
        ended = 0;
      }
    }
    if (ended) break;
  }
}

orc_tree *orc_tree_build(int64_t n, const double *xyz, uint32_t thr) { return orc_tree_build_box(n, xyz, thr, 0); }

/* cubic != 0: the bounding box is widened to the bounding cube centred on it
 * (the library's cubic_box option); 0: the reference's box (tree.cpp:20-30) */
orc_tree *orc_tree_build_box(int64_t n, const double *xyz, uint32_t thr, int cubic) {
  if (n < 1 || !xyz || thr < 1) return NULL;
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!isfinite(xyz[i])) return NULL;
  orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
  t->n = n;
  t->thr = thr;
  /* tree.cpp:20-30 */
  double a[3], b[3];
  for (int d = 0; d < 3; ++d) a[d] = b[d] = xyz[d];
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      double v = xyz[3 * i + d];
      if (v < a[d])
        a[d] = v;
      else if (v > b[d])
        b[d] = v;
    }
  if (cubic) {
    double ext = 0;
    for (int d = 0; d < 3; ++d)
      if (b[d] - a[d] > ext) ext = b[d] - a[d];
    for (int d = 0; d < 3; ++d) {
      const double c = 0.5 * (a[d] + b[d]);
      double lo = c - 0.5 * ext, hi = c + 0.5 * ext;
      if (lo > a[d]) lo = a[d];
      if (hi < b[d]) hi = b[d];
      a[d] = lo;
      b[d] = hi;
    }
  }
  memcpy(t->lo, a, sizeof(a));
  memcpy(t->hi, b, sizeof(b));

  /* tree.cpp:32-69 depth loop */
  const unsigned dim = 3;
  const unsigned max_depth = ((64 / dim) >> dim) << dim; /* 16 */
  uint32_t maxcnt = (uint32_t)n;
  unsigned depth = 0;
  double ls[3];
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  do {
    uint64_t nl = (uint64_t)((maxcnt - 1u) / thr + 1u);
    depth += (unsigned)bit_width_u64(nl) / dim + 1;
    if (depth > max_depth) depth = max_depth;
    uint64_t nld = (uint64_t)1 << depth;
    for (int d = 0; d < 3; ++d) ls[d] = (b[d] - a[d]) / (double)nld;
    compute_keys(n, xyz, a, ls, nld, keys);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    radix_sort_pairs(n, keys, idx, 3 * (int)depth);
    maxcnt = 0;
    int64_t run = 0;
    for (int64_t i = 0; i < n; ++i) {
      run = (i > 0 && keys[i] == keys[i - 1]) ? run + 1 : 1;
      if ((uint32_t)run > maxcnt) maxcnt = (uint32_t)run;
    }
    t->trace_depth[t->niter] = (int)depth;
    t->trace_maxcnt[t->niter] = maxcnt;
    t->niter++;
  } while (depth < max_depth && maxcnt > thr);
  t->depth = (int)depth;
  memcpy(t->leaf_size, ls, sizeof(ls));
  /* the last iteration's sort is stable over input order: leaves hold their
   * particles in input order, as Leaf does (tree.cpp:55-62, 71-74) */
  t->perm = idx;
  t->pkey = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) t->pkey[idx[i]] = keys[i];

  /* cells of every level: unique prefixes of the sorted leaf keys
   * (tree.cpp:71-84 creates exactly the ancestors of occupied leaves) */
  int D = t->depth;
  int64_t *count = (int64_t *)calloc((size_t)D + 1, sizeof(int64_t));
  for (int lev = 0; lev <= D; ++lev) {
    int sh = 3 * (D - lev);
    for (int64_t i = 0; i < n; ++i)
      if (i == 0 || (keys[i] >> sh) != (keys[i - 1] >> sh)) count[lev]++;
  }
  t->level_start = (int64_t *)malloc(sizeof(int64_t) * ((size_t)D + 2));
  t->level_start[0] = 0;
  for (int lev = 0; lev <= D; ++lev) t->level_start[lev + 1] = t->level_start[lev] + count[lev];
  int64_t nc = t->level_start[D + 1];
  t->ncells = nc;
  t->key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nc);
  t->pbeg = (int64_t *)malloc(sizeof(int64_t) * (size_t)nc);
  t->pend = (int64_t *)malloc(sizeof(int64_t) * (size_t)nc);
  t->parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)nc);
  for (int lev = 0; lev <= D; ++lev) {
    int sh = 3 * (D - lev);
    int64_t c = t->level_start[lev];
    for (int64_t i = 0; i < n; ++i) {
      if (i == 0 || (keys[i] >> sh) != (keys[i - 1] >> sh)) {
        t->key[c] = keys[i] >> sh;
        t->pbeg[c] = i;
        if (c > t->level_start[lev]) t->pend[c - 1] = i;
        ++c;
      }
    }
    t->pend[c - 1] = n;
  }
  for (int lev = 0; lev <= D; ++lev)
    for (int64_t c = t->level_start[lev]; c < t->level_start[lev + 1]; ++c)
      t->parent[c] = lev == 0 ? -1 : orc_tree_find(t, lev - 1, t->key[c] >> 3);
  free(count);
  free(keys);

  /* lists (tree.cpp:86-90 -> node_i.cpp:23-99) */
  t->nbr_off = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nc + 1));
  t->il_off = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nc + 1));
  int64_t cap_n = nc * 4 + 64, cap_i = nc * 16 + 64;
  t->nbr = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap_n);
  t->il = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap_i);
  int64_t tn = 0, ti = 0;
  int64_t bn[26], bi[216];
  for (int lev = 0; lev <= D; ++lev) {
    for (int64_t c = t->level_start[lev]; c < t->level_start[lev + 1]; ++c) {
      int64_t nn, ni;
      cell_lists(t, lev, c, bn, &nn, bi, &ni);
      t->nbr_off[c] = tn;
      t->il_off[c] = ti;
      if (tn + nn > cap_n) {
        cap_n = 2 * (tn + nn);
        t->nbr = (int64_t *)realloc(t->nbr, sizeof(int64_t) * (size_t)cap_n);
      }
      if (ti + ni > cap_i) {
        cap_i = 2 * (ti + ni);
        t->il = (int64_t *)realloc(t->il, sizeof(int64_t) * (size_t)cap_i);
      }
      memcpy(t->nbr + tn, bn, sizeof(int64_t) * (size_t)nn);
      memcpy(t->il + ti, bi, sizeof(int64_t) * (size_t)ni);
      tn += nn;
      ti += ni;
    }
  }
  t->nbr_off[nc] = tn;
  t->il_off[nc] = ti;
  return t;
}

void orc_tree_free(orc_tree *t) {
  if (!t) return;
  free(t->perm);
  free(t->pkey);
  free(t->level_start);
  free(t->key);
  free(t->pbeg);
  free(t->pend);
  free(t->parent);
  free(t->nbr_off);
  free(t->nbr);
  free(t->il_off);
  free(t->il);
  free(t);
}

int orc_tree_depth(const orc_tree *t) { return t->depth; }
int64_t orc_tree_ncells(const orc_tree *t) { return t->ncells; }
void orc_tree_geometry(const orc_tree *t, double *lo, double *hi, double *leaf_size) {
  memcpy(lo, t->lo, sizeof(t->lo));
  memcpy(hi, t->hi, sizeof(t->hi));
  memcpy(leaf_size, t->leaf_size, sizeof(t->leaf_size));
}
int orc_tree_depth_trace(const orc_tree *t, int *depths, uint32_t *maxcnts) {
  for (int i = 0; i < t->niter; ++i) {
    if (depths) depths[i] = t->trace_depth[i];
    if (maxcnts) maxcnts[i] = t->trace_maxcnt[i];
  }
  return t->niter;
}
void orc_tree_levels(const orc_tree *t, int64_t *level_start) {
  memcpy(level_start, t->level_start, sizeof(int64_t) * ((size_t)t->depth + 2));
}
void orc_tree_cells(const orc_tree *t, uint64_t *key, int64_t *pbeg, int64_t *pend,
                    int64_t *parent) {
  size_t nc = (size_t)t->ncells;
  if (key) memcpy(key, t->key, sizeof(uint64_t) * nc);
  if (pbeg) memcpy(pbeg, t->pbeg, sizeof(int64_t) * nc);
  if (pend) memcpy(pend, t->pend, sizeof(int64_t) * nc);
  if (parent) memcpy(parent, t->parent, sizeof(int64_t) * nc);
}
void orc_tree_perm(const orc_tree *t, int64_t *perm) {
  memcpy(perm, t->perm, sizeof(int64_t) * (size_t)t->n);
}
void orc_tree_particle_keys(const orc_tree *t, uint64_t *keys) {
  memcpy(keys, t->pkey, sizeof(uint64_t) * (size_t)t->n);
}
int64_t orc_tree_nbr_total(const orc_tree *t) { return t->nbr_off[t->ncells]; }
int64_t orc_tree_il_total(const orc_tree *t) { return t->il_off[t->ncells]; }
void orc_tree_lists(const orc_tree *t, int64_t *nbr_off, int64_t *nbr, int64_t *il_off,
                    int64_t *il) {
  size_t nc = (size_t)t->ncells;
  if (nbr_off) memcpy(nbr_off, t->nbr_off, sizeof(int64_t) * (nc + 1));
  if (nbr) memcpy(nbr, t->nbr, sizeof(int64_t) * (size_t)t->nbr_off[nc]);
  if (il_off) memcpy(il_off, t->il_off, sizeof(int64_t) * (nc + 1));
  if (il) memcpy(il, t->il, sizeof(int64_t) * (size_t)t->il_off[nc]);
}
