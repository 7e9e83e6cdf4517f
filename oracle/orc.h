/*
 * orc.h -- CPU oracle for the 3D gravitational FMM force evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  It may be loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * else.  The product (libnbfmm.so) never links, loads or calls it.
 *
 * It is a plain-C (C11, FP64) restatement of the algorithm in
 * rezucch1/nbody-FMM.  Every function cites the reference file:line it
 * follows.  Two operator sets are provided:
 *   - "refcompat": the reference's own solid-harmonic operators
 *     (PowerSet<3>, Irregular, MultipoleSet<3>, LocalSet<3>), restated loop
 *     for loop, including their conventions, with mass centres as expansion
 *     centres.  Used to prove parity with the reference itself.
 *   - "exact": correct operators (Yoshida-normalised complex solid
 *     harmonics) with geometric or mass centres, any order p.  Used to prove
 *     convergence to the direct sum and as the checker of the GPU path.
 *
 * Parity pins (what checks each part; the tests under tests/ run them):
 *   Morton encode/decode ........ closed form + brute force + reference build
 *   bbox / depth loop / leaves .. reference test_tree golden (glibc rand) +
 *                                 reference build (oracle/_ref) at depth <= 9
 *   neighbour / interaction lists brute force over occupied coordinates +
 *                                 reference build
 *   refcompat operators ......... reference build at several orders L
 *   exact harmonics ............. associated-Legendre closed forms, addition
 *                                 theorem 1/|x-y| = sum conj(R(y)) S(x)
 *   exact P2M/M2M ............... M2M(P2M(child)) == P2M(parent) identity
 *   exact M2L/L2L/L2P ........... convergence to the direct sum as p grows,
 *                                 L2L composition, finite differences
 *   direct sum .................. two-body closed form, reference build
 *   integrators ................. closed-form one step + reference build
 */
#ifndef NBFMM_ORACLE_H
#define NBFMM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Morton keys (morton_code.cpp:9-36), 64-bit correct ---------------- */
uint64_t orc_morton_encode3(uint32_t x, uint32_t y, uint32_t z);
void orc_morton_decode3(uint64_t key, uint32_t *x, uint32_t *y, uint32_t *z);

/* ---- Tree (tree.cpp:15-90, node_i.cpp:23-99) --------------------------- */
typedef struct orc_tree orc_tree;

/* Builds the reference's uniform-depth sparse octree.  xyz is AoS (n x 3).
 * thr = leaf threshold (reference hardcodes 20, tree.cpp:32).
 * Returns NULL on bad input (n < 1, non-finite coordinates). */
orc_tree *orc_tree_build(int64_t n, const double *xyz, uint32_t thr);
orc_tree *orc_tree_build_box(int64_t n, const double *xyz, uint32_t thr, int cubic);
void orc_tree_free(orc_tree *t);
int orc_tree_depth(const orc_tree *t);
int64_t orc_tree_ncells(const orc_tree *t);
/* bbox lo/hi and leaf-level cell size (3 each) */
void orc_tree_geometry(const orc_tree *t, double *lo, double *hi, double *leaf_size);
/* depth-loop trace: number of iterations, depth and max count per iteration */
int orc_tree_depth_trace(const orc_tree *t, int *depths, uint32_t *maxcnts);
/* level_start has depth+2 entries: cells of level d are [level_start[d], level_start[d+1]) */
void orc_tree_levels(const orc_tree *t, int64_t *level_start);
/* per cell: key, particle range [pbeg,pend) in sorted order, parent index (-1 for root) */
void orc_tree_cells(const orc_tree *t, uint64_t *key, int64_t *pbeg, int64_t *pend, int64_t *parent);
/* sorted position -> original particle index (stable within a leaf = input order) */
void orc_tree_perm(const orc_tree *t, int64_t *perm);
/* per-particle leaf key at the final depth */
void orc_tree_particle_keys(const orc_tree *t, uint64_t *keys);
/* CSR lists in the reference's enumeration order (x fastest over the 6x6x6 window). */
int64_t orc_tree_nbr_total(const orc_tree *t);
int64_t orc_tree_il_total(const orc_tree *t);
void orc_tree_lists(const orc_tree *t, int64_t *nbr_off, int64_t *nbr, int64_t *il_off, int64_t *il);

/* ---- Reference-compatible operators (multipole_utils: 3D sets, irregular) */
/* Regular solid harmonics PowerSet<3> (power_set3D.cpp:10-47): (L+1)^2 doubles. */
void orc_ref_powerset(uint32_t L, const double r[3], double *out);
/* Irregular<3> (irregular.cpp:10-41): (L+1)^2 doubles. */
void orc_ref_irregular(uint32_t L, const double r[3], double *out);
/* P2M accumulate q*PowerSet into a MultipoleSet<3> (2L^2+4L+1 doubles)
 * (leaf.cpp:15-35, multipole_set3D.cpp:31-43). */
void orc_ref_p2m_add(uint32_t L, const double r[3], double q, double *M);
/* MultipoleSet<3>::weigh_children_with_distance (multipole_set3D.cpp:45-86) */
void orc_ref_m2m(uint32_t L, const double *M, const double d[3], double *out);
/* MultipoleSet<3>::to_local (multipole_set3D.cpp:88-118): out has 2(L+1)^2 doubles */
void orc_ref_m2l(uint32_t L, const double *M, const double d[3], double *out);
/* LocalSet<3>::distribute_parent_with_distance (local_set3D.cpp:51-81) */
void orc_ref_l2l(uint32_t L, const double *Lc, const double d[3], double *out);
/* LocalSet<3>::get_gradient (local_set3D.cpp:83-210) */
void orc_ref_l2p(uint32_t L, const double *Lc, const double d[3], double grad[3]);

/* Reference FMM::update (FMM.cpp:9-12, tree.cpp:93-111) with the reference
 * operators and mass centres; L and thr are parameters (reference: 2 / 20).
 * acc = G * grad (particle.cpp:16-21).  Returns 0 on success. */
int orc_fmm_refcompat(int64_t n, const double *xyz, const double *m, uint32_t L, uint32_t thr,
                      double G, double *acc);

/* ---- Exact operators (Yoshida normalisation, complex, m >= 0 stored) ---- */
/* R_n^m, S_n^m for 0<=m<=n<=p, complex interleaved: index (n(n+1)/2+m)*2 */
void orc_exact_regular(uint32_t p, const double r[3], double *out);
void orc_exact_singular(uint32_t p, const double r[3], double *out);
/* P2M add: M_n^m += q conj(R_n^m(r)) */
void orc_exact_p2m_add(uint32_t p, const double r[3], double q, double *M);
/* M2M: shift multipole from centre c to c' ; d = c - c' */
void orc_exact_m2m(uint32_t p, const double *M, const double d[3], double *out);
/* M2L: D = c_local - c_multipole */
void orc_exact_m2l(uint32_t p, const double *M, const double D[3], double *out);
/* L2L: d = c_child - c_parent */
void orc_exact_l2l(uint32_t p, const double *Lc, const double d[3], double *out);
/* L2P potential and gradient at t = x - c */
void orc_exact_l2p(uint32_t p, const double *Lc, const double t[3], double *phi, double grad[3]);
/* M2P potential and gradient at x - c (far field of a multipole) */
void orc_exact_m2p(uint32_t p, const double *M, const double t[3], double *phi, double grad[3]);

/* Exact FMM on the reference tree.  center_mode 0 = geometric cell centres,
 * 1 = mass centres.  An interaction-list pair (t, s) is evaluated by M2L, or
 * by direct P2P between the two cells when n_t * n_s < p2p_cut (0 = always
 * M2L, as the reference does).  P2P excludes coincident pairs (r^2 == 0).
 * acc = G * grad.  Returns 0 on success. */
int orc_fmm_exact(int64_t n, const double *xyz, const double *m, uint32_t p, uint32_t thr,
                  double G, int center_mode, int64_t p2p_cut, double *acc);
/* the same on the cubic box of orc_tree_build_box(..., cubic = 1) when cubic != 0 */
int orc_fmm_exact_box(int64_t n, const double *xyz, const double *m, uint32_t p, uint32_t thr, double G,
                      int center_mode, int64_t p2p_cut, int cubic, double *acc);
/* with distance-dependent M2L orders (class_order_tol > 0; 0 = p everywhere) */
int orc_fmm_exact_opt(int64_t n, const double *xyz, const double *m, uint32_t p, uint32_t thr, double G,
                      int center_mode, int64_t p2p_cut, int cubic, double class_order_tol, double *acc);
/* the order of each of the 343 offset classes (index (dz+3)*49 + (dy+3)*7 + (dx+3)) */
void orc_class_orders(const double ext[3], int p, double tol, int *q);

/* ---- Direct sums (naive_update.cpp:9-35) ------------------------------- */
/* NaiveUpdate formula: grad -= m_j d / ((r^2 + eps2) * r), acc = G * grad.
 * Self excluded by index, as the reference excludes by address. */
void orc_direct_naive(int64_t n, const double *xyz, const double *m, double G, double eps2,
                      double *acc);
/* Metric oracle: unsoftened when eps2 == 0, per-component Neumaier-compensated
 * sums, evaluated for the targets listed (indices into xyz); coincident pairs
 * skipped.  acc has 3*ntgt entries. */
void orc_direct_targets(int64_t n, const double *xyz, const double *m, double G, double eps2,
                        int64_t ntgt, const int64_t *tgt, double *acc);

/* ---- Integrators (symplectic_euler.cpp:8-15, runge_kutta_4.cpp:10-93) --- */
enum { ORC_UPD_NAIVE = 0, ORC_UPD_FMM_REF = 1, ORC_UPD_FMM_EXACT = 2 };
typedef struct {
  int updater;          /* ORC_UPD_* */
  uint32_t order;       /* L or p */
  uint32_t thr;         /* leaf threshold */
  double G, eps2;       /* eps2 used by ORC_UPD_NAIVE */
  int center_mode;      /* for ORC_UPD_FMM_EXACT */
  int64_t p2p_cut;      /* for ORC_UPD_FMM_EXACT */
} orc_upd_cfg;
/* acc holds the particles' carried acceleration (in/out), like Particle::acceleration. */
int orc_step_se(int64_t n, double *xyz, double *vel, double *acc, const double *m,
                const orc_upd_cfg *cfg, double dt);
/* rk4_mode 0 = reference semantics (stage velocities from the carried
 * acceleration, runge_kutta_4.cpp:41-79; carried acceleration unchanged),
 * 1 = classic RK4 (stage slopes; carried acceleration set to a(x_n)). */
int orc_step_rk4(int64_t n, double *xyz, double *vel, double *acc, const double *m,
                 const orc_upd_cfg *cfg, double dt, int rk4_mode);

#ifdef __cplusplus
}
#endif
#endif
