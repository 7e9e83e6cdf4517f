// ref_driver.cpp -- C-ABI entry points over the reference's own classes.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
// unmodified reference sources (read from /root/reference/src by
// oracle/Makefile) into oracle/_ref/libnbref.so, which the CPU tests use to
// pin the oracle restatement.  No reference source is copied; this file only
// calls the reference's public API (and reads the protected members that the
// reference's own tests read, through derived accessor types in the same way
// as tests/test_tree.cpp:14-40 does).
#include <cstdint>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <vector>
#include <algorithm>

#include "particle.hpp"
#include "quadratures/runge_kutta_4.hpp"
#include "quadratures/symplectic_euler.hpp"
#include "system_updaters/FMM.hpp"
#include "system_updaters/multipole_utils/irregular.hpp"
#include "system_updaters/multipole_utils/multipole_set.hpp"
#include "system_updaters/naive_update.hpp"
#include "system_updaters/tree_components/leaf.hpp"
#include "system_updaters/tree_components/tree.hpp"

namespace {

std::vector<Particle> make_particles(int64_t n, const double *xyz, const double *m,
                                     const double *vel) {
  std::vector<Particle> P;
  P.reserve((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    double vx = vel ? vel[3 * i] : 0.0, vy = vel ? vel[3 * i + 1] : 0.0, vz = vel ? vel[3 * i + 2] : 0.0;
    P.push_back(Particle(m[i], {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, {vx, vy, vz}));
  }
  return P;
}

struct TreeT : public Tree {
  template <class It>
  TreeT(It b, It e) : Tree(b, e) {}
  std::vector<std::unordered_map<u_int64_t, std::unique_ptr<NodeI>>> &levels() { return nodes_vector; }
};
struct LeafT : public Leaf {
  const std::vector<const Particle *> &parts() const { return particles; }
};
struct NodeIT : public NodeI {
  const std::vector<NodeI *> &nbrs() const { return neighbours_list; }
  const std::vector<NodeI *> &ils() const { return interaction_list; }
  uint64_t key() const { return id_child; }
};

struct TreeState {
  std::vector<Particle> P;
  std::unique_ptr<TreeT> tree;
};
TreeState g_tree;

void copy_acc(const std::vector<Particle> &P, double *acc) {
  for (size_t i = 0; i < P.size(); ++i)
    for (int d = 0; d < 3; ++d) acc[3 * i + d] = P[i].get_acceleration()[d];
}

}  // namespace

extern "C" {

double ref_grav_const() { return GRAV_CONST; }

// FMM::update (FMM.cpp:9-12): returns each particle's acceleration
int ref_fmm_update(int64_t n, const double *xyz, const double *m, double *acc) {
  auto P = make_particles(n, xyz, m, nullptr);
  FMM f;
  f.update(P);
  copy_acc(P, acc);
  return 0;
}

// NaiveUpdate::update (naive_update.cpp:9-35)
int ref_naive_update(int64_t n, const double *xyz, const double *m, double *acc) {
  auto P = make_particles(n, xyz, m, nullptr);
  NaiveUpdate u;
  u.update(P);
  copy_acc(P, acc);
  return 0;
}

// Runs `nsteps` integrator steps (kind 0 = SyplecticEuler, 1 = RungeKutta4)
// with updater 0 = NaiveUpdate, 1 = FMM.  With preupdate != 0 the updater
// runs once before the first step, as tests/test_naive_update.cpp:27 does,
// so the particles carry a(x0).  xyz/vel/acc are updated in place; acc
// returns each particle's carried acceleration after the last step.
int ref_integrate(int64_t n, double *xyz, double *vel, double *acc, const double *m, int kind,
                  int updater, double dt, int nsteps, int preupdate) {
  auto P = make_particles(n, xyz, m, vel);
  NaiveUpdate nu;
  FMM fm;
  SystemUpdateMethod *u = updater == 0 ? (SystemUpdateMethod *)&nu : (SystemUpdateMethod *)&fm;
  if (preupdate) u->update(P);
  SyplecticEuler se;
  RungeKutta4 rk;
  QuadratureMethod *q = kind == 0 ? (QuadratureMethod *)&se : (QuadratureMethod *)&rk;
  for (int s = 0; s < nsteps; ++s) q->integrate(u, P, dt);
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      xyz[3 * i + d] = P[i].get_position()[d];
      vel[3 * i + d] = P[i].get_velocity()[d];
      acc[3 * i + d] = P[i].get_acceleration()[d];
    }
  return 0;
}

// ---- tree inspection (the seams of tests/test_tree.cpp:14-40) -------------
int ref_tree_build(int64_t n, const double *xyz, const double *m) {
  g_tree.tree.reset();
  g_tree.P = make_particles(n, xyz, m, nullptr);
  g_tree.tree = std::make_unique<TreeT>(g_tree.P.begin(), g_tree.P.end());
  return (int)g_tree.tree->levels().size() - 1;
}
void ref_tree_free() {
  g_tree.tree.reset();
  g_tree.P.clear();
}
int64_t ref_tree_level_count(int lev) { return (int64_t)g_tree.tree->levels()[lev].size(); }
void ref_tree_level_keys(int lev, uint64_t *keys) {
  std::vector<uint64_t> k;
  for (auto &kv : g_tree.tree->levels()[lev]) k.push_back(kv.first);
  std::sort(k.begin(), k.end());
  std::copy(k.begin(), k.end(), keys);
}
static NodeI *node_at(int lev, uint64_t key) {
  auto &mp = g_tree.tree->levels()[lev];
  auto it = mp.find(key);
  return it == mp.end() ? nullptr : it->second.get();
}
// particle indices of the leaf `key` at the leaf level, in leaf order
int64_t ref_tree_leaf_particles(uint64_t key, int64_t *idx) {
  int D = (int)g_tree.tree->levels().size() - 1;
  NodeI *nd = node_at(D, key);
  if (!nd) return -1;
  const auto &ps = ((LeafT *)nd)->parts();
  if (idx)
    for (size_t i = 0; i < ps.size(); ++i) idx[i] = ps[i] - g_tree.P.data();
  return (int64_t)ps.size();
}
// list keys of node `key` at level lev in list order; returns counts
void ref_tree_lists(int lev, uint64_t key, int64_t *nn, int64_t *ni, uint64_t *nbr_keys,
                    uint64_t *il_keys) {
  NodeI *nd = node_at(lev, key);
  const auto &nb = ((NodeIT *)nd)->nbrs();
  const auto &il = ((NodeIT *)nd)->ils();
  *nn = (int64_t)nb.size();
  *ni = (int64_t)il.size();
  if (nbr_keys)
    for (size_t i = 0; i < nb.size(); ++i) nbr_keys[i] = ((NodeIT *)nb[i])->key();
  if (il_keys)
    for (size_t i = 0; i < il.size(); ++i) il_keys[i] = ((NodeIT *)il[i])->key();
}

// ---- operators at any order L --------------------------------------------
void ref_powerset(unsigned L, const double *r, double *out) {
  PowerSet<3> z(L, Tensor{r[0], r[1], r[2]});
  const auto &e = z.get_elements();
  std::copy(e.begin(), e.end(), out);
}
void ref_irregular(unsigned L, const double *r, double *out) {
  Irregular z(L, Tensor{r[0], r[1], r[2]});
  const auto &e = z.get_elements();
  std::copy(e.begin(), e.end(), out);
}
// P2M of n particles at offsets rel (n x 3) with charges q (leaf.cpp:15-35)
void ref_p2m(unsigned L, int64_t n, const double *rel, const double *q, double *M) {
  auto ms = std::make_unique<MultipoleSet<3>>(L);
  *ms = 0;
  for (int64_t i = 0; i < n; ++i) {
    PowerSet<3> z(L, Tensor{rel[3 * i], rel[3 * i + 1], rel[3 * i + 2]});
    z *= q[i];
    *ms += z;
  }
  const auto &e = ms->get_elements();
  std::copy(e.begin(), e.end(), M);
}
static std::unique_ptr<MultipoleSet<3>> ms_from(unsigned L, const double *M) {
  auto ms = std::make_unique<MultipoleSet<3>>(L);
  ms->set_elements(std::vector<double>(M, M + 2 * L * L + 4 * L + 1));
  return ms;
}
static std::unique_ptr<LocalSet<3>> ls_from(unsigned L, const double *Lc) {
  auto ls = std::make_unique<LocalSet<3>>(L);
  ls->BaseMultipoleSetI::set_elements(std::vector<double>(Lc, Lc + 2 * (L + 1) * (L + 1)));
  return ls;
}
void ref_m2m(unsigned L, const double *M, const double *d, double *out) {
  auto r = ms_from(L, M)->weigh_children_with_distance(Tensor{d[0], d[1], d[2]});
  const auto &e = r->get_elements();
  std::copy(e.begin(), e.end(), out);
}
void ref_m2l(unsigned L, const double *M, const double *d, double *out) {
  auto r = ms_from(L, M)->to_local(Tensor{d[0], d[1], d[2]});
  const auto &e = r->get_elements();
  std::copy(e.begin(), e.end(), out);
}
void ref_l2l(unsigned L, const double *Lc, const double *d, double *out) {
  auto r = ls_from(L, Lc)->distribute_parent_with_distance(Tensor{d[0], d[1], d[2]});
  const auto &e = r->get_elements();
  std::copy(e.begin(), e.end(), out);
}
void ref_l2p(unsigned L, const double *Lc, const double *d, double *grad) {
  Tensor g = ls_from(L, Lc)->get_gradient(Tensor{d[0], d[1], d[2]});
  for (int i = 0; i < 3; ++i) grad[i] = g[i];
}

}  // extern "C"
