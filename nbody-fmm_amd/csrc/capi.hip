// capi.hip -- the extern "C" boundary of libnbfmm.so (declared in include/nbfmm.h).
#include "nbfmm_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>

using namespace nbfmm;

namespace {
thread_local std::string g_err;

int fail(int status, const std::string &msg) {
  g_err = msg;
  return status;
}

#define API_BEGIN try {
#define API_END                                        \
  }                                                    \
  catch (const Error &e) {                             \
    return fail(e.status, e.what());                   \
  }                                                    \
  catch (const std::exception &e) {                    \
    return fail(NBFMM_ERR_HIP, e.what());              \
  }                                                    \
  return NBFMM_OK;

int check_params(const nbfmm_params *p) {
  if (!p) return fail(NBFMM_ERR_ARG, "params is null");
  if (p->order < 1 || p->order > 14) return fail(NBFMM_ERR_ARG, "order must be in [1, 14]");
  if (p->leaf_capacity < 1) return fail(NBFMM_ERR_ARG, "leaf_capacity must be >= 1");
  if (!std::isfinite(p->G)) return fail(NBFMM_ERR_ARG, "G must be finite");
  if (!(p->eps2 >= 0.0) || !std::isfinite(p->eps2)) return fail(NBFMM_ERR_ARG, "eps2 must be finite and >= 0");
  if (p->p2p_cut < 0) return fail(NBFMM_ERR_ARG, "p2p_cut must be >= 0");
  if (p->m2l_pass_pairs < 0) return fail(NBFMM_ERR_ARG, "m2l_pass_pairs must be >= 0");
  if (p->precision != NBFMM_PLAIN_FP64 && p->precision != NBFMM_MIXED)
    return fail(NBFMM_ERR_ARG, "unknown precision");
  if (p->cubic_box != 0 && p->cubic_box != 1) return fail(NBFMM_ERR_ARG, "cubic_box must be 0 or 1");
  if (!(p->class_order_tol >= 0.0 && p->class_order_tol <= 1.0))
    return fail(NBFMM_ERR_ARG, "class_order_tol must be in [0, 1]");
  if (p->class_order_tol > 0.0 && (p->order < (uint32_t)kOrderLo || p->order > (uint32_t)kOrderHi))
    return fail(NBFMM_ERR_ARG, "class orders need 4 <= order <= 14");
  if (p->mixed_fp64_levels < 0) return fail(NBFMM_ERR_ARG, "mixed_fp64_levels must be >= 0");
  if (p->order_select != NBFMM_ORDER_FIXED && p->order_select != NBFMM_ORDER_LEARNED)
    return fail(NBFMM_ERR_ARG, "unknown order_select");
  if (p->order_select == NBFMM_ORDER_LEARNED) {
    if (p->order_min < kOrderLo || p->order_min > (int32_t)p->order)
      return fail(NBFMM_ERR_ARG, "order_min must be in [4, order]");
    if (p->order < (uint32_t)kOrderLo) return fail(NBFMM_ERR_ARG, "learned order selection needs order >= 4");
    if (!(p->order_conf > 0.0 && p->order_conf <= 1.0)) return fail(NBFMM_ERR_ARG, "order_conf must be in (0, 1]");
  }
  return NBFMM_OK;
}

void eval_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, double *d_acc,
              double *d_phi = nullptr) {
  c->plan.n = n;
  c->phi_out = d_phi;
  c->nev = 0;
  const bool planned = c->cap.on;
  if (!planned) {
    // a direct evaluation: its sizes set the capacities of later planned ones
    c->last = PlanSizes{};
    fill_i32(c->stream, &dplan(c)->ovf, 1, 0);
  }
  mark(c, 0);
  build_tree(c, d_xyz, d_m);
  mark(c, 1);
  try {
    build_lists(c);
    mark(c, 2);
    run_expansions(c); // marks 3, 4, 5 (7, 8 inside); starts P2P on the second stream (marks 9, 10)
  } catch (...) {
    (void)hipStreamSynchronize(c->stream2); // the particle gather or P2P still reads buffers
    c->phi_out = nullptr;
    throw;
  }
  run_finish(c, d_acc);
  c->phi_out = nullptr;
  mark(c, 6);
  if (planned) return; // the sizes stay on the device
  Plan &P = c->plan;
  {
    PlanSizes &L = c->last;
    L.depth = P.geo.depth;
    L.n = n;
    L.nc = P.n_cells;
    L.parents = P.lvl_start[P.geo.depth];
    for (int d = 0; d <= P.geo.depth; ++d) L.lev[d] = P.lvl_start[d + 1] - P.lvl_start[d];
    L.n_m2l = P.n_m2l;
    L.n_sub = P.n_sub;
    L.n_tgt = P.n_tgt;
    L.ntask = P.n_tasks;
    L.nnarrow = P.n_tasks_narrow;
  }
  nbfmm_stats &S = c->stats;
  std::memset(&S, 0, sizeof(S));
  S.depth = P.geo.depth;
  S.depth_iterations = P.iterations;
  S.depth_sorts = P.depth_sorts;
  S.n_cells = P.n_cells;
  S.n_leaves = P.lvl_start[P.geo.depth + 1] - P.lvl_start[P.geo.depth];
  S.n_il = P.n_il;
  S.n_m2l = P.n_m2l;
  S.n_m2l_targets = P.n_tgt;
  S.n_m2l_sources = P.n_src;
  S.n_p2p_pairs = P.n_p2p_pairs;
  S.n_p2m_terms = P.n_p2m_terms;
  S.n_l2p_terms = P.n_l2p_terms;
  S.n_p2m_cells = P.n_p2m_cells;
  S.n_m2m_cells = P.n_m2m_cells;
  S.n_l2l_cells = P.n_l2l_cells;
  S.n_l2p_cells = P.n_l2p_cells;
  S.isolated = c->isolated ? 1 : 0;
  if (c->timing) {
    NB_HIP(hipEventSynchronize(c->ev[6]));
    auto span = [&](int a, int b) {
      float t = 0;
      NB_HIP(hipEventElapsedTime(&t, c->ev[a], c->ev[b]));
      return (double)t;
    };
    // event order on the main stream: 0 tree 1 lists 2 up-pass 14 M2M 3 M2L
    // 13 L2L 4 L2P 5 finish 6
    S.ms_tree = span(0, 1);
    S.ms_lists = span(1, 2);
    S.ms_p2m = span(2, 14);
    S.ms_m2m = span(14, 3);
    S.ms_m2l = span(3, 13);
    S.ms_l2l = span(13, 4);
    S.ms_l2p = span(4, 5);
    S.ms_p2p = span(5, 6);
    float tot = 0;
    NB_HIP(hipEventElapsedTime(&tot, c->ev[0], c->ev[6]));
    S.ms_total = tot;
    float km = 0, kp = 0;
    if (P.n_tgt > 0) NB_HIP(hipEventElapsedTime(&km, c->ev[7], c->ev[8]));
    float kp2 = 0;
    NB_HIP(hipEventElapsedTime(&kp, c->ev[9], c->ev[10]));
    NB_HIP(hipEventElapsedTime(&kp2, c->ev[11], c->ev[12]));
    kp += kp2;
    S.ms_m2l_kernel = km;
    S.ms_p2p_kernel = kp;
  }
  S.flop_m2l = P.flop_m2l;
  S.flop_p2p = 20.0 * (double)P.n_p2p_pairs;
  S.mean_order = P.mean_order;
}

// the evaluation with the parameters `q` in place of the context's
void eval_with(nbfmm_ctx *c, const nbfmm_params &q, int64_t n, const double *x, const double *m, double *a) {
  const nbfmm_params save = c->prm;
  c->prm = q;
  try {
    eval_dev(c, n, x, m, a);
  } catch (...) {
    c->prm = save;
    throw;
  }
  c->prm = save;
}

// accelerations a(x) of the particle set being integrated
using Eval = std::function<void(const double *x, double *a)>;

// SyplecticEuler::integrate (symplectic_euler.cpp:8-15), one step
void step_se(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const Eval &ev, double dt) {
  ev(x, a);
  se_update(c, n, x, v, a, dt);
}
void step_se(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const double *m, double dt) {
  step_se(c, n, x, v, a, [&](const double *xx, double *aa) { eval_dev(c, n, xx, m, aa); }, dt);
}

// RungeKutta4::integrate (runge_kutta_4.cpp:10-93), one step; mode 0 the
// reference's semantics, 1 classic
void step_rk4(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const Eval &ev, double dt, int mode) {
  const size_t m3 = (size_t)n * 3;
  double *xk = dget<double>(c->xk, m3), *vk = dget<double>(c->vk, m3), *ak = dget<double>(c->ak, m3);
  double *asum = dget<double>(c->asum, m3), *vsum = dget<double>(c->vsum, m3);
  const double h2 = dt / 2, h6 = dt / 6;
  copy_d(c, vk, v, (int64_t)m3);         // velocity_k = v      (runge_kutta_4.cpp:28-30)
  ev(x, ak);                             // stage 1              (:32)
  rk4_accum(c, n, asum, vsum, ak, vk, 1.0, 1);
  if (mode == 1) copy_d(c, a, ak, (int64_t)m3);
  for (int stage = 2; stage <= 4; ++stage) {
    const double h = stage == 4 ? dt : h2;
    const double *avel = mode == 0 ? a : ak;  // reference: carried acceleration of the input (:46-48)
    rk4_stage(c, n, x, v, vk, avel, h, xk, vk);
    ev(xk, ak);
    rk4_accum(c, n, asum, vsum, ak, vk, stage == 4 ? 1.0 : 2.0, 0);
  }
  rk4_final(c, n, x, v, asum, vsum, h6);  // (:89-92)
}
void step_rk4(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const double *m, double dt, int mode) {
  step_rk4(c, n, x, v, a, [&](const double *xx, double *aa) { eval_dev(c, n, xx, m, aa); }, dt, mode);
}

// the batched evaluation of nsys systems (offsets on the device)
Eval batch_evaluator(nbfmm_ctx *c, int nsys, const int64_t *d_off, int64_t n, const double *m) {
  return [=](const double *x, double *a) {
    batch_eval(c, nsys, d_off, n, x, m, a,
               [c](int64_t nb, const double *xb, const double *mb, double *ab) { eval_dev(c, nb, xb, mb, ab); });
  };
}

// ---- integration steps, direct or replayed from a captured graph ----------------------
// kind: 0 SyplecticEuler, 1 RungeKutta4 with the reference's semantics, 2 classic RK4
struct StepSpec {
  int kind;
  int64_t n;
  double *x, *v, *a;
  const double *m;
  double dt;
  int nsys;             // 0: one system
  const int64_t *d_off; // a batch's offsets (device)
};

// one step; the plan's overflow flag is cleared first (a planned step's
// kernels leave at once once it is set, and the integrator does not update)
void step_one(nbfmm_ctx *c, const StepSpec &sp, const Eval &ev) {
  fill_i32(c->stream, &dplan(c)->ovf, 1, 0);
  if (sp.kind == 0)
    step_se(c, sp.n, sp.x, sp.v, sp.a, ev, sp.dt);
  else
    step_rk4(c, sp.n, sp.x, sp.v, sp.a, ev, sp.dt, sp.kind == 1 ? 0 : 1);
}

bool same_key(const nbfmm_ctx::GraphKey &k, const StepSpec &sp, const nbfmm_ctx *c) {
  return k.kind == sp.kind && k.n == sp.n && k.x == sp.x && k.v == sp.v && k.a == sp.a && k.m == sp.m &&
         k.dt == sp.dt && k.nsys == sp.nsys && k.off == sp.d_off && std::memcmp(&k.prm, &c->prm, sizeof(k.prm)) == 0 &&
         std::memcmp(&k.shapes, &c->shapes, sizeof(k.shapes)) == 0 && k.gen == dbuf_generation();
}

void drop_graph(nbfmm_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->gkey = nbfmm_ctx::GraphKey{};
}

// Captures one step as a graph: the capacities come from the last direct
// evaluation (a quarter more); a planned evaluation outside the capture
// allocates every buffer at them, so the capture allocates nothing.
void capture_step(nbfmm_ctx *c, const StepSpec &sp, const Eval &ev) {
  drop_graph(c);
  if (c->isolated) throw Error(NBFMM_ERR_UNSUPPORTED, "graph steps do not run with isolated timing");
  if (c->prm.order_select != NBFMM_ORDER_FIXED) throw Error(NBFMM_ERR_UNSUPPORTED, "graph steps take a fixed order");
  if (c->keep_lists) throw Error(NBFMM_ERR_UNSUPPORTED, "graph steps keep no lists");
  if (c->last.depth < 0) throw Error(NBFMM_ERR_HIP, "no direct evaluation to plan the graph step from");
  const bool timing = c->timing;
  struct Restore {
    nbfmm_ctx *c;
    bool timing;
    ~Restore() {
      c->cap.on = false;
      c->timing = timing;
    }
  } restore{c, timing};
  c->cap = caps_from(c->last);
  c->cap.on = true;
  c->timing = false;
  // the allocation pass: the step's own scratch and one planned evaluation
  const size_t m3 = (size_t)sp.n * 3;
  double *scratch = dget<double>(c->ak, m3);
  if (sp.kind != 0) {
    dget<double>(c->xk, m3);
    dget<double>(c->vk, m3);
    dget<double>(c->asum, m3);
    dget<double>(c->vsum, m3);
  }
  fill_i32(c->stream, &dplan(c)->ovf, 1, 0);
  ev(sp.x, scratch);
  NB_HIP(hipStreamSynchronize(c->stream));
  const uint64_t gen = dbuf_generation();
  hipGraph_t g = nullptr;
  set_capturing(true);
  NB_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    step_one(c, sp, ev);
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    set_capturing(false);
    throw;
  }
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  set_capturing(false);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    throw Error(NBFMM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    c->gexec = nullptr;
    throw Error(NBFMM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  if (dbuf_generation() != gen) { // nothing may move between the allocation pass and the replays
    drop_graph(c);
    throw Error(NBFMM_ERR_HIP, "a buffer grew during the graph capture");
  }
  nbfmm_ctx::GraphKey &k = c->gkey;
  k.kind = sp.kind;
  k.n = sp.n;
  k.x = sp.x;
  k.v = sp.v;
  k.a = sp.a;
  k.m = sp.m;
  k.dt = sp.dt;
  k.nsys = sp.nsys;
  k.off = sp.d_off;
  k.prm = c->prm;
  k.shapes = c->shapes;
  k.gen = gen;
  c->g_captures++;
}

// nsteps steps; with graphs on, a step replays the captured graph of the
// same step (same arrays, sizes, dt and parameters), or runs directly and
// captures one for the next step (of this call or a later one).  A replay whose planned sizes overflowed left
// the state unchanged: the step runs again directly (and is captured anew).
void run_steps(nbfmm_ctx *c, const StepSpec &sp, int nsteps, const Eval &ev,
               const std::function<void(int)> &after = nullptr) {
  for (int st = 1; st <= nsteps; ++st) {
    if (!c->graph_on) {
      step_one(c, sp, ev);
    } else if (c->gexec && same_key(c->gkey, sp, c)) {
      NB_HIP(hipGraphLaunch(c->gexec, c->stream));
      c->g_launches++;
      read_plan_field(c, &DevPlan::ovf, c->stream);
      if (c->hplan->ovf) {
        c->g_redone++;
        drop_graph(c);
        step_one(c, sp, ev);
        c->g_direct++;
        capture_step(c, sp, ev); // at this step's sizes
      }
    } else {
      // the graph is kept across calls: one step per call replays it too
      step_one(c, sp, ev);
      c->g_direct++;
      capture_step(c, sp, ev);
    }
    if (after) after(st);
  }
}

// the one-system step of `kind` (0 SE, 1 RK4 reference semantics, 2 classic RK4)
StepSpec one_system(int kind, int64_t n, double *x, double *v, double *a, const double *m, double dt) {
  return StepSpec{kind, n, x, v, a, m, dt, 0, nullptr};
}
Eval one_system_eval(nbfmm_ctx *c, int64_t n, const double *m) {
  return [c, n, m](const double *xx, double *aa) { eval_dev(c, n, xx, m, aa); };
}

void check_batch(int nsys, const int64_t *off) {
  if (nsys < 1 || !off) throw Error(NBFMM_ERR_ARG, "nsys >= 1 and offsets are required");
  if (off[0] != 0) throw Error(NBFMM_ERR_ARG, "offsets[0] must be 0");
  for (int k = 0; k < nsys; ++k)
    if (off[k + 1] <= off[k]) throw Error(NBFMM_ERR_ARG, "every system needs at least one particle");
  if (off[nsys] > ((int64_t)1 << 30) - 2) throw Error(NBFMM_ERR_ARG, "too many particles");
}

void check_n(int64_t n) {
  if (n < 1) throw Error(NBFMM_ERR_ARG, "n must be >= 1");
  if (n > (int64_t)1 << 30) throw Error(NBFMM_ERR_ARG, "n must be < 2^30");
}

} // namespace

namespace nbfmm {
namespace {
__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (((uintptr_t)p & 15) == 0) { // four at a time where aligned
    int4 *q = reinterpret_cast<int4 *>(p);
    const int64_t n4 = n / 4;
    for (int64_t j = i; j < n4; j += stride) q[j] = make_int4(v, v, v, v);
    for (int64_t j = 4 * n4 + i; j < n; j += stride) p[j] = v;
  } else {
    for (; i < n; i += stride) p[i] = v;
  }
}
std::atomic<uint64_t> g_dbuf_gen{0}; // (re)allocations of any context's buffers
thread_local bool g_capturing = false;
} // namespace
uint64_t dbuf_generation() { return g_dbuf_gen.load(); }
void fill_i32(hipStream_t s, void *p, int64_t count, int32_t v) {
  if (count <= 0) return;
  const int64_t nb = std::min<int64_t>(4096, (count / 4 + 255) / 256 + 1);
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)nb), dim3(256), 0, s, (int32_t *)p, count, v);
  NB_LAUNCH();
}
void set_capturing(bool on) { g_capturing = on; }
void *dbuf_ensure(DBuf &b, size_t bytes) {
  if (b.cap >= bytes && b.p) return b.p;
  // a captured graph holds the addresses: nothing may move while capturing
  if (g_capturing) throw Error(NBFMM_ERR_HIP, "buffer growth during a graph capture");
  g_dbuf_gen++;
  if (b.p) NB_HIP(hipFree(b.p));
  b.p = nullptr;
  b.cap = 0;
  size_t want = bytes + bytes / 4 + 256;
  NB_HIP(hipMalloc(&b.p, want));
  b.cap = want;
  return b.p;
}
void dbuf_free(DBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}
void read_plan(nbfmm_ctx *c, const void *field, size_t bytes, hipStream_t s) {
  const size_t off = (size_t)((const char *)field - (const char *)c->hplan);
  NB_HIP(hipMemcpyAsync((char *)c->hplan + off, (const char *)c->dplan.p + off, bytes, hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
}
Caps caps_from(const PlanSizes &s) {
  auto up = [](int64_t v) { return v + v / 4 + 64; };
  Caps K;
  K.depth = s.depth;
  K.nc = up(s.nc);
  K.parents = 0;
  for (int d = 0; d <= kMaxDepth; ++d) {
    K.lev[d] = d <= s.depth ? (int32_t)up(s.lev[d]) : 0;
    if (d < s.depth) K.parents += K.lev[d];
    K.m2m_lev[d] = (int32_t)up(s.m2m_lev[d]);
    K.l2l_lev[d] = (int32_t)up(s.l2l_lev[d]);
  }
  K.n_m2l = up(s.n_m2l);
  K.n_sub = up(s.n_sub);
  K.n_tgt = up(s.n_tgt);
  K.ntask = up(s.ntask);
  K.nnarrow = up(s.nnarrow);
  K.nM = up(s.nM);
  K.nL = up(s.nL);
  K.nP2M = up(s.nP2M);
  K.items = up(s.items) + kMaxBins;
  for (int oi = 0; oi < kNumOrders; ++oi) K.items_ord[oi] = (int32_t)up(s.items_ord[oi]);
  return K;
}
void mark(nbfmm_ctx *c, int slot) {
  if (c->timing) NB_HIP(hipEventRecord(c->ev[slot], c->stream));
}
void mark2(nbfmm_ctx *c, int slot) {
  if (c->timing) NB_HIP(hipEventRecord(c->ev[slot], c->stream2));
}
} // namespace nbfmm

nbfmm_ctx::~nbfmm_ctx() {
  if (stream2) (void)hipStreamSynchronize(stream2);
  if (stream) (void)hipStreamSynchronize(stream);
  DBuf *all[] = {&xyz_in, &m_in, &acc_out, &key[0], &key[1], &idx[0], &idx[1], &sort_hist, &scan_tmp, &act_par, &pos,
                 &pcell, &grad, &grad_near, &red, &ckey, &cbeg, &cend, &cparent, &child8, &nbr27, &flags, &fscan, &closed, &cnt_il,
                 &cnt_m2l, &cnt_sub, &off_il, &off_m2l, &off_sub, &near_pairs, &il, &m2l_src, &m2l_cls, &sub_src,
                 &src_slot, &tgt_slot, &src_cells, &tgt_cells, &is_src, &sub_pairs, &tasks, &task_off, &task_cnt, &dpart, &p2p_ctr,
                 &tiles, &stab, &tmat, &tmat32, &mexp64, &lexp64, &trans, &trans32, &mode_m, &mode_l, &mslot, &lslot, &eslot, &up_cells, &mexp, &lexp, &consts, &ypair, &cls_pairs, &cls_cnt, &m2l_items, &m2l_tab, &ops_key, &xk, &vk, &ak, &asum, &vsum, &diag_acc, &diag_phi,
                 &ord, &msum, &feat, &rmax, &mlp_w, &mlp_wb, &mlp_m, &mlp_v, &mlp_act, &mlp_perm, &ds_feat, &ds_label, &lab_tgt,
                 &lab_dir, &lab_sum, &lab_flag, &lab_lab, &chk_acc, &bat_off, &bat_bb, &bat_st, &bat_x, &bat_m,
                 &bat_a, &dplan};
  for (DBuf *b : all) dbuf_free(*b);
  if (hred.p) (void)hipHostFree(hred.p);
  if (hplan) (void)hipHostFree(hplan);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_fork2) (void)hipEventDestroy(ev_fork2);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (ev_pos) (void)hipEventDestroy(ev_pos);
  if (stream2) (void)hipStreamDestroy(stream2);
  if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

void nbfmm_default_params(nbfmm_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->order = 2;
  p->leaf_capacity = 20;
  p->G = 6.6743e-11;
  p->eps2 = 1.0e-4;
  p->precision = NBFMM_PLAIN_FP64;
  p->p2p_cut = 0;
  p->order_select = NBFMM_ORDER_FIXED;
  p->order_min = 4;
  p->order_conf = 0.9;
  p->mixed_fp64_levels = 0;
}

const char *nbfmm_last_error(void) { return g_err.c_str(); }

int nbfmm_create(const nbfmm_params *params, nbfmm_ctx **out) {
  if (!out) return fail(NBFMM_ERR_ARG, "ctx is null");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  nbfmm_ctx *c = nullptr;
  API_BEGIN
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(NBFMM_ERR_NO_DEVICE, "no HIP device");
  c = new nbfmm_ctx();
  c->prm = *params;
  NB_HIP(hipGetDevice(&c->device));
  NB_HIP(hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  NB_HIP(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device));
  NB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  NB_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  NB_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  NB_HIP(hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming));
  NB_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  NB_HIP(hipEventCreateWithFlags(&c->ev_pos, hipEventDisableTiming));
  for (auto &e : c->ev) NB_HIP(hipEventCreate(&e));
  scratch_init(c->device); // this device's copy of the fixed scratch (once per device)
  NB_HIP(hipHostMalloc(&c->hred.p, 65536, hipHostMallocDefault));
  c->hred.cap = 65536;
  dbuf_ensure(c->dplan, sizeof(DevPlan));
  NB_HIP(hipMemset(c->dplan.p, 0, sizeof(DevPlan)));
  NB_HIP(hipHostMalloc((void **)&c->hplan, sizeof(DevPlan), hipHostMallocDefault));
  std::memset((void *)c->hplan, 0, sizeof(DevPlan));
  upload_consts(c);
  *out = c;
  c = nullptr;
  }
  catch (const Error &e) {
    delete c;
    return fail(e.status, e.what());
  }
  return NBFMM_OK;
}

int nbfmm_destroy(nbfmm_ctx *c) {
  if (!c) return NBFMM_OK;
  API_BEGIN
  (void)hipStreamSynchronize(c->stream);
  delete c;
  API_END
}

int nbfmm_set_params(nbfmm_ctx *c, const nbfmm_params *params) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  int rc = check_params(params);
  if (rc) return rc;
  c->prm = *params;
  return NBFMM_OK;
}

void *nbfmm_stream(nbfmm_ctx *c) { return c ? (void *)c->stream : nullptr; }

int nbfmm_set_tiles(nbfmm_ctx *c, const nbfmm_tiles *t) {
  if (!c || !t) return fail(NBFMM_ERR_ARG, "null argument");
  const nbfmm_tiles d = default_tiles();
  nbfmm_tiles r = *t;
  if (r.p2p_group_f32 == 0) r.p2p_group_f32 = d.p2p_group_f32;
  if (r.p2p_group_f64 == 0) r.p2p_group_f64 = d.p2p_group_f64;
  if (r.p2p_narrow_max_f32 == 0) r.p2p_narrow_max_f32 = d.p2p_narrow_max_f32;
  if (r.p2p_narrow_max_f64 == 0) r.p2p_narrow_max_f64 = d.p2p_narrow_max_f64;
  if (r.p2p_early_share_f32 < 0.0) r.p2p_early_share_f32 = d.p2p_early_share_f32;
  if (r.p2p_early_share_f64 < 0.0) r.p2p_early_share_f64 = d.p2p_early_share_f64;
  if (r.p2p_early_launches_f32 == 0) r.p2p_early_launches_f32 = d.p2p_early_launches_f32;
  if (r.p2p_early_launches_f64 == 0) r.p2p_early_launches_f64 = d.p2p_early_launches_f64;
  if (r.m2l_slices == 0) r.m2l_slices = d.m2l_slices;
  if (r.p2p_queue_waves0 == 0) r.p2p_queue_waves0 = d.p2p_queue_waves0;
  if (r.p2p_queue_chunk == 0) r.p2p_queue_chunk = d.p2p_queue_chunk;
  if (r.m2l_item_pairs == 0) r.m2l_item_pairs = d.m2l_item_pairs;
  if (r.p2p_rest_launches == 0) r.p2p_rest_launches = d.p2p_rest_launches;
  if (r.m2l_mode == 0) r.m2l_mode = d.m2l_mode;
  std::string why;
  if (!tiles_ok(r, why)) return fail(NBFMM_ERR_ARG, why);
  c->shapes = r;
  return NBFMM_OK;
}

int nbfmm_get_tiles(nbfmm_ctx *c, nbfmm_tiles *t) {
  if (!c || !t) return fail(NBFMM_ERR_ARG, "null argument");
  *t = c->shapes;
  return NBFMM_OK;
}

int nbfmm_set_timing(nbfmm_ctx *c, int enabled) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  if (enabled < 0 || enabled > 2) return fail(NBFMM_ERR_ARG, "timing mode must be 0, 1 or 2");
  c->timing = enabled != 0;
  c->isolated = enabled == 2;
  return NBFMM_OK;
}

int nbfmm_keep_lists(nbfmm_ctx *c, int enabled) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  c->keep_lists = enabled != 0;
  return NBFMM_OK;
}

int nbfmm_eval_accel_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, double *d_acc) {
  if (!c || !d_xyz || !d_m || !d_acc) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  check_n(n);
  eval_dev(c, n, d_xyz, d_m, d_acc);
  API_END
}

int nbfmm_eval_accel(nbfmm_ctx *c, int64_t n, const double *xyz, const double *m, double *acc) {
  if (!c || !xyz || !m || !acc) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  check_n(n);
  double *dx = dget<double>(c->xyz_in, (size_t)n * 3), *dm = dget<double>(c->m_in, (size_t)n);
  double *da = dget<double>(c->acc_out, (size_t)n * 3);
  NB_HIP(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  NB_HIP(hipMemcpyAsync(dm, m, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  eval_dev(c, n, dx, dm, da);
  NB_HIP(hipMemcpyAsync(acc, da, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  API_END
}

int nbfmm_direct_accel_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, int64_t ntgt,
                           const int64_t *d_tgt, int mode, double *d_acc) {
  if (!c || !d_xyz || !d_m || !d_acc) return fail(NBFMM_ERR_ARG, "null argument");
  if (mode != 0 && mode != 1) return fail(NBFMM_ERR_ARG, "mode must be 0 or 1");
  API_BEGIN
  check_n(n);
  if (ntgt < 1) throw Error(NBFMM_ERR_ARG, "ntgt must be >= 1");
  run_direct(c, n, d_xyz, d_m, ntgt, d_tgt, mode, d_acc);
  API_END
}

int nbfmm_direct_accel_noalloc_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, int64_t ntgt,
                                   const int64_t *d_tgt, int mode, double *d_acc) {
  if (!c || !d_xyz || !d_m || !d_acc) return fail(NBFMM_ERR_ARG, "null argument");
  if (mode != 0 && mode != 1) return fail(NBFMM_ERR_ARG, "mode must be 0 or 1");
  API_BEGIN
  check_n(n);
  if (ntgt < 1) throw Error(NBFMM_ERR_ARG, "ntgt must be >= 1");
  // the same kernels as nbfmm_direct_accel_dev; the partial sums go to this
  // device's fixed scratch, so nothing is allocated
  run_direct(c, n, d_xyz, d_m, ntgt, d_tgt, mode, d_acc, scratch_ptr(c->device));
  API_END
}

int nbfmm_direct_accel(nbfmm_ctx *c, int64_t n, const double *xyz, const double *m, double *acc) {
  if (!c || !xyz || !m || !acc) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  check_n(n);
  double *dx = dget<double>(c->xyz_in, (size_t)n * 3), *dm = dget<double>(c->m_in, (size_t)n);
  double *da = dget<double>(c->acc_out, (size_t)n * 3);
  NB_HIP(hipMemcpyAsync(dx, xyz, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  NB_HIP(hipMemcpyAsync(dm, m, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  run_direct(c, n, dx, dm, n, nullptr, 0, da);
  NB_HIP(hipMemcpyAsync(acc, da, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  API_END
}

int nbfmm_step_se_dev(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const double *m, double dt,
                      int nsteps) {
  if (!c || !x || !v || !a || !m) return fail(NBFMM_ERR_ARG, "null argument");
  if (nsteps < 0) return fail(NBFMM_ERR_ARG, "nsteps must be >= 0");
  API_BEGIN
  check_n(n);
  run_steps(c, one_system(0, n, x, v, a, m, dt), nsteps, one_system_eval(c, n, m));
  API_END
}

int nbfmm_step_rk4_dev(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const double *m, double dt,
                       int nsteps, int mode) {
  if (!c || !x || !v || !a || !m) return fail(NBFMM_ERR_ARG, "null argument");
  if (nsteps < 0) return fail(NBFMM_ERR_ARG, "nsteps must be >= 0");
  if (mode != 0 && mode != 1) return fail(NBFMM_ERR_ARG, "mode must be 0 or 1");
  API_BEGIN
  check_n(n);
  run_steps(c, one_system(mode == 0 ? 1 : 2, n, x, v, a, m, dt), nsteps, one_system_eval(c, n, m));
  API_END
}

int nbfmm_eval_accel_pot_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, double *d_acc,
                             double *d_phi) {
  if (!c || !d_xyz || !d_m || !d_acc || !d_phi) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  check_n(n);
  eval_dev(c, n, d_xyz, d_m, d_acc, d_phi);
  API_END
}

int nbfmm_diagnostics_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_vel, const double *d_m,
                          const double *d_phi, double *out) {
  if (!c || !d_xyz || !d_vel || !d_m || !out) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  check_n(n);
  diagnostics(c, n, d_xyz, d_vel, d_m, d_phi, out);
  API_END
}

int nbfmm_run_dev(nbfmm_ctx *c, int64_t n, double *x, double *v, double *a, const double *m, double dt, int nsteps,
                  int integrator, int every, double *h_diag) {
  if (!c || !x || !v || !a || !m) return fail(NBFMM_ERR_ARG, "null argument");
  if (nsteps < 0) return fail(NBFMM_ERR_ARG, "nsteps must be >= 0");
  if (integrator < 0 || integrator > 2) return fail(NBFMM_ERR_ARG, "integrator must be 0, 1 or 2");
  if (h_diag && every < 1) return fail(NBFMM_ERR_ARG, "every must be >= 1");
  API_BEGIN
  check_n(n);
  // the monitor evaluates into its own buffers, so the state the integrator
  // carries (the acceleration in d_acc) is not touched
  double *da = h_diag ? dget<double>(c->diag_acc, (size_t)n * 3) : nullptr;
  double *dp = h_diag ? dget<double>(c->diag_phi, (size_t)n) : nullptr;
  int rec = 0;
  auto monitor = [&]() {
    eval_dev(c, n, x, m, da, dp);
    diagnostics(c, n, x, v, m, dp, h_diag + 10 * rec++);
  };
  if (h_diag) monitor();
  run_steps(c, one_system(integrator, n, x, v, a, m, dt), nsteps, one_system_eval(c, n, m), [&](int s) {
    if (h_diag && s % every == 0) monitor();
  });
  NB_HIP(hipStreamSynchronize(c->stream));
  API_END
}

int nbfmm_order_model_init(nbfmm_ctx *c, uint64_t seed) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  API_BEGIN
  mlp_init(c, seed);
  API_END
}

int nbfmm_order_model_set(nbfmm_ctx *c, const float *w, int64_t nw) {
  if (!c || !w) return fail(NBFMM_ERR_ARG, "null argument");
  if (nw != kMlpWeights) return fail(NBFMM_ERR_ARG, "the weight vector has NBFMM_MLP_WEIGHTS entries");
  API_BEGIN
  for (int64_t i = 0; i < nw; ++i)
    if (!std::isfinite(w[i])) throw Error(NBFMM_ERR_ARG, "non-finite weight");
  for (int i = 0; i < kMlpIn; ++i)
    if (!(w[kMlpParams + kMlpIn + i] > 0.0f)) throw Error(NBFMM_ERR_ARG, "feature deviations must be > 0");
  mlp_set(c, w);
  API_END
}

int nbfmm_order_model_get(nbfmm_ctx *c, float *w, int64_t nw) {
  if (!c || !w) return fail(NBFMM_ERR_ARG, "null argument");
  if (nw < kMlpWeights) return fail(NBFMM_ERR_ARG, "the weight vector has NBFMM_MLP_WEIGHTS entries");
  API_BEGIN
  mlp_get(c, w);
  API_END
}

int nbfmm_order_dataset_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, int64_t stride,
                            uint64_t seed, double tol, int64_t *added) {
  if (!c || !d_xyz || !d_m) return fail(NBFMM_ERR_ARG, "null argument");
  if (stride < 1) return fail(NBFMM_ERR_ARG, "stride must be >= 1");
  if (!(tol > 0.0)) return fail(NBFMM_ERR_ARG, "tol must be > 0");
  API_BEGIN
  check_n(n);
  double *acc = dget<double>(c->chk_acc, (size_t)n * 3);
  label_begin(c, n, stride, seed, d_xyz, d_m);
  nbfmm_params q = c->prm;
  q.order_select = NBFMM_ORDER_FIXED;
  for (int qi = 0; qi < kNumOrders; ++qi) {
    q.order = (uint32_t)(kOrderLo + qi);
    eval_with(c, q, n, d_xyz, d_m, acc);
    label_add(c, qi, acc);
  }
  const int64_t a = label_finish(c, tol);
  if (added) *added = a;
  API_END
}

int nbfmm_order_dataset_get(nbfmm_ctx *c, float *feat, int8_t *label, int64_t cap, int64_t *n) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  API_BEGIN
  if (n) *n = c->ds_n;
  if ((feat || label) && cap > 0) dataset_get(c, feat, label, cap);
  API_END
}

int nbfmm_order_dataset_put(nbfmm_ctx *c, int64_t n, const float *feat, const int8_t *label) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  API_BEGIN
  if (n == 0 && !feat && !label) {
    c->ds_n = 0;
  } else {
    if (n < 1 || !feat || !label) throw Error(NBFMM_ERR_ARG, "n >= 1 rows with features and labels");
    for (int64_t i = 0; i < n; ++i)
      if (label[i] < 0 || label[i] >= kNumOrders) throw Error(NBFMM_ERR_ARG, "labels are order - 4 in [0, 10]");
    dataset_put(c, n, feat, label);
  }
  API_END
}

int nbfmm_order_train(nbfmm_ctx *c, int32_t epochs, int32_t batch, double lr, uint64_t seed, float *loss) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  if (epochs < 0 || batch < 1 || !(lr > 0.0)) return fail(NBFMM_ERR_ARG, "epochs >= 0, batch >= 1, lr > 0");
  API_BEGIN
  mlp_train(c, epochs, batch, lr, seed, loss);
  API_END
}

int nbfmm_order_grad(nbfmm_ctx *c, int32_t nb, const float *feat, const int8_t *label, float *grad, float *loss) {
  if (!c || !feat || !label || !grad || !loss) return fail(NBFMM_ERR_ARG, "null argument");
  if (nb < 1) return fail(NBFMM_ERR_ARG, "nb must be >= 1");
  API_BEGIN
  for (int i = 0; i < nb; ++i)
    if (label[i] < 0 || label[i] >= kNumOrders) throw Error(NBFMM_ERR_ARG, "labels are order - 4 in [0, 10]");
  mlp_grad(c, nb, feat, label, grad, loss);
  API_END
}

int nbfmm_order_predict(nbfmm_ctx *c, int64_t nb, const float *feat, float *prob) {
  if (!c || !feat || !prob) return fail(NBFMM_ERR_ARG, "null argument");
  if (nb < 1) return fail(NBFMM_ERR_ARG, "nb must be >= 1");
  API_BEGIN
  mlp_predict(c, nb, feat, prob);
  API_END
}

int nbfmm_eval_accel_checked_dev(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, double *d_acc,
                                 int64_t ntgt, uint64_t seed, double gate, int32_t fallback_order, double *err,
                                 int32_t *fell_back) {
  if (!c || !d_xyz || !d_m || !d_acc) return fail(NBFMM_ERR_ARG, "null argument");
  if (ntgt < 1 || !(gate > 0.0)) return fail(NBFMM_ERR_ARG, "ntgt >= 1 and gate > 0");
  if (fallback_order < 1 || fallback_order > 14) return fail(NBFMM_ERR_ARG, "fallback_order must be in [1, 14]");
  API_BEGIN
  check_n(n);
  eval_dev(c, n, d_xyz, d_m, d_acc);
  double e = sampled_error(c, n, d_xyz, d_m, d_acc, ntgt, seed);
  int fb = 0;
  if (!(e <= gate)) {
    nbfmm_params q = c->prm;
    q.order_select = NBFMM_ORDER_FIXED;
    q.order = (uint32_t)fallback_order;
    eval_with(c, q, n, d_xyz, d_m, d_acc);
    e = sampled_error(c, n, d_xyz, d_m, d_acc, ntgt, seed);
    fb = 1;
  }
  c->stats.fallback = fb;
  if (err) *err = e;
  if (fell_back) *fell_back = fb;
  API_END
}

int nbfmm_batch_eval_accel_dev(nbfmm_ctx *c, int32_t nsys, const int64_t *offsets, const double *d_xyz,
                               const double *d_m, double *d_acc) {
  if (!c || !d_xyz || !d_m || !d_acc) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  check_batch(nsys, offsets);
  const int64_t *d_off = upload_offsets(c, nsys, offsets);
  batch_evaluator(c, nsys, d_off, offsets[nsys], d_m)(d_xyz, d_acc);
  API_END
}

int nbfmm_batch_step_se_dev(nbfmm_ctx *c, int32_t nsys, const int64_t *offsets, double *x, double *v, double *a,
                            const double *m, double dt, int nsteps) {
  if (!c || !x || !v || !a || !m) return fail(NBFMM_ERR_ARG, "null argument");
  if (nsteps < 0) return fail(NBFMM_ERR_ARG, "nsteps must be >= 0");
  API_BEGIN
  check_batch(nsys, offsets);
  const int64_t *d_off = upload_offsets(c, nsys, offsets);
  const int64_t n = offsets[nsys];
  run_steps(c, StepSpec{0, n, x, v, a, m, dt, nsys, d_off}, nsteps, batch_evaluator(c, nsys, d_off, n, m));
  API_END
}

int nbfmm_batch_step_rk4_dev(nbfmm_ctx *c, int32_t nsys, const int64_t *offsets, double *x, double *v, double *a,
                             const double *m, double dt, int nsteps, int mode) {
  if (!c || !x || !v || !a || !m) return fail(NBFMM_ERR_ARG, "null argument");
  if (nsteps < 0) return fail(NBFMM_ERR_ARG, "nsteps must be >= 0");
  if (mode != 0 && mode != 1) return fail(NBFMM_ERR_ARG, "mode must be 0 or 1");
  API_BEGIN
  check_batch(nsys, offsets);
  const int64_t *d_off = upload_offsets(c, nsys, offsets);
  const int64_t n = offsets[nsys];
  run_steps(c, StepSpec{mode == 0 ? 1 : 2, n, x, v, a, m, dt, nsys, d_off}, nsteps,
            batch_evaluator(c, nsys, d_off, n, m));
  API_END
}

int nbfmm_set_graph(nbfmm_ctx *c, int enabled) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  if (enabled != 0 && enabled != 1) return fail(NBFMM_ERR_ARG, "enabled must be 0 or 1");
  API_BEGIN
  if (enabled && c->prm.order_select != NBFMM_ORDER_FIXED)
    throw Error(NBFMM_ERR_UNSUPPORTED, "graph steps take a fixed order (order_select = NBFMM_ORDER_FIXED)");
  if (enabled && c->prm.m2l_pass_pairs > 0)
    throw Error(NBFMM_ERR_UNSUPPORTED, "graph steps run the M2L in one pass (m2l_pass_pairs = 0)");
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  c->graph_on = enabled != 0;
  API_END
}

int nbfmm_graph_stats(nbfmm_ctx *c, int64_t *out4) {
  if (!c || !out4) return fail(NBFMM_ERR_ARG, "null argument");
  out4[0] = c->g_captures;
  out4[1] = c->g_launches;
  out4[2] = c->g_redone;
  out4[3] = c->g_direct;
  return NBFMM_OK;
}

int nbfmm_get_stats(nbfmm_ctx *c, nbfmm_stats *s) {
  if (!c || !s) return fail(NBFMM_ERR_ARG, "null argument");
  API_BEGIN
  Plan &P = c->plan;
  if (P.m2l_stats_dev) {
    // the target-major M2L of the last direct evaluation copies its flops and
    // order sum to the pinned plan mirror (in stream order, no wait)
    NB_HIP(hipStreamSynchronize(c->stream));
    P.flop_m2l = c->hplan->flop_m2l;
    P.mean_order = P.n_m2l > 0 ? c->hplan->order_sum / (double)P.n_m2l : P.mean_order;
    P.m2l_stats_dev = false;
    c->stats.flop_m2l = P.flop_m2l;
    c->stats.mean_order = P.mean_order;
  }
  *s = c->stats;
  API_END
}

int nbfmm_tree_sizes(nbfmm_ctx *c, int32_t *depth, int64_t *n_cells, int64_t *n_nbr, int64_t *n_il) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  API_BEGIN
  const Plan &P = c->plan;
  if (P.n_cells == 0) throw Error(NBFMM_ERR_ARG, "no evaluation has run");
  if (depth) *depth = P.geo.depth;
  if (n_cells) *n_cells = P.n_cells;
  if (n_nbr) {
    nbr_leaf_rows(c, nullptr); // every leaf's row
    NB_HIP(hipStreamSynchronize(c->stream));
    std::vector<int32_t> nb((size_t)P.n_cells * 27);
    NB_HIP(hipMemcpy(nb.data(), c->nbr27.p, nb.size() * 4, hipMemcpyDeviceToHost));
    int64_t k = 0;
    for (size_t i = 0; i < nb.size(); ++i)
      if (i % 27 != 13 && nb[i] >= 0) ++k;
    *n_nbr = k;
  }
  if (n_il) *n_il = P.n_il;
  API_END
}

int nbfmm_tree_get(nbfmm_ctx *c, int64_t *level_start, uint64_t *key, int32_t *pbeg, int32_t *pend,
                   int32_t *parent, int32_t *perm, int32_t *nbr_off, int32_t *nbr, int32_t *il_off, int32_t *il) {
  if (!c) return fail(NBFMM_ERR_ARG, "ctx is null");
  API_BEGIN
  const Plan &P = c->plan;
  if (P.n_cells == 0) throw Error(NBFMM_ERR_ARG, "no evaluation has run");
  NB_HIP(hipStreamSynchronize(c->stream));
  const size_t nc = (size_t)P.n_cells;
  if (level_start)
    for (int d = 0; d <= P.geo.depth + 1; ++d) level_start[d] = P.lvl_start[d];
  if (key) NB_HIP(hipMemcpy(key, c->ckey.p, nc * 8, hipMemcpyDeviceToHost));
  if (pbeg) NB_HIP(hipMemcpy(pbeg, c->cbeg.p, nc * 4, hipMemcpyDeviceToHost));
  if (pend) NB_HIP(hipMemcpy(pend, c->cend.p, nc * 4, hipMemcpyDeviceToHost));
  if (parent) NB_HIP(hipMemcpy(parent, c->cparent.p, nc * 4, hipMemcpyDeviceToHost));
  if (perm) NB_HIP(hipMemcpy(perm, c->idx[0].p, (size_t)P.n * 4, hipMemcpyDeviceToHost));
  if (nbr_off || nbr) {
    nbr_leaf_rows(c, nullptr); // every leaf's row
    NB_HIP(hipStreamSynchronize(c->stream));
    std::vector<int32_t> nb(nc * 27);
    NB_HIP(hipMemcpy(nb.data(), c->nbr27.p, nb.size() * 4, hipMemcpyDeviceToHost));
    int32_t k = 0;
    for (size_t cc = 0; cc < nc; ++cc) {
      if (nbr_off) nbr_off[cc] = k;
      for (int o = 0; o < 27; ++o) {
        int32_t q = nb[cc * 27 + o];
        if (o != 13 && q >= 0) {
          if (nbr) nbr[k] = q;
          ++k;
        }
      }
    }
    if (nbr_off) nbr_off[nc] = k;
  }
  if (il_off || il) {
    if (!c->keep_lists || !c->il.p) throw Error(NBFMM_ERR_ARG, "interaction lists were not kept");
    if (il_off) NB_HIP(hipMemcpy(il_off, c->off_il.p, (nc + 1) * 4, hipMemcpyDeviceToHost));
    if (il && P.n_il) NB_HIP(hipMemcpy(il, c->il.p, (size_t)P.n_il * 4, hipMemcpyDeviceToHost));
  }
  API_END
}

} // extern "C"
