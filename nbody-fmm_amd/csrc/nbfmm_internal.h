// nbfmm_internal.h -- internal declarations of the gfx950 FMM library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbfmm.h"

namespace nbfmm {

// ---- errors ----------------------------------------------------------------
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define NB_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::nbfmm::Error(e_ == hipErrorOutOfMemory ? NBFMM_ERR_OOM : NBFMM_ERR_HIP,   \
                           std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)
#define NB_LAUNCH() NB_HIP(hipGetLastError())

// ---- device buffers grown on demand ------------------------------------------
struct DBuf {
  void *p = nullptr;
  size_t cap = 0;
};
void *dbuf_ensure(DBuf &b, size_t bytes);
void dbuf_free(DBuf &b);
// counts every (re)allocation: a captured graph is valid while it is unchanged
uint64_t dbuf_generation();
// while set, a buffer that would have to grow throws instead (graph capture)
void set_capturing(bool on);
template <class T>
inline T *dget(DBuf &b, size_t count) {
  return static_cast<T *>(dbuf_ensure(b, count * sizeof(T) + 16));
}

// ---- geometry of one evaluation ----------------------------------------------
constexpr int kMaxDepth = 16;   // ((64/3)>>3)<<3, tree.cpp:38
constexpr int kNumClasses = 343; // offsets in [-3,3]^3, class = (dz+3)*49+(dy+3)*7+(dx+3)
constexpr int kMaxOrder = 16;
constexpr int kTaskTargets = 256; // targets of one P2P sub-task (one wavefront, 4 per lane)
constexpr int kNarrowCap = 64;    // the largest narrow threshold (sets the narrow kernel's LDS)
constexpr int kMaxList = 192;     // longest source-cell list of a P2P task (interaction list <= 189)
// P2P blocks of at most this many targets go to the sources-across-lanes
// kernel (a pass over the sources per target group), larger ones to the
// four-targets-per-lane kernel, which leaves target slots empty below 64
// targets (measured at 4M, mixed: 64 against 32 gives P2P alone 10.27 ->
// 9.13 ms; FP64 12.07 -> 11.42 ms)
#ifndef NBFMM_P2P_NARROW_F32
#define NBFMM_P2P_NARROW_F32 64
#endif
#ifndef NBFMM_P2P_NARROW_F64
#define NBFMM_P2P_NARROW_F64 64
#endif
// the built-in tiles (nbfmm_tiles, every field resolved); tiles_ok checks a
// resolved set (why: the reason when it fails)
nbfmm_tiles default_tiles();
bool tiles_ok(const nbfmm_tiles &t, std::string &why);
// orders a learned selection may give a cell
constexpr int kOrderLo = 4, kOrderHi = 14, kNumOrders = kOrderHi - kOrderLo + 1;
// the order-selection MLP: kMlpIn -> kMlpHidden -> kMlpHidden -> kNumOrders,
// kMlpParams trainable FP32 values then the input mean and deviation
constexpr int kMlpIn = 16, kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden * kNumOrders + kNumOrders;
constexpr int kMlpWeights = kMlpParams + 2 * kMlpIn;

// One P2P sub-task: targets [cbeg[tcell] + toff, + kTaskTargets) of cell
// tcell against the particles [f0, f1) of the concatenated particle ranges
// of a list of source cells: list >= 0 names sub_src[list, list + nlist),
// list < 0 row (-list - 1) of the 27-neighbour table (nlist = 27, empty
// entries -1).
struct P2PTask {
  int32_t tcell, toff, list, nlist, f0, f1;
};

struct Geometry {
  double lo[3], hi[3];
  int depth = 0;
  // cell size per dimension at level l is (hi-lo)/2^l; rho_l = max of the three
  double h(int l, int d) const { return (hi[d] - lo[d]) / (double)(1ull << l); }
  double rho(int l) const {
    double r = h(l, 0);
    if (h(l, 1) > r) r = h(l, 1);
    if (h(l, 2) > r) r = h(l, 2);
    return r > 0 ? r : 1.0;
  }
};

// Per-level geometry passed to kernels by value.
struct LevelGeo {
  double lo[3];
  double h[kMaxDepth + 1][3];
  double inv_rho[kMaxDepth + 1];
  int32_t lvl_start[kMaxDepth + 2];
  int depth;
};

// geometric centre of the cell with Morton key `key` at level lev
__device__ __forceinline__ void cell_centre(uint64_t key, int lev, const LevelGeo &g, double c[3]) {
  uint64_t k = key;
  uint32_t q[3] = {0, 0, 0};
  for (int b = 0; b < lev; ++b) {
    q[0] |= (uint32_t)((k >> (3 * b)) & 1u) << b;
    q[1] |= (uint32_t)((k >> (3 * b + 1)) & 1u) << b;
    q[2] |= (uint32_t)((k >> (3 * b + 2)) & 1u) << b;
  }
  for (int d = 0; d < 3; ++d) c[d] = g.lo[d] + ((double)q[d] + 0.5) * g.h[lev][d];
}

__device__ __forceinline__ int level_of(int32_t cell, const LevelGeo &g) {
  int l = 0;
  while (l < g.depth && cell >= g.lvl_start[l + 1]) ++l;
  return l;
}

// Offset classes whose M2L does not converge fast enough for the cell shape
// (bit k: class k); their pairs are evaluated directly (lists.hip).  Built
// on the device from the tree's box (k_tree_plan, tree.hip).
struct ClsMask {
  uint32_t w[(kNumClasses + 31) / 32];
};
constexpr double kMacRatio = 0.9;
constexpr int64_t kBadDirect = (int64_t)1 << 20; // largest n_t n_s evaluated directly for such classes
// M2L order of every offset class (class_order_tol > 0), built beside the mask;
// kOrderAlpha: observed decay per degree over the worst-case rate
constexpr double kOrderAlpha = 0.55;
struct ClsOrd {
  uint8_t q[kNumClasses + 1];
};

// ---- the counts of one evaluation, in device memory ---------------------------------
// Every size an evaluation derives from the particles (the box, the level
// starts, list and task totals, the up-pass planning, the M2L schedule) is
// written by kernels into one DevPlan, and the kernels that depend on a size
// read it there.  A direct evaluation also reads the sizes on the host, to
// launch and allocate exactly; a planned evaluation (Caps::on) launches and
// allocates by capacities fixed beforehand and reads nothing back, so that a
// whole integration step can be captured in a HIP graph.  A planned size
// beyond its capacity, or a depth other than the planned one, sets `ovf`;
// every later kernel of the step then leaves at once, the integrator does
// not update the state, and the host runs the step again directly.
constexpr int kMaxBins = kNumClasses * kNumOrders; // M2L bins: (order, class)
constexpr int kMaxChunks0 = 64;                    // P2P first-share launches at most
// P2P task ranges [begin, end) in DevPlan::p2p_rng: the first share's narrow
// tasks in chunks (chunk ch at 2 ch) and whole, its wide tasks, the rest's
// narrow and wide tasks, and all narrow and all wide tasks
constexpr int kMaxChunks1 = 8;                      // launches of the rest of the narrow P2P tasks at most
constexpr int kRngNarrow0 = 2 * kMaxChunks0, kRngWide0 = kRngNarrow0 + 2, kRngNarrow1 = kRngNarrow0 + 4,
              kRngWide1 = kRngNarrow0 + 6, kRngNarrowAll = kRngNarrow0 + 8, kRngWideAll = kRngNarrow0 + 10,
              kRngRest0 = kRngNarrow0 + 12, kRngCount = kRngRest0 + 2 * kMaxChunks1;
struct DevPlan {
  double lo[3], hi[3]; // the tree's box (tree.cpp:20-30; the cube around it with cubic_box)
  double hs[3];        // leaf edges over rho at the depth (operator geometry)
  int32_t bad;         // a non-finite coordinate or mass
  int32_t ovf;         // planned evaluation: a size beyond its capacity (bit 1), another depth (bit 0), bad input (bit 2)
  int32_t depth, nc;
  int32_t iterations;                    // iterations of the reference's depth loop
  int32_t one_nc[2];                     // {1, nc}: the cells that have a parent
  int32_t lvl_start[kMaxDepth + 2];      // cells concatenated from the root
  uint32_t lcount[2 * (kMaxDepth + 1)];  // the last count: cells per level, then each level's fullest cell
  LevelGeo g;
  ClsMask bad_cls;
  ClsOrd qc;
  // lists and tasks
  int32_t nactive;                       // parents with a child to classify
  int32_t nst[2];                        // M2L sources, M2L targets
  int32_t nnarrow;                       // tasks of narrow blocks (numbered first)
  unsigned long long tot[3];             // interaction-list, M2L and substituted entries
  unsigned long long ntask, npairs;
  int32_t p2p_rng[kRngCount];
  // up-pass planning: M cells, P2M cells, M2M cells, L cells, L2L cells; the
  // M2M cells before level d at 8 + d, the L2L cells before level d at 32 + d
  int32_t reads[64];
  unsigned long long terms[2];           // particles of the P2M cells, of the L2P cells
  // M2L schedule (single pass): items of order index oi at [ord_items[oi][0], + ord_items[oi][1])
  int32_t ord_items[kNumOrders][2];
  int32_t n_items;
  double flop_m2l, order_sum;
  int32_t bin_cnt[kMaxBins];
  int32_t bin_cur[kMaxBins];
  int32_t ops_valid; // the operators built by an earlier evaluation fit this one (k_ops_check)
};

// Geometry and order the operator buffers were last built for (device side).
struct OpsKey {
  double hs[3];
  int32_t p, kp, f32, pad;
};

// Sizes of a direct evaluation, kept to set the capacities of planned ones.
struct PlanSizes {
  int depth = -1;
  int64_t n = 0, nc = 0, parents = 0;
  int32_t lev[kMaxDepth + 1] = {};
  int64_t n_m2l = 0, n_sub = 0, n_tgt = 0, ntask = 0, nnarrow = 0;
  int64_t nM = 0, nL = 0, nP2M = 0;
  int32_t m2m_lev[kMaxDepth + 1] = {}, l2l_lev[kMaxDepth + 1] = {};
  int64_t items = 0;
  int32_t items_ord[kNumOrders] = {};
};

// Capacities of a planned evaluation (every field a bound the kernels check).
struct Caps {
  bool on = false;
  int depth = 0;
  int64_t nc = 0, parents = 0;
  int32_t lev[kMaxDepth + 1] = {};
  int64_t n_m2l = 0, n_sub = 0, n_tgt = 0, ntask = 0, nnarrow = 0;
  int64_t nM = 0, nL = 0, nP2M = 0;
  int32_t m2m_lev[kMaxDepth + 1] = {}, l2l_lev[kMaxDepth + 1] = {};
  int64_t items = 0;
  int32_t items_ord[kNumOrders] = {};
};
Caps caps_from(const PlanSizes &s);

// capacities as kernel arguments of the checks (by value)
struct CapLevels {
  int32_t v[kMaxDepth + 1];
};

// the planned evaluation of a step has overflowed: kernels after a check leave
__device__ __forceinline__ bool plan_dead(const DevPlan *dp) {
  return __builtin_amdgcn_readfirstlane(*(const volatile int32_t *)&dp->ovf) != 0;
}

// ---- the device-side state of one evaluation -------------------------------------
struct Plan {
  int64_t n = 0;
  Geometry geo;
  int iterations = 0;  // iterations of the reference's depth loop
  int depth_sorts = 0; // key sorts run to find the depth
  std::vector<int32_t> lvl_start; // depth+2, cells concatenated from level 0
  int64_t n_cells = 0, n_leaves = 0;
  int64_t n_il = 0, n_m2l = 0, n_sub = 0;
  int64_t n_src = 0, n_tgt = 0, n_p2p_pairs = 0, n_p2m_terms = 0, n_l2p_terms = 0;
  int64_t n_p2m_items = 0, n_m2l_tiles = 0, n_tasks = 0;
  int64_t n_tasks_narrow = 0; // tasks [0, n_tasks_narrow) have blocks of at most the tiles' narrow_max targets
  int64_t n_p2m_cells = 0, n_m2m_cells = 0, n_l2l_cells = 0, n_l2p_cells = 0; // cells of each pass
  int kp = 0; // padded coefficient count (multiple of 16)
  double flop_m2l = 0; // algorithmic M2L flops of the evaluation
  double mean_order = 0; // mean order over the M2L pairs (the fixed order without selection)
  bool m2l_stats_dev = false; // flop_m2l and mean_order are still in the device plan (read with the stats)
};

} // namespace nbfmm

// The context (definition shared by the host translation units).
struct nbfmm_ctx {
  nbfmm_params prm{};
  nbfmm_tiles shapes = nbfmm::default_tiles(); // tile shapes, every field resolved
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;            // P2P, concurrent with the expansion passes
  hipEvent_t ev_fork = nullptr, ev_fork2 = nullptr, ev_join = nullptr;
  hipEvent_t ev_pos = nullptr; // the sorted particles (gathered on stream2) are complete
  int device = 0;
  int lds_max = 65536; // LDS bytes one workgroup may allocate on this device
  int n_cu = 256;      // compute units of this device
  bool timing = true;
  // phases one after another (P2P alone at the end, after L2P) so that each
  // phase's events time its kernels alone; the results are the same
  bool isolated = false;
  bool keep_lists = false;
  int depth_guess = 0; // the last tree's depth: the first depth sorted at by the next
  int nbr_leaf = 0; // rows of the leaf level in nbr27: 0 none, 1 leaves that are not closed, 2 all
  double *phi_out = nullptr; // potential output of the current evaluation (device, input order) or null
  nbfmm::Plan plan;
  nbfmm_stats stats{};
  // particle arrays
  nbfmm::DBuf xyz_in, m_in, acc_out;           // staging for host-pointer calls
  nbfmm::DBuf key[2], idx[2], sort_hist, scan_tmp, act_par;
  nbfmm::DBuf pos;                             // double4 sorted (x,y,z,m)
  nbfmm::DBuf pcell;                           // leaf cell of each sorted particle
  nbfmm::DBuf grad;                            // 3 doubles per sorted particle (far field)
  nbfmm::DBuf grad_near;                       // 3 doubles per sorted particle (P2P)
  nbfmm::DBuf red;                             // small reductions (bbox, flags, counts)
  nbfmm::DBuf hred;                            // pinned host mirror of `red`
  // cells
  nbfmm::DBuf ckey, cbeg, cend, cparent, child8, nbr27, flags, fscan, closed;
  nbfmm::DBuf cnt_il, cnt_m2l, cnt_sub, off_il, off_m2l, off_sub, near_pairs;
  nbfmm::DBuf il, m2l_src, m2l_cls, sub_src;
  nbfmm::DBuf src_slot, tgt_slot, src_cells, tgt_cells, is_src;
  nbfmm::DBuf sub_pairs, tiles;
  nbfmm::DBuf tasks, task_off, task_cnt, dpart, p2p_ctr;
  // expansions
  nbfmm::DBuf stab, tmat, tmat32, trans, trans32, mode_m, mode_l, mslot, lslot, eslot, up_cells, mexp, lexp, consts, ypair, cls_pairs, cls_cnt, m2l_items, m2l_tab;
  nbfmm::DBuf mexp64, lexp64; // FP64 copies for the coarse M2M / L2L levels of mixed precision
  std::vector<int32_t> h_cls_off; // host schedule of the class-sorted M2L
  std::vector<int4> h_m2l_items;
  // integrator scratch
  nbfmm::DBuf xk, vk, ak, asum, vsum, diag_acc, diag_phi;
  // learned order selection: per-cell orders and features, the model, its Adam
  // state and training scratch, the training set and the label runs
  nbfmm::DBuf ord, msum, feat, rmax, mlp_w, mlp_wb, mlp_m, mlp_v, mlp_act, mlp_perm;
  nbfmm::DBuf ds_feat, ds_label, lab_tgt, lab_dir, lab_sum, lab_flag, lab_lab, chk_acc;
  int64_t n_local = 0; // cells with a local expansion in the last evaluation
  int ord_max = 0;     // highest order selected in the last evaluation
  bool mlp_ok = false, mlp_fresh = false;
  int64_t mlp_step = 0, ds_n = 0;
  int64_t lab_ns = 0, lab_stride = 1, lab_nl = -1;
  uint64_t lab_seed = 0;
  int fallback = 0; // the last checked evaluation fell back to the fixed order
  // batched systems: offsets, boxes, placement, placed copies; first level with lists
  nbfmm::DBuf bat_off, bat_bb, bat_st, bat_x, bat_m, bat_a;
  int list_level = 1;
  bool p2p_started = false; // P2P of the current evaluation is enqueued
  // the sizes of the current evaluation on the device, their pinned mirror
  nbfmm::DBuf dplan;
  nbfmm::DBuf ops_key;                 // OpsKey of the operators in tmat, trans, tmat32, trans32
  const void *ops_bufs[4] = {};        // those buffers' addresses when the key was written
  nbfmm::DevPlan *hplan = nullptr;
  nbfmm::Caps cap;        // capacities of planned evaluations (cap.on: this evaluation is planned)
  nbfmm::PlanSizes last;  // sizes of the last direct evaluation
  // graph-captured integration steps: the instantiated graph, what it was captured for
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  struct GraphKey {
    int kind = -1; // 0 SE, 1 RK4 reference semantics, 2 RK4 classic; + 3 for a batch
    int64_t n = 0;
    const void *x = nullptr, *v = nullptr, *a = nullptr, *m = nullptr, *off = nullptr;
    double dt = 0;
    int nsys = 0;
    nbfmm_params prm{};
    nbfmm_tiles shapes{};
    uint64_t gen = 0; // buffer generation at the capture
  } gkey;
  int64_t g_captures = 0, g_launches = 0, g_redone = 0, g_direct = 0;
  // events
  hipEvent_t ev[16] = {};
  int nev = 0;
  ~nbfmm_ctx();
};

namespace nbfmm {
// scan.hip
void scan_excl_i32(const int32_t *in, int32_t *out, int64_t n, DBuf &tmp, hipStream_t s);
void scan_excl_i64(const int64_t *in, int64_t *out, int64_t n, DBuf &tmp, hipStream_t s);
void scan_excl_f64(const double *in, double *out, int64_t n, DBuf &tmp, hipStream_t s);
// tree.hip
void build_tree(nbfmm_ctx *c, const double *d_xyz, const double *d_m);
// lists.hip
void build_lists(nbfmm_ctx *c);
void nbr_leaf_rows(nbfmm_ctx *c, const int32_t *closed); // leaf-level rows of nbr27 (all when closed is null)
// expansions.hip
void run_expansions(nbfmm_ctx *c);
void upload_consts(nbfmm_ctx *c); // the recurrence constants, once per context
int padded_k(int p); // pad16((p+1)^2)
LevelGeo level_geo(const Plan &P); // the host's copy of the geometry (direct evaluations)
// the device plan of the context
inline DevPlan *dplan(nbfmm_ctx *c) { return static_cast<DevPlan *>(c->dplan.p); }
// cells of level d: exact (direct evaluation) or the capacity (planned)
inline int64_t level_cells(nbfmm_ctx *c, int d) {
  return c->cap.on ? c->cap.lev[d] : (int64_t)(c->plan.lvl_start[d + 1] - c->plan.lvl_start[d]);
}
// reads [field, field + bytes) of the device plan into its pinned mirror (stream-ordered, then waits)
void read_plan(nbfmm_ctx *c, const void *field, size_t bytes, hipStream_t s);
template <class F>
inline void read_plan_field(nbfmm_ctx *c, F DevPlan::*f, hipStream_t s) {
  read_plan(c, &(c->hplan->*f), sizeof(F), s);
}
// p[0, count) = v (a kernel: graph captures take no memset nodes)
void fill_i32(hipStream_t s, void *p, int64_t count, int32_t v);
inline void fill_zero(hipStream_t s, void *p, size_t bytes) { fill_i32(s, p, (int64_t)(bytes / 4), 0); }
// order.hip
void order_features(nbfmm_ctx *c, const LevelGeo &g);
const int8_t *order_select(nbfmm_ctx *c, const LevelGeo &g);
void mlp_set(nbfmm_ctx *c, const float *w);
void mlp_get(nbfmm_ctx *c, float *w);
void mlp_init(nbfmm_ctx *c, uint64_t seed);
void mlp_predict(nbfmm_ctx *c, int64_t nb, const float *feat_h, float *prob_h);
void mlp_grad(nbfmm_ctx *c, int nb, const float *feat_h, const int8_t *lab_h, float *grad_h, float *loss_h);
void mlp_train(nbfmm_ctx *c, int epochs, int batch, double lr, uint64_t seed, float *loss_h);
void label_begin(nbfmm_ctx *c, int64_t n, int64_t stride, uint64_t seed, const double *d_xyz, const double *d_m);
void label_add(nbfmm_ctx *c, int qi, const double *d_acc);
int64_t label_finish(nbfmm_ctx *c, double tol);
double sampled_error(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, const double *d_acc,
                     int64_t ntgt, uint64_t seed);
void dataset_get(nbfmm_ctx *c, float *feat_h, int8_t *lab_h, int64_t cap);
void dataset_put(nbfmm_ctx *c, int64_t n, const float *feat_h, const int8_t *lab_h);
// m2l.hip
void run_m2l(nbfmm_ctx *c, int KP, bool f32, const int8_t *ord, int p);
// M2M (l2l false) or L2L (true) for the parents cells[rng[0], rng[1]) (a
// device range, one level's), launched for at most grid_cells of them
void run_trans(nbfmm_ctx *c, bool l2l, int KP, bool f32, const int32_t *cells, const int32_t *rng,
               int64_t grid_cells);
// the same in FP64 on the expansions X (FP64, slot-indexed like the context's)
void run_trans64(nbfmm_ctx *c, bool l2l, int KP, const int32_t *cells, const int32_t *rng, int64_t grid_cells,
                 double *X);
// count values between FP32 and FP64 arrays (stream-ordered)
void widen(nbfmm_ctx *c, const float *in, int64_t count, double *out);
void narrow(nbfmm_ctx *c, const double *in, int64_t count, float *out);
void trans_to_f32(nbfmm_ctx *c, int KP);
// the FP32 / bf16 operator buffers (allocated at KP): tmat32, and trans32 through *trans32
const void *ops_f32_buffers(nbfmm_ctx *c, int KP, const void **trans32);
// p2p.hip
// P2P on the second stream: fork_p2p marks the point of the main stream it
// waits for (its inputs are complete there).  launch_p2p(c, 0) enqueues the
// first share of the tasks, to run beside the up-pass and its host reads;
// launch_p2p(c, 1) the rest, after the first M2L contraction, to run beside
// the sums, L2L and L2P.
void fork_p2p(nbfmm_ctx *c);
void launch_p2p(nbfmm_ctx *c, int part);
// joins the P2P stream and writes acc = G * (far + near) in input order
void run_finish(nbfmm_ctx *c, double *d_acc_out);
// part_fixed: the partial sums go there (the fixed scratch) instead of a
// context buffer; sizes beyond it are rejected
void run_direct(nbfmm_ctx *c, int64_t n, const double *d_xyz, const double *d_m, int64_t ntgt,
                const int64_t *d_tgt, int mode, double *d_acc, double *part_fixed = nullptr);
// the fixed scratch of the allocation-free direct sum: scratch_init(device)
// looks up its address on `device` (current), once per device; scratch_ptr
// returns it
void scratch_init(int device);
double *scratch_ptr(int device);
// batch.hip: the systems [off[k], off[k+1]) of one concatenated set of n
// particles evaluated as one forest (d_off: the offsets on the device, from
// upload_offsets); eval(n, x, m, a) runs the evaluation of the placed copy
void batch_eval(nbfmm_ctx *c, int nsys, const int64_t *d_off, int64_t n, const double *d_xyz, const double *d_m,
                double *d_acc, const std::function<void(int64_t, const double *, const double *, double *)> &eval);
int64_t *upload_offsets(nbfmm_ctx *c, int nsys, const int64_t *h_off);
// integrate.hip
void se_update(nbfmm_ctx *c, int64_t n, double *x, double *v, const double *a, double dt);
void rk4_stage(nbfmm_ctx *c, int64_t n, const double *x, const double *v, const double *vk_prev,
               const double *a_vel, double h, double *xk, double *vk_out);
void rk4_accum(nbfmm_ctx *c, int64_t n, double *asum, double *vsum, const double *a,
               const double *vk, double w, int first);
void rk4_final(nbfmm_ctx *c, int64_t n, double *x, double *v, const double *asum,
               const double *vsum, double h6);
void copy_d(nbfmm_ctx *c, double *dst, const double *src, int64_t count);
// 10 FP64 sums of a state (host out): kinetic, potential, total energy, momentum, angular momentum, mass
void diagnostics(nbfmm_ctx *c, int64_t n, const double *x, const double *v, const double *m, const double *phi,
                 double *h_out);
// timing helpers (capi.cpp)
void mark(nbfmm_ctx *c, int slot);
void mark2(nbfmm_ctx *c, int slot); // on the P2P stream
} // namespace nbfmm
