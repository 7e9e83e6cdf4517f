/*
 * nbfmm.h -- C ABI of the MI355X (gfx950) 3D gravitational FMM library.
 *
 * The library replaces the force-evaluation path behind the reference's
 * system-updater API (rezucch1/nbody-FMM):
 *   class SystemUpdateMethod { virtual void update(const std::vector<Particle>&) = 0; }
 *     (src/system_updaters/system_update_method.hpp:24-31)
 *   FMM::update          (src/system_updaters/FMM.cpp:9-12)        -> nbfmm_eval_accel
 *   NaiveUpdate::update  (src/system_updaters/naive_update.cpp:9-35) -> nbfmm_direct_accel
 *   SyplecticEuler::integrate (src/quadratures/symplectic_euler.cpp:8-15) -> nbfmm_step_se_dev
 *   RungeKutta4::integrate    (src/quadratures/runge_kutta_4.cpp:10-93)   -> nbfmm_step_rk4_dev
 * Particle data crosses the boundary as plain arrays: positions AoS (n x 3
 * doubles, x y z per particle, the order of Particle::position), masses
 * (n doubles, Particle::weight) and accelerations AoS (n x 3 doubles,
 * Particle::acceleration = G * grad, particle.cpp:16-21).
 *
 * Conventions
 *  - Every function returns an nbfmm_status; 0 is success.  The message of the
 *    last failure on the calling thread is returned by nbfmm_last_error().
 *  - A context owns its device buffers and one HIP stream; it is not
 *    thread-safe.  Different contexts may be used concurrently.
 *  - *_dev functions take device pointers and are ordered on the context's
 *    stream; they return before the GPU finishes unless stated otherwise.
 *    Host-pointer functions synchronise before returning.
 */
#ifndef NBFMM_H
#define NBFMM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nbfmm_status {
  NBFMM_OK = 0,
  NBFMM_ERR_ARG = 1,         /* invalid argument (null pointer, n < 1, bad parameter) */
  NBFMM_ERR_DEGENERATE = 2,  /* non-finite coordinates or masses, zero total mass */
  NBFMM_ERR_HIP = 3,         /* a HIP runtime call failed */
  NBFMM_ERR_OOM = 4,         /* device allocation failed */
  NBFMM_ERR_UNSUPPORTED = 5, /* parameter combination not implemented */
  NBFMM_ERR_NO_DEVICE = 6    /* no usable GPU */
} nbfmm_status;

typedef enum nbfmm_precision {
  NBFMM_PLAIN_FP64 = 0, /* every phase in FP64 */
  NBFMM_MIXED = 1       /* FP32 P2P (compensated) and FP32 MFMA M2L; FP64 tree/integrator */
} nbfmm_precision;

typedef enum nbfmm_order_select {
  NBFMM_ORDER_FIXED = 0,  /* every cell uses `order` (the reference's single L) */
  NBFMM_ORDER_LEARNED = 1 /* evaluation cells take the order the context's model predicts */
} nbfmm_order_select;

typedef struct nbfmm_params {
  uint32_t order;         /* expansion order p (reference L, hardcoded 2 at tree.cpp:94) */
  uint32_t leaf_capacity; /* depth-loop threshold (reference 20, tree.cpp:32) */
  double G;               /* gravitational constant (reference 6.6743e-11, system_update_method.hpp:17) */
  double eps2;            /* softening of the direct sum (reference NaiveUpdate 1e-4, naive_update.cpp:27) */
  int32_t precision;      /* nbfmm_precision */
  /* M2L pairs per pass of the translation (0: as many as 16 GiB of per-pair
   * partials hold).  Larger interaction sets are done in several passes. */
  int32_t m2l_pass_pairs;
  /* An interaction-list pair (t, s) is evaluated by M2L, or by direct P2P
   * between the two cells when n_t * n_s < p2p_cut.  0 = always M2L, as the
   * reference does (node_i.cpp:105-120).  P2P is exact, so this only changes
   * cost and removes truncation error from the substituted pairs. */
  int64_t p2p_cut;
  /* nbfmm_order_select.  With NBFMM_ORDER_LEARNED, `order` is the highest
   * order (multipoles, L2L and the cells that push their local to their
   * children use it) and every cell that evaluates its local at its particles
   * gets the lowest order in [order_min, order] whose cumulative probability
   * under the model reaches order_conf.  No counterpart in the reference
   * (one L for all cells, tree.cpp:94). */
  int32_t order_select;
  int32_t order_min;  /* >= 4 */
  double order_conf;  /* in (0, 1] */
  /* 0: the reference's bounding box (tree.cpp:20-30, one extent per axis:
   * cells have the box's aspect ratio).  1: the bounding cube centred on it
   * (cubic cells).  With the reference's box, the nearest interaction classes
   * along the short axes of a box longer than sqrt(2) : 1 converge slowly or
   * not at all; their pairs are evaluated directly up to n_t * n_s < 2^20
   * and by M2L above that, so such inputs reach high accuracy only with
   * cubic_box = 1. */
  int32_t cubic_box;
  /* Distance-dependent M2L orders (no counterpart in the reference, which uses
   * one L everywhere).  0: every M2L pair at `order`.  tol > 0: a pair of
   * offset class d is truncated at the lowest q >= 4 (at most `order`) with
   * rate_d^(q+1) <= tol * rate_0^(order+1), where rate_d = (cell diagonal) /
   * |d| and rate_0 is the largest rate of the M2L classes (sqrt(3)/2 for
   * cubic cells): far classes converge faster and need fewer terms.  Needs
   * 4 <= order <= 14. */
  double class_order_tol;
  /* Mixed precision only: the M2M and L2L translations whose parents lie on
   * the levels 0 .. mixed_fp64_levels - 1 (the coarsest, whose cells sum the
   * most particles) run in FP64 on FP64 copies of the expansions, the others
   * in FP32.  0 (the default): every level in FP32; a value at or above the
   * tree depth: every level in FP64.  Measured on 4M Plummer (p = 11): the
   * six coarsest levels in FP64 leave the error unchanged and cost 1.1 ms per
   * evaluation; all levels in FP64 lower it from 7.7e-7 to 7.2e-7 and cost
   * 1.7 ms. */
  int32_t mixed_fp64_levels;
} nbfmm_params;

/* Reference-compatible defaults: p = 2, leaf capacity 20, G = 6.6743e-11,
 * eps2 = 1e-4 (NaiveUpdate), FP64, p2p_cut = 0, fixed order (order_min 4,
 * order_conf 0.9 for a later learned selection), mixed_fp64_levels 0. */
void nbfmm_default_params(nbfmm_params *params);

typedef struct nbfmm_ctx nbfmm_ctx;

/* Creates a context on the current HIP device.  The context keeps its device
 * buffers across calls and grows them as needed. */
int nbfmm_create(const nbfmm_params *params, nbfmm_ctx **ctx);
int nbfmm_destroy(nbfmm_ctx *ctx);
/* Replaces the parameters used by later calls. */
int nbfmm_set_params(nbfmm_ctx *ctx, const nbfmm_params *params);
/* The context's HIP stream (hipStream_t), for callers that order their own work. */
void *nbfmm_stream(nbfmm_ctx *ctx);

/* FMM::update (FMM.cpp:9-12): accelerations of all n particles.
 * Host pointers; synchronous.  acc = G * grad, grad = sum_j m_j (x_j-x_i)/|x_j-x_i|^3
 * (leaf.cpp:119-142; coincident pairs contribute nothing). */
int nbfmm_eval_accel(nbfmm_ctx *ctx, int64_t n, const double *xyz, const double *m, double *acc);
/* Same on device pointers, ordered on the context's stream. */
int nbfmm_eval_accel_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_m,
                         double *d_acc);

/* NaiveUpdate::update (naive_update.cpp:9-35): grad -= m_j d/((r^2+eps2) r),
 * acc = G * grad, computed on the GPU in FP64.  Host pointers; synchronous. */
int nbfmm_direct_accel(nbfmm_ctx *ctx, int64_t n, const double *xyz, const double *m, double *acc);
/* Direct sum for the targets listed in d_tgt (indices into d_xyz), written to
 * d_acc (3 per target).  mode 0 = the NaiveUpdate formula above; mode 1 =
 * grad = sum m_j d/(r^2+eps2)^{3/2} with compensated (Neumaier) accumulation,
 * the FP64 oracle of the accuracy metric.  Device pointers, stream-ordered. */
int nbfmm_direct_accel_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_m,
                           int64_t ntgt, const int64_t *d_tgt, int mode, double *d_acc);

/* Fixed device memory of nbfmm_direct_accel_noalloc_dev, per GPU, and the
 * most targets one mode-1 call of it may ask for (their partial sums fill
 * the array at most). */
#define NBFMM_NOALLOC_SCRATCH_BYTES ((size_t)48 << 20)
#define NBFMM_NOALLOC_MAX_TARGETS ((int64_t)1 << 20)
/* nbfmm_direct_accel_dev without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results (bit for bit) as nbfmm_direct_accel_dev.
 * Mode 1 keeps its partial sums (nchunk x ntgt x 6 doubles) in a fixed array
 * of NBFMM_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends,
 * set up by nbfmm_create; a mode-1 call with more than
 * NBFMM_NOALLOC_MAX_TARGETS targets is rejected with NBFMM_ERR_ARG (mode 0
 * uses no scratch and takes any ntgt).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (from any context or stream) race.
 * Device pointers, stream-ordered on the context's stream. */
int nbfmm_direct_accel_noalloc_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_m,
                                   int64_t ntgt, const int64_t *d_tgt, int mode, double *d_acc);

/* SyplecticEuler::integrate (symplectic_euler.cpp:8-15), nsteps times:
 * acc = FMM(x); v += dt*acc; x += dt*v.  d_acc receives the last acceleration.
 * Device pointers, stream-ordered. */
int nbfmm_step_se_dev(nbfmm_ctx *ctx, int64_t n, double *d_xyz, double *d_vel, double *d_acc,
                      const double *d_m, double dt, int nsteps);
/* RungeKutta4::integrate (runge_kutta_4.cpp:10-93), nsteps times.
 * mode 0 reproduces the reference: every stage starts from a copy of the
 * input particles, so stage velocities are v + h*a_in with a_in the carried
 * acceleration in d_acc (lines 41-79), and d_acc is left unchanged.
 * mode 1 is classic RK4 and leaves a(x_n) of the last step in d_acc.
 * Device pointers, stream-ordered. */
int nbfmm_step_rk4_dev(nbfmm_ctx *ctx, int64_t n, double *d_xyz, double *d_vel, double *d_acc,
                       const double *d_m, double dt, int nsteps, int mode);

/* Accelerations and potentials phi_i = -G sum_{j != i} m_j / |x_i - x_j| of
 * all n particles (the FMM of nbfmm_eval_accel_dev with the potential formed
 * beside the gradient: P2P sums m_j / r_ij, L2P evaluates sum L R).  The
 * reference computes no potential; this serves the energy monitor.  Device
 * pointers, stream-ordered. */
int nbfmm_eval_accel_pot_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_m, double *d_acc,
                             double *d_phi);

/* Sums of a particle state in FP64 with a fixed summation order, written to
 * the host array out[10]: out[0] kinetic energy sum m v^2 / 2, out[1]
 * potential energy sum m phi / 2 (0 when d_phi is NULL; d_phi from
 * nbfmm_eval_accel_pot_dev at the same positions), out[2] total energy,
 * out[3..5] momentum sum m v, out[6..8] angular momentum sum m x cross v about
 * the origin, out[9] total mass.  Device pointers; synchronous. */
int nbfmm_diagnostics_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_vel, const double *d_m,
                          const double *d_phi, double *out);

/* Multi-step driver: advances nsteps with integrator 0 = SyplecticEuler,
 * 1 = RungeKutta4 with the reference's semantics, 2 = classic RK4 (as
 * nbfmm_step_se_dev / nbfmm_step_rk4_dev), and, when h_diag is not NULL,
 * records the 10 diagnostics of nbfmm_diagnostics_dev before the first step
 * and after every `every` steps into h_diag (host memory, room for
 * 10 * (1 + nsteps / every) doubles).  The monitor evaluates into its own
 * buffers, so d_acc carries the integrator's acceleration as without it.
 * Device pointers; synchronous. */
int nbfmm_run_dev(nbfmm_ctx *ctx, int64_t n, double *d_xyz, double *d_vel, double *d_acc, const double *d_m,
                  double dt, int nsteps, int integrator, int every, double *h_diag);

/* Learned order selection (no counterpart in the reference).
 *
 * The model is an MLP 16 -> 64 -> 64 -> 11 (ReLU) over features of a cell,
 * whose 11 outputs are the orders 4..14; FP32 weights, used as BF16 in the
 * forward passes, FP32 accumulation.  Weight vector (NBFMM_MLP_WEIGHTS floats):
 * W1[16][64], b1[64], W2[64][64], b2[64], W3[64][11], b3[11] (the
 * NBFMM_MLP_PARAMS trainable values), then the input mean[16] and
 * deviation[16] (features are used as (f - mean) / deviation). */
#define NBFMM_MLP_FEATURES 16
#define NBFMM_MLP_ORDERS 11
#define NBFMM_MLP_ORDER_LO 4
#define NBFMM_MLP_PARAMS 5963
#define NBFMM_MLP_WEIGHTS 5995
/* He-uniform initial weights from `seed`; the normalisation is taken from
 * the training set by the first nbfmm_order_train call. */
int nbfmm_order_model_init(nbfmm_ctx *ctx, uint64_t seed);
int nbfmm_order_model_set(nbfmm_ctx *ctx, const float *w, int64_t nw);
int nbfmm_order_model_get(nbfmm_ctx *ctx, float *w, int64_t nw);
/* Training data from one particle set (device pointers): the FMM with the
 * context's parameters is run at every order 4..14 (fixed selection, same
 * tree); one particle in every `stride` (a seeded pick in each block of
 * `stride`) is compared with the FP64 direct sum (compensated, no softening),
 * and every cell that evaluates its local expansion gets the label "lowest
 * order whose mean squared force error over its sampled particles is within
 * tol^2 times the mean squared acceleration of all sampled particles" (the
 * highest when none is; cells within it keep the global relative L2 error
 * within tol).  The labelled cells' features are appended to
 * the context's training set; *added receives their number.  Synchronous. */
int nbfmm_order_dataset_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_m, int64_t stride,
                            uint64_t seed, double tol, int64_t *added);
/* Size of the training set; copies of its first `cap` rows (features
 * [cap][16], labels [cap] as order - 4) when the pointers are not NULL. */
int nbfmm_order_dataset_get(nbfmm_ctx *ctx, float *feat, int8_t *label, int64_t cap, int64_t *n);
/* Appends rows given on the host; clears the set when n == 0 and both pointers are NULL. */
int nbfmm_order_dataset_put(nbfmm_ctx *ctx, int64_t n, const float *feat, const int8_t *label);
/* Minibatch Adam (lr, beta 0.9 / 0.999) on the mean cross entropy over the
 * training set, shuffled by `seed`; loss[e] (host, may be NULL) receives the
 * mean loss of epoch e.  Runs on the GPU; synchronous. */
int nbfmm_order_train(nbfmm_ctx *ctx, int32_t epochs, int32_t batch, double lr, uint64_t seed, float *loss);
/* The GPU forward and backward pass over nb samples given on the host
 * (features [nb][16], labels as order - 4) with the current weights:
 * grad[NBFMM_MLP_PARAMS] of the mean cross entropy and *loss.  Synchronous. */
int nbfmm_order_grad(nbfmm_ctx *ctx, int32_t nb, const float *feat, const int8_t *label, float *grad, float *loss);
/* Predicted probabilities prob[nb][11] of nb feature rows (host).  Synchronous. */
int nbfmm_order_predict(nbfmm_ctx *ctx, int64_t nb, const float *feat, float *prob);
/* nbfmm_eval_accel_dev followed by the sampled check: the relative L2 error
 * of the accelerations against the FP64 direct sum (compensated, no
 * softening) on about ntgt sampled targets.  When it exceeds `gate`, the
 * evaluation is repeated with a fixed order `fallback_order` and checked
 * again; *err receives the error of the accelerations left in d_acc and
 * *fell_back whether the fallback ran.  Synchronous. */
int nbfmm_eval_accel_checked_dev(nbfmm_ctx *ctx, int64_t n, const double *d_xyz, const double *d_m, double *d_acc,
                                 int64_t ntgt, uint64_t seed, double gate, int32_t fallback_order, double *err,
                                 int32_t *fell_back);

/* Batched systems (no counterpart in the reference, which advances one
 * system).  nsys independent systems stored one after the other in the
 * arrays: system k is particles [offsets[k], offsets[k+1]) (offsets: host
 * array of nsys + 1, offsets[0] = 0); no system acts on another.  All systems
 * are evaluated together, every phase of the FMM in one launch for all: each
 * is scaled by a power of two into its own cell of a grid of isolated cells
 * and the interaction lists stop at those cells.  Systems of any size (one
 * particle or more) are accepted: the tree is refined at least one level
 * below the system cells.  Device pointers, stream-ordered. */
int nbfmm_batch_eval_accel_dev(nbfmm_ctx *ctx, int32_t nsys, const int64_t *offsets, const double *d_xyz,
                               const double *d_m, double *d_acc);
/* nbfmm_step_se_dev / nbfmm_step_rk4_dev over a batch (each system advanced
 * as by itself; the update of every stage is one launch for all). */
int nbfmm_batch_step_se_dev(nbfmm_ctx *ctx, int32_t nsys, const int64_t *offsets, double *d_xyz, double *d_vel,
                            double *d_acc, const double *d_m, double dt, int nsteps);
int nbfmm_batch_step_rk4_dev(nbfmm_ctx *ctx, int32_t nsys, const int64_t *offsets, double *d_xyz, double *d_vel,
                             double *d_acc, const double *d_m, double dt, int nsteps, int mode);

/* Integration steps replayed from a HIP graph (no counterpart in the
 * reference, which calls RungeKutta4::integrate step by step,
 * runge_kutta_4.cpp:32,50,66,81 evaluating four times per step).  With
 * enabled = 1, the step functions above (and nbfmm_run_dev) capture one whole
 * step -- every evaluation of it (tree, lists, expansions, P2P on both
 * streams) and the integrator updates -- as a graph, and replay it for the
 * following steps with the same arrays, n, dt, parameters and tiles.  A
 * replayed evaluation reads no size back to the host: its launches and
 * buffers are sized by capacities a quarter above the sizes of the last
 * direct evaluation, and the kernels read the actual sizes on the device.
 * When a size exceeds its capacity (or the tree depth changes) the replayed
 * step leaves the state untouched, and the step runs again directly and is
 * captured anew.  Results equal the direct steps' to rounding (the near-field
 * sums meet in FP64 atomics).  The first step of a run, and every step after
 * a change, runs directly.  Requires order_select = NBFMM_ORDER_FIXED and
 * m2l_pass_pairs = 0 (NBFMM_ERR_UNSUPPORTED otherwise); graph steps take no
 * timing events, and the statistics (nbfmm_get_stats, nbfmm_tree_get) keep
 * describing the last direct evaluation. */
int nbfmm_set_graph(nbfmm_ctx *ctx, int enabled);
/* out[0] graphs captured, out[1] steps replayed, out[2] replayed steps whose
 * sizes overflowed and that ran again directly, out[3] steps run directly. */
int nbfmm_graph_stats(nbfmm_ctx *ctx, int64_t *out4);

/* Tile shapes and launch split of the P2P and M2L kernels (no counterpart
 * in the reference, whose serial code has none; the tuner in
 * nbody-fmm_amd/tune.py searches them per workload, as it searches the
 * order, the P2P cut and the leaf capacity that replace the reference's
 * hardcoded L and threshold at tree.cpp:32,94).  A zero (or, for the share,
 * a negative) field keeps the built-in default.  They change speed only: the
 * results agree to rounding (the near-field sums meet in FP64 atomics). */
typedef struct nbfmm_tiles {
  int32_t p2p_group_f32;      /* targets per source pass of the narrow FP32 P2P kernel: 8, 12, 16 (default) or 24 */
  int32_t p2p_group_f64;      /* the same for FP64: 4, 8 (default) or 12 */
  int32_t p2p_narrow_max_f32; /* blocks of at most this many targets go to the narrow kernel, larger ones to the
                                 four-targets-per-lane kernel: 1..64 (default 64) */
  int32_t p2p_narrow_max_f64; /* the same for FP64 (default 64) */
  double p2p_early_share_f32; /* share of the P2P tasks started beside the up-pass, in [0, 1] (default 0.5) */
  double p2p_early_share_f64; /* the same for FP64 (default 0.2) */
  int32_t p2p_early_launches_f32; /* launches of that share's narrow tasks, 1..64 (default 16) */
  int32_t p2p_early_launches_f64; /* the same for FP64 (default 1) */
  int32_t m2l_slices;         /* row-block slices of a class operator per M2L contraction workgroup: 1 = the
                                 fewest that fit the LDS (default), 2 = one more (half the LDS, more workgroups) */
  int32_t p2p_queue;          /* P2P scheduling: 0 (default) one workgroup per task; 1 a persistent launch over an
                                 atomic work queue (wide tasks first) for the share after the up-pass; 2 the same
                                 for the first share too (one launch per share, every system of a batch) */
  int32_t p2p_queue_waves0;   /* wavefronts per CU of the first share's persistent launch (p2p_queue 2), 1..16
                                 (default 8; the later share takes all the device holds, 16) */
  int32_t p2p_queue_chunk;    /* tasks a wavefront takes per atomic grab of the queue, 1..64 (default 4) */
  int32_t m2l_item_pairs;     /* most M2L pairs of one offset class per contraction workgroup (the workgroup
                                 stages the class operator once for them): a multiple of 128 in [128, 16384]
                                 (default 1024; items hold npair / 2048 pairs, at least 128, up to this) */
  int32_t p2p_rest_launches;  /* launches of the narrow P2P tasks not started beside the up-pass, 1..8: between
                                 two, the main stream's sums, L2L and L2P get compute units (default 4) */
  int32_t m2l_mode;           /* mixed-precision M2L: 1 (default) class-major items with per-pair partials and a
                                 per-target sum (always used in FP64 and with learned orders); 2 target-major, each
                                 workgroup keeps 128 targets' locals on chip over the 56 class operators up to
                                 reflection (no partials; slower on gfx950: an operator fills the LDS, leaving two
                                 wavefronts per SIMD to hide the multipole gathers) */
} nbfmm_tiles;
/* Tiles in force for later evaluations; fields are validated (NBFMM_ERR_ARG). */
int nbfmm_set_tiles(nbfmm_ctx *ctx, const nbfmm_tiles *tiles);
/* The tiles in force, every default resolved to its value. */
int nbfmm_get_tiles(nbfmm_ctx *ctx, nbfmm_tiles *tiles);

/* Statistics of the last evaluation: sizes for the roofline and per-phase
 * GPU times measured with HIP events on the context's stream.  The phases
 * follow each other on that stream; P2P runs on a second stream beside them
 * (beside the up-pass and after the first M2L contraction), so with the
 * default timing a phase's time includes waiting for CUs that P2P holds.
 * With nbfmm_set_timing(ctx, 2) the phases, P2P included, run one after
 * another, and each time is its kernels' own. */
typedef struct nbfmm_stats {
  int32_t depth;            /* leaf depth of the reference tree rule */
  int32_t depth_iterations; /* iterations of the depth loop (tree.cpp:43-69) */
  int64_t n_cells;          /* cells over all levels */
  int64_t n_leaves;
  int64_t n_il;             /* interaction-list entries over all levels */
  int64_t n_m2l;            /* entries evaluated by M2L */
  int64_t n_m2l_targets;    /* cells with at least one M2L source */
  int64_t n_m2l_sources;    /* cells used as an M2L source */
  int64_t n_p2p_pairs;      /* particle pairs evaluated directly (near + substituted) */
  int64_t n_p2m_terms;      /* particle-cell pairs of the P2M passes */
  int64_t n_l2p_terms;      /* particle-cell pairs of the L2P passes */
  double ms_tree;           /* bbox, depth loop, sort, levels, neighbour tables */
  double ms_lists;          /* interaction lists, classification, P2P lists */
  double ms_p2m;            /* up-pass planning, operator build, P2M */
  double ms_m2l;            /* M2L: class schedule, contraction, per-target sums */
  double ms_l2p;
  double ms_p2p;            /* waiting for P2P, then acc = G (far + near) */
  double ms_total;          /* first to last event of the evaluation */
  double ms_m2l_kernel;     /* the first pass's MFMA M2L contraction launches */
  double ms_p2p_kernel;     /* the P2P launches (on the second stream) */
  double flop_m2l;          /* algorithmic M2L flops: 2 (q+1)^4 per M2L pair of order q */
  double flop_p2p;          /* algorithmic P2P flops: 20 per particle pair */
  double mean_order;        /* mean order over the M2L pairs */
  int32_t fallback;         /* nbfmm_eval_accel_checked_dev fell back to the fixed order */
  int32_t isolated;         /* the phases ran one after another (timing mode 2) */
  int64_t n_p2m_cells;      /* cells forming their multipole from particles */
  int64_t n_m2m_cells;      /* cells forming their multipole from their children */
  int64_t n_l2l_cells;      /* cells pushing their local to their children */
  int64_t n_l2p_cells;      /* cells evaluating their local at their particles */
  double ms_m2m;            /* M2M, every level */
  double ms_l2l;            /* L2L, every level */
  int32_t depth_sorts;      /* key sorts the depth loop ran: one when the previous evaluation's depth was
                               confirmed from every level's fullest cell, else one per iteration (or one more) */
  int32_t reserved1;
} nbfmm_stats;
int nbfmm_get_stats(nbfmm_ctx *ctx, nbfmm_stats *stats);
/* Per-phase events: 0 off, 1 on (the default), 2 on with the phases run one
 * after another (P2P alone after L2P) so that each phase is timed alone;
 * mode 2 is for measurement and is slower than mode 1. */
int nbfmm_set_timing(nbfmm_ctx *ctx, int enabled);
/* Keeps (1) the full interaction lists of later evaluations for
 * nbfmm_tree_get (off by default: only the lists the evaluation needs are built). */
int nbfmm_keep_lists(nbfmm_ctx *ctx, int enabled);

/* Tree of the last evaluation, for inspection and parity tests (host copies).
 * nbfmm_tree_sizes returns the depth, the number of cells and list entries.
 * nbfmm_tree_get fills (any pointer may be NULL):
 *   level_start[depth+2]  cells of level d are [level_start[d], level_start[d+1])
 *   key[n_cells]          Morton key of each cell at its level (64-bit)
 *   pbeg/pend[n_cells]    particle range in sorted order
 *   parent[n_cells]       parent cell (-1 for the root)
 *   perm[n]               sorted position -> input index (stable: input order inside a leaf)
 *   nbr_off[n_cells+1], nbr[n_nbr]  neighbour lists (node_i.cpp:23-99 order)
 *   il_off[n_cells+1], il[n_il]     interaction lists (node_i.cpp:23-99 order) */
int nbfmm_tree_sizes(nbfmm_ctx *ctx, int32_t *depth, int64_t *n_cells, int64_t *n_nbr,
                     int64_t *n_il);
int nbfmm_tree_get(nbfmm_ctx *ctx, int64_t *level_start, uint64_t *key, int32_t *pbeg,
                   int32_t *pend, int32_t *parent, int32_t *perm, int32_t *nbr_off, int32_t *nbr,
                   int32_t *il_off, int32_t *il);

/* Message of the last failure on this thread ("" if none). */
const char *nbfmm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
