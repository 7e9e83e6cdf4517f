// test_dropin.cpp -- the reference's own tests, run unchanged in spirit against
// the drop-in classes (the GPU FMM / NaiveUpdate behind SystemUpdateMethod).
//   TwoParticlesGetCloserUnderGravity   tests/test_naive_update.cpp:13-55
//   CompareAccelerations3D              tests/test_fmm_vs_naive_3d.cpp:17-208
//   RK4 reference semantics / classic   src/quadratures/runge_kutta_4.cpp:10-93
//   frames written every step (src/main.cpp:78-93) and a run resumed from one
// Prints one line per check and exits nonzero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nbfmm/frame_export.hpp"
#include "nbfmm/quadratures.hpp"
#include "nbfmm/updaters.hpp"

static int failures = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      std::printf("FAIL: " __VA_ARGS__);      \
      std::printf("\n");                      \
      ++failures;                             \
    }                                         \
  } while (0)

static void two_particles() {
  double mass = 1.0e9;
  Particle p1(mass, {-10.0, 0.0, 0.0}, {0.0, 0.0, 0.0});
  Particle p2(mass, {+10.0, 0.0, 0.0}, {0.0, 0.0, 0.0});
  std::vector<Particle> particles = {p1, p2};
  double prev = (particles[1].get_position() - particles[0].get_position()).norm();
  CHECK(prev == 20.0, "initial distance");
  NaiveUpdate solver;
  solver.update(particles);
  SyplecticEuler integrator;
  for (int step = 1; step <= 20; ++step) {
    integrator.integrate(&solver, particles, 0.5);
    double cur = (particles[1].get_position() - particles[0].get_position()).norm();
    CHECK(cur < prev, "distance must decrease at step %d", step);
    CHECK(particles[0].get_position()[0] > -10.0, "particle 1 moves right");
    CHECK(particles[1].get_position()[0] < 10.0, "particle 2 moves left");
    prev = cur;
  }
  std::printf("two_particles: final distance %.12f\n", prev);
}

static void fmm_vs_naive() {
  const unsigned N = 100;
  const double B = 50.0, M = 1.0e9;
  srand(42);
  std::vector<Particle> pn;
  for (unsigned i = 0; i < N; ++i) {
    double x = (rand() * 2.0 * B) / RAND_MAX - B;
    double y = (rand() * 2.0 * B) / RAND_MAX - B;
    double z = (rand() * 2.0 * B) / RAND_MAX - B;
    double vx = (rand() * 2.0) / RAND_MAX - 1.0;
    double vy = (rand() * 2.0) / RAND_MAX - 1.0;
    double vz = (rand() * 2.0) / RAND_MAX - 1.0;
    pn.push_back(Particle(M, {x, y, z}, {vx, vy, vz}));
  }
  std::vector<Particle> pf = pn;
  NaiveUpdate naive;
  naive.update(pn);
  FMM fmm;  // reference settings: order 2, leaf threshold 20
  fmm.update(pf);
  double maxerr = 0;
  for (unsigned i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) {
      double e = std::fabs(pf[i].get_acceleration()[d] - pn[i].get_acceleration()[d]);
      maxerr = e > maxerr ? e : maxerr;
      CHECK(e <= 1.0e-2, "particle %u component %d: |a_fmm - a_naive| = %g", i, d, e);
    }
  // order 10 with the reference's tree rule: far field converged
  nbfmm_params p;
  nbfmm_default_params(&p);
  p.order = 10;
  p.G = GRAV_CONST;
  FMM fmm10(p);
  std::vector<Particle> p10 = pn;
  fmm10.update(p10);
  double num = 0, den = 0;
  for (unsigned i = 0; i < N; ++i) {
    num += (p10[i].get_acceleration() - pn[i].get_acceleration()).squared_norm();
    den += pn[i].get_acceleration().squared_norm();
  }
  std::printf("fmm_vs_naive: max |a_fmm(p=2) - a_naive| = %.3e (bound 1e-2), rel L2 (p=10) = %.3e\n", maxerr,
              std::sqrt(num / den));
  CHECK(std::sqrt(num / den) < 1e-4, "order-10 FMM vs NaiveUpdate");
}

static void rk4_modes() {
  std::vector<Particle> P;
  srand(7);
  for (int i = 0; i < 64; ++i) {
    double x = rand() / (double)RAND_MAX, y = rand() / (double)RAND_MAX, z = rand() / (double)RAND_MAX;
    P.push_back(Particle(1.0, {x, y, z}, {0.0, 0.0, 0.0}));
  }
  nbfmm_params p;
  nbfmm_default_params(&p);
  p.G = 1.0;
  p.eps2 = 0.0;
  NaiveUpdate solver(p);
  solver.update(P);  // particles carry a(x0)
  std::vector<Particle> Pr = P, Pc = P;
  RungeKutta4 ref_rk4, classic(true);
  const double dt = 1e-3;
  ref_rk4.integrate(&solver, Pr, dt);
  classic.integrate(&solver, Pc, dt);
  double dref = 0, dcls = 0;
  for (size_t i = 0; i < P.size(); ++i) {
    // reference semantics: x1 = x0 + dt v0 + dt^2/2 a_in, exactly in this arithmetic
    Tensor want = P[i].get_position() + dt / 6 * (P[i].get_velocity() + 2 * (P[i].get_velocity() + dt / 2 * P[i].get_acceleration()) +
                                                  2 * (P[i].get_velocity() + dt / 2 * P[i].get_acceleration()) +
                                                  (P[i].get_velocity() + dt * P[i].get_acceleration()));
    dref = std::fmax(dref, (Pr[i].get_position() - want).norm());
    CHECK(Pr[i].get_acceleration() == P[i].get_acceleration(), "reference RK4 keeps the carried acceleration");
    dcls = std::fmax(dcls, (Pc[i].get_position() - Pr[i].get_position()).norm());
  }
  std::printf("rk4: reference-semantics position deviation %.3e, classic-vs-reference %.3e\n", dref, dcls);
  CHECK(dref == 0.0, "reference RK4 positions");
}

// main.cpp's loop (update, integrate, export) for 6 steps, then the run
// resumed from the frame after step 3 must reach the same state after 3 more
static void export_and_resume(const char *dir) {
  std::vector<Particle> P;
  srand(11);
  for (int i = 0; i < 500; ++i) {
    double x = rand() / (double)RAND_MAX, y = rand() / (double)RAND_MAX, z = rand() / (double)RAND_MAX;
    P.push_back(Particle(0.5 + rand() / (double)RAND_MAX, {x, y, z}, {0.0, 0.0, 0.0}));
  }
  nbfmm_params p;
  nbfmm_default_params(&p);
  p.order = 8;
  p.leaf_capacity = 16;
  p.G = 1.0;
  p.eps2 = 0.0;
  FMM fmm(p);
  RungeKutta4 rk4;
  const double dt = 1e-3;
  HDF5Exporter ex(dir, "nbody_fmm");
  fmm.update(P);
  for (int step = 0; step < 6; ++step) {
    ex.write_frame(P, step, step * dt);
    rk4.integrate(&fmm, P, dt);
  }
  ex.write_frame(P, 6, 6 * dt);
  auto f = nbfmm_io::read_frame(ex.index_path(), 3);
  CHECK(f.step == 3 && f.time == 3 * dt, "frame 3");
  std::vector<Particle> R = nbfmm_io::particles_from_frame(f);
  for (int step = 3; step < 6; ++step) rk4.integrate(&fmm, R, dt);
  double dev = 0, scale = 0;
  for (size_t i = 0; i < P.size(); ++i) {
    dev = std::fmax(dev, (R[i].get_position() - P[i].get_position()).norm());
    scale = std::fmax(scale, P[i].get_position().norm());
  }
  std::printf("export_and_resume: %zu frames, resumed-run position deviation %.3e\n",
              nbfmm_io::read_index(ex.index_path()).size(), dev);
  CHECK(nbfmm_io::read_index(ex.index_path()).size() == 7, "7 frames indexed");
  // the near-field sums run in any order on the GPU: equal to rounding
  CHECK(dev <= 1e-13 * scale, "resumed run");
}

int main(int argc, char **argv) {
  try {
    two_particles();
    fmm_vs_naive();
    rk4_modes();
    export_and_resume(argc > 1 ? argv[1] : "dropin_frames");
  } catch (const std::exception &e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 2;
  }
  std::printf(failures ? "FAILED (%d)\n" : "PASSED\n", failures);
  return failures ? 1 : 0;
}
