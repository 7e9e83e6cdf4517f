// quadratures.cpp -- SyplecticEuler and RungeKutta4 over any SystemUpdateMethod.
#include "nbfmm/quadratures.hpp"

// symplectic_euler.cpp:8-15
void SyplecticEuler::integrate(SystemUpdateMethod *method, std::vector<Particle> &P, const double &dt) {
  method->update(P);
  for (Particle &p : P) {
    velocity(p) += dt * p.get_acceleration();
    position(p) += dt * velocity(p);
  }
}

// runge_kutta_4.cpp:10-93 (reference semantics) or classic RK4.
void RungeKutta4::integrate(SystemUpdateMethod *method, std::vector<Particle> &P, const double &dt) {
  const size_t n = P.size();
  std::vector<Particle> pk(P);
  std::vector<Tensor> vk(n), asum(n), vsum(n), a1(n);
  for (size_t i = 0; i < n; ++i) vk[i] = P[i].get_velocity();
  method->update(pk);
  for (size_t i = 0; i < n; ++i) {
    asum[i] = pk[i].get_acceleration();
    vsum[i] = vk[i];
    a1[i] = pk[i].get_acceleration();
  }
  for (int stage = 2; stage <= 4; ++stage) {
    const double h = stage == 4 ? dt : dt / 2;
    // acceleration used for the stage velocity: the carried one of the input
    // particles (reference) or the previous stage's (classic)
    std::vector<Tensor> avel(n);
    for (size_t i = 0; i < n; ++i) avel[i] = classic_ ? pk[i].get_acceleration() : P[i].get_acceleration();
    pk = P;
    for (size_t i = 0; i < n; ++i) position(pk[i]) += h * vk[i];
    for (size_t i = 0; i < n; ++i) vk[i] = P[i].get_velocity() + h * avel[i];
    method->update(pk);
    for (size_t i = 0; i < n; ++i) {
      if (stage == 4) {
        asum[i] += pk[i].get_acceleration();
        vsum[i] += vk[i];
      } else {
        asum[i] += 2 * pk[i].get_acceleration();
        vsum[i] += 2 * vk[i];
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    velocity(P[i]) += dt / 6 * asum[i];
    position(P[i]) += dt / 6 * vsum[i];
    if (classic_) acceleration(P[i]) = a1[i];
  }
}
