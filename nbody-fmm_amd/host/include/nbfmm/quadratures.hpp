// quadratures.hpp -- the integrators that drive an updater
// (src/quadratures/quadrature_method.hpp:16-47, symplectic_euler.hpp,
// runge_kutta_4.hpp).
#pragma once
#include <vector>

#include "nbfmm/system_update_method.hpp"

class QuadratureMethod {
 public:
  virtual ~QuadratureMethod() = default;
  virtual void integrate(SystemUpdateMethod *method, std::vector<Particle> &particles,
                         const double &delta_t) = 0;

 protected:
  Tensor &position(Particle &p) { return p.position; }
  Tensor &velocity(Particle &p) { return p.velocity; }
  Tensor &acceleration(Particle &p) { return p.acceleration; }
};

// v += dt * a; x += dt * v after one update (symplectic_euler.cpp:8-15).
class SyplecticEuler : public QuadratureMethod {
 public:
  void integrate(SystemUpdateMethod *method, std::vector<Particle> &particles,
                 const double &delta_t) override;
};

// Reference semantics by default (runge_kutta_4.cpp:10-93): every stage
// starts from a copy of the input particles, so stage velocities use the
// acceleration the particles carry into the step, and that carried
// acceleration is left unchanged.  classic = true gives textbook RK4.
class RungeKutta4 : public QuadratureMethod {
 public:
  explicit RungeKutta4(bool classic = false) : classic_(classic) {}
  void integrate(SystemUpdateMethod *method, std::vector<Particle> &particles,
                 const double &delta_t) override;

 private:
  bool classic_;
};
