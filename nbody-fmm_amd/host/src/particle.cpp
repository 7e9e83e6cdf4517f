// particle.cpp -- Particle (interface of src/particle.hpp:15-69).
#include "nbfmm/particle.hpp"

#include "nbfmm/system_update_method.hpp"

Particle::Particle(double w, Tensor &&pos, Tensor &&vel)
    : weight(w), position(pos), velocity(vel), acceleration(0 * position) {}

const Tensor &Particle::compute_new_accelaration(const Tensor &grad) const {
  acceleration = GRAV_CONST * grad;
  return acceleration;
}
