// system_update_method.hpp -- the updater plugin interface
// (src/system_updaters/system_update_method.hpp:17-31).
#pragma once
#include <vector>

#include "nbfmm/particle.hpp"

constexpr double GRAV_CONST = 6.6743e-11;

class SystemUpdateMethod {
 public:
  virtual ~SystemUpdateMethod() = default;
  // Computes every particle's acceleration and stores it in the particle.
  virtual void update(const std::vector<Particle> &particles) = 0;
};
