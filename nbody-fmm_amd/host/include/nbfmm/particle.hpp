// particle.hpp -- body state of the particle API.
// Interface of the reference's Particle (src/particle.hpp:15-69): weight,
// position, velocity and a mutable acceleration written by the updaters.
#pragma once
#include "nbfmm/tensor.hpp"

class QuadratureMethod;
class GpuUpdater;
struct FrameRestore;

class Particle {
 public:
  Particle() = default;
  Particle(double weight, Tensor &&pos, Tensor &&vel);
  // acceleration = GRAV_CONST * grad (particle.cpp:16-21)
  const Tensor &compute_new_accelaration(const Tensor &potential_gradient) const;
  const Tensor &get_acceleration() const { return acceleration; }
  double get_weight() const { return weight; }
  const Tensor &get_position() const { return position; }
  const Tensor &get_velocity() const { return velocity; }

  friend class QuadratureMethod;
  friend class GpuUpdater;
  friend struct FrameRestore; // restores the acceleration of a saved frame

 protected:
  double weight = 0.0;
  Tensor position;
  Tensor velocity;
  mutable Tensor acceleration;
};
