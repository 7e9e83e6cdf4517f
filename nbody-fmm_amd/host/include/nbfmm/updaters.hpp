// updaters.hpp -- the GPU updaters behind SystemUpdateMethod.
//   FMM          replaces FMM::update (src/system_updaters/FMM.cpp:9-12)
//   NaiveUpdate  replaces NaiveUpdate::update (src/system_updaters/naive_update.cpp:9-35)
// Both own an nbfmm context (device buffers and a HIP stream) and call the
// C ABI in include/nbfmm.h; a nonzero status becomes std::runtime_error.
// FMM uses the reference tree rule and correct solid-harmonic operators
// (the reference's operators do not converge to the direct sum).
#pragma once
#include <memory>
#include <vector>

#include "../../../../include/nbfmm.h"
#include "nbfmm/system_update_method.hpp"

class GpuUpdater : public SystemUpdateMethod {
 public:
  explicit GpuUpdater(const nbfmm_params &p);
  ~GpuUpdater() override;
  GpuUpdater(const GpuUpdater &) = delete;
  GpuUpdater &operator=(const GpuUpdater &) = delete;
  nbfmm_ctx *context() const { return ctx_; }
  nbfmm_stats stats() const;

 protected:
  void pack(const std::vector<Particle> &particles);
  void unpack(const std::vector<Particle> &particles);
  static void check(int status);
  nbfmm_ctx *ctx_ = nullptr;
  std::vector<double> xyz_, m_, acc_;
};

class FMM : public GpuUpdater {
 public:
  // Reference settings: order 2, leaf threshold 20, G = GRAV_CONST, FP64.
  FMM();
  explicit FMM(const nbfmm_params &p);
  void update(const std::vector<Particle> &particles) override;
};

class NaiveUpdate : public GpuUpdater {
 public:
  // Reference settings: G = GRAV_CONST, softening 1e-4 (naive_update.cpp:27).
  NaiveUpdate();
  explicit NaiveUpdate(const nbfmm_params &p);
  void update(const std::vector<Particle> &particles) override;
};
