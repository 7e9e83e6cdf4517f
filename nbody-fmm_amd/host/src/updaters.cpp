// updaters.cpp -- FMM and NaiveUpdate over the C ABI of libnbfmm.so.
#include "nbfmm/updaters.hpp"

#include <stdexcept>
#include <string>

void GpuUpdater::check(int status) {
  if (status != NBFMM_OK)
    throw std::runtime_error("nbfmm error " + std::to_string(status) + ": " + nbfmm_last_error());
}

GpuUpdater::GpuUpdater(const nbfmm_params &p) { check(nbfmm_create(&p, &ctx_)); }

GpuUpdater::~GpuUpdater() { nbfmm_destroy(ctx_); }

nbfmm_stats GpuUpdater::stats() const {
  nbfmm_stats s{};
  check(nbfmm_get_stats(ctx_, &s));
  return s;
}

void GpuUpdater::pack(const std::vector<Particle> &P) {
  const size_t n = P.size();
  xyz_.resize(3 * n);
  m_.resize(n);
  acc_.resize(3 * n);
  for (size_t i = 0; i < n; ++i) {
    const Tensor &x = P[i].get_position();
    if (x.dim != 3) throw std::invalid_argument("only 3D particles are supported");
    xyz_[3 * i] = x[0];
    xyz_[3 * i + 1] = x[1];
    xyz_[3 * i + 2] = x[2];
    m_[i] = P[i].get_weight();
  }
}

void GpuUpdater::unpack(const std::vector<Particle> &P) {
  for (size_t i = 0; i < P.size(); ++i) P[i].acceleration = Tensor{acc_[3 * i], acc_[3 * i + 1], acc_[3 * i + 2]};
}

static nbfmm_params reference_params() {
  nbfmm_params p;
  nbfmm_default_params(&p);
  p.G = GRAV_CONST;
  return p;
}

FMM::FMM() : GpuUpdater(reference_params()) {}
FMM::FMM(const nbfmm_params &p) : GpuUpdater(p) {}

void FMM::update(const std::vector<Particle> &P) {
  if (P.empty()) return;
  pack(P);
  check(nbfmm_eval_accel(ctx_, (int64_t)P.size(), xyz_.data(), m_.data(), acc_.data()));
  unpack(P);
}

NaiveUpdate::NaiveUpdate() : GpuUpdater(reference_params()) {}
NaiveUpdate::NaiveUpdate(const nbfmm_params &p) : GpuUpdater(p) {}

void NaiveUpdate::update(const std::vector<Particle> &P) {
  if (P.empty()) return;
  pack(P);
  check(nbfmm_direct_accel(ctx_, (int64_t)P.size(), xyz_.data(), m_.data(), acc_.data()));
  unpack(P);
}
