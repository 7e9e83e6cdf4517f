// test_frames.cpp -- frame export and read-back (no GPU): the datasets of the
// reference's HDF5Exporter (src/io/hdf5_exporter.cpp:15-128) written and read
// back bitwise, the XDMF index listing every frame, 2D states padded with
// z = 0 as the reference pads them.
// Usage: test_frames <output dir>.  Prints PASSED and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "nbfmm/frame_export.hpp"

static int failures = 0;
#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
      ++failures;                        \
    }                                    \
  } while (0)

static std::vector<Particle> make(int step, size_t n) {
  std::vector<Particle> ps;
  for (size_t i = 0; i < n; ++i) {
    const double b = (double)i + 0.125 * step;
    ps.emplace_back(1.0 + 0.5 * (double)i, Tensor{b, -b, 2 * b}, Tensor{0.25 * b, 1.0 / (1.0 + b), -3.0});
  }
  return ps;
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : "frames_out";
  const size_t n = 37;
  {
    HDF5Exporter ex(dir, "nbody_fmm");
    for (int step = 0; step < 3; ++step) ex.write_frame(make(step, n), step * 10, 0.001 * step);
    ex.write_frame(std::vector<Particle>{}, 99, 1.0); // nothing written
  }
  const auto idx = nbfmm_io::read_index(dir + "/nbody_fmm.xmf");
  CHECK(idx.size() == 3, "index lists %zu frames", idx.size());
  for (int k = 0; k < 3 && k < (int)idx.size(); ++k) {
    char name[64];
    std::snprintf(name, sizeof(name), "nbody_fmm_%05d.bin", 10 * k);
    CHECK(idx[k].first == name, "frame %d file %s", k, idx[k].first.c_str());
    CHECK(idx[k].second == 0.001 * k, "frame %d time", k);
    const auto f = nbfmm_io::read_frame(dir + "/nbody_fmm.xmf", k);
    const auto ref = make(k, n);
    CHECK(f.size() == n && f.step == 10 * k, "frame %d size/step", k);
    for (size_t i = 0; i < n && f.size() == n; ++i)
      for (int d = 0; d < 3; ++d) {
        CHECK(f.position[3 * i + d] == ref[i].get_position()[d], "position %zu", i);
        CHECK(f.velocity[3 * i + d] == ref[i].get_velocity()[d], "velocity %zu", i);
      }
  }
  // the last frame restored as particles, acceleration included
  auto f = nbfmm_io::read_frame(dir + "/nbody_fmm.xmf");
  for (size_t i = 0; i < n; ++i) f.acceleration[3 * i + 1] = -(double)i;
  const auto ps = nbfmm_io::particles_from_frame(f);
  CHECK(ps.size() == n && ps[5].get_acceleration()[1] == -5.0 && ps[5].get_weight() == 3.5, "restore");
  // 2D particles: z written as 0
  {
    HDF5Exporter ex(dir, "plane");
    std::vector<Particle> p2;
    p2.emplace_back(2.0, Tensor{1.0, 2.0}, Tensor{3.0, 4.0});
    ex.write_frame(p2, 0, 0.0);
    const auto g = nbfmm_io::read_frame(dir + "/plane.xmf");
    CHECK(g.position[2] == 0.0 && g.velocity[1] == 4.0 && g.velocity[2] == 0.0 && g.weight[0] == 2.0, "2D frame");
  }
  std::printf(failures ? "FAILED (%d)\n" : "PASSED\n", failures);
  return failures ? 1 : 0;
}
