// frame_export.hpp -- trajectory frames for ParaView, and reading them back
// to resume a run.
//
// Interface of the reference's HDF5Exporter (src/io/hdf5_exporter.hpp:20-49,
// hdf5_exporter.cpp:10-128): write_frame(particles, step, time) stores the
// datasets "position" [n][3], "velocity" [n][3], "acceleration" [n][3] and
// "weight" [n] (float64) of one frame in <dir>/<base>_<step, 5 digits>.<ext>
// and rewrites the XDMF 2.0 index <dir>/<base>.xmf: one temporal collection
// of Polyvertex grids with the Geometry (XYZ) and the Velocity,
// Acceleration and Weight attributes, as the reference writes it.
//
// This image has no HDF5 library, so the frame container is raw
// little-endian float64 (extension .bin): the four datasets one after the
// other, referenced from the index as XDMF Binary data items with their
// byte offsets (Seek); ParaView's XDMF reader loads them as it loads the
// HDF5 items.  read_frame parses the index and a frame back for a restart.
#pragma once
#include <string>
#include <vector>

#include "nbfmm/particle.hpp"

namespace nbfmm_io {

struct Frame {
  int step = 0;
  double time = 0.0;
  std::vector<double> position, velocity, acceleration, weight; // n*3, n*3, n*3, n
  size_t size() const { return weight.size(); }
};

class FrameExporter {
 public:
  explicit FrameExporter(const std::string &dir = "output", const std::string &base = "particles");
  // hdf5_exporter.cpp:15-81 (empty particle sets write nothing, as there)
  void write_frame(const std::vector<Particle> &particles, int step, double time);
  // the same from plain arrays (x y z per particle), e.g. a device-resident state copied to the host
  void write_frame(size_t n, const double *pos, const double *vel, const double *acc, const double *weight, int step,
                   double time);
  const std::string &index_path() const { return xmf_path_; }

 private:
  void write_index(size_t n) const; // hdf5_exporter.cpp:83-128
  std::string dir_, base_, xmf_path_;
  std::vector<double> times_;
  std::vector<std::string> files_;
};

// Frames listed in an index written by FrameExporter: (file, time), in order.
std::vector<std::pair<std::string, double>> read_index(const std::string &xmf_path);
// Frame `index` of the index (negative: counted from the end, -1 = last).
Frame read_frame(const std::string &xmf_path, int index = -1);
// Particles of a frame, acceleration included (the carried acceleration of
// RungeKutta4's reference semantics).
std::vector<Particle> particles_from_frame(const Frame &f);

} // namespace nbfmm_io

// the reference's class name
using HDF5Exporter = nbfmm_io::FrameExporter;
