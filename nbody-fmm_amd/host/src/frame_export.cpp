// frame_export.cpp -- see nbfmm/frame_export.hpp.
#include "nbfmm/frame_export.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace nbfmm_io {
namespace {

constexpr const char *kExt = ".bin";

void write_all(std::FILE *f, const double *p, size_t count, const std::string &path) {
  if (count && std::fwrite(p, sizeof(double), count, f) != count)
    throw std::runtime_error("short write to " + path);
}

std::string dir_of(const std::string &path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

// value of attribute `name` in the tag text `tag`
std::string attr(const std::string &tag, const std::string &name) {
  const std::string key = name + "=\"";
  const size_t a = tag.find(key);
  if (a == std::string::npos) return "";
  const size_t b = tag.find('"', a + key.size());
  return tag.substr(a + key.size(), b - a - key.size());
}

} // namespace

FrameExporter::FrameExporter(const std::string &dir, const std::string &base)
    : dir_(dir), base_(base), xmf_path_(dir + "/" + base + ".xmf") {
  if (::mkdir(dir_.c_str(), 0755) != 0 && errno != EEXIST)
    throw std::runtime_error("cannot create directory " + dir_ + ": " + std::strerror(errno));
}

void FrameExporter::write_frame(const std::vector<Particle> &particles, int step, double time) {
  const size_t n = particles.size();
  if (n == 0) return;
  std::vector<double> pos(n * 3, 0.0), vel(n * 3, 0.0), acc(n * 3, 0.0), w(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const Particle &p = particles[i];
    // 2D states are written with z = 0 (hdf5_exporter.cpp:29-44)
    for (size_t d = 0; d < 3; ++d) {
      pos[3 * i + d] = d < p.get_position().dim ? p.get_position()[d] : 0.0;
      vel[3 * i + d] = d < p.get_velocity().dim ? p.get_velocity()[d] : 0.0;
      acc[3 * i + d] = d < p.get_acceleration().dim ? p.get_acceleration()[d] : 0.0;
    }
    w[i] = p.get_weight();
  }
  write_frame(n, pos.data(), vel.data(), acc.data(), w.data(), step, time);
}

void FrameExporter::write_frame(size_t n, const double *pos, const double *vel, const double *acc,
                                const double *weight, int step, double time) {
  if (n == 0) return;
  std::ostringstream ss;
  ss << base_ << "_" << std::setfill('0') << std::setw(5) << step << kExt;
  const std::string file = ss.str(), path = dir_ + "/" + file;
  std::FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot create " + path + ": " + std::strerror(errno));
  try {
    write_all(f, pos, n * 3, path);
    write_all(f, vel, n * 3, path);
    write_all(f, acc, n * 3, path);
    write_all(f, weight, n, path);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  if (std::fclose(f) != 0) throw std::runtime_error("cannot close " + path);
  times_.push_back(time);
  files_.push_back(file);
  write_index(n);
}

void FrameExporter::write_index(size_t n) const {
  std::ofstream x(xmf_path_);
  if (!x) throw std::runtime_error("cannot write " + xmf_path_);
  x << std::setprecision(std::numeric_limits<double>::max_digits10);
  const size_t b3 = n * 3 * sizeof(double);
  auto item = [&](const char *dims, const std::string &file, size_t seek) {
    x << "       <DataItem Dimensions=\"" << dims << "\" NumberType=\"Float\" Precision=\"8\" Format=\"Binary\""
      << " Endian=\"Little\" Seek=\"" << seek << "\">\n         " << file << "\n       </DataItem>\n";
  };
  const std::string d3 = std::to_string(n) + " 3", d1 = std::to_string(n);
  x << "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n <Domain>\n";
  x << "  <Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
  for (size_t t = 0; t < times_.size(); ++t) {
    x << "   <Grid Name=\"Step_" << t << "\" GridType=\"Uniform\">\n";
    x << "     <Time Value=\"" << times_[t] << "\"/>\n";
    x << "     <Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" << n << "\"/>\n";
    x << "     <Geometry GeometryType=\"XYZ\">\n";
    item(d3.c_str(), files_[t], 0);
    x << "     </Geometry>\n";
    x << "     <Attribute Name=\"Velocity\" AttributeType=\"Vector\" Center=\"Node\">\n";
    item(d3.c_str(), files_[t], b3);
    x << "     </Attribute>\n";
    x << "     <Attribute Name=\"Acceleration\" AttributeType=\"Vector\" Center=\"Node\">\n";
    item(d3.c_str(), files_[t], 2 * b3);
    x << "     </Attribute>\n";
    x << "     <Attribute Name=\"Weight\" AttributeType=\"Scalar\" Center=\"Node\">\n";
    item(d1.c_str(), files_[t], 3 * b3);
    x << "     </Attribute>\n";
    x << "   </Grid>\n";
  }
  x << "  </Grid>\n </Domain>\n</Xdmf>\n";
  if (!x) throw std::runtime_error("cannot write " + xmf_path_);
}

std::vector<std::pair<std::string, double>> read_index(const std::string &xmf_path) {
  std::ifstream in(xmf_path);
  if (!in) throw std::runtime_error("cannot read " + xmf_path);
  std::stringstream buf;
  buf << in.rdbuf();
  const std::string s = buf.str();
  std::vector<std::pair<std::string, double>> out;
  size_t at = 0;
  for (;;) {
    const size_t t = s.find("<Time Value=\"", at);
    if (t == std::string::npos) break;
    const double time = std::stod(attr(s.substr(t, s.find('>', t) - t), "Value"));
    // the geometry item of this grid names the frame file
    const size_t g = s.find("<DataItem", t);
    const size_t b = s.find('>', g) + 1, e = s.find("</DataItem>", b);
    std::string file = s.substr(b, e - b);
    file.erase(0, file.find_first_not_of(" \t\r\n"));
    file.erase(file.find_last_not_of(" \t\r\n") + 1);
    out.emplace_back(file, time);
    at = e;
  }
  return out;
}

Frame read_frame(const std::string &xmf_path, int index) {
  const auto frames = read_index(xmf_path);
  const int k = index < 0 ? (int)frames.size() + index : index;
  if (k < 0 || k >= (int)frames.size()) throw std::out_of_range("frame index out of range");
  std::ifstream in(xmf_path);
  std::stringstream buf;
  buf << in.rdbuf();
  const std::string s = buf.str();
  const size_t tp = s.find("TopologyType=\"Polyvertex\"");
  const size_t n = std::stoull(attr(s.substr(tp, s.find('>', tp) - tp), "NumberOfElements"));
  const std::string path = dir_of(xmf_path) + "/" + frames[k].first;
  std::FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot read " + path);
  Frame fr;
  fr.time = frames[k].second;
  // the step number is the file name's 5-digit suffix
  const size_t us = frames[k].first.find_last_of('_'), dot = frames[k].first.find_last_of('.');
  fr.step = std::stoi(frames[k].first.substr(us + 1, dot - us - 1));
  fr.position.resize(n * 3);
  fr.velocity.resize(n * 3);
  fr.acceleration.resize(n * 3);
  fr.weight.resize(n);
  const bool ok = std::fread(fr.position.data(), sizeof(double), n * 3, f) == n * 3 &&
                  std::fread(fr.velocity.data(), sizeof(double), n * 3, f) == n * 3 &&
                  std::fread(fr.acceleration.data(), sizeof(double), n * 3, f) == n * 3 &&
                  std::fread(fr.weight.data(), sizeof(double), n, f) == n;
  std::fclose(f);
  if (!ok) throw std::runtime_error("short frame file " + path);
  return fr;
}

} // namespace nbfmm_io

// Particle's acceleration is written by the updaters only; restoring a frame
// sets it through this friend.
struct FrameRestore {
  static void set(const Particle &p, const double a[3]) { p.acceleration = Tensor{a[0], a[1], a[2]}; }
};

namespace nbfmm_io {
std::vector<Particle> particles_from_frame(const Frame &f) {
  std::vector<Particle> ps;
  ps.reserve(f.size());
  for (size_t i = 0; i < f.size(); ++i) {
    ps.emplace_back(f.weight[i], Tensor{f.position[3 * i], f.position[3 * i + 1], f.position[3 * i + 2]},
                    Tensor{f.velocity[3 * i], f.velocity[3 * i + 1], f.velocity[3 * i + 2]});
    FrameRestore::set(ps.back(), &f.acceleration[3 * i]);
  }
  return ps;
}
} // namespace nbfmm_io
