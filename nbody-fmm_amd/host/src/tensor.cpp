// tensor.cpp -- Tensor (interface of src/math_utils/tensor.hpp:18-218).
#include "nbfmm/tensor.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

Tensor::Tensor(size_t d) : dim(d) {
  if (d > kMaxDim) throw std::invalid_argument("Tensor dimension above 3");
}

Tensor::Tensor(std::initializer_list<double> init) : Tensor(init.size()) {
  size_t i = 0;
  for (double v : init) data_[i++] = v;
}

const double &Tensor::operator[](size_t i) const {
  if (i >= dim) throw std::out_of_range("Tensor index " + std::to_string(i) + " outside dimension " + std::to_string(dim));
  return data_[i];
}
double &Tensor::operator[](size_t i) {
  if (i >= dim) throw std::out_of_range("Tensor index " + std::to_string(i) + " outside dimension " + std::to_string(dim));
  return data_[i];
}

void Tensor::same_dim(const Tensor &o) const {
  if (o.dim != dim) throw std::invalid_argument("Tensor dimensions differ");
}

Tensor &Tensor::operator+=(const Tensor &o) {
  same_dim(o);
  for (size_t i = 0; i < dim; ++i) data_[i] += o.data_[i];
  return *this;
}
Tensor &Tensor::operator-=(const Tensor &o) {
  same_dim(o);
  for (size_t i = 0; i < dim; ++i) data_[i] -= o.data_[i];
  return *this;
}
Tensor &Tensor::operator*=(double s) {
  for (size_t i = 0; i < dim; ++i) data_[i] *= s;
  return *this;
}
Tensor &Tensor::operator/=(double s) {
  for (size_t i = 0; i < dim; ++i) data_[i] /= s;
  return *this;
}
Tensor Tensor::operator+(const Tensor &o) const {
  Tensor r(*this);
  return r += o;
}
Tensor Tensor::operator-(const Tensor &o) const {
  Tensor r(*this);
  return r -= o;
}
Tensor Tensor::operator-() const {
  Tensor r(dim);
  for (size_t i = 0; i < dim; ++i) r.data_[i] = -data_[i];
  return r;
}
Tensor Tensor::operator*(double s) const {
  Tensor r(*this);
  return r *= s;
}
Tensor Tensor::operator/(double s) const {
  Tensor r(*this);
  return r /= s;
}
bool Tensor::operator==(const Tensor &o) const {
  if (dim != o.dim) return false;
  for (size_t i = 0; i < dim; ++i)
    if (!(data_[i] == o.data_[i])) return false;
  return true;
}
double Tensor::squared_norm() const {
  double r = 0.0;
  for (size_t i = 0; i < dim; ++i) r += data_[i] * data_[i];
  return r;
}
double Tensor::norm() const { return std::sqrt(squared_norm()); }

Tensor operator*(double s, const Tensor &v) { return v * s; }

std::ostream &operator<<(std::ostream &os, const Tensor &v) {
  os << "(";
  for (size_t i = 0; i < v.dim; ++i) os << (i ? ", " : "") << v[i];
  return os << ")";
}
