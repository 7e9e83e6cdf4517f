// tensor.hpp -- small dense vector type of the particle API.
// Same interface as the reference's Tensor (src/math_utils/tensor.hpp:18-218):
// runtime dimension, element access, +, -, unary -, scalar * and /, compound
// assignment, ==, squared_norm, norm, stream output.  Storage is inline (no
// heap) for dimensions up to 3 (the reference's std::vector takes any; a
// larger dimension throws std::invalid_argument here).  Errors throw instead of ending the process:
// out-of-range access -> std::out_of_range (reference: std::exit(1),
// tensor.cpp:47-63); dimension mismatch -> std::invalid_argument (reference:
// std::abort, tensor.cpp:65-69).
#pragma once
#include <cstddef>
#include <initializer_list>
#include <ostream>


class Tensor {
 public:
  static constexpr size_t kMaxDim = 3;
  Tensor() = default;
  Tensor(size_t dim); // implicit, as the reference's (tensor.hpp:35)
  Tensor(std::initializer_list<double> init);

  const double &operator[](size_t i) const;
  double &operator[](size_t i);

  Tensor operator+(const Tensor &o) const;
  Tensor operator-(const Tensor &o) const;
  Tensor operator-() const;
  Tensor &operator+=(const Tensor &o);
  Tensor &operator-=(const Tensor &o);
  Tensor &operator*=(double s);
  Tensor &operator/=(double s);
  Tensor operator*(double s) const;
  Tensor operator/(double s) const;
  bool operator==(const Tensor &o) const;

  double squared_norm() const;
  double norm() const;

  size_t dim = 0;

 private:
  void same_dim(const Tensor &o) const;
  double data_[kMaxDim] = {0.0, 0.0, 0.0};
};

Tensor operator*(double s, const Tensor &v);
std::ostream &operator<<(std::ostream &os, const Tensor &v);
