// OperatorValue: the value an operator returns (src/operator/src/traits.rs:
// 14-260), Scalar, Vector or Matrix, with the reference's arithmetic kept as
// it is, quirks included:
//   Vector / Scalar = Scalar ./ Vector   (traits.rs:150, 245)
//   Vector - Scalar = Scalar - Vector, Vector - Vector reversed (traits.rs:102-103, 197-198)
//   Vector op Scalar for + and * commutes; unimplemented combinations panic.
#pragma once

#include <cmath>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "errors.hpp"

namespace mole {

struct Matrix {
  size_t rows = 0, cols = 0;
  std::vector<double> data;  // row-major
  double &operator()(size_t i, size_t j) { return data[i * cols + j]; }
  double operator()(size_t i, size_t j) const { return data[i * cols + j]; }
};

class OperatorValue {
 public:
  enum class Kind { Scalar, Vector, Matrix };

  OperatorValue() : OperatorValue(Kind::Scalar, {0.0}, 1, 1) {}  // Scalar(0.0)

  static OperatorValue scalar(double v) { return OperatorValue(Kind::Scalar, {v}, 1, 1); }
  static OperatorValue vector(std::vector<double> v) {
    const size_t n = v.size();
    return OperatorValue(Kind::Vector, std::move(v), n, 1);
  }
  static OperatorValue matrix(Matrix m) { return OperatorValue(Kind::Matrix, std::move(m.data), m.rows, m.cols); }

  Kind kind() const { return kind_; }
  // OperatorValue::get_scalar / get_vector / get_matrix (traits.rs:24-59)
  Result<double> get_scalar() const {
    if (kind_ != Kind::Scalar) return Error{ErrorKind::OperatorValueAccessError, "not a scalar"};
    return v_[0];
  }
  Result<std::vector<double>> get_vector() const {
    if (kind_ != Kind::Vector) return Error{ErrorKind::OperatorValueAccessError, "not a vector"};
    return v_;
  }
  Result<Matrix> get_matrix() const {
    if (kind_ != Kind::Matrix) return Error{ErrorKind::OperatorValueAccessError, "not a matrix"};
    return Matrix{rows_, cols_, v_};
  }
  // OperatorValue::map (traits.rs:61-63)
  OperatorValue map(const std::function<double(double)> &f) const {
    OperatorValue r = *this;
    for (double &x : r.v_) x = f(x);
    return r;
  }

  friend OperatorValue operator+(const OperatorValue &a, const OperatorValue &b) {
    return combine(a, b, '+', [](double x, double y) { return x + y; }, false);
  }
  friend OperatorValue operator-(const OperatorValue &a, const OperatorValue &b) {
    return combine(a, b, '-', [](double x, double y) { return x - y; }, true);
  }
  friend OperatorValue operator*(const OperatorValue &a, const OperatorValue &b) {
    return combine(a, b, '*', [](double x, double y) { return x * y; }, false);
  }
  friend OperatorValue operator/(const OperatorValue &a, const OperatorValue &b) {
    return combine(a, b, '/', [](double x, double y) { return x / y; }, true);
  }

 private:
  OperatorValue(Kind k, std::vector<double> v, size_t r, size_t c) : kind_(k), v_(std::move(v)), rows_(r), cols_(c) {}

  static OperatorValue like(const OperatorValue &shape, std::vector<double> v) {
    return OperatorValue(shape.kind_, std::move(v), shape.rows_, shape.cols_);
  }

  // reversed: the operation is not commutative and the reference swaps the
  // operands of Vector/Matrix op Scalar and Vector - Vector
  template <typename F>
  static OperatorValue combine(const OperatorValue &a, const OperatorValue &b, char op, F f, bool reversed) {
    using K = Kind;
    std::vector<double> out;
    if (a.kind_ == K::Scalar && b.kind_ == K::Scalar) return scalar(f(a.v_[0], b.v_[0]));
    if (a.kind_ == K::Scalar) {  // Scalar op Vector/Matrix: elementwise s op x
      for (double x : b.v_) out.push_back(f(a.v_[0], x));
      return like(b, std::move(out));
    }
    if (b.kind_ == K::Scalar) {  // Vector/Matrix op Scalar: the reference computes s op x
      for (double x : a.v_) out.push_back(reversed ? f(b.v_[0], x) : f(x, b.v_[0]));
      return like(a, std::move(out));
    }
    if (a.kind_ == b.kind_ && a.v_.size() == b.v_.size()) {
      // Vector: +, - (reversed), * ; Matrix: + only; the rest is unimplemented!()
      if (op == '/' || (a.kind_ == K::Matrix && op != '+')) throw Panic("not implemented");
      for (size_t i = 0; i < a.v_.size(); ++i) out.push_back(reversed ? f(b.v_[i], a.v_[i]) : f(a.v_[i], b.v_[i]));
      return like(a, std::move(out));
    }
    throw Panic("not implemented");  // unimplemented!() in traits.rs
  }

  Kind kind_;
  std::vector<double> v_;
  size_t rows_, cols_;
};

// HashMap<String, Vec<OperatorValue>> and HashMap<String, OperatorValue>
using DataMap = std::map<std::string, std::vector<OperatorValue>>;
using AverageMap = std::map<std::string, OperatorValue>;

}  // namespace mole
