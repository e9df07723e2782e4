// qudrip/matrix.hpp -- small dense complex matrix for gate algebra.
//
// Replaces Eigen::MatrixXcd (matrix.hpp:10 of the reference) where the
// drop-in API needs a dense matrix: gate definitions such as
// `Matrix Uc = 0.5 * (pauli_x + II * pauli_y)` (programs/spin.cpp:23) and
// the comma initialiser `(Matrix(2, 2) << 0, 1, 1, 0).finished()`
// (operators/bit.hpp:42-45).  Element arithmetic is std::complex<double>,
// the same operations Eigen performs element-wise.
#pragma once

#include <cassert>
#include <complex>
#include <cstddef>
#include <ostream>
#include <vector>

namespace qudrip {

using value_type = std::complex<double>;

class Matrix {
  std::size_t rows_ = 0, cols_ = 0;
  std::vector<value_type> d_;  // column-major

 public:
  Matrix() = default;
  Matrix(std::size_t r, std::size_t c) : rows_(r), cols_(c), d_(r * c, value_type(0.0, 0.0)) {}

  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  value_type& operator()(std::size_t i, std::size_t j) { return d_[j * rows_ + i]; }
  const value_type& operator()(std::size_t i, std::size_t j) const { return d_[j * rows_ + i]; }
  value_type* data() { return d_.data(); }
  const value_type* data() const { return d_.data(); }

  static Matrix Zero(std::size_t r, std::size_t c) { return Matrix(r, c); }
  static Matrix Identity(std::size_t n) {
    Matrix m(n, n);
    for (std::size_t i = 0; i < n; ++i) m(i, i) = value_type(1.0, 0.0);
    return m;
  }

  // comma initialiser, row-major fill order like Eigen's
  class Init {
    Matrix& m_;
    std::size_t k_;

   public:
    Init(Matrix& m, std::size_t k) : m_(m), k_(k) {}
    template <class T>
    Init& operator,(const T& v) {
      m_(k_ / m_.cols_, k_ % m_.cols_) = value_type(v);
      ++k_;
      return *this;
    }
    Matrix& finished() { return m_; }
  };
  template <class T>
  Init operator<<(const T& v) {
    (*this)(0, 0) = value_type(v);
    return Init(*this, 1);
  }

  Matrix operator+(const Matrix& o) const {
    assert(rows_ == o.rows_ && cols_ == o.cols_);
    Matrix r(rows_, cols_);
    for (std::size_t k = 0; k < d_.size(); ++k) r.d_[k] = d_[k] + o.d_[k];
    return r;
  }
  Matrix operator-(const Matrix& o) const {
    assert(rows_ == o.rows_ && cols_ == o.cols_);
    Matrix r(rows_, cols_);
    for (std::size_t k = 0; k < d_.size(); ++k) r.d_[k] = d_[k] - o.d_[k];
    return r;
  }
  Matrix operator*(const Matrix& o) const {
    assert(cols_ == o.rows_);
    Matrix r(rows_, o.cols_);
    for (std::size_t j = 0; j < o.cols_; ++j)
      for (std::size_t l = 0; l < cols_; ++l)
        for (std::size_t i = 0; i < rows_; ++i) r(i, j) += (*this)(i, l) * o(l, j);
    return r;
  }
  Matrix adjoint() const {
    Matrix r(cols_, rows_);
    for (std::size_t i = 0; i < rows_; ++i)
      for (std::size_t j = 0; j < cols_; ++j) r(j, i) = std::conj((*this)(i, j));
    return r;
  }
  friend Matrix operator*(const value_type& s, const Matrix& m) {
    Matrix r(m.rows_, m.cols_);
    for (std::size_t k = 0; k < m.d_.size(); ++k) r.d_[k] = s * m.d_[k];
    return r;
  }
  friend Matrix operator*(double s, const Matrix& m) {
    Matrix r(m.rows_, m.cols_);
    for (std::size_t k = 0; k < m.d_.size(); ++k) r.d_[k] = s * m.d_[k];
    return r;
  }
  friend Matrix operator*(const Matrix& m, const value_type& s) { return s * m; }
  friend Matrix operator*(const Matrix& m, double s) { return s * m; }
  friend std::ostream& operator<<(std::ostream& os, const Matrix& m) {
    for (std::size_t i = 0; i < m.rows_; ++i) {
      for (std::size_t j = 0; j < m.cols_; ++j) os << (j ? " " : "") << m(i, j);
      os << "\n";
    }
    return os;
  }
};

}  // namespace qudrip
