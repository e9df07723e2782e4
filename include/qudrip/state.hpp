// qudrip/state.hpp -- states, expectation values, lazy products, start vectors.
//
// State<T> keeps the reference's layout (ndim x nt complex128, column-major,
// state.hpp:8-96) but lives in device memory; a host mirror is kept
// coherent per column for cursor access (`psi[t]`, state.hpp:52-60).
//   getState(index, nt)          state.hpp:99-102
//   psi(t), psi[t], eval(t, M)   state.hpp:23-74 (eval: <psi|M|psi>, unnormalised)
//   Hermitian::norm(psi)         sparse_algorithm.hpp:7-16
//   shuffle_state(psi)           utils.hpp:129-140, with a reproducible seeded
//                                generator instead of random_device
//   psi2 = ham * psi1            the lazy form (OPsiType, operator.hpp:57-66;
//                                OperatorSum::map, operator_abel.hpp:44-49),
//                                evaluated on the fly by the matrix-free kernel.
#pragma once

#include <cmath>
#include <complex>
#include <iostream>
#include <memory>
#include <vector>

#include "operator.hpp"

namespace qudrip {

template <class STATE, class OP>
struct OPsiType;

template <class T>
class State {
  T& idx_;
  size_t nt_, ndim_;
  size_t t_ = 0;
  std::shared_ptr<detail::DeviceBuffer> dev_;
  mutable std::vector<value_type> host_;
  mutable std::vector<char> host_ok_, dev_ok_;
  value_type null_position_ = 0.0;

  void to_device(size_t t) const {
    if (!dev_ok_[t]) {
      detail::check(qd_memcpy_h2d(detail::ctx(), dev_->p + 2 * ndim_ * t, host_.data() + ndim_ * t,
                                  sizeof(value_type) * ndim_),
                    "upload state");
      dev_ok_[t] = 1;
    }
  }
  void to_host(size_t t) const {
    if (!host_ok_[t]) {
      detail::check(qd_memcpy_d2h(detail::ctx(), host_.data() + ndim_ * t, dev_->p + 2 * ndim_ * t,
                                  sizeof(value_type) * ndim_),
                    "download state");
      host_ok_[t] = 1;
    }
  }

 public:
  State(T& idx, size_t nt)
      : idx_(idx),
        nt_(nt),
        ndim_(idx.range()),
        dev_(std::make_shared<detail::DeviceBuffer>(idx.range() * nt)),
        host_(idx.range() * nt, value_type(0.0, 0.0)),
        host_ok_(nt, 1),
        dev_ok_(nt, 0) {}

  // device column t (uploading pending host writes); `write` invalidates the host copy
  double* device(size_t t, bool write) {
    to_device(t);
    if (write) host_ok_[t] = 0;
    return dev_->p + 2 * ndim_ * t;
  }
  const double* device(size_t t) const {
    to_device(t);
    return dev_->p + 2 * ndim_ * t;
  }
  // host column t (downloading pending device writes); `write` invalidates the device copy
  value_type* host(size_t t, bool write) {
    to_host(t);
    if (write) dev_ok_[t] = 0;
    return host_.data() + ndim_ * t;
  }

  State& operator()(int t) {
    t_ = (size_t)t;
    return *this;
  }
  int t() const { return (int)t_; }
  int nt() const { return (int)nt_; }
  int ndim() const { return (int)ndim_; }
  T& get_index() const { return idx_; }

  // cursor access (state.hpp:52-60): element at the index's current rank
  value_type& operator[](int tstp) {
    t_ = (size_t)tstp;
    idxv_type r = (idxv_type)idx_;
    if (r < idx_.range()) return host(tstp, true)[r];
    null_position_ = 0.0;
    return null_position_;
  }

  // <psi_t|M|psi_t>, unnormalised (state.hpp:33-44)
  value_type eval(int tstp, const SpMatrix& M) const {
    if ((size_t)M.cols() != ndim_) {
      std::cout << "State: evaluation failure due to dimension mismatch. Range = " << ndim_
                << " and Matrix col = " << M.cols() << std::endl;
      throw std::runtime_error("State::eval: dimension mismatch");
    }
    double out[2] = {0.0, 0.0};
    detail::check(qd_expect(detail::ctx(), M.linop(), device(tstp), out), "qd_expect");
    return value_type(out[0], out[1]);
  }

  State& operator*=(value_type a) {
    value_type* h = host(t_, true);
    for (size_t i = 0; i < ndim_; ++i) h[i] *= a;
    return *this;
  }
  State& operator/=(value_type a) {
    value_type* h = host(t_, true);
    for (size_t i = 0; i < ndim_; ++i) h[i] /= a;
    return *this;
  }

  // psi(t1) = op * psi2(t2), lazily evaluated (matrix-free)
  template <class OPS>
  State& operator=(const OPsiType<State, OPS>& opsi);
};

template <class T>
State<T> getState(T& index, int nt) {
  return State<T>(index, (size_t)nt);
}

namespace Hermitian {
// Re(psi^H psi) of the active column (sparse_algorithm.hpp:13-16)
template <class T>
double norm(const State<T>& psi) {
  double out = 0.0;
  detail::check(qd_norm2(detail::ctx(), psi.ndim(), psi.device(psi.t()), &out), "qd_norm2");
  return out;
}
}  // namespace Hermitian

// Seeded start vector in column 0, normalised (utils.hpp:129-140 fills
// column 0 through the cursor, then normalises).  Each call uses
// shuffle_seed() and then increments it.
template <class T>
void shuffle_state(State<T>& psi) {
  unsigned long long seed = shuffle_seed()++;
  detail::check(qd_seeded_state(detail::ctx(), psi.ndim(), seed, psi.device(0, true)), "qd_seeded_state");
}

// ------------------------------------------------------------------ lazy apply
// The source column is captured when the product is formed, so
// `psi(1) = ham * psi(0)` on one State reads column 0 (the reference's
// OPsiType reads psi.t() later, state.hpp:334-338).
template <class STATE, class OP>
struct OPsiType {
  const OP& op;
  STATE& psi;
  int t;
};
template <class T>
OPsiType<State<T>, OperatorSum> operator*(const OperatorSum& op, State<T>& psi) {
  return OPsiType<State<T>, OperatorSum>{op, psi, psi.t()};
}

// OperatorSum::map_acc (operator_abel.hpp:36-42): to(t1) += c * H from(t2),
// H evaluated on the fly from the term program (no stored matrix); over a
// qubits-times-mode index through a temporary stored matrix (the same leaf
// filter as `>>`).
template <class T>
void map_acc(const OperatorSum& op, State<T>& to, int t1, State<T>& from, int t2, value_type c = 1.0) {
  auto& sub = from.get_index();
  const int mdim = sub.basis()->mdim;
  qd_termprog* p = op.make_program(sub.nbits(), mdim);
  qd_linop* lo = nullptr;
  qd_csr* csr = nullptr;
  int rc = mdim > 1 ? qd_csr_build(detail::ctx(), sub.basis()->b, p, &csr) : QD_OK;
  if (rc == QD_OK)
    rc = mdim > 1 ? qd_linop_from_csr(detail::ctx(), csr, &lo) : qd_linop_matfree(detail::ctx(), sub.basis()->b, p, &lo);
  if (rc == QD_OK) {
    double cc[2] = {c.real(), c.imag()};
    rc = qd_spmv_acc(detail::ctx(), lo, cc, from.device(t2), to.device(t1, true));
  }
  if (lo) qd_linop_destroy(lo);
  if (csr) qd_csr_destroy(csr);
  qd_termprog_destroy(p);
  detail::check(rc, "map_acc");
}

template <class T>
template <class OPS>
State<T>& State<T>::operator=(const OPsiType<State<T>, OPS>& opsi) {
  // OperatorSum::map: zero the target column, then accumulate
  value_type* h = host(t_, true);
  for (size_t i = 0; i < ndim_; ++i) h[i] = 0.0;
  map_acc(opsi.op, *this, (int)t_, opsi.psi, opsi.t);
  return *this;
}

}  // namespace qudrip
