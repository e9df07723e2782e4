// qudrip/operator.hpp -- gates, operator products and sums, `>>`.
//
// Same algebra as the reference (operator.hpp:15-132, operator_abel.hpp:9-158,
// operators/bit.hpp, operators/hopbit.hpp):
//   gate << U stores U_[i*2+j] = U(i,j); gate(site) / gate(site2, site1)
//   copies the gate (with its current U) into a fresh shared instance;
//   Operator * Operator concatenates op lists; OperatorSum keeps terms in
//   construction order, `append` merges only identical gate instances;
//   (sum) * (sum) expands i-major.
// `ham >> index` flattens the sum into a term program and builds the
// reference's ColMajor matrix on the GPU (qd_csr_build), bit-identical to
// the reference's arrays.  The result is a device-resident SpMatrix.
#pragma once

#include <complex>
#include <memory>
#include <stdexcept>
#include <vector>

#include "index.hpp"
#include "matrix.hpp"

namespace qudrip {

const value_type II = value_type(0.0, 1.0);  // qudrip.hpp:20

// Pauli matrices (operators/bit.hpp:42-45)
inline const Matrix& pauli_x_m() {
  static Matrix m = (Matrix(2, 2) << 0, 1, 1, 0).finished();
  return m;
}
inline const Matrix& pauli_y_m() {
  static Matrix m = (Matrix(2, 2) << 0, -II, II, 0).finished();
  return m;
}
inline const Matrix& pauli_z_m() {
  static Matrix m = (Matrix(2, 2) << 1, 0, 0, -1).finished();
  return m;
}
inline const Matrix& id_m() {
  static Matrix m = (Matrix(2, 2) << 1, 0, 0, 1).finished();
  return m;
}
static const Matrix& pauli_x = pauli_x_m();
static const Matrix& pauli_y = pauli_y_m();
static const Matrix& pauli_z = pauli_z_m();
static const Matrix& id = id_m();

struct OpRec {
  int kind;
  int site_a, site_b;
  int string;
  value_type U[4];
  std::vector<value_type> M;  // mode gate: U_[i*mdim+j] = U(i,j)
};
using op_ptr = std::shared_ptr<const OpRec>;

class OperatorSum;

class Operator {
  std::vector<op_ptr> ops_;

 public:
  explicit Operator(op_ptr p) { ops_.push_back(std::move(p)); }
  Operator() = default;
  const std::vector<op_ptr>& ops() const { return ops_; }
  Operator operator*(const Operator& o) const {  // operator.hpp:82-91
    Operator r(*this);
    r.ops_.insert(r.ops_.end(), o.ops_.begin(), o.ops_.end());
    return r;
  }
  bool operator==(const Operator& o) const {  // :123-131, instance identity
    if (ops_.size() != o.ops_.size()) return false;
    for (size_t i = 0; i < ops_.size(); ++i)
      if (ops_[i] != o.ops_[i]) return false;
    return true;
  }
};

template <class T = idx_size_t>
class bitOp {
  int string_;
  value_type U_[4] = {};

 public:
  bitOp(QbitsIndex<T>&, int string) : string_(string) {}
  void operator<<(const Matrix& U) {  // operator.hpp:24-29
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) U_[i * 2 + j] = U(i, j);
  }
  Operator operator()(size_t site) const {  // bit.hpp:17-21
    return Operator(std::make_shared<OpRec>(OpRec{QD_OP_BIT, (int)site, (int)site, string_, {U_[0], U_[1], U_[2], U_[3]}}));
  }
};

template <class T = idx_size_t>
class hopBitOp {
  int string_;

 public:
  hopBitOp(QbitsIndex<T>&, int string) : string_(string) {}
  void operator<<(const Matrix&) {}
  Operator operator()(size_t site2, size_t site1) const {  // hopbit.hpp:17-22
    return Operator(std::make_shared<OpRec>(OpRec{QD_OP_HOP, (int)site2, (int)site1, string_, {}}));
  }
};

// Gate on a boson mode (modeOp, operators/mode.hpp:6-35): `gate << U`
// stores the mdim x mdim matrix, gate(site) a fresh instance (site unused);
// the instance takes mode m to sum_j U(m, j) |j> (mode.hpp:80-103).
template <class T = idx_size_t>
class modeOp {
  int mdim_;
  std::vector<value_type> U_;

 public:
  explicit modeOp(SingleModeIndex<T>& index) : mdim_((int)index.range()), U_((size_t)mdim_ * mdim_) {}
  void operator<<(const Matrix& U) {  // operator.hpp:24-29
    for (int i = 0; i < mdim_; ++i)
      for (int j = 0; j < mdim_; ++j) U_[(size_t)i * mdim_ + j] = U(i, j);
  }
  Operator operator()(size_t = 0) const {
    auto r = std::make_shared<OpRec>(OpRec{QD_OP_MODE, 0, 0, 1, {}, U_});
    return Operator(r);
  }
};
template <class T = idx_size_t>
modeOp<T> getModeOp(SingleModeIndex<T>& index) {  // mode.hpp:31-34
  return modeOp<T>(index);
}

template <class T = idx_size_t>
bitOp<T> getFermiGate(QbitsIndex<T>& index) {
  return bitOp<T>(index, -1);
}
template <class T = idx_size_t>
bitOp<T> getBoseGate(QbitsIndex<T>& index) {
  return bitOp<T>(index, 1);
}
template <class T = idx_size_t>
hopBitOp<T> getHopFermiGate(QbitsIndex<T>& index) {
  return hopBitOp<T>(index, -1);
}
template <class T = idx_size_t>
hopBitOp<T> getHopBoseGate(QbitsIndex<T>& index) {
  return hopBitOp<T>(index, 1);
}

// ------------------------------------------------------------------ SpMatrix
namespace detail {
struct CsrHolder {
  std::shared_ptr<detail::BasisHolder> basis;
  qd_termprog* prog = nullptr;
  qd_csr* csr = nullptr;
  qd_linop* op = nullptr;
  std::vector<int32_t> outer, inner;
  std::vector<value_type> vals;
  ~CsrHolder() {
    if (op) qd_linop_destroy(op);
    if (csr) qd_csr_destroy(csr);
    if (prog) qd_termprog_destroy(prog);
  }
  void download() {
    if (!outer.empty()) return;
    int64_t n = 0, nnz = 0;
    check(qd_csr_info(csr, &n, &nnz, nullptr), "qd_csr_info");
    outer.resize(n + 1);
    inner.resize(nnz);
    vals.resize(nnz);
    check(qd_csr_download(csr, outer.data(), inner.data(), reinterpret_cast<double*>(vals.data())),
          "qd_csr_download");
  }
};
}  // namespace detail

// Device-resident matrix returned by `>>` (the reference's SpMatrix,
// Eigen ColMajor with int32 indices, matrix.hpp:11).  The arithmetic the
// reference's callers write on it (programs/hubbard_diss.cpp:106-159,
// 443-446, 514-529) runs on the device with Eigen's element rules:
//   SpMatrix(n, n), setZero(), setIdentity()
//   A + B, A - B, A += B, A -= B, -A      union of the patterns, a lone
//                                         entry is a + 0 / 0 + b / 0 - b
//   c * A, A * c (complex or double c)    per entry; nothing pruned
//   A.adjoint()                           conjugate transpose
//   A * B                                 conservative sparse product
// Every operation returns a new matrix; operands are never modified except
// by the compound assignments, which rebind the left-hand side.
class SpMatrix {
  std::shared_ptr<detail::CsrHolder> h_;

  static SpMatrix wrap(qd_csr* c, const std::shared_ptr<detail::BasisHolder>& basis) {
    auto h = std::make_shared<detail::CsrHolder>();
    h->basis = basis;
    h->csr = c;
    return SpMatrix(h);
  }
  const std::shared_ptr<detail::BasisHolder>& basis_of(const SpMatrix& o) const { return h_->basis ? h_->basis : o.h_->basis; }
  SpMatrix scaled(value_type c, bool real) const {
    qd_csr* r = nullptr;
    detail::check(qd_csr_scale(detail::ctx(), h_->csr, c.real(), real ? 0.0 : c.imag(), real ? 1 : 0, &r),
                  "qd_csr_scale");
    return wrap(r, h_->basis);
  }
  SpMatrix combined(const SpMatrix& o, int sign) const {
    qd_csr* r = nullptr;
    detail::check(qd_csr_add(detail::ctx(), h_->csr, o.h_->csr, sign, &r), "qd_csr_add");
    return wrap(r, basis_of(o));
  }

 public:
  SpMatrix() = default;
  explicit SpMatrix(std::shared_ptr<detail::CsrHolder> h) : h_(std::move(h)) {}
  // SparseMatrix(rows, cols): square, no stored entries
  SpMatrix(long rows, long cols) {
    if (rows != cols) throw std::runtime_error("SpMatrix: only square matrices are supported");
    qd_csr* r = nullptr;
    detail::check(qd_csr_zero(detail::ctx(), rows, &r), "qd_csr_zero");
    *this = wrap(r, nullptr);
  }
  long rows() const {
    int64_t n = 0;
    if (h_) qd_csr_info(h_->csr, &n, nullptr, nullptr);
    return (long)n;
  }
  long cols() const { return rows(); }
  long nonZeros() const {
    int64_t n = 0, nnz = 0;
    if (h_) qd_csr_info(h_->csr, &n, &nnz, nullptr);
    return (long)nnz;
  }
  // Eigen-style raw storage (host copies, downloaded on first use)
  const int32_t* outerIndexPtr() const {
    h_->download();
    return h_->outer.data();
  }
  const int32_t* innerIndexPtr() const {
    h_->download();
    return h_->inner.data();
  }
  const value_type* valuePtr() const {
    h_->download();
    return h_->vals.data();
  }
  qd_linop* linop() const {
    if (!h_->op) detail::check(qd_linop_from_csr(detail::ctx(), h_->csr, &h_->op), "qd_linop_from_csr");
    return h_->op;
  }

  // ---- algebra
  SpMatrix& setZero() {
    *this = SpMatrix(rows(), rows());
    return *this;
  }
  SpMatrix& setIdentity() {
    qd_csr* r = nullptr;
    detail::check(qd_csr_identity(detail::ctx(), rows(), &r), "qd_csr_identity");
    *this = wrap(r, h_->basis);
    return *this;
  }
  SpMatrix adjoint() const {
    qd_csr* r = nullptr;
    detail::check(qd_csr_adjoint(detail::ctx(), h_->csr, &r), "qd_csr_adjoint");
    return wrap(r, h_->basis);
  }
  SpMatrix operator+(const SpMatrix& o) const { return combined(o, 1); }
  SpMatrix operator-(const SpMatrix& o) const { return combined(o, -1); }
  SpMatrix& operator+=(const SpMatrix& o) { return *this = combined(o, 1); }
  SpMatrix& operator-=(const SpMatrix& o) { return *this = combined(o, -1); }
  SpMatrix operator-() const { return scaled(value_type(-1.0, 0.0), true); }  // exact negation
  SpMatrix operator*(const SpMatrix& o) const {
    qd_csr* r = nullptr;
    detail::check(qd_csr_multiply(detail::ctx(), h_->csr, o.h_->csr, &r), "qd_csr_multiply");
    return wrap(r, basis_of(o));
  }
  friend SpMatrix operator*(value_type c, const SpMatrix& m) { return m.scaled(c, false); }
  friend SpMatrix operator*(double c, const SpMatrix& m) { return m.scaled(value_type(c, 0.0), true); }
  friend SpMatrix operator*(const SpMatrix& m, value_type c) { return m.scaled(c, false); }
  friend SpMatrix operator*(const SpMatrix& m, double c) { return m.scaled(value_type(c, 0.0), true); }
};

// ------------------------------------------------------------------ sums
class OperatorSum {
  std::vector<Operator> ops_;
  std::vector<value_type> coeffs_;

 public:
  explicit OperatorSum(int reserve = 2) {
    ops_.reserve(reserve);
    coeffs_.reserve(reserve);
  }
  OperatorSum(const Operator& op, value_type c = 1.0) : ops_(1, op), coeffs_(1, c) {}
  int size() const { return (int)ops_.size(); }
  const std::vector<Operator>& ops() const { return ops_; }
  const std::vector<value_type>& coeffs() const { return coeffs_; }

  OperatorSum& append(const Operator& op, value_type coeff) {  // operator_abel.hpp:115-127
    for (size_t i = 0; i < ops_.size(); ++i)
      if (ops_[i] == op) {
        coeffs_[i] += coeff;
        return *this;
      }
    ops_.push_back(op);
    coeffs_.push_back(coeff);
    return *this;
  }
  OperatorSum operator*(const Operator& op) const {  // :51-58
    OperatorSum r(*this);
    for (auto& o : r.ops_) o = o * op;
    return r;
  }
  OperatorSum operator*(value_type m) const {  // :65-71
    OperatorSum r(*this);
    for (auto& c : r.coeffs_) c *= m;
    return r;
  }
  OperatorSum operator*(const OperatorSum& o) const {  // :73-81
    OperatorSum r((int)(ops_.size() * o.ops_.size()));
    for (size_t i = 0; i < ops_.size(); ++i)
      for (size_t j = 0; j < o.ops_.size(); ++j) r.append(ops_[i] * o.ops_[j], coeffs_[i] * o.coeffs_[j]);
    return r;
  }
  OperatorSum operator+(const Operator& op) const {  // :83-87
    OperatorSum r(*this);
    r.append(op, 1);
    return r;
  }
  OperatorSum operator+(const OperatorSum& o) const {  // :89-96
    OperatorSum r(*this);
    r.ops_.insert(r.ops_.end(), o.ops_.begin(), o.ops_.end());
    r.coeffs_.insert(r.coeffs_.end(), o.coeffs_.begin(), o.coeffs_.end());
    return r;
  }
  OperatorSum operator-(const OperatorSum& o) const {  // :98-107
    OperatorSum r(*this);
    r.ops_.insert(r.ops_.end(), o.ops_.begin(), o.ops_.end());
    for (auto c : o.coeffs_) r.coeffs_.push_back(-c);
    return r;
  }
  OperatorSum operator-(const Operator& op) const {  // :109-113
    OperatorSum r(*this);
    r.append(op, -1);
    return r;
  }

  // flattened term program (include/qudrip_hip.h data format); the matrices
  // of mode gates go to mats (a mode gate's site_a is its position there)
  void program(std::vector<qd_term>& terms, std::vector<qd_op>& ops, int* nbits_used,
               std::vector<double>* mats = nullptr) const {
    terms.clear();
    ops.clear();
    if (mats) mats->clear();
    int nmat = 0;
    int nb = 0;
    for (size_t k = 0; k < ops_.size(); ++k) {
      qd_term t{};
      t.coeff[0] = coeffs_[k].real();
      t.coeff[1] = coeffs_[k].imag();
      t.op_begin = (int32_t)ops.size();
      t.n_ops = (int32_t)ops_[k].ops().size();
      for (const auto& r : ops_[k].ops()) {
        qd_op o{};
        o.kind = r->kind;
        o.site_a = r->site_a;
        o.site_b = r->site_b;
        o.string = r->string;
        if (r->kind == QD_OP_MODE) {
          o.site_a = o.site_b = nmat++;
          if (mats)
            for (const auto& u : r->M) {
              mats->push_back(u.real());
              mats->push_back(u.imag());
            }
          ops.push_back(o);
          continue;
        }
        for (int q = 0; q < 4; ++q) {
          o.U[2 * q] = r->U[q].real();
          o.U[2 * q + 1] = r->U[q].imag();
        }
        nb = std::max(nb, std::max(r->site_a, r->site_b) + 1);
        ops.push_back(o);
      }
      terms.push_back(t);
    }
    if (nbits_used) *nbits_used = nb;
  }

  qd_termprog* make_program(int nbits, int mdim = 1) const {
    std::vector<qd_term> terms;
    std::vector<qd_op> ops;
    std::vector<double> mats;
    int nb = 0;
    program(terms, ops, &nb, &mats);
    qd_termprog* p = nullptr;
    if (mdim > 1)
      detail::check(qd_termprog_create_modes(detail::ctx(), nbits, mdim, terms.data(), (int)terms.size(), ops.data(),
                                             (int)ops.size(), mats.data(), (int)(mats.size() / (2 * mdim * mdim)),
                                             &p),
                    "qd_termprog_create_modes");
    else
      detail::check(qd_termprog_create(detail::ctx(), nbits, terms.data(), (int)terms.size(), ops.data(),
                                       (int)ops.size(), &p),
                    "qd_termprog_create");
    return p;
  }

  // `ham >> index` (operator_abel.hpp:25-34)
  template <class IDX>
  SpMatrix operator>>(IDX& index) const {
    auto h = std::make_shared<detail::CsrHolder>();
    h->basis = index.basis();
    h->prog = make_program(index.nbits(), index.basis()->mdim);
    detail::check(qd_csr_build(detail::ctx(), h->basis->b, h->prog, &h->csr), "qd_csr_build");
    return SpMatrix(h);
  }
};

inline OperatorSum operator*(value_type c, const Operator& op) { return OperatorSum(op, c); }  // :143
inline OperatorSum operator*(const Operator& op, value_type c) { return OperatorSum(op, c); }  // :147
inline OperatorSum operator*(value_type c, const OperatorSum& s) { return s * c; }             // :151
inline OperatorSum operator+(const Operator& a, const Operator& b) {                           // :157-161
  OperatorSum r(a, 1.0);
  r.append(b, 1.0);
  return r;
}
inline OperatorSum operator-(const Operator& a, const Operator& b) {  // :163-167
  OperatorSum r(a, 1.0);
  r.append(b, -1.0);
  return r;
}
inline OperatorSum operator+(const Operator& a, const OperatorSum& b) { return OperatorSum(a, 1.0) + b; }
inline OperatorSum getEmptyOperator() { return OperatorSum(2); }  // :170

// `op >> index` for a single product (operator.hpp:235-267)
template <class IDX>
SpMatrix operator>>(const Operator& op, IDX& index) {
  return OperatorSum(op, 1.0) >> index;
}

// Basis over the whole QbitsIndex (no constraint) for `op >> qbits`.
inline SubIndex<> FullIndex(QbitsIndex<>& q) {
  SubIndex<> s(q, detail::make_basis(q.nbits(), {}));
  return s;
}

}  // namespace qudrip
