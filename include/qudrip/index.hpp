// qudrip/index.hpp -- Hilbert-space indices and conserved quantities.
//
// Same call surface as the reference:
//   getQbits(N)                         index.hpp:300 (QbitsIndex :80-95)
//   getNset / getNup / getNdo, `cq = v`  conservation.hpp:89-115, 233-246
//   getNd / getNh                         conservation.hpp:117-151, 248-256
//   SingleModeIndex, Indices (q * m)      index.hpp:94-184
//   Constrain(idx, cq = v, ...)          conservation.hpp:212-231 -> SubIndex
//   cursor protocol idx[i], ++idx, (idxv_type)idx, idx.range()
// The sector table is produced on the GPU (qd_basis_create); the host keeps
// a lazily downloaded copy only for cursor-style element access.
#pragma once

#include <algorithm>
#include <bitset>
#include <cstdint>
#include <iostream>
#include <memory>
#include <vector>

#include "runtime.hpp"

namespace qudrip {

using idx_size_t = unsigned long long;
using idxv_type = unsigned long long;
static const double eps = 1e-10;  // qudrip.hpp:18
constexpr short BIT_LIMIT = 64;
using bits_type = std::bitset<BIT_LIMIT>;

template <class T = idx_size_t>
class QbitsIndex {
  int nbits_;
  T range_;
  T idx_value_ = 0;

 public:
  struct is_index {
    using type = int;
  };
  explicit QbitsIndex(T N) : nbits_((int)N), range_(T(1) << N) {}
  int nbits() const { return nbits_; }
  T range() const { return range_; }
  operator T() const { return idx_value_; }
  void operator[](T v) { idx_value_ = v; }
  QbitsIndex& operator++() {
    ++idx_value_;
    return *this;
  }
};

inline QbitsIndex<> getQbits(int N) { return QbitsIndex<>(N); }

// One boson mode with n_mode levels (SingleModeIndex, index.hpp:94-110).
template <class T = idx_size_t>
class SingleModeIndex {
  T n_;
  T v_ = 0;

 public:
  struct is_index {
    using type = int;
  };
  explicit SingleModeIndex(int n_mode) : n_((T)n_mode) {}
  T range() const { return n_; }
  operator T() const { return v_; }
  void operator[](T v) { v_ = v; }
  SingleModeIndex& operator++() {
    ++v_;
    return *this;
  }
};

// Tensor product of a QbitsIndex and a SingleModeIndex (Indices,
// index.hpp:112-177): value q * mdim + m, the mode the least significant
// digit; `q * m` (:177-184).
template <class L, class R>
class Indices {
  L& lhs_;
  R& rhs_;

 public:
  struct is_index {
    using type = int;
  };
  Indices(L& l, R& r) : lhs_(l), rhs_(r) {}
  idxv_type range() const { return (idxv_type)lhs_.range() * (idxv_type)rhs_.range(); }
  operator idxv_type() const { return (idxv_type)lhs_ * (idxv_type)rhs_.range() + (idxv_type)rhs_; }
  void operator[](idxv_type i) {  // :158-161
    lhs_[i / rhs_.range()];
    rhs_[i % rhs_.range()];
  }
  Indices& operator++() {  // :164-172
    if ((idxv_type)rhs_ < (idxv_type)rhs_.range() - 1) {
      ++rhs_;
    } else {
      rhs_[0];
      ++lhs_;
    }
    return *this;
  }
  int nbits() const { return lhs_.nbits(); }
  int mdim() const { return (int)rhs_.range(); }
};
template <class T>
Indices<QbitsIndex<T>, SingleModeIndex<T>> operator*(QbitsIndex<T>& q, SingleModeIndex<T>& m) {
  return Indices<QbitsIndex<T>, SingleModeIndex<T>>(q, m);
}

// popcount(x & mask) == value  (Nset, conservation.hpp:89-115)
template <class T1 = idx_size_t>
class Nset {
  uint64_t mask_ = 0;
  long long val_ = 0;

 public:
  using quant_type = T1;
  struct is_cq {};
  Nset(const QbitsIndex<>& index, int start = -1, int end = -1) {
    if (start == -1) start = 0;
    if (end == -1) end = (int)bits_type(index.range() - 1).count();  // N: masks N+1 bits (:112)
    for (int i = start; i <= end && i < 64; ++i) mask_ |= (1ull << i);
  }
  Nset& operator=(quant_type v) {  // `cq = value` (:78-81)
    val_ = (long long)v;
    return *this;
  }
  uint64_t mask() const { return mask_; }
  long long val() const { return val_; }
  qd_quantity quantity() const { return qd_quantity{mask_, 0, 0, 0, (int64_t)val_}; }
};

// Doubly occupied (hole = 0) or empty (hole = 1) sites of a spinful index
// (Nd, conservation.hpp:117-151): popcount(((x >> Nsite) ^ m) & (x ^ m)),
// m = bits 0..Nsite-1 when hole, else 0.
template <class T1 = idx_size_t>
class Nd {
  int nsite_;
  uint64_t flip_ = 0;
  long long val_ = 0;

 public:
  using quant_type = T1;
  struct is_cq {};
  Nd(const QbitsIndex<>& index, int Nsite, bool hole = 0) : nsite_(Nsite) {
    (void)index;
    for (int i = 0; i < Nsite && i < 64; ++i)
      if (hole) flip_ |= (1ull << i);
  }
  Nd& operator=(quant_type v) {
    val_ = (long long)v;
    return *this;
  }
  long long val() const { return val_; }
  qd_quantity quantity() const { return qd_quantity{~0ull, flip_, nsite_, 0, (int64_t)val_}; }
};

inline Nset<> getNset(const QbitsIndex<>& index) { return Nset<>(index, -1, -1); }
inline Nset<> getNup(const QbitsIndex<>& index, int Nsite) { return Nset<>(index, 0, Nsite - 1); }
inline Nset<> getNdo(const QbitsIndex<>& index, int Nsite) { return Nset<>(index, Nsite, 2 * Nsite - 1); }
inline Nd<> getNd(const QbitsIndex<>& index, int Nsite) { return Nd<>(index, Nsite, 0); }  // :248-251
inline Nd<> getNh(const QbitsIndex<>& index, int Nsite) { return Nd<>(index, Nsite, 1); }  // :253-256

namespace detail {
// device sector table shared by every object built on it
struct BasisHolder {
  qd_basis* b = nullptr;
  int nbits = 0;
  int mdim = 1;  // >= 2: composite index with a boson mode
  std::vector<uint64_t> host;  // lazily downloaded mapping
  ~BasisHolder() {
    if (b) qd_basis_destroy(b);
  }
  const std::vector<uint64_t>& mapping() {
    if (host.empty() && qd_basis_size(b) > 0) {
      host.resize((size_t)qd_basis_size(b));
      check(qd_basis_download(b, host.data()), "qd_basis_download");
    }
    return host;
  }
};
inline std::shared_ptr<BasisHolder> make_basis(int nbits, const std::vector<qd_quantity>& q) {
  auto h = std::make_shared<BasisHolder>();
  h->nbits = nbits;
  check(qd_basis_create_q(ctx(), nbits, q.data(), (int)q.size(), &h->b), "qd_basis_create_q");
  return h;
}
inline std::shared_ptr<BasisHolder> make_basis_modes(int nbits, int mdim, const std::vector<qd_quantity>& q) {
  auto h = std::make_shared<BasisHolder>();
  h->nbits = nbits;
  h->mdim = mdim;
  check(qd_basis_create_modes(ctx(), nbits, mdim, q.data(), (int)q.size(), &h->b), "qd_basis_create_modes");
  return h;
}
}  // namespace detail

// Constrained index (SubIndex, index.hpp:187-261): ascending sector table,
// unrank via idx[i], rank via (idxv_type)idx, a miss ranks to range().
template <class E = QbitsIndex<>>
class SubIndex {
  E& earless_;
  std::shared_ptr<detail::BasisHolder> basis_;
  idxv_type idx_value_ = 0;

 public:
  struct is_index {
    using type = int;
  };
  SubIndex(E& idx, std::shared_ptr<detail::BasisHolder> b) : earless_(idx), basis_(std::move(b)) {}

  idxv_type range() const { return (idxv_type)qd_basis_size(basis_->b); }
  E& earless() const { return earless_; }
  const std::shared_ptr<detail::BasisHolder>& basis() const { return basis_; }
  int nbits() const { return basis_->nbits; }

  operator idxv_type() const {
    const auto& m = basis_->mapping();
    auto v = (uint64_t)(idxv_type)earless_;
    auto it = std::lower_bound(m.begin(), m.end(), v);
    if (it == m.end() || *it != v) return range();
    return (idxv_type)(it - m.begin());
  }
  void operator[](idxv_type i) {
    idx_value_ = i;
    const auto& m = basis_->mapping();
    if (i < m.size())
      earless_[m[i]];
    else
      earless_[earless_.range()];
  }
  SubIndex& operator++() {
    (*this)[idx_value_ + 1];
    return *this;
  }
};

template <class... CQs>
SubIndex<> Constrain(QbitsIndex<>& index, CQs&&... cqs) {
  std::cout << "Constrain: original range " << index.range() << std::endl;  // :214
  std::vector<qd_quantity> q;
  (q.push_back(cqs.quantity()), ...);
  SubIndex<> sub(index, detail::make_basis(index.nbits(), q));
  sub[0];  // :223
  return sub;
}

// Constrain over a qubits-times-mode index: the quantities act on the qubits.
template <class L, class R, class... CQs>
SubIndex<Indices<L, R>> Constrain(Indices<L, R>& index, CQs&&... cqs) {
  std::cout << "Constrain: original range " << index.range() << std::endl;
  std::vector<qd_quantity> q;
  (q.push_back(cqs.quantity()), ...);
  SubIndex<Indices<L, R>> sub(index, detail::make_basis_modes(index.nbits(), index.mdim(), q));
  sub[0];
  return sub;
}

}  // namespace qudrip
