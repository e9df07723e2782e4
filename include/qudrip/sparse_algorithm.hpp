// qudrip/sparse_algorithm.hpp -- Krylov solvers of the drop-in API.
//
//   Arnoldi::getSolver(idx, k)        -> ArnoldiSolver<T, 0>   (sparse_algorithm.hpp:596-599)
//   Arnoldi::getLanczosSolver(idx, k) -> ArnoldiSolver<T, 1>   (:601-604)
//   getGroundState(H, psi, tstp, add_dim)                     (:564-585)
//   evolve(H, psi, tstp, h, kry_dim_rt)                       (:548-562)
//   error(), act_dim(), convergeEV(N, iv), storeE(fname)      (:492-500, :587-593)
// Both solvers keep their whole basis on the GPU with full (CGS2)
// reorthogonalisation; the tridiagonal eigenproblem and the Hessenberg
// exponential are solved in FP64 on the host.  Like the reference, the
// solver methods are noexcept and abort() when the small eigenproblem fails.
//
//   Hermitian::getSolver(idx, k) -> Hermitian::LanczosSolver<T> (:236-313)
//   getGroundState(H, psi, tstp), evolve(H, psi, tstp, h), error(), storeE
// The two-vector Lanczos of the reference: no stored basis and no
// reorthogonalisation, two passes of the three-term recurrence per solve
// (qd_hlanczos_*).
#pragma once

#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "state.hpp"

namespace qudrip {
namespace Arnoldi {

template <class T, bool herm = false>
class ArnoldiSolver {
  int kry_dim_;
  std::shared_ptr<qd_krylov> kr_;

  static void fail(int rc) noexcept {
    if (rc == QD_E_EIGEN) std::cout << "failed to solve the Krylov matrix" << std::endl;
    else
      std::fprintf(stderr, "qudrip solver: %s\n", qd_last_error(detail::ctx()));
    std::abort();
  }

 public:
  ArnoldiSolver(T& index, int kry_dim) : kry_dim_(kry_dim) {
    qd_krylov* k = nullptr;
    detail::check(qd_krylov_create(detail::ctx(), (int64_t)index.range(), kry_dim, herm ? 1 : 0, &k),
                  "qd_krylov_create");
    kr_.reset(k, [](qd_krylov* p) { qd_krylov_destroy(p); });
  }

  double error() const {
    qd_krylov_info inf{};
    qd_krylov_get_info(kr_.get(), &inf);
    return inf.error;
  }
  int act_dim() const {
    qd_krylov_info inf{};
    qd_krylov_get_info(kr_.get(), &inf);
    return inf.act_dim;
  }
  void convergeEV(int N, value_type iv = 0.) {
    detail::check(qd_krylov_converge_ev(kr_.get(), N, iv.real(), iv.imag()), "convergeEV");
  }

  void getGroundState(const SpMatrix& H, State<T>& psi, int tstp, int add_dim = -1) noexcept {
    int rc = qd_krylov_ground_state(kr_.get(), H.linop(), psi.device(tstp, true), add_dim, nullptr);
    if (rc != QD_OK) fail(rc);
  }

  void evolve(const SpMatrix& H, State<T>& psi, int tstp, double h, int kry_dim_rt = 0) noexcept {
    const double* in = psi.device(tstp);
    int rc = qd_krylov_evolve(kr_.get(), H.linop(), in, psi.device(tstp + 1, true), h, kry_dim_rt);
    if (rc != QD_OK) fail(rc);
  }

  std::vector<double> eigenvalues() const {
    std::vector<double> ev(kry_dim_ + 1);
    int cnt = 0;
    qd_krylov_eigenvalues(kr_.get(), ev.data(), (int)ev.size(), &cnt);
    ev.resize(cnt);
    return ev;
  }

  void storeE(std::string fname) {  // one eigenvalue per line (:587-593)
    std::ofstream ef(fname);
    for (double e : eigenvalues()) ef << e << std::endl;
  }
};

template <class T>
ArnoldiSolver<T, false> getSolver(T& index, int kry_dim) {
  return ArnoldiSolver<T, false>(index, kry_dim);
}
template <class T>
ArnoldiSolver<T, true> getLanczosSolver(T& index, int kry_dim) {
  return ArnoldiSolver<T, true>(index, kry_dim);
}

}  // namespace Arnoldi

namespace Hermitian {

template <class T>
class LanczosSolver {
  int kry_dim_;
  std::shared_ptr<qd_hlanczos> lz_;

  static void fail(int rc) noexcept {
    if (rc == QD_E_EIGEN) std::cout << "failed to solve the Krylov matrix" << std::endl;
    else
      std::fprintf(stderr, "qudrip solver: %s\n", qd_last_error(detail::ctx()));
    std::abort();
  }

 public:
  LanczosSolver(const T& index, int kry_dim) : kry_dim_(kry_dim) {
    qd_hlanczos* p = nullptr;
    detail::check(qd_hlanczos_create(detail::ctx(), (int64_t)index.range(), kry_dim, &p), "qd_hlanczos_create");
    lz_.reset(p, [](qd_hlanczos* q) { qd_hlanczos_destroy(q); });
  }

  double error() const {  // |r| after the last recurrence step (0 before any solve)
    qd_hlanczos_info inf{};
    qd_hlanczos_get_info(lz_.get(), &inf);
    return inf.error;
  }

  // psi(tstp + 1) = exp(-i h H) psi(tstp) / |psi(tstp)| in the Krylov space (:257-283)
  void evolve(const SpMatrix& H, State<T>& psi, int tstp, double h) noexcept {
    assert(tstp < (int)psi.nt() - 1);
    const double* in = psi.device(tstp);
    int rc = qd_hlanczos_evolve(lz_.get(), H.linop(), in, psi.device(tstp + 1, true), h);
    if (rc != QD_OK) fail(rc);
  }

  // psi(tstp) <- lowest Ritz vector of its Krylov space (:285-300)
  void getGroundState(const SpMatrix& H, State<T>& psi, int tstp) noexcept {
    assert(tstp < (int)psi.nt());
    int rc = qd_hlanczos_ground_state(lz_.get(), H.linop(), psi.device(tstp, true), nullptr);
    if (rc != QD_OK) fail(rc);
  }

  std::vector<double> eigenvalues() const {
    std::vector<double> ev(kry_dim_ + 2);
    int cnt = 0;
    qd_hlanczos_eigenvalues(lz_.get(), ev.data(), (int)ev.size(), &cnt);
    ev.resize(cnt);
    return ev;
  }

  void storeE(std::string fname) {  // one eigenvalue per line (:302-308)
    std::ofstream ef(fname);
    for (double e : eigenvalues()) ef << e << std::endl;
  }
};

template <class T>
LanczosSolver<T> getSolver(const T& index, int kry_dim) {
  return LanczosSolver<T>(index, kry_dim);
}

}  // namespace Hermitian
}  // namespace qudrip
