// qudrip/mc_manager.hpp -- Monte-Carlo observable accumulator (mc_manager.hpp:13-235).
//
// Host-side accumulation of <psi|O|psi> per time step, averaged over
// trajectories, with the reference's text output format.  The per-sample
// expectation values come from the GPU (State::eval).  The MPI reduction of
// the reference (USE_MPI, :130-145) is not part of this one-GPU build.
#pragma once

#include <fstream>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

#include "state.hpp"

namespace qudrip {
#define MC_REAL 0
#define MC_COMPLEX 1

using functor_type = std::function<value_type(const value_type&)>;

template <typename TS>
class McManager {
  TS* p_psi_;
  int nt_;
  std::unordered_map<std::string, std::vector<value_type>> data_;  // nt x dim, row-major
  std::unordered_map<std::string, int> dim_;
  std::unordered_map<std::string, std::vector<SpMatrix>> obs_;
  std::unordered_map<std::string, int> type_;
  std::unordered_map<std::string, functor_type> func_;
  std::unordered_map<std::string, std::vector<std::string>> aux_;
  std::vector<size_t> Nmc_;

 public:
  McManager(int nt, TS* p) : p_psi_(p), nt_(nt), Nmc_(nt, 0) {}
  explicit McManager(TS* p) : McManager(p->nt(), p) {}

  void addObs(const std::string& name, int dim, int type = MC_COMPLEX) {  // :43-48
    data_[name].assign((size_t)nt_ * dim, value_type(0.0, 0.0));
    dim_[name] = dim;
    obs_[name] = std::vector<SpMatrix>(dim);
    type_[name] = type;
    aux_[name] = {};
  }
  void addAux(const std::string& name, const std::string& name2, const functor_type& f, int dim,
              int aux_type = MC_COMPLEX) {  // :50-59
    if (obs_.find(name2) == obs_.end()) throw std::string{"addAux: wrong dependent variable!"};
    data_[name].assign((size_t)nt_ * dim, value_type(0.0, 0.0));
    dim_[name] = dim;
    func_[name] = f;
    type_[name] = aux_type;
    aux_[name2].push_back(name);
  }
  void updateMatrix(const std::string& name, int idx, const SpMatrix& m) { obs_.at(name)[idx] = m; }  // :61-64

  void incrAllObs(int tstp, int psi_tstp = -1) {  // :88-106
    if (psi_tstp == -1) psi_tstp = tstp;
    for (auto& o : obs_) {
      auto& v = data_.at(o.first);
      int dim = dim_.at(o.first);
      for (int d = 0; d < dim; ++d) {
        value_type term = p_psi_->eval(psi_tstp, o.second[d]);
        v[(size_t)tstp * dim + d] += term;
        for (const auto& n : aux_.at(o.first)) data_.at(n)[(size_t)tstp * dim + d] += func_.at(n)(term);
      }
    }
    ++Nmc_[tstp];
  }
  void clearAll() {
    for (auto& d : data_)
      for (auto& x : d.second) x = 0.0;
  }
  void normAll() {  // :122-128
    for (auto& d : data_) {
      int dim = dim_.at(d.first);
      for (int t = 0; t < nt_; ++t)
        for (int k = 0; k < dim; ++k) d.second[(size_t)t * dim + k] *= 1. / Nmc_[t];
    }
  }
  void collectAll() {}
  value_type obsValue(const std::string& name, int tstp, int idx) {
    return data_.at(name)[(size_t)tstp * dim_.at(name) + idx];
  }
  int report() { return 0; }

  // text table (:147-196): "#time[\tindex]" then real[/imag] columns per name
  void outFile(double h, const std::string& fname, const std::vector<std::string>& names, int = 6) {
    std::ofstream f(fname);
    if (data_.find(names[0]) == data_.end()) throw std::string{"outFile: wrong name!!"};
    int dim = dim_.at(names[0]);
    std::string label = dim > 1 ? "#time\tindex" : "#time";
    for (auto& n : names) {
      if (type_.find(n) == type_.end()) throw std::string{"outFile: wrong name!!"};
      label += (type_.at(n) == MC_REAL) ? "\t" + n : "\t" + n + "(real)" + "\t" + n + "(imag)";
      if (dim_.at(n) != dim) {
        std::cout << "mismatched dimensions" << std::endl;
        return;
      }
    }
    f << label << std::endl;
    for (int t = 0; t < nt_; ++t) {
      for (int idx = 0; idx < dim; ++idx) {
        label = dim > 1 ? std::to_string(t * h) + "\t" + std::to_string(idx) : std::to_string(t * h);
        for (auto& n : names) {
          value_type v = data_.at(n)[(size_t)t * dim + idx];
          label += "\t" + std::to_string(v.real());
          if (type_.at(n) == MC_COMPLEX) label += "\t" + std::to_string(v.imag());
        }
        f << label << std::endl;
      }
      if (dim > 1) f << std::endl;
    }
  }
};

template <typename TS>
McManager<TS> getMC(TS& psi) {
  return McManager<TS>(&psi);
}
template <typename TS>
McManager<TS> getMC(int nt, TS& psi) {
  return McManager<TS>(nt, &psi);
}
}  // namespace qudrip
