// hubbard_quench.cpp -- the dissipative Hubbard driver's matrix algebra and
// real-time loop through the drop-in API.
//
// Follows programs/hubbard_diss.cpp of the reference: the stored matrices
// Tp, Tm, docc, h0, Id, qj_tot and Pc are built with SpMatrix arithmetic
// (:106-159), the ground state of Htot = Tp + Tm + U0 docc with
// Hermitian::getSolver (:377-384), then each time step forms
// Htot = Tdr[i] Tp + conj(Tdr[i]) Tm + (ndr[i] + U0) docc (:443), the jump
// weights <psi| L_i^+ L_i |psi> (:519-529), H_eff = Htot - 0.5 i (qj_tot -
// gamma w Id) (QJ_operator, :514-517) and the Arnoldi step
// l_solver.evolve(Htot, psi, 0, h) with renormalisation (:467-470).  The
// jump branch (:461-465) applies a dense Eigen product and is not part of
// this driver: every step takes the no-jump branch.  Open chain of L sites.
//   usage: hubbard_quench <L> <Nup> <Ndo> <kry_dim> <kry_dim_rt> <nt> <h> <U0> <gamma> <out-prefix>
// Files: <prefix>.H1.{outer,inner,vals} (H_eff of step 1), <prefix>.qj.*,
// <prefix>.pc.* (the same arrays of qj_tot and Pc), <prefix>.gs (ground
// state), <prefix>.psi (state after nt steps), <prefix>.obs (per step:
// t, <docc>, w_tot as text).
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "qudrip/qudrip.hpp"

using namespace qudrip;
using cplx = std::complex<double>;

template <typename TG, typename TI>
SpMatrix lindblad(int i, int Nsite, TG& hopgate, TI& index) {  // hubbard_diss.cpp:15-18
  return ((hopgate(i, i) + hopgate(i + Nsite, i + Nsite)) >> index);
}

void QJ_operator(SpMatrix& H, const SpMatrix& qj_tot, const SpMatrix& Id, double wt) {  // :514-517
  H -= 0.5 * II * (qj_tot - wt * Id);
}

template <typename T, typename Tg>
std::vector<double> lindblad_weight(int tstp, int Nsite, const State<T>& psi, Tg& hopgate, T& hubbard) {  // :519-529
  std::vector<double> res(Nsite);
  for (int i = 0; i < Nsite; ++i) {
    auto ni = lindblad(i, Nsite, hopgate, hubbard);
    res[i] = psi.eval(tstp, SpMatrix(ni.adjoint()) * ni).real();
  }
  return res;
}

template <class T>
static void dump(const std::string& path, const T* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), sizeof(T) * n);
}

static void dump_matrix(const std::string& prefix, const SpMatrix& M) {
  dump(prefix + ".outer", M.outerIndexPtr(), (size_t)M.rows() + 1);
  dump(prefix + ".inner", M.innerIndexPtr(), (size_t)M.nonZeros());
  dump(prefix + ".vals", M.valuePtr(), (size_t)M.nonZeros());
}

int main(int argc, char** argv) {
  if (argc < 11) {
    std::cerr << "usage: hubbard_quench L Nup Ndo kry_dim kry_dim_rt nt h U0 gamma out-prefix" << std::endl;
    return 2;
  }
  const int Nsite = std::stoi(argv[1]), Nup = std::stoi(argv[2]), Ndo = std::stoi(argv[3]);
  const int kry_dim = std::stoi(argv[4]), kry_dim_rt = std::stoi(argv[5]), nt = std::stoi(argv[6]);
  const double h = std::stod(argv[7]), U0 = std::stod(argv[8]), gamma = std::stod(argv[9]);
  const std::string out = argv[10];
  const double thop = -1.0;
  const int OPEN = 1;

  // time-dependent drive, as the parameter file would give it (:79-94)
  std::vector<cplx> Tdr(nt + 1, 0);
  std::vector<double> ndr(nt + 1, 0);
  for (int i = 0; i <= nt; ++i) {
    const double Adr = 0.7 * std::sin(0.9 * i * h);
    Tdr[i] = cplx(std::cos(Adr), std::sin(Adr)) * cplx(1.0, 0.15);
    ndr[i] = 0.3 * std::cos(2.0 * i * h);
  }

  auto fermi = getQbits(2 * Nsite);
  auto nup = getNup(fermi, Nsite);
  auto ndo = getNdo(fermi, Nsite);
  auto hubbard = Constrain(fermi, nup = Nup, ndo = Ndo);
  auto gate = getHopFermiGate(fermi);
  auto ndim = hubbard.range();
  std::cout << "dimension " << ndim << std::endl;
  SpMatrix Tp(ndim, ndim), Tm(ndim, ndim), Htot(ndim, ndim), h0(ndim, ndim), Id(ndim, ndim), qj_tot(ndim, ndim);
  h0.setZero();
  Tp.setZero();
  Tm.setZero();
  Htot.setZero();
  qj_tot.setZero();
  Id.setIdentity();
  SpMatrix docc(ndim, ndim);
  docc.setZero();
  SpMatrix Pc(ndim, ndim);
  Pc.setIdentity();
  for (auto site : range(Nsite)) {  // :125-159
    int site1u = site;
    int site2u = (site + 1) % Nsite;
    auto hop21u = gate(site2u, site1u);
    auto hop12u = gate(site1u, site2u);
    if (!(site == (size_t)Nsite - 1 && OPEN)) {
      Tp += thop * hop21u >> hubbard;
      Tm += thop * hop12u >> hubbard;
    }
    int site1d = site + Nsite;
    int site2d = (site + 1) % Nsite + Nsite;
    auto hop21d = gate(site2d, site1d);
    auto hop12d = gate(site1d, site2d);
    if (!(site == (size_t)Nsite - 1 && OPEN)) {
      Tp += thop * hop21d >> hubbard;
      Tm += thop * hop12d >> hubbard;
    }
    auto nupg = gate(site1u, site1u);
    auto ndog = gate(site1d, site1d);
    docc += nupg * ndog >> hubbard;
    h0 += (nupg + ndog) >> hubbard;
    auto ni = lindblad(site, Nsite, gate, hubbard);
    qj_tot += gamma * SpMatrix(ni.adjoint()) * ni;
    Pc = Pc * (Id - ((nupg + ndog) >> hubbard) + 2. * ((nupg * ndog) >> hubbard));
  }
  dump_matrix(out + ".qj", qj_tot);
  dump_matrix(out + ".pc", Pc);

  auto psi = getState(hubbard, 2);
  shuffle_state(psi);
  Htot = Tp + Tm + U0 * docc;
  auto solver = Hermitian::getSolver(hubbard, kry_dim);
  solver.getGroundState(Htot, psi, 0);
  std::cout << std::setprecision(16) << "ground state energy: " << psi.eval(0, Htot).real() << std::endl;
  dump(out + ".gs", psi.host(0, false), (size_t)ndim);

  std::ofstream obs(out + ".obs");
  obs << std::setprecision(17);
  auto l_solver = Arnoldi::getSolver(hubbard, kry_dim_rt);
  for (int i = 1; i <= nt; ++i) {
    Htot = Tdr[i] * Tp + std::conj(Tdr[i]) * Tm + (ndr[i] + U0) * docc;  // :443
    auto weights = lindblad_weight(0, Nsite, psi, gate, hubbard);
    double wtot = 0.0;
    for (double w : weights) wtot += w;
    QJ_operator(Htot, qj_tot, Id, gamma * wtot);  // :467
    if (i == 1) dump_matrix(out + ".H1", Htot);
    l_solver.evolve(Htot, psi, 0, h);             // :469
    psi(1) /= std::sqrt(Hermitian::norm(psi(1)));  // :470
    obs << i * h << "\t" << psi.eval(1, docc).real() << "\t" << wtot << "\n";
    const value_type* src = psi.host(1, false);
    value_type* dst = psi.host(0, true);
    for (size_t k = 0; k < (size_t)ndim; ++k) dst[k] = src[k];  // psi(0) = psi(1), :480
  }
  obs.close();
  dump(out + ".psi", psi.host(0, false), (size_t)ndim);
  std::cout << "done " << nt << " steps" << std::endl;
  return 0;
}
