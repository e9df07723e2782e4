// spin.cpp -- XY chain ground state through the drop-in API.
//
// Exercises the reference's canonical call sequence (programs/spin.cpp:13-48
// of the reference): getQbits -> getNset -> Bose gates -> Constrain ->
// gate << U -> OperatorSum algebra -> `ham >> chain` -> getState ->
// shuffle_state -> Arnoldi::getLanczosSolver(...).getGroundState -> psi.eval.
// Only the include and the link line differ from a reference build.
//   usage: spin <Nsite> [krylov_dim]
#include <chrono>
#include <iostream>
#include <string>

#include "qudrip/qudrip.hpp"

using namespace qudrip;

int main(int argc, char** argv) {
  const int nsite = argc > 1 ? std::stoi(argv[1]) : 16;
  const int kdim = argc > 2 ? std::stoi(argv[2]) : 100;

  auto qubits = getQbits(nsite);
  auto nset = getNset(qubits);
  auto raise = getBoseGate(qubits);
  auto lower = getBoseGate(qubits);

  std::cout << "Constraining..." << std::endl;
  auto chain = Constrain(qubits, nset = nsite / 2);
  std::cout << "dimension of Hil space = " << chain.range() << std::endl;

  raise << 0.5 * (pauli_x + II * pauli_y);
  lower << 0.5 * (pauli_x - II * pauli_y);

  auto t0 = std::chrono::steady_clock::now();
  auto ham = getEmptyOperator();
  for (int i = 0; i + 1 < nsite; ++i) ham = ham + raise(i + 1) * lower(i) + raise(i) * lower(i + 1);
  auto hmat = ham >> chain;
  double build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "nnz = " << hmat.nonZeros() << ", build " << build_ms << " ms" << std::endl;

  auto psi = getState(chain, 1);
  shuffle_state(psi(0));
  auto solver = Arnoldi::getLanczosSolver(chain, kdim);
  solver.getGroundState(hmat, psi(0), 0);

  std::cout.precision(16);
  auto e = psi.eval(0, hmat);
  std::cout << "E = " << e << std::endl;
  std::cout << "norm = " << Hermitian::norm(psi(0)) << std::endl;
  std::cout << "act_dim = " << solver.act_dim() << std::endl;
  return 0;
}
