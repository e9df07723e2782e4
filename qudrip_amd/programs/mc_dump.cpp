// mc_dump.cpp -- McManager output through the drop-in API (mc_manager.hpp,
// the accumulation and text tables a reference driver writes): XY chain of N
// qubits, half filling, a seeded start evolved nt-1 Arnoldi steps of h; at
// every step the energy (complex observable, one value) and the site
// occupations (real observable, N values) are accumulated for two samples
// (the state and the same state again), averaged with normAll and written by
// outFile.
//   usage: mc_dump <N> <nt> <h> <out-prefix>   ->  <prefix>.E.txt, <prefix>.n.txt
#include <iostream>
#include <string>

#include "qudrip/qudrip.hpp"

using namespace qudrip;

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "usage: mc_dump <N> <nt> <h> <out-prefix>" << std::endl;
    return 2;
  }
  const int N = std::stoi(argv[1]), nt = std::stoi(argv[2]);
  const double h = std::stod(argv[3]);
  const std::string out = argv[4];
  auto q = getQbits(N);
  auto gc = getBoseGate(q), ga = getBoseGate(q), gn = getBoseGate(q);
  gc << 0.5 * (pauli_x + II * pauli_y);
  ga << 0.5 * (pauli_x - II * pauli_y);
  gn << (Matrix(2, 2) << 0, 0, 0, 1).finished();
  auto ham = getEmptyOperator();
  for (int i = 0; i + 1 < N; ++i) ham = ham + gc(i + 1) * ga(i) + gc(i) * ga(i + 1);
  auto nset = getNset(q);
  auto sub = Constrain(q, nset = N / 2);
  auto H = ham >> sub;
  auto psi = getState(sub, nt);
  shuffle_seed() = 5;
  shuffle_state(psi(0));
  auto solver = Arnoldi::getSolver(sub, 12);
  for (int t = 0; t + 1 < nt; ++t) solver.evolve(H, psi, t, h, 12);
  auto mc = getMC(psi);
  mc.addObs("E", 1, MC_COMPLEX);
  mc.addObs("n", N, MC_REAL);
  mc.updateMatrix("E", 0, H);
  for (int i = 0; i < N; ++i) mc.updateMatrix("n", i, gn(i) >> sub);
  for (int sample = 0; sample < 2; ++sample)
    for (int t = 0; t < nt; ++t) mc.incrAllObs(t);
  mc.normAll();
  mc.outFile(h, out + ".E.txt", {"E"});
  mc.outFile(h, out + ".n.txt", {"n"});
  std::cout << "written " << out << std::endl;
  return 0;
}
