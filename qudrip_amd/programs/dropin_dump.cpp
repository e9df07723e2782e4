// dropin_dump.cpp -- builds a model through the drop-in C++ API and writes
// its term program and `>>` matrix as raw binary files, so the tests can
// compare them with the Python mirror and the CPU oracle.
//   usage: dropin_dump <xy|xxz|hubchain|hubnd|jc|lazy> <N> <out-prefix> [Jxy Jz | t U | g D]
//          [herm <kry_dim> <h>]
// Files: <prefix>.terms (qd_term[]), <prefix>.ops (qd_op[]), <prefix>.outer,
// <prefix>.inner (int32), <prefix>.vals (complex128), <prefix>.y (H x for the
// seeded x of seed 7; for "lazy" computed as psi2 = ham * psi1).  "hubnd" is
// the Hubbard chain in the sector that also fixes N/4 doubly occupied sites
// (getNd).  "jc" is one qubit times a boson mode of N levels (Jaynes-
// Cummings, SingleModeIndex / Indices / getModeOp), its mode matrices in
// <prefix>.mats.  With
// "herm k h" the two-vector solver (Hermitian::getSolver(sub, k)) then
// evolves x by h into <prefix>.ev and replaces x by its ground state,
// <prefix>.gs, with the Ritz values in <prefix>.E (storeE) and
// "error = ..." on stdout.
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "qudrip/qudrip.hpp"

using namespace qudrip;

template <class T>
static void dump(const std::string& path, const T* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), sizeof(T) * n);
}

static int jaynes_cummings(int nd, double g, double D, const std::string& out) {
  auto q = getQbits(1);
  SingleModeIndex<> md(nd);
  auto idx = q * md;
  auto sp = getBoseGate(q), sm = getBoseGate(q), ne = getBoseGate(q);
  sp << (Matrix(2, 2) << 0, 1, 0, 0).finished();
  sm << (Matrix(2, 2) << 0, 0, 1, 0).finished();
  ne << (Matrix(2, 2) << 0, 0, 0, 1).finished();
  auto ad = getModeOp(md), a = getModeOp(md);
  ad << LadderMatrix(nd, Mode::Lower);
  a << LadderMatrix(nd, Mode::Upper);
  auto ham = getEmptyOperator();
  ham = ham + ad(0) * a(0);
  ham = ham + D * ne(0);
  ham = ham + g * (sp(0) * a(0)) + g * (sm(0) * ad(0));
  std::vector<qd_term> terms;
  std::vector<qd_op> ops;
  std::vector<double> mats;
  int nb = 0;
  ham.program(terms, ops, &nb, &mats);
  dump(out + ".terms", terms.data(), terms.size());
  dump(out + ".ops", ops.data(), ops.size());
  dump(out + ".mats", mats.data(), mats.size());
  auto sub = Constrain(idx);
  auto H = ham >> sub;
  dump(out + ".outer", H.outerIndexPtr(), (size_t)H.rows() + 1);
  dump(out + ".inner", H.innerIndexPtr(), (size_t)H.nonZeros());
  dump(out + ".vals", H.valuePtr(), (size_t)H.nonZeros());
  auto psi = getState(sub, 2);
  shuffle_seed() = 7;
  shuffle_state(psi(0));
  qd_spmv(detail::ctx(), H.linop(), psi.device(0), psi.device(1, true));
  dump(out + ".x", psi.host(0, false), sub.range());
  dump(out + ".y", psi.host(1, false), sub.range());
  psi(1) = ham * psi(0);  // the lazy product on the composite index
  dump(out + ".ylazy", psi.host(1, false), sub.range());
  std::cout << "dumped " << terms.size() << " terms, n = " << sub.range() << std::endl;
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: dropin_dump <xy|xxz|hubchain|lazy> <N> <out-prefix> [a b]" << std::endl;
    return 2;
  }
  std::string model = argv[1];
  int N = std::stoi(argv[2]);
  std::string out = argv[3];
  double a = argc > 4 ? std::stod(argv[4]) : 1.0, b = argc > 5 ? std::stod(argv[5]) : 1.0;
  if (model == "jc") return jaynes_cummings(N, a, b, out);

  bool hub = (model == "hubchain" || model == "hubnd" || model == "lazy");
  auto q = getQbits(hub ? 2 * N : N);
  OperatorSum ham = getEmptyOperator();
  if (!hub) {
    auto gc = getBoseGate(q), ga = getBoseGate(q), gz = getBoseGate(q);
    gc << 0.5 * (pauli_x + II * pauli_y);
    ga << 0.5 * (pauli_x - II * pauli_y);
    gz << 0.5 * pauli_z;
    if (model == "xy") {
      for (int i = 0; i + 1 < N; ++i) ham = ham + gc(i + 1) * ga(i) + gc(i) * ga(i + 1);
    } else {
      for (int i = 0; i < N; ++i) {
        int j = (i + 1) % N;
        ham = ham + (0.5 * a) * (gc(j) * ga(i) + gc(i) * ga(j)) + b * (gz(j) * gz(i));
      }
    }
  } else {
    auto g = getHopFermiGate(q);
    for (int i = 0; i + 1 < N; ++i)
      for (int off : {0, N}) ham = ham + (-a) * g(i + off, i + 1 + off) + (-a) * g(i + 1 + off, i + off);
    for (int i = 0; i < N; ++i) ham = ham + b * (g(i, i) * g(i + N, i + N));
  }
  std::vector<qd_term> terms;
  std::vector<qd_op> ops;
  int nb = 0;
  ham.program(terms, ops, &nb);
  dump(out + ".terms", terms.data(), terms.size());
  dump(out + ".ops", ops.data(), ops.size());

  auto nset = getNset(q);
  auto nup = getNup(q, N);
  auto ndo = getNdo(q, N);
  auto nd = getNd(q, N);
  auto sub = model == "hubnd" ? Constrain(q, nup = N / 2, ndo = N / 2, nd = N / 4)
             : hub            ? Constrain(q, nup = N / 2, ndo = N / 2)
                              : Constrain(q, nset = N / 2);
  auto psi = getState(sub, 2);
  shuffle_seed() = 7;
  shuffle_state(psi(0));
  if (model == "lazy") {
    psi(1) = ham * psi(0);  // lazy product, evaluated matrix-free
  } else {
    auto H = ham >> sub;
    dump(out + ".outer", H.outerIndexPtr(), (size_t)H.rows() + 1);
    dump(out + ".inner", H.innerIndexPtr(), (size_t)H.nonZeros());
    dump(out + ".vals", H.valuePtr(), (size_t)H.nonZeros());
    std::vector<double> y(2 * sub.range());
    qd_spmv(detail::ctx(), H.linop(), psi.device(0), psi.device(1, true));
  }
  dump(out + ".x", psi.host(0, false), sub.range());
  dump(out + ".y", psi.host(1, false), sub.range());
  if (argc > 8 && std::string(argv[6]) == "herm" && model != "lazy") {
    auto H = ham >> sub;
    auto solver = Hermitian::getSolver(sub, std::stoi(argv[7]));
    solver.evolve(H, psi, 0, std::stod(argv[8]));
    dump(out + ".ev", psi.host(1, false), sub.range());
    solver.getGroundState(H, psi, 0);
    dump(out + ".gs", psi.host(0, false), sub.range());
    solver.storeE(out + ".E");
    std::cout.precision(17);
    std::cout << "error = " << solver.error() << std::endl;
  }
  std::cout << "dumped " << terms.size() << " terms, n = " << sub.range() << std::endl;
  return 0;
}
