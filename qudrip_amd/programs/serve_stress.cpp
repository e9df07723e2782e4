// serve_stress.cpp -- the solve server under concurrent load, from C++ alone.
//
// Registers XXZ chain models (N = 16, 18, 20; ground states) and the 3x4
// Hubbard (6,6) model (two-block form; real-time requests), submits a mixed
// stream of requests at once to W workers, drains, and repeats the stream
// `rounds` times on the same server.  Checks every status, and that requests
// with the same model, coefficient and seed give the same energy and
// observable in every round and position of the stream (a served result
// does not depend on what else runs).  Built also against the library with
// AddressSanitizer on its host code (`make asan`, bin/serve_stress_asan),
// to check the server's host memory use while several workers launch at
// once.
//   usage: serve_stress <workers> <requests> [rounds] [rt_steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>

#include "qudrip/qudrip.hpp"

using namespace qudrip;

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
// the HIP runtime keeps its allocations until exit: no leak report at exit
extern "C" const char* __asan_default_options() { return "detect_leaks=0"; }
#endif
#endif

namespace {

struct Program {
  std::vector<qd_term> terms;
  std::vector<qd_op> ops;
};

Program flatten(const OperatorSum& h) {
  Program p;
  int nb = 0;
  h.program(p.terms, p.ops, &nb);
  return p;
}

void check(int rc, qd_server* s, const char* what) {
  if (rc != QD_OK) {
    std::fprintf(stderr, "%s failed: %d (%s)\n", what, rc, s ? qd_server_last_error(s) : "");
    std::exit(1);
  }
}

// XXZ chain split as H(Jz) = A + Jz D (periodic), sector Sz = 0
int add_xxz(qd_server* s, int N) {
  auto q = getQbits(N);
  auto gc = getBoseGate(q), ga = getBoseGate(q), gz = getBoseGate(q);
  gc << 0.5 * (pauli_x + II * pauli_y);
  ga << 0.5 * (pauli_x - II * pauli_y);
  gz << 0.5 * pauli_z;
  auto A = getEmptyOperator();
  auto D = getEmptyOperator();
  for (int i = 0; i < N; ++i) {
    const int j = (i + 1) % N;
    A = A + 0.5 * (gc(j) * ga(i) + gc(i) * ga(j));
    D = D + gz(j) * gz(i);
  }
  const Program pa = flatten(A), pd = flatten(D);
  const qd_constraint c{(1ull << N) - 1ull, N / 2};
  int id = -1;
  check(qd_server_add_model(s, N, &c, 1, pa.terms.data(), (int)pa.terms.size(), pa.ops.data(), (int)pa.ops.size(),
                            pd.terms.data(), (int)pd.terms.size(), pd.ops.data(), (int)pd.ops.size(), 0, &id),
        s, "qd_server_add_model (XXZ)");
  return id;
}

// Hubbard model on the 3x4 periodic lattice split as H(U) = T + U D, sector (6, 6)
int add_hubbard(qd_server* s) {
  const int Lx = 3, Ly = 4, nsite = Lx * Ly;
  auto q = getQbits(2 * nsite);
  auto g = getHopFermiGate(q);
  auto T = getEmptyOperator();
  auto D = getEmptyOperator();
  std::vector<std::pair<int, int>> bonds;
  for (int y = 0; y < Ly; ++y)
    for (int x = 0; x < Lx; ++x) {
      const int i = y * Lx + x;
      bonds.push_back({i, y * Lx + (x + 1) % Lx});
      bonds.push_back({i, ((y + 1) % Ly) * Lx + x});
    }
  for (auto [a, b] : bonds)
    for (int off : {0, nsite}) T = T + (-1.0) * g(a + off, b + off) + (-1.0) * g(b + off, a + off);
  for (int i = 0; i < nsite; ++i) D = D + g(i, i) * g(i + nsite, i + nsite);
  const Program pt = flatten(T), pd = flatten(D);
  const qd_constraint c[2] = {{(1ull << nsite) - 1ull, 6}, {((1ull << nsite) - 1ull) << nsite, 6}};
  int id = -1;
  check(qd_server_add_model(s, 2 * nsite, c, 2, pt.terms.data(), (int)pt.terms.size(), pt.ops.data(),
                            (int)pt.ops.size(), pd.terms.data(), (int)pd.terms.size(), pd.ops.data(),
                            (int)pd.ops.size(), 1, &id),
        s, "qd_server_add_model (Hubbard)");
  return id;
}

}  // namespace

int main(int argc, char** argv) {
  const int workers = argc > 1 ? std::atoi(argv[1]) : 6;
  const int nreq = argc > 2 ? std::atoi(argv[2]) : 48;
  const int rounds = argc > 3 ? std::atoi(argv[3]) : 2;
  const int rt_steps = argc > 4 ? std::atoi(argv[4]) : 20;
  qd_server* s = nullptr;
  check(qd_server_create(0, size_t(64) << 30, workers, &s), nullptr, "qd_server_create");
  std::map<int, int> xxz;
  for (int N : {16, 18, 20}) xxz[N] = add_xxz(s, N);
  const int hub = add_hubbard(s);
  // identical requests recur: key (kind, N, coefficient index, seed)
  std::map<std::tuple<int, int, int, int>, std::pair<double, double>> seen;
  int failures = 0, done = 0;
  for (int r = 0; r < rounds; ++r) {
    std::vector<std::tuple<int, int, int, int>> keys(nreq);
    for (int i = 0; i < nreq; ++i) {
      const int pos = (i * 7 + r * 3) % nreq;  // a different order every round
      qd_request q{};
      q.tag = (int64_t)r * nreq + i;
      // kind from pos % 3, then model, coefficient and seed from the higher
      // digits: every (model, coefficient, seed) recurs within and across rounds
      const int ci = (pos / 9) % 3, seed = 100 + (pos / 27) % 2;
      if (pos % 3 == 2) {
        q.kind = QD_REQ_EVOLVE;
        q.model = hub;
        q.coeff = 2.0 + 2.0 * ci;
        q.k = 30;
        q.chunk = 16;
        q.tol = 1e-13;
        q.steps = rt_steps;
        q.dt = 0.05;
        keys[i] = {1, 12, ci, seed};
      } else {
        const int N = 16 + 2 * ((pos / 3) % 3);
        q.kind = QD_REQ_GROUND_STATE;
        q.model = xxz[N];
        q.coeff = 0.5 * ci;
        q.k = 150;
        q.chunk = 4;
        q.tol = 1e-12;
        keys[i] = {0, N, ci, seed};
      }
      q.seed = (uint64_t)seed;
      check(qd_server_submit(s, &q), s, "qd_server_submit");
    }
    check(qd_server_drain(s), s, "qd_server_drain");
    std::vector<qd_result> res(nreq);
    int got = 0;
    check(qd_server_poll(s, res.data(), nreq, &got), s, "qd_server_poll");
    if (got != nreq) {
      std::fprintf(stderr, "round %d: %d results for %d requests\n", r, got, nreq);
      return 1;
    }
    for (const qd_result& x : res) {
      const int i = (int)(x.tag - (int64_t)r * nreq);
      if (x.status != QD_OK) {
        std::fprintf(stderr, "request %lld: status %d\n", (long long)x.tag, x.status);
        ++failures;
        continue;
      }
      auto it = seen.find(keys[i]);
      if (it == seen.end()) {
        seen[keys[i]] = {x.energy, x.observable};
      } else if (std::fabs(x.energy - it->second.first) > 1e-12 * std::fabs(it->second.first) ||
                 std::fabs(x.observable - it->second.second) > 1e-12 * (1.0 + std::fabs(it->second.second))) {
        std::fprintf(stderr, "request %lld differs from an identical one: %.17g vs %.17g\n", (long long)x.tag,
                     x.energy, it->second.first);
        ++failures;
      }
      ++done;
    }
    std::printf("round %d: %d requests served with %d workers\n", r, got, workers);
    std::fflush(stdout);
  }
  qd_server_stats st{};
  check(qd_server_get_stats(s, &st), s, "qd_server_get_stats");
  check(qd_server_destroy(s), nullptr, "qd_server_destroy");
  std::printf("serve_stress: %d results, %zu distinct requests, %d failures, max running %d\n", done, seen.size(),
              failures, st.max_running);
  return failures ? 1 : 0;
}
