// qudrip/utils.hpp -- small host utilities of the reference's public API.
//
//   range(N), range(M, N), range(M, N, STEP)   qudrip.hpp:20-40: index vectors
//                                               for range-for loops
//   sec_clock<T>, get_clock<T>()               utils.hpp:142-163: a stopwatch;
//                                               note() restarts it and returns
//                                               the start time, release()
//                                               returns the elapsed count of T
//                                               times 1e-3 (seconds for the
//                                               default milliseconds)
// Pure host code: no device call.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <numeric>
#include <vector>

namespace qudrip {

using range_type = std::vector<std::size_t>;

inline range_type range(const std::size_t N) {
  range_type r(N);
  std::iota(r.begin(), r.end(), std::size_t(0));
  return r;
}

inline range_type range(const std::size_t M, const std::size_t N) {
  range_type r(N - M);
  std::iota(r.begin(), r.end(), M);
  return r;
}

// (N - M + 1) / STEP values M, M + STEP, ... (the reference's count)
inline range_type range(const std::size_t M, const std::size_t N, const std::size_t STEP) {
  range_type r((N - M + 1) / STEP);
  for (std::size_t k = 0; k < r.size(); ++k) r[k] = M + k * STEP;
  return r;
}

template <typename T>
class sec_clock {
  std::chrono::time_point<std::chrono::high_resolution_clock> t_;

 public:
  sec_clock() : t_(std::chrono::high_resolution_clock::now()) {}

  auto note() {
    t_ = std::chrono::high_resolution_clock::now();
    return t_;
  }

  auto release() {
    const auto t1 = std::chrono::high_resolution_clock::now();
    return std::chrono::duration_cast<T>(t1 - t_).count() * 1e-3;
  }
};

template <typename T = std::chrono::milliseconds>
auto get_clock() {
  return sec_clock<T>();
}

}  // namespace qudrip
