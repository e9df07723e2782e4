// qudrip/mode.hpp -- matrices for boson-mode gates (operators/mode.hpp:36-78):
// LadderMatrix, CoherentProjector, FockProjector, as `modeOp << M` takes
// them.  A mode gate with matrix U takes level m to sum_j U(m, j) |j>
// (mode.hpp:80-103), so LadderMatrix(n, Mode::Upper), with sqrt(i+1) at
// (i+1, i), lowers the level and Mode::Lower raises it.
#pragma once

#include <cmath>
#include <complex>

#include "matrix.hpp"

namespace qudrip {

enum class Mode { Upper = 1, Lower = -1 };  // mode.hpp:36

inline Matrix LadderMatrix(std::size_t ndim, Mode mode = Mode::Upper) {  // :38-52
  Matrix m(ndim, ndim);
  for (std::size_t i = 0; i + 1 < ndim; ++i) {
    if (mode == Mode::Upper)
      m(i + 1, i) = std::sqrt((double)(i + 1));
    else
      m(i, i + 1) = std::sqrt((double)(i + 1));
  }
  return m;
}

inline Matrix CoherentProjector(std::size_t ndim, value_type amp) {  // :55-67
  std::vector<value_type> a(ndim);
  a[0] = 1.0;
  for (std::size_t i = 1; i < ndim; ++i) a[i] = a[i - 1] * amp / std::sqrt((double)i);
  const double s = std::exp(-std::abs(amp) * std::abs(amp) / 2.0);
  Matrix m(ndim, ndim);
  for (std::size_t i = 0; i < ndim; ++i)
    for (std::size_t j = 0; j < ndim; ++j) m(i, j) = (a[i] * s) * std::conj(a[j] * s);
  return m;
}

inline Matrix FockProjector(std::size_t ndim, std::size_t idx) {  // :70-77
  Matrix m(ndim, ndim);
  m(idx, idx) = 1.0;
  return m;
}

}  // namespace qudrip
