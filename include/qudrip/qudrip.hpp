// qudrip/qudrip.hpp -- umbrella header of the MI355X drop-in API.
//
// `#include "qudrip/qudrip.hpp"` in place of the reference's qudrip.hpp
// (qudrip.hpp:59-67 include order) gives the same names in namespace qudrip
// for the Krylov path: indices and sectors, gates and operator algebra,
// `ham >> index`, states, Arnoldi/Lanczos solvers.  Link with
// qudrip_amd/libqudrip_hip.so.
#pragma once

#include <complex>

#include "matrix.hpp"
#include "utils.hpp"
#include "runtime.hpp"
#include "index.hpp"
#include "operator.hpp"
#include "mode.hpp"
#include "state.hpp"
#include "sparse_algorithm.hpp"
#include "mc_manager.hpp"

namespace qudrip {
using cplx = std::complex<double>;
}
