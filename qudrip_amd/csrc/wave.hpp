// wave.hpp -- wavefront (64-lane) reductions for the HIP kernels.
#pragma once

#include "qd_internal.hpp"

// Wavefront sums with DPP lane moves (vector ALU, no LDS round trips):
// quad xor 1 and 2, half-row and row mirrors (every lane of a 16-lane row
// holds the row's sum), then row broadcasts 15 and 31 accumulate the rows
// into lane 63, which is broadcast to every lane.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum_d(double v) {
  v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xF>(v);  // row_mirror
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}
__device__ inline cplx wave_sum(cplx s) { return cmake(wave_sum_d(s.x), wave_sum_d(s.y)); }
