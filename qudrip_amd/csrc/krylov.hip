// krylov.hip -- K7/K8/K9: Krylov expansion with full reorthogonalisation,
// ground state and real-time step (sparse_algorithm.hpp:372-604).
//
// Basis V (kry_dim + 1 columns of length n, column-major, device-resident).
// One expansion step j turns v_j into v_{j+1}:
//     w      = H v_j                          (K5 or K6, written into V[:, j+1])
//     h1     = V_{0..j}^H w ;  w -= V h1       (classical Gram-Schmidt)
//     h2     = V_{0..j}^H w ;  w -= V h2       (second pass: CGS2)
//     beta   = |w| ;  v_{j+1} = w / beta
// and column j of the Hessenberg matrix is h1 + h2 (rows 0..j) and beta
// (row j+1).  That is method 0; the default (method 1) is the delayed form
// below, two passes per step instead of four, and Hermitian operators can
// use method 2, a local window with partial reorthogonalisation.  The
// reference's Lanczos (expandBasis<1>, :395-425) keeps only the last two
// vectors; full reorthogonalisation changes no exact-arithmetic quantity
// (same Krylov space, same tridiagonal) and keeps the Ritz values free of
// ghost copies.  Every reduction has a fixed tree (block partials summed in
// index order), so a solve gives the same bits whatever else runs.
//
// Breakdown (beta < 1e-10, :387-390, :421) sets a device flag: the remaining
// kernels of the expansion see it and return, so the whole expansion is
// enqueued without a host round trip and read back once at the end.
#include <algorithm>
#include <cmath>

#include "qd_internal.hpp"
#include "wave.hpp"

int qd_linop_apply(qd_linop* op, const cplx* x, cplx* y, const double* d_xscale, int xscale_idx,
                   const cplx* d_diag, cplx shift);

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));
constexpr int KT = 256;

// Basis columns are streamed once per pass and are far larger than the
// caches; loading (and writing) them non-temporally leaves L2/MALL to the
// vectors that are read again (u, w, the product's x).
__device__ __forceinline__ cplx ld_nt(const cplx* p) {
  const dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(p));
  return cmake(v.x, v.y);
}  // threads per block for the tall-skinny kernels


// gfx950 cross-half swaps: permlane32 trades lanes 32-63 of a with lanes
// 0-31 of b; permlane16 trades the odd 16-lane rows of a with the even rows
// of b.
__device__ inline void swap_lanes32(double& a, double& b) {
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
  a = __hiloint2double((int)hi[0], (int)lo[0]);
  b = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ inline void swap_lanes16(double& a, double& b) {
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
  a = __hiloint2double((int)hi[0], (int)lo[0]);
  b = __hiloint2double((int)hi[1], (int)lo[1]);
}
// Eight wavefront sums in one butterfly (each exchange step halves the
// values a lane carries): afterwards lanes 8v .. 8v+7 hold the sum of d[v]
// over the wave.  About 34 instructions against 160 for eight wave_sum_d.
__device__ inline double wave_sum8(const double (&d)[8]) {
  double e[4], f[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {  // lane bit 5 picks d[i] or d[i + 4]
    double a = d[i], b = d[i + 4];
    swap_lanes32(a, b);
    e[i] = a + b;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {  // lane bit 4 picks e[i] or e[i + 2]
    double a = e[i], b = e[i + 2];
    swap_lanes16(a, b);
    f[i] = a + b;
  }
  const bool hi8 = (threadIdx.x & 8) != 0;  // lane bit 3 picks f[0] or f[1]
  const double own = hi8 ? f[1] : f[0], send = hi8 ? f[0] : f[1];
  double g = own + dpp_move<0x128, 0xF>(send);  // row_ror:8 = lane xor 8
  g += dpp_move<0xB1, 0xF>(g);
  g += dpp_move<0x4E, 0xF>(g);
  g += dpp_move<0x141, 0xF>(g);
  return g;
}

// Columns [lo, ncols) of the basis that a projection covers.  arr == nullptr:
// the whole basis.  Otherwise arr[j] is the first column for step j: >= 0 as
// set by the step before (0 when partial reorthogonalisation asks for a full
// pass), -1 for the local window max(0, j - W).
struct Win {
  const int* arr;
  int j;
  int W;
  __device__ int lo() const {
    if (!arr) return 0;
    const int v = arr[j];
    return v >= 0 ? v : max(0, j - W);
  }
};

// part[b * ncols + c] = sum over rows of block b of conj(V[r, c]) * w[r]
template <int RPT>
__global__ void __launch_bounds__(KT)
    gemv_t_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ w,
                  cplx* __restrict__ part, const int* __restrict__ flag, Win win = {nullptr, 0, 0}) {
  if (flag && *flag) return;
  __shared__ cplx sh[KT / 64][64];
  const int64_t base = (int64_t)blockIdx.x * KT * RPT;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  cplx wr[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    wr[q] = (r < n) ? w[r] : cmake(0.0, 0.0);
  }
  for (int c0 = win.lo(); c0 < ncols; c0 += 64) {
    const int cend = min(ncols, c0 + 64);
    for (int c = c0; c < cend; ++c) {
      const cplx* __restrict__ Vc = V + (int64_t)c * ld;
      cplx s = cmake(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        int64_t r = base + (int64_t)q * KT + threadIdx.x;
        if (r < n) {
          cplx v = Vc[r];
          s.x += v.x * wr[q].x + v.y * wr[q].y;
          s.y += v.x * wr[q].y - v.y * wr[q].x;
        }
      }
      s = wave_sum(s);
      if (lane == 0) sh[wv][c - c0] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < cend - c0) {
      cplx t = sh[0][threadIdx.x];
      for (int q = 1; q < KT / 64; ++q) t = cadd(t, sh[q][threadIdx.x]);
      part[(int64_t)blockIdx.x * ncols + c0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// h[c] = sum_b part[b * ncols + c]; kcol[c] (+)= h[c].  One block per column.
// With a window (win.arr set), entries c < nrv * jcols whose basis column
// c % jcols lies below the window are zero (no block wrote them).
__global__ void __launch_bounds__(KT)
    reduce_cols_kernel(const cplx* __restrict__ part, int nb, int ncols, cplx* __restrict__ h,
                       cplx* __restrict__ kcol, int accumulate, const int* __restrict__ flag, int nk = 1 << 30,
                       Win win = {nullptr, 0, 0}, int jcols = 1, int nrv = 0) {
  if (flag && *flag) return;
  __shared__ cplx sh[KT / 64];
  const int c = blockIdx.x;
  if (win.arr && c < nrv * jcols && c % jcols < win.lo()) {
    if (threadIdx.x == 0) {
      h[c] = cmake(0.0, 0.0);
      if (kcol && c < nk && !accumulate) kcol[c] = cmake(0.0, 0.0);
    }
    return;
  }
  cplx s = cmake(0.0, 0.0);
  for (int b = threadIdx.x; b < nb; b += KT) s = cadd(s, part[(int64_t)b * ncols + c]);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    cplx t = sh[0];
    for (int q = 1; q < KT / 64; ++q) t = cadd(t, sh[q]);
    h[c] = t;
    if (kcol && c < nk) kcol[c] = accumulate ? cadd(kcol[c], t) : t;
  }
}

// w[r] -= sum_c V[r, c] h[c]; optional block partial of |w|^2
template <int RPT>
__global__ void __launch_bounds__(KT)
    gemv_n_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ h,
                  cplx* __restrict__ w, double* __restrict__ norm_part, const int* __restrict__ flag,
                  Win win = {nullptr, 0, 0}) {
  if (flag && *flag) return;
  __shared__ double sh[KT / 64];
  const int64_t base = (int64_t)blockIdx.x * KT * RPT;
  cplx acc[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    acc[q] = (r < n) ? w[r] : cmake(0.0, 0.0);
  }
  for (int c = win.lo(); c < ncols; ++c) {
    const cplx hc = h[c];
    const cplx* __restrict__ Vc = V + (int64_t)c * ld;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      int64_t r = base + (int64_t)q * KT + threadIdx.x;
      if (r < n) {
        cplx v = Vc[r];
        acc[q].x -= v.x * hc.x - v.y * hc.y;
        acc[q].y -= v.x * hc.y + v.y * hc.x;
      }
    }
  }
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    if (r < n) {
      w[r] = acc[q];
      s += acc[q].x * acc[q].x + acc[q].y * acc[q].y;
    }
  }
  if (norm_part) {
    s = wave_sum_d(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = sh[0];
      for (int q = 1; q < KT / 64; ++q) t += sh[q];
      norm_part[blockIdx.x] = t;
    }
  }
}

// beta = sqrt(sum norm_part); kcol[j+1] = beta; inv = 1/beta; breakdown flag
__global__ void __launch_bounds__(KT)
    finalize_kernel(const double* __restrict__ norm_part, int nb, cplx* __restrict__ kcol_next, double* __restrict__ inv,
                    int* __restrict__ flag, int* __restrict__ bd_step, int step, double eps) {
  if (*flag) return;
  __shared__ double sh[KT / 64];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += KT) s += norm_part[b];
  s = wave_sum_d(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int q = 1; q < KT / 64; ++q) t += sh[q];
    double beta = sqrt(t);
    *kcol_next = cmake(beta, 0.0);
    *inv = 1.0 / beta;
    if (!(beta >= eps)) {
      *flag = 1;
      *bd_step = step;
    }
  }
}

__global__ void scale_kernel(int64_t n, cplx* __restrict__ w, const double* __restrict__ inv,
                             const int* __restrict__ flag) {
  if (flag && *flag) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = *inv;
  cplx v = w[i];
  w[i] = cmake(v.x * s, v.y * s);
}

// out[r] = sum_c V[r, c] y[c]
template <int RPT>
__global__ void __launch_bounds__(KT)
    combine_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ y,
                   cplx* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * KT * RPT;
  cplx acc[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) acc[q] = cmake(0.0, 0.0);
  for (int c = 0; c < ncols; ++c) {
    const cplx yc = y[c];
    const cplx* __restrict__ Vc = V + (int64_t)c * ld;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      int64_t r = base + (int64_t)q * KT + threadIdx.x;
      if (r < n) {
        const cplx v = ld_nt(Vc + r);
        acc[q].x += v.x * yc.x - v.y * yc.y;
        acc[q].y += v.x * yc.y + v.y * yc.x;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    if (r < n) out[r] = acc[q];
  }
}

__global__ void normalise_copy_kernel(int64_t n, const cplx* __restrict__ src, cplx* __restrict__ dst,
                                      const double* __restrict__ nrm2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = sqrt(*nrm2);
  cplx v = src[i];
  dst[i] = cmake(v.x / s, v.y / s);
}

__global__ void norm_partial_kernel(int64_t n, const cplx* __restrict__ a, double* __restrict__ part) {
  __shared__ double sh[KT / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)gridDim.x * KT) {
    cplx v = a[i];
    s += v.x * v.x + v.y * v.y;
  }
  s = wave_sum_d(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int q = 1; q < KT / 64; ++q) t += sh[q];
    part[blockIdx.x] = t;
  }
}
__global__ void sum_double_kernel(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[KT / 64];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += KT) s += part[b];
  s = wave_sum_d(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int q = 1; q < KT / 64; ++q) t += sh[q];
    *out = t;
  }
}


// ---------------------------------------------------------------- DCGS2
// Delayed classical Gram-Schmidt with reorthogonalisation (after Bielich,
// Langou, Thomas, Swirydowicz, Yamazaki, Boman, "Low-synch Gram-Schmidt with
// delayed reorthogonalization for Krylov solvers", Parallel Computing 2022).
// Step j holds u_j = A q_{j-1} projected ONCE against Q_{j-1}.  One pass over
// Q_{j-1} computes c = Q^H u_j, d = Q^H w (w = A u_j), u^H u, u^H w; then
//   alpha = sqrt(u^H u - |c|^2),  q_j = (u_j - Q c) / alpha      (2nd pass of u_j)
//   H[:, j-1] += c, H[j, j-1] = alpha
//   A q_j = (w - Q_j e) / alpha,  e = H[0:j+1, 0:j] c            (Arnoldi relation)
//   f = Q_j^H w / alpha = [d ; (u^H w - c^H d) / alpha] / alpha
//   H[0:j+1, j] = f - e / alpha                                   (1st pass of A q_j)
//   u_{j+1} = w / alpha - Q_{j-1} (f[:j] - c f[j] / alpha) - u_j f[j] / alpha
// and a second pass over Q_{j-1} writes q_j and u_{j+1}: 2 passes per step
// instead of CGS2's 4.

// The reduction pass of a DCGS2 step in a form whose blocks all stay
// resident: gridDim.x blocks (at most what the device holds at once) take
// contiguous ranges of rows_pb rows (a multiple of 64) and walk them in tiles
// of KT * RPT rows; each wave adds the butterfly sums of a column pair into
// its own LDS accumulators.  Columns go in groups of cmax (u and w are read
// again for each further group).  Output:
// part[b * P + c] = Q_c^H u, part[b * P + ncols + c] = Q_c^H w for the
// window's columns, part[b * P + P - 2] = u^H u, P - 1 -> u^H w.
template <int RPT>
__global__ void __launch_bounds__(KT)
    dc_red_rows_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ u,
                       const cplx* __restrict__ w, cplx* __restrict__ part, const int* __restrict__ flag, Win win,
                       int64_t rows_pb, int cmax) {
  if (*flag) return;
  extern __shared__ double acc[];  // [KT / 64][cmax][4]: Re, Im of Q^H u, then of Q^H w
  __shared__ cplx sx[KT / 64][2];
  const int P = 2 * ncols + 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * rows_pb;
  const int64_t r1 = min(n, r0 + rows_pb);
  double* __restrict__ wacc = acc + (size_t)wv * cmax * 4;
  cplx uu = cmake(0.0, 0.0), uw = cmake(0.0, 0.0);
  const int lo = win.lo();
  // the first group always runs (it also accumulates u^H u and u^H w)
  for (int cg = lo; cg == lo || cg < ncols; cg += cmax) {
    const int cend = min(ncols, cg + cmax);
    for (int i = lane; i < (cend - cg) * 4; i += 64) wacc[i] = 0.0;
    __syncthreads();
    for (int64_t t0 = r0; t0 < r1; t0 += (int64_t)KT * RPT) {
      cplx ur[RPT], wr[RPT];
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int64_t r = t0 + (int64_t)q * KT + threadIdx.x;
        ur[q] = (r < r1) ? u[r] : cmake(0.0, 0.0);
        wr[q] = (r < r1) ? w[r] : cmake(0.0, 0.0);
        if (cg == lo) {
          uu = cadd(uu, cdotmul(ur[q], ur[q]));
          uw = cadd(uw, cdotmul(ur[q], wr[q]));
        }
      }
      for (int c = cg; c < cend; c += 2) {
        const bool two = c + 1 < cend;  // a lone last column repeats in the second slot, unused
        const cplx* __restrict__ Vc = V + (int64_t)c * ld;
        const cplx* __restrict__ Vd = two ? Vc + ld : Vc;
        cplx a0 = cmake(0.0, 0.0), b0 = cmake(0.0, 0.0), a1 = cmake(0.0, 0.0), b1 = cmake(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const int64_t r = t0 + (int64_t)q * KT + threadIdx.x;
          if (r < r1) {
            const cplx v0 = ld_nt(Vc + r), v1 = ld_nt(Vd + r);
            a0 = cadd(a0, cdotmul(v0, ur[q]));
            a1 = cadd(a1, cdotmul(v1, ur[q]));
            b0 = cadd(b0, cdotmul(v0, wr[q]));
            b1 = cadd(b1, cdotmul(v1, wr[q]));
          }
        }
        const double d8[8] = {a0.x, a0.y, a1.x, a1.y, b0.x, b0.y, b1.x, b1.y};
        const double g = wave_sum8(d8);
        if ((lane & 7) == 0) {  // lane 8v holds d8[v]
          const int v = lane >> 3, second = (v >> 1) & 1;
          if (two || !second) wacc[(c - cg + second) * 4 + (v >> 2) * 2 + (v & 1)] += g;
        }
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < cend - cg; t += KT) {
      double sm[4] = {0.0, 0.0, 0.0, 0.0};
      for (int q = 0; q < KT / 64; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) sm[k] += acc[((size_t)q * cmax + t) * 4 + k];
      part[(int64_t)blockIdx.x * P + cg + t] = cmake(sm[0], sm[1]);
      part[(int64_t)blockIdx.x * P + ncols + cg + t] = cmake(sm[2], sm[3]);
    }
    __syncthreads();
  }
  uu = wave_sum(uu);
  uw = wave_sum(uw);
  if (lane == 0) {
    sx[wv][0] = uu;
    sx[wv][1] = uw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    cplx t = sx[0][0], t2 = sx[0][1];
    for (int q = 1; q < KT / 64; ++q) {
      t = cadd(t, sx[q][0]);
      t2 = cadd(t2, sx[q][1]);
    }
    part[(int64_t)blockIdx.x * P + P - 2] = t;
    part[(int64_t)blockIdx.x * P + P - 1] = t2;
  }
}

// The reduction pass of the step that finalises a basis (one vector u):
// part[b * P + k] = Q_k^H u for the window's columns k < ncols,
// part[b * P + P - 2] = u^H u, P - 1 -> 0; P = ncols + 2; one tile of KT * RPT
// rows per block.
template <int RPT>
__global__ void __launch_bounds__(KT)
    dc_red_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ u,
                  cplx* __restrict__ part, const int* __restrict__ flag, Win win) {
  if (*flag) return;
  __shared__ cplx sh[KT / 64][64];
  __shared__ cplx sx[KT / 64];
  const int P = ncols + 2;
  const int64_t base = (int64_t)blockIdx.x * KT * RPT;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  cplx ur[RPT];
  cplx uu = cmake(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    ur[q] = (r < n) ? u[r] : cmake(0.0, 0.0);
    uu = cadd(uu, cdotmul(ur[q], ur[q]));
  }
  uu = wave_sum(uu);
  if (lane == 0) sx[wv] = uu;
  for (int c0 = win.lo(); c0 < ncols; c0 += 64) {
    const int cend = min(ncols, c0 + 64);
    int c = c0;
    // two columns at a time: twice the loads in flight per wave and two
    // independent reduction chains
    for (; c + 1 < cend; c += 2) {
      const cplx* __restrict__ Vc = V + (int64_t)c * ld;
      const cplx* __restrict__ Vd = Vc + ld;
      cplx a0 = cmake(0.0, 0.0), a1 = cmake(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        int64_t r = base + (int64_t)q * KT + threadIdx.x;
        if (r < n) {
          a0 = cadd(a0, cdotmul(ld_nt(Vc + r), ur[q]));
          a1 = cadd(a1, cdotmul(ld_nt(Vd + r), ur[q]));
        }
      }
      a0 = wave_sum(a0);
      a1 = wave_sum(a1);
      if (lane == 0) {
        sh[wv][c - c0] = a0;
        sh[wv][c + 1 - c0] = a1;
      }
    }
    for (; c < cend; ++c) {
      const cplx* __restrict__ Vc = V + (int64_t)c * ld;
      cplx a = cmake(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        int64_t r = base + (int64_t)q * KT + threadIdx.x;
        if (r < n) a = cadd(a, cdotmul(ld_nt(Vc + r), ur[q]));
      }
      a = wave_sum(a);
      if (lane == 0) sh[wv][c - c0] = a;
    }
    __syncthreads();
    if ((int)threadIdx.x < cend - c0) {
      const int c = threadIdx.x;
      cplx t = sh[0][c];
      for (int q = 1; q < KT / 64; ++q) t = cadd(t, sh[q][c]);
      part[(int64_t)blockIdx.x * P + c0 + c] = t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cplx t = sx[0];
    for (int q = 1; q < KT / 64; ++q) t = cadd(t, sx[q]);
    part[(int64_t)blockIdx.x * P + P - 2] = t;
    part[(int64_t)blockIdx.x * P + P - 1] = cmake(0.0, 0.0);
  }
}

// Partial reorthogonalisation (Simon, "The Lanczos algorithm with partial
// reorthogonalization", Math. Comp. 42 (1984) 115-142) on top of the delayed
// scheme, for Hermitian operators: each step projects against the local
// window of W columns only, and omega_{j,k} ~ q_j^H q_k is estimated from the
// tridiagonal entries by Simon's recurrence
//   beta_j w_{j,k} = beta_{k+1} w_{j-1,k+1} + (alpha_k - alpha_{j-1}) w_{j-1,k}
//                    + beta_k w_{j-1,k-1} - beta_{j-1} w_{j-2,k} +- eps1 |T|
// (entries inside the window reset to eps1, w_{j,j} = 1).  When max |w_{j,k}|
// passes the trigger level the next two steps project against the whole
// basis (the pending vector's second pass and its successor's first and
// second), which keeps the basis semi-orthogonal: the tridiagonal then
// agrees with that of a fully orthogonal basis to O(eps |T|).
struct Pro {
  int* lo;      // first column per step (Win::arr); nullptr: no estimate
  double* om;   // three rotating rows of the estimate, stride ldo
  double* prm;  // [0] |T| estimate, [1] eps1, [2] trigger level, [3] full passes
  int ldo;
  int W;
};

__device__ void pro_step(const Pro& P, const cplx* __restrict__ KR, int ldk, int j, double beta) {
  __shared__ double smax[KT / 64];
  __syncthreads();  // column j - 1 is final
  const int lo_here = Win{P.lo, j, P.W}.lo();
  double* rj = P.om + (size_t)(j % 3) * P.ldo;
  const double* r1 = P.om + (size_t)((j + 2) % 3) * P.ldo;  // row j - 1
  const double* r2 = P.om + (size_t)((j + 1) % 3) * P.ldo;  // row j - 2
  const double aj1 = KR[(size_t)(j - 1) * ldk + j - 1].x;
  const double bj1 = j >= 2 ? KR[(size_t)(j - 2) * ldk + j - 1].x : 0.0;
  const double eps1 = P.prm[1];
  const double T = eps1 * P.prm[0];
  double mx = 0.0;
  for (int k = threadIdx.x; k < j; k += KT) {
    double v = eps1;
    if (k < lo_here) {  // lo_here <= j - W, so k + 1 < j - 1
      const double ak = KR[(size_t)k * ldk + k].x;
      const double bk1 = KR[(size_t)k * ldk + k + 1].x;
      double t = bk1 * r1[k + 1] + (ak - aj1) * r1[k] - bj1 * r2[k];
      if (k > 0) t += KR[(size_t)(k - 1) * ldk + k].x * r1[k - 1];
      v = (t + copysign(T, t)) / beta;
      mx = fmax(mx, fabs(v));
    }
    rj[k] = v;
  }
  if (threadIdx.x == 0) rj[j] = 1.0;
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = smax[0];
    for (int q = 1; q < KT / 64; ++q) m = fmax(m, smax[q]);
    P.prm[0] = fmax(P.prm[0], fabs(aj1) + bj1 + beta);
    if (lo_here == 0 && j > P.W) P.prm[3] += 1.0;
    if (m > P.prm[2]) {
      P.lo[j + 1] = 0;
      P.lo[j + 2] = 0;
    } else if (P.lo[j + 1] != 0) {
      P.lo[j + 1] = max(0, j + 1 - P.W);
    }
  }
}

__global__ void pro_init_kernel(int* __restrict__ lo, int nlo, double* __restrict__ om, double* __restrict__ prm,
                                double eps1, double trigger) {
  for (int i = threadIdx.x; i < nlo; i += blockDim.x) lo[i] = -1;
  if (threadIdx.x == 0) {
    om[0] = 1.0;  // w_{0,0}
    prm[0] = 0.0;
    prm[1] = eps1;
    prm[2] = trigger;
    prm[3] = 0.0;
  }
}

// small algebra of step j (one block).  red: reduced partials (layout above).
// KR column k at KR + k * ldk, rows 0..k+1.  Outputs coefA (= c), coefB and
// scal = {1/alpha, Re, Im of f[j]/alpha}.  last != 0: finalisation (no w).
__global__ void __launch_bounds__(KT)
    dc_small_kernel(const cplx* __restrict__ red, int j, int last, cplx* __restrict__ KR, int ldk,
                    cplx* __restrict__ coefA, cplx* __restrict__ coefB, double* __restrict__ scal,
                    int* __restrict__ flag, int* __restrict__ bd_step, double eps, Pro pro) {
  if (*flag) return;
  __shared__ double sred[KT / 64];
  __shared__ cplx cred[KT / 64];
  __shared__ double s_alpha;
  __shared__ cplx s_fj;
  const int NR = last ? 1 : 2;
  const int P = NR * j + 2;
  const cplx* c = red;
  const cplx* d = red + j;
  const double uu = red[P - 2].x;
  const cplx uw = red[P - 1];
  // |c|^2 and c^H d
  double cc = 0.0;
  cplx chd = cmake(0.0, 0.0);
  for (int i = threadIdx.x; i < j; i += KT) {
    cc += c[i].x * c[i].x + c[i].y * c[i].y;
    if (!last) chd = cadd(chd, cdotmul(c[i], d[i]));
  }
  cc = wave_sum_d(cc);
  chd = wave_sum(chd);
  if ((threadIdx.x & 63) == 0) {
    sred[threadIdx.x >> 6] = cc;
    cred[threadIdx.x >> 6] = chd;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    cplx t2 = cmake(0.0, 0.0);
    for (int q = 0; q < KT / 64; ++q) {
      t += sred[q];
      t2 = cadd(t2, cred[q]);
    }
    double a2 = uu - t;
    double alpha = a2 > 0.0 ? sqrt(a2) : 0.0;
    s_alpha = alpha;
    if (!(alpha >= eps)) {
      *flag = 1;
      *bd_step = j;
    } else {
      // f[j] = (u^H w - c^H d) / alpha^2
      s_fj = last ? cmake(0.0, 0.0) : crmul(1.0 / (alpha * alpha), csub(uw, t2));
    }
  }
  __syncthreads();
  const double alpha = s_alpha;
  // finalise column j-1 (also on breakdown: its entries up to row j-1 are final)
  cplx* col = KR + (size_t)(j - 1) * ldk;
  for (int i = threadIdx.x; i < j; i += KT) col[i] = cadd(col[i], c[i]);
  if (threadIdx.x == 0) col[j] = cmake(alpha, 0.0);
  if (!(alpha >= eps)) return;
  if (pro.lo) pro_step(pro, KR, ldk, j, alpha);
  const double inv = 1.0 / alpha;
  for (int i = threadIdx.x; i < j; i += KT) coefA[i] = c[i];
  if (threadIdx.x == 0) scal[0] = inv;
  if (last) return;
  __threadfence_block();
  __syncthreads();
  const cplx fj = s_fj;
  // e = H[0:j+1, 0:j] c;  column j first pass = f - e/alpha
  cplx* colj = KR + (size_t)j * ldk;
  for (int i = threadIdx.x; i <= j; i += KT) {
    cplx e = cmake(0.0, 0.0);
    for (int k = (i > 0 ? i - 1 : 0); k < j; ++k) e = cadd(e, cmul(KR[(size_t)k * ldk + i], c[k]));
    cplx f = (i < j) ? crmul(inv, d[i]) : fj;
    colj[i] = csub(f, crmul(inv, e));
    if (i < j) coefB[i] = csub(f, crmul(inv, cmul(c[i], fj)));
  }
  if (threadIdx.x == 0) {
    scal[1] = fj.x * inv;
    scal[2] = fj.y * inv;
  }
}

// writes q_j = (u - Q coefA) / alpha into V[:, j]; unless FINAL also
// u <- w / alpha - Q coefB - u (f[j] / alpha)
template <int RPT, bool FINAL>
__global__ void __launch_bounds__(KT)
    dc_update_kernel(int64_t n, cplx* __restrict__ V, int64_t ld, int j, cplx* __restrict__ u,
                     const cplx* __restrict__ w, const cplx* __restrict__ coefA, const cplx* __restrict__ coefB,
                     const double* __restrict__ scal, const int* __restrict__ flag, Win win) {
  if (*flag) return;
  const int64_t base = (int64_t)blockIdx.x * KT * RPT;
  cplx s1[RPT], s2[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) s1[q] = s2[q] = cmake(0.0, 0.0);
  for (int c = win.lo(); c < j; ++c) {
    const cplx a = coefA[c];
    const cplx bcf = FINAL ? cmake(0.0, 0.0) : coefB[c];
    const cplx* __restrict__ Vc = V + (int64_t)c * ld;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      int64_t r = base + (int64_t)q * KT + threadIdx.x;
      if (r < n) {
        const cplx v = ld_nt(Vc + r);
        s1[q].x += v.x * a.x - v.y * a.y;
        s1[q].y += v.x * a.y + v.y * a.x;
        if (!FINAL) {
          s2[q].x += v.x * bcf.x - v.y * bcf.y;
          s2[q].y += v.x * bcf.y + v.y * bcf.x;
        }
      }
    }
  }
  const double inv = scal[0];
  const cplx us = cmake(scal[1], scal[2]);
  cplx* __restrict__ Vj = V + (int64_t)j * ld;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    if (r < n) {
      cplx ur = u[r];
      __builtin_nontemporal_store((ur.x - s1[q].x) * inv, &Vj[r].x);
      __builtin_nontemporal_store((ur.y - s1[q].y) * inv, &Vj[r].y);
      if (!FINAL) {
        cplx wr = w[r];
        cplx t = cmul(ur, us);
        u[r] = cmake(wr.x * inv - s2[q].x - t.x, wr.y * inv - s2[q].y - t.y);
      }
    }
  }
}

int pick_rpt(int64_t n, int ncu) {
  if (n >= (int64_t)KT * 8 * ncu * 4) return 8;
  if (n >= (int64_t)KT * 4 * ncu * 2) return 4;
  return 1;
}

}  // namespace

// ------------------------------------------------------------------ state
struct KrylovBufs {
  // device scalar area layout (doubles):
  //   [0] flag (int), [1] bd_step (int), [2] inv, [3] nrm2 scratch
  //   h buffer: kry_dim + 2 complex; KR: (kry_dim + 2) * (kry_dim + 1) complex
  //   partials: nb * (kry_dim + 1) complex; norm partials nb doubles
};

static int64_t kr_nb(qd_krylov* kr, int rpt) { return (kr->n + (int64_t)KT * rpt - 1) / ((int64_t)KT * rpt); }

struct KrylovDev {
  int* flag;
  int* bd_step;
  double* inv;
  double* nrm2;
  cplx* h;
  cplx* KR;
  cplx* part;
  double* npart;
  cplx* y;
  cplx* coefA;
  cplx* coefB;
  double* scal;
  cplx* red;
  double* prm;  // partial reorthogonalisation state (Pro)
  double* om;
  int* lo;
};

constexpr int PRO_W = 2;  // local window of the partial scheme

static KrylovDev kr_layout(qd_krylov* kr) {
  KrylovDev d;
  char* p = (char*)kr->d_scal;
  d.flag = (int*)p;
  d.bd_step = (int*)(p + 8);
  d.inv = (double*)(p + 16);
  d.nrm2 = (double*)(p + 24);
  p += 64;
  int K = kr->kry_dim;
  d.h = (cplx*)p;
  p += sizeof(cplx) * (K + 2);
  d.y = (cplx*)p;
  p += sizeof(cplx) * (K + 2);
  d.KR = (cplx*)p;
  p += sizeof(cplx) * (size_t)(K + 2) * (K + 1);
  d.coefA = (cplx*)p;
  p += sizeof(cplx) * (K + 2);
  d.coefB = (cplx*)p;
  p += sizeof(cplx) * (K + 2);
  d.red = (cplx*)p;
  p += sizeof(cplx) * (2 * K + 8);
  d.scal = (double*)p;
  p += sizeof(double) * 8;
  d.prm = (double*)p;
  p += sizeof(double) * 8;
  d.om = (double*)p;
  p += sizeof(double) * 3 * (size_t)(K + 2);
  d.lo = (int*)p;
  p += sizeof(int) * (size_t)(K + 4);
  d.npart = (double*)p;
  int64_t nb = kr_nb(kr, 1);
  p += sizeof(double) * (size_t)(nb + 1);
  p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
  d.part = (cplx*)p;
  return d;
}

static size_t kr_scal_bytes(int64_t n, int K) {
  int64_t nb = (n + KT - 1) / KT;  // rpt = 1 is the largest block count
  return 64 + sizeof(cplx) * (size_t)(2 * (K + 2)) + sizeof(cplx) * (size_t)(K + 2) * (K + 1) +
         sizeof(cplx) * (size_t)(2 * (K + 2) + 2 * K + 8) + sizeof(double) * 8 + sizeof(double) * (size_t)(nb + 1) +
         sizeof(double) * (8 + 3 * (size_t)(K + 2)) + sizeof(int) * (size_t)(K + 4) +
         256 + sizeof(cplx) * (size_t)nb * (2 * K + 4);
}

extern "C" int qd_krylov_create(qd_ctx* ctx, int64_t n, int kry_dim, int herm, qd_krylov** out) {
  if (!ctx || !out || n < 1 || kry_dim < 1 || kry_dim > 4096)
    return qd_set_error(ctx, QD_E_ARG, "qd_krylov_create: bad arguments");
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_krylov* kr = new qd_krylov();
  kr->ctx = ctx;
  kr->n = n;
  kr->kry_dim = kry_dim;
  kr->herm = herm ? 1 : 0;
  kr->ld = (n + 31) / 32 * 32;
  size_t vbytes = sizeof(cplx) * (size_t)kr->ld * (size_t)(kry_dim + 1);
  if (qd_dev_alloc(ctx, vbytes, (void**)&kr->d_V) != QD_OK) {
    delete kr;
    return qd_set_error(ctx, QD_E_OOM, "Krylov basis: %zu bytes", vbytes);
  }
  if (qd_dev_alloc(ctx, kr_scal_bytes(n, kry_dim), (void**)&kr->d_scal) != QD_OK ||
      qd_dev_alloc(ctx, sizeof(cplx) * 2 * (size_t)kr->ld, (void**)&kr->d_w) != QD_OK) {
    qd_dev_free(ctx, kr->d_V);
    qd_dev_free(ctx, kr->d_scal);
    delete kr;
    return qd_set_error(ctx, QD_E_OOM, "Krylov scalars");
  }
  kr->KR.assign((size_t)kry_dim * kry_dim, cmake(0.0, 0.0));
  kr->evs_prev.assign(1, cmake(0.0, 0.0));
  *out = kr;
  return QD_OK;
}

extern "C" int qd_krylov_destroy(qd_krylov* kr) {
  if (!kr) return QD_OK;
  hipSetDevice(kr->ctx->device);
  hipStreamSynchronize(kr->ctx->stream);
  qd_dev_free(kr->ctx, kr->d_V);
  qd_dev_free(kr->ctx, kr->d_scal);
  qd_dev_free(kr->ctx, kr->d_w);
  delete kr;
  return QD_OK;
}

// Gram-Schmidt variant: 0 = CGS2 (four passes over the basis per step),
// 1 = DCGS2 (two passes, the default), 2 = DCGS2 over a local window with
// partial reorthogonalisation (Hermitian only), 3 = local window only.
extern "C" int qd_krylov_set_method(qd_krylov* kr, int method) {
  if (!kr || method < 0 || method > 3) return QD_E_ARG;
  if (method >= 2 && !kr->herm)
    return qd_set_error(kr->ctx, QD_E_ARG, "qd_krylov_set_method: windowed schemes need a Hermitian operator");
  kr->method = method;
  return QD_OK;
}

extern "C" int qd_krylov_reset(qd_krylov* kr) {
  if (!kr) return QD_E_ARG;
  kr->act_dim = 0;
  kr->breakdown = 0;
  kr->pending = false;
  kr->beta_last = 0.0;
  std::fill(kr->KR.begin(), kr->KR.end(), cmake(0.0, 0.0));
  return QD_OK;
}

// V[:, 0] = psi / |psi|  (initKrylov, :514-519)
static int kr_init(qd_krylov* kr, const cplx* psi) {
  qd_ctx* ctx = kr->ctx;
  KrylovDev d = kr_layout(kr);
  int nb = (int)std::min<int64_t>(1024, (kr->n + KT - 1) / KT);
  norm_partial_kernel<<<nb, KT, 0, ctx->stream>>>(kr->n, psi, d.npart);
  sum_double_kernel<<<1, KT, 0, ctx->stream>>>(d.npart, nb, d.nrm2);
  normalise_copy_kernel<<<qd_grid(kr->n, 256), 256, 0, ctx->stream>>>(kr->n, psi, kr->d_V, d.nrm2);
  // partial scheme: eps1 = sqrt(n) eps / 2 models the rounding of one step;
  // trigger at a quarter of sqrt(eps / kry_dim) (semi-orthogonality), one
  // step ahead of the full passes it schedules
  const double eps = 2.220446049250313e-16;
  const double trigger = kr->method == 3 ? INFINITY : 0.25 * std::sqrt(eps / kr->kry_dim);
  pro_init_kernel<<<1, KT, 0, ctx->stream>>>(d.lo, kr->kry_dim + 3, d.om, d.prm, 0.5 * std::sqrt((double)kr->n) * eps,
                                             trigger);
  QD_CHECK_LAUNCH(ctx);
  kr->full_steps = 0;
  kr->act_dim = 0;
  kr->breakdown = 0;
  kr->pending = false;
  std::fill(kr->KR.begin(), kr->KR.end(), cmake(0.0, 0.0));
  return QD_OK;
}

template <int RPT>
static void launch_gs_step(qd_krylov* kr, const KrylovDev& d, int j, cplx* w) {
  qd_ctx* ctx = kr->ctx;
  hipStream_t st = ctx->stream;
  int64_t n = kr->n;
  int ncols = j + 1;
  int nb = (int)kr_nb(kr, RPT);
  cplx* kcol = d.KR + (size_t)j * (kr->kry_dim + 2);
  // pass 1
  gemv_t_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, ncols, w, d.part, d.flag);
  reduce_cols_kernel<<<ncols, KT, 0, st>>>(d.part, nb, ncols, d.h, kcol, 0, d.flag);
  gemv_n_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, ncols, d.h, w, nullptr, d.flag);
  // pass 2
  gemv_t_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, ncols, w, d.part, d.flag);
  reduce_cols_kernel<<<ncols, KT, 0, st>>>(d.part, nb, ncols, d.h, kcol, 1, d.flag);
  gemv_n_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, ncols, d.h, w, d.npart, d.flag);
  finalize_kernel<<<1, KT, 0, st>>>(d.npart, nb, kcol + (j + 1), d.inv, d.flag, d.bd_step, j, QD_LEAF_EPS);
  scale_kernel<<<qd_grid(n, 256), 256, 0, st>>>(n, w, d.inv, d.flag);
}

// cont: U holds the pending vector u_{s+1} left by the previous expansion
// (no bootstrap); fin: the last step second-projects the pending vector into
// V[:, s + add] (no product), else it runs as a full step and leaves
// u_{s+add+1} pending.  Either way V[:, 0..s+add] and Hessenberg columns
// 0..s+add-1 are final afterwards.
// Launch shape of dc_red_rows_kernel<RT>: every block takes the same number
// k of whole tiles, k the fewest that lets all blocks be resident at once
// (occupancy at the kernel's LDS size, measured once).  Whole tiles because
// each tile costs a full pass of the column loop (latency-bound when its
// lanes are mostly idle).  The count never exceeds kr_nb(kr, 1), the size of
// the partials buffer.
struct RowsGrid {
  static constexpr int kCmax = 160;  // columns per LDS group: 20 KB of accumulators
  int blocks;
  int64_t rows_pb;
  size_t lds;
};
template <int RT>
static RowsGrid rows_grid(int64_t n, int ncu) {
  constexpr size_t lds = (size_t)(KT / 64) * RowsGrid::kCmax * 4 * sizeof(double);
  // queried once per process; the initialisation of a function-local static
  // is thread-safe, so concurrent solves (server workers) never see a
  // half-written value
  static const int per_cu = []() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dc_red_rows_kernel<RT>, KT, lds) != hipSuccess || b < 1) b = 1;
    return b;
  }();
  const int64_t tile = (int64_t)KT * RT;
  const int64_t tiles = std::max<int64_t>(1, (n + tile - 1) / tile);
  const int64_t slots = std::max<int64_t>(1, (int64_t)per_cu * ncu);
  const int64_t k = (tiles + slots - 1) / slots;
  return RowsGrid{(int)((tiles + k - 1) / k), k * tile, lds};
}

template <int RPT>
static int dc_expand_launch(qd_krylov* kr, qd_linop* op, const KrylovDev& d, int s, int add, bool cont, bool fin) {
  qd_ctx* ctx = kr->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = kr->n;
  const int nb = (int)kr_nb(kr, RPT);
  const int ldk = kr->kry_dim + 2;
  cplx* U = kr->d_w;
  cplx* W = kr->d_w + kr->ld;
  const bool windowed = kr->method >= 2;
  const Pro pro = windowed ? Pro{d.lo, d.om, d.prm, kr->kry_dim + 2, PRO_W} : Pro{nullptr, nullptr, nullptr, 0, 0};
  auto win = [&](int j) { return windowed ? Win{d.lo, j, PRO_W} : Win{nullptr, 0, 0}; };
  constexpr int RT = RPT >= 4 ? 4 : RPT;  // rows per thread of a reduction tile
  const RowsGrid rg = rows_grid<RT>(n, ctx->num_cus);
  int rc;
  if (!cont) {
    // bootstrap: u_{s+1} = A q_s - Q_s (Q_s^H A q_s); H[0:s+1, s] = first pass
    rc = qd_linop_apply(op, kr->d_V + (size_t)s * kr->ld, U, nullptr, 0, nullptr, cmake(0.0, 0.0));
    if (rc) return rc;
    const Win w0 = win(s + 1);
    gemv_t_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, s + 1, U, d.part, d.flag, w0);
    reduce_cols_kernel<<<s + 1, KT, 0, st>>>(d.part, nb, s + 1, d.h, d.KR + (size_t)s * ldk, 0, d.flag, 1 << 30, w0,
                                             s + 1, 1);
    gemv_n_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, s + 1, d.h, U, nullptr, d.flag, w0);
  }
  const int jend = fin ? s + add : s + add + 1;
  for (int j = s + 1; j < jend; ++j) {
    rc = qd_linop_apply(op, U, W, nullptr, 0, nullptr, cmake(0.0, 0.0));
    if (rc) return rc;
    const int P = 2 * j + 2;
    const Win wj = win(j);
    dc_red_rows_kernel<RT><<<rg.blocks, KT, rg.lds, st>>>(n, kr->d_V, kr->ld, j, U, W, d.part, d.flag, wj,
                                                          rg.rows_pb, RowsGrid::kCmax);
    reduce_cols_kernel<<<P, KT, 0, st>>>(d.part, rg.blocks, P, d.red, nullptr, 0, d.flag, 1 << 30, wj, j, 2);
    dc_small_kernel<<<1, KT, 0, st>>>(d.red, j, 0, d.KR, ldk, d.coefA, d.coefB, d.scal, d.flag, d.bd_step, QD_LEAF_EPS,
                                      pro);
    dc_update_kernel<RPT, false><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, j, U, W, d.coefA, d.coefB, d.scal, d.flag,
                                                    wj);
  }
  if (!fin) {
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  }
  // finalise: reorthogonalise the pending vector into V[:, s + add]
  const int j = s + add;
  const Win wj = win(j);
  dc_red_kernel<RPT><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, j, U, d.part, d.flag, wj);
  reduce_cols_kernel<<<j + 2, KT, 0, st>>>(d.part, nb, j + 2, d.red, nullptr, 0, d.flag, 1 << 30, wj, j, 1);
  dc_small_kernel<<<1, KT, 0, st>>>(d.red, j, 1, d.KR, ldk, d.coefA, d.coefB, d.scal, d.flag, d.bd_step, QD_LEAF_EPS,
                                    pro);
  dc_update_kernel<RPT, true><<<nb, KT, 0, st>>>(n, kr->d_V, kr->ld, j, U, nullptr, d.coefA, d.coefB, d.scal, d.flag,
                                                 wj);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

// DCGS2 expansion by add columns from act_dim (two passes over the basis per
// step).  Breakdown at step j leaves act_dim = j.
// may_continue: a later call will probably extend this basis (chunked
// ground states), so the last step leaves the pending vector for it instead
// of finalising (saves the finalisation and the next bootstrap: two passes
// over the basis).
static int kr_expand_dc(qd_krylov* kr, qd_linop* op, int add, bool may_continue) {
  qd_ctx* ctx = kr->ctx;
  if (kr->breakdown || add <= 0) return QD_OK;
  KrylovDev d = kr_layout(kr);
  QD_HIP(ctx, hipMemsetAsync(d.flag, 0, 16, ctx->stream));
  const int s = kr->act_dim;
  const bool cont = kr->pending;
  const bool fin = !may_continue || s + add >= kr->kry_dim;
  kr->pending = false;
  int rpt = pick_rpt(kr->n, ctx->num_cus);
  int rc = (rpt == 8)   ? dc_expand_launch<8>(kr, op, d, s, add, cont, fin)
           : (rpt == 4) ? dc_expand_launch<4>(kr, op, d, s, add, cont, fin)
                        : dc_expand_launch<1>(kr, op, d, s, add, cont, fin);
  if (rc) return rc;
  const int K = kr->kry_dim, ldk = K + 2;
  std::vector<cplx> cols((size_t)add * ldk);
  int hflag[4] = {0, 0, 0, 0};  // flag at +0, breakdown step at +8
  double prm[4] = {0.0, 0.0, 0.0, 0.0};
  QD_HIP(ctx, hipMemcpyAsync(hflag, d.flag, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipMemcpyAsync(cols.data(), d.KR + (size_t)s * ldk, sizeof(cplx) * cols.size(), hipMemcpyDeviceToHost,
                             ctx->stream));
  if (kr->method >= 2)
    QD_HIP(ctx, hipMemcpyAsync(prm, d.prm, sizeof(prm), hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  kr->full_steps = (int)prm[3];
  int end = s + add;
  if (hflag[0]) {
    end = hflag[2];  // columns 0..end-1 span an invariant subspace
    kr->breakdown = 1;
  }
  kr->n_spmv += add + (cont ? 0 : 1) - (fin ? 1 : 0);
  kr->pending = !fin && !hflag[0];
  for (int j = s; j < end; ++j) {
    const cplx* c = cols.data() + (size_t)(j - s) * ldk;
    for (int i = 0; i <= j && i < K; ++i) kr->KR[(size_t)j * K + i] = c[i];
    if (j + 1 < K) kr->KR[(size_t)j * K + j + 1] = c[j + 1];
    kr->beta_last = c[j + 1].x;
  }
  kr->n_last = cmake(kr->beta_last, 0.0);
  kr->act_dim = end;
  return QD_OK;
}

// extend the basis from act_dim by add columns; reads KR back.
static int kr_expand_cgs2(qd_krylov* kr, qd_linop* op, int add);
static int kr_expand(qd_krylov* kr, qd_linop* op, int add, bool may_continue = false) {
  if (kr->method == 0) {
    kr->pending = false;
    return kr_expand_cgs2(kr, op, add);
  }
  return kr_expand_dc(kr, op, add, may_continue);
}

static int kr_expand_cgs2(qd_krylov* kr, qd_linop* op, int add) {
  qd_ctx* ctx = kr->ctx;
  if (kr->breakdown || add <= 0) return QD_OK;
  KrylovDev d = kr_layout(kr);
  QD_HIP(ctx, hipMemsetAsync(d.flag, 0, 16, ctx->stream));
  int start = kr->act_dim;
  int rpt = pick_rpt(kr->n, ctx->num_cus);
  for (int j = start; j < start + add; ++j) {
    cplx* vj = kr->d_V + (size_t)j * kr->ld;
    cplx* w = kr->d_V + (size_t)(j + 1) * kr->ld;
    int rc = qd_linop_apply(op, vj, w, nullptr, 0, nullptr, cmake(0.0, 0.0));
    if (rc) return rc;
    if (rpt == 8)
      launch_gs_step<8>(kr, d, j, w);
    else if (rpt == 4)
      launch_gs_step<4>(kr, d, j, w);
    else
      launch_gs_step<1>(kr, d, j, w);
    QD_CHECK_LAUNCH(ctx);
  }
  // read back flag, breakdown step and the new Hessenberg columns
  int K = kr->kry_dim;
  std::vector<cplx> cols((size_t)add * (K + 2));
  int hflag[4] = {0, 0, 0, 0};  // flag at +0, breakdown step at +8
  QD_HIP(ctx, hipMemcpyAsync(hflag, d.flag, 4 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipMemcpyAsync(cols.data(), d.KR + (size_t)start * (K + 2), sizeof(cplx) * cols.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int end = start + add;
  if (hflag[0]) {
    end = hflag[2] + 1;  // column bd_step is complete; the space is invariant
    kr->breakdown = 1;
  }
  kr->n_spmv += (end - start) + (hflag[0] ? 0 : 0);
  for (int j = start; j < end; ++j) {
    const cplx* c = cols.data() + (size_t)(j - start) * (K + 2);
    for (int i = 0; i <= j && i < K; ++i) kr->KR[(size_t)j * K + i] = c[i];
    if (j + 1 < K) kr->KR[(size_t)j * K + j + 1] = c[j + 1];
    kr->beta_last = c[j + 1].x;
  }
  kr->n_last = cmake(kr->beta_last, 0.0);
  kr->act_dim = end;
  return QD_OK;
}

extern "C" int qd_krylov_ground_state(qd_krylov* kr, qd_linop* op, double* d_psi, int add_dim, double* h_E0) {
  if (!kr || !op || !d_psi || op->n != kr->n) return qd_set_error(kr ? kr->ctx : nullptr, QD_E_ARG, "qd_krylov_ground_state: bad arguments");
  if (kr->act_dim >= kr->kry_dim && !kr->evals.empty()) {  // :567
    if (h_E0) *h_E0 = kr->evals[0];
    return QD_OK;
  }
  return qd_krylov_gs_step(kr, op, d_psi, add_dim, h_E0, true);
}

// Extends the basis by add_dim (-1: to kry_dim; clamped) and solves the Ritz
// problem; the lowest Ritz vector overwrites d_psi only if want_vector (a
// caller that checks convergence between extensions asks for it once, at the
// end, instead of paying a pass over the basis per check).
int qd_krylov_gs_step(qd_krylov* kr, qd_linop* op, double* d_psi, int add_dim, double* h_E0, bool want_vector) {
  if (!kr || !op || !d_psi || op->n != kr->n) return qd_set_error(kr ? kr->ctx : nullptr, QD_E_ARG, "qd_krylov_gs_step: bad arguments");
  qd_ctx* ctx = kr->ctx;
  int rc;
  if (kr->act_dim == 0) {
    rc = kr_init(kr, (const cplx*)d_psi);
    if (rc) return rc;
  }
  int add = add_dim;
  if (add == -1 || kr->kry_dim < kr->act_dim + add) add = kr->kry_dim - kr->act_dim;  // :523-524
  rc = kr_expand(kr, op, add, true);  // chunked callers extend the same basis again
  if (rc) return rc;
  int m = kr->act_dim;
  if (m < 1) return qd_set_error(ctx, QD_E_BREAKDOWN, "Krylov space is empty");
  // lower triangle: diag Re(alpha), sub-diagonal beta (SelfAdjointEigenSolver, :571)
  std::vector<double> dg(m), e(m > 1 ? m - 1 : 1), w(m), z((size_t)m * m);
  int K = kr->kry_dim;
  for (int i = 0; i < m; ++i) dg[i] = kr->KR[(size_t)i * K + i].x;
  for (int i = 0; i + 1 < m; ++i) e[i] = kr->KR[(size_t)i * K + i + 1].x;
  rc = qd_tridiag_eig(m, dg.data(), e.data(), w.data(), z.data(), 1);
  if (rc) return qd_set_error(ctx, QD_E_EIGEN, "failed to solve the Krylov matrix");
  kr->evals = w;
  // residual of the lowest Ritz pair: beta_m |z_{m-1,0}| (exact for the Krylov relation)
  kr->ritz_resid = kr->beta_last * std::fabs(z[(size_t)m - 1]);
  kr->ritz_gap = m > 1 ? w[1] - w[0] : 0.0;
  if (!want_vector) {
    if (h_E0) *h_E0 = w[0];
    return QD_OK;
  }
  // psi = V[:, :m] y0  (getState, :450-473)
  KrylovDev d = kr_layout(kr);
  std::vector<cplx> y(m);
  for (int i = 0; i < m; ++i) y[i] = cmake(z[i], 0.0);  // column 0 of Z
  QD_HIP(ctx, hipMemcpyAsync(d.y, y.data(), sizeof(cplx) * m, hipMemcpyHostToDevice, ctx->stream));
  int rpt = pick_rpt(kr->n, ctx->num_cus);
  int nb = (int)kr_nb(kr, rpt);
  if (rpt == 8)
    combine_kernel<8><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi);
  else if (rpt == 4)
    combine_kernel<4><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi);
  else
    combine_kernel<1><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi);
  QD_CHECK_LAUNCH(ctx);
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // err_ (:578-584): change of the tracked Ritz values
  int N = (int)std::min<size_t>(kr->evs_prev.size(), (size_t)m);
  double s = 0.0;
  for (int i = 0; i < N; ++i) {
    double dr = w[i] - kr->evs_prev[i].x, di = -kr->evs_prev[i].y;
    s += dr * dr + di * di;
  }
  kr->err = std::sqrt(s);
  for (int i = 0; i < N; ++i) kr->evs_prev[i] = cmake(w[i], 0.0);
  if (h_E0) *h_E0 = w[0];
  return QD_OK;
}

extern "C" int qd_krylov_evolve(qd_krylov* kr, qd_linop* op, const double* d_psi_in, double* d_psi_out, double h,
                                int kry_dim_rt) {
  if (!kr || !op || !d_psi_in || !d_psi_out || op->n != kr->n)
    return qd_set_error(kr ? kr->ctx : nullptr, QD_E_ARG, "qd_krylov_evolve: bad arguments");
  qd_ctx* ctx = kr->ctx;
  if (kry_dim_rt == 0) kry_dim_rt = kr->kry_dim;  // :552
  int rc = kr_init(kr, (const cplx*)d_psi_in);
  if (rc) return rc;
  int add = kry_dim_rt - 1;
  if (add < 1) return qd_set_error(ctx, QD_E_ARG, "kry_dim_rt must be >= 2");
  if (kr->kry_dim < add) add = kr->kry_dim;
  rc = kr_expand(kr, op, add);
  if (rc) return rc;
  int m = kr->act_dim;
  int K = kr->kry_dim;
  // E = expm(-i h K_m) (:559)
  std::vector<cplx> A((size_t)m * m), E((size_t)m * m);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      cplx kv = kr->KR[(size_t)j * K + i];
      A[(size_t)j * m + i] = cmake(h * kv.y, -h * kv.x);
    }
  rc = qd_expm_pade13(m, A.data(), E.data());
  if (rc) return qd_set_error(ctx, QD_E_EIGEN, "expm of the Krylov matrix failed");
  KrylovDev d = kr_layout(kr);
  QD_HIP(ctx, hipMemcpyAsync(d.y, E.data(), sizeof(cplx) * m, hipMemcpyHostToDevice, ctx->stream));
  int rpt = pick_rpt(kr->n, ctx->num_cus);
  int nb = (int)kr_nb(kr, rpt);
  if (rpt == 8)
    combine_kernel<8><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi_out);
  else if (rpt == 4)
    combine_kernel<4><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi_out);
  else
    combine_kernel<1><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi_out);
  QD_CHECK_LAUNCH(ctx);
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

// Real-time step with the Krylov dimension chosen per step: the basis grows
// from m_min by 2 vectors until the a-posteriori estimate of the step error,
// h_{m+1,m} |[expm(-i h K_m)]_{m,1}| (Saad, SIAM J. Numer. Anal. 29 (1992)
// 209, section 5), is at most tol, or the basis reaches kry_max - 1 vectors
// or an invariant subspace.  psi_out = V_m expm(-i h K_m) e_1 as in evolve.
// m_enough receives the smallest leading dimension whose estimate met tol
// (the next step's m_min for a caller that steps repeatedly), or m_used.
static double step_error_estimate(const qd_krylov* kr, int mp, double h, std::vector<cplx>& A, std::vector<cplx>& E,
                                  double beta) {
  const int K = kr->kry_dim;
  A.assign((size_t)mp * mp, cmake(0.0, 0.0));
  E.assign((size_t)mp * mp, cmake(0.0, 0.0));
  for (int j = 0; j < mp; ++j)
    for (int i = 0; i < mp; ++i) {
      cplx kv = kr->KR[(size_t)j * K + i];
      A[(size_t)j * mp + i] = cmake(h * kv.y, -h * kv.x);
    }
  if (qd_expm_pade13(mp, A.data(), E.data())) return -1.0;
  const cplx em = E[(size_t)mp - 1];  // row mp-1 of column 0
  return beta * std::sqrt(em.x * em.x + em.y * em.y);
}

extern "C" int qd_krylov_evolve_tol(qd_krylov* kr, qd_linop* op, const double* d_psi_in, double* d_psi_out,
                                    double h, int m_min, int kry_max, double tol, int* m_used, int* m_enough) {
  if (!kr || !op || !d_psi_in || !d_psi_out || op->n != kr->n || m_min < 1 || kry_max < 2 || !(tol >= 0.0))
    return qd_set_error(kr ? kr->ctx : nullptr, QD_E_ARG, "qd_krylov_evolve_tol: bad arguments");
  qd_ctx* ctx = kr->ctx;
  int rc = kr_init(kr, (const cplx*)d_psi_in);
  if (rc) return rc;
  const int mmax = std::min(kry_max - 1, kr->kry_dim);
  int target = std::min(m_min, mmax);
  const int K = kr->kry_dim;
  std::vector<cplx> A, E;
  int m = 0;
  for (;;) {
    rc = kr_expand(kr, op, target - kr->act_dim);
    if (rc) return rc;
    m = kr->act_dim;
    if (m < 1) return qd_set_error(ctx, QD_E_BREAKDOWN, "Krylov space is empty");
    const double est = step_error_estimate(kr, m, h, A, E, kr->beta_last);
    if (est < 0.0) return qd_set_error(ctx, QD_E_EIGEN, "expm of the Krylov matrix failed");
    if (kr->breakdown || est <= tol || m >= mmax) break;
    target = std::min(mmax, m + 2);
  }
  if (m_used) *m_used = m;
  if (m_enough) {
    // smallest leading block whose own estimate meets tol (E of the full m kept)
    std::vector<cplx> A2, E2;
    int me = m;
    for (int mp = m - 1; mp >= 2; --mp) {
      const double beta = kr->KR[(size_t)(mp - 1) * K + mp].x;  // h_{mp+1, mp}
      const double est = step_error_estimate(kr, mp, h, A2, E2, beta);
      if (est < 0.0 || est > tol) break;
      me = mp;
    }
    *m_enough = me;
  }
  KrylovDev d = kr_layout(kr);
  QD_HIP(ctx, hipMemcpyAsync(d.y, E.data(), sizeof(cplx) * m, hipMemcpyHostToDevice, ctx->stream));
  int rpt = pick_rpt(kr->n, ctx->num_cus);
  int nb = (int)kr_nb(kr, rpt);
  if (rpt == 8)
    combine_kernel<8><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi_out);
  else if (rpt == 4)
    combine_kernel<4><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi_out);
  else
    combine_kernel<1><<<nb, KT, 0, ctx->stream>>>(kr->n, kr->d_V, kr->ld, m, d.y, (cplx*)d_psi_out);
  QD_CHECK_LAUNCH(ctx);
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

extern "C" int qd_krylov_get_info(qd_krylov* kr, qd_krylov_info* info) {
  if (!kr || !info) return QD_E_ARG;
  info->act_dim = kr->act_dim;
  info->breakdown = kr->breakdown;
  info->error = kr->err;
  info->beta_last = kr->beta_last;
  info->n_spmv = kr->n_spmv;
  info->ritz_resid = kr->ritz_resid;
  info->ritz_gap = kr->ritz_gap;
  info->full_steps = kr->full_steps;
  return QD_OK;
}

extern "C" int qd_krylov_eigenvalues(qd_krylov* kr, double* h_evals, int cap, int* count) {
  if (!kr || (cap > 0 && !h_evals)) return QD_E_ARG;
  int c = (int)std::min<size_t>(kr->evals.size(), (size_t)(cap > 0 ? cap : 0));
  for (int i = 0; i < c; ++i) h_evals[i] = kr->evals[i];
  if (count) *count = c;
  return QD_OK;
}

extern "C" int qd_krylov_converge_ev(qd_krylov* kr, int N, double iv_re, double iv_im) {
  if (!kr || N < 1 || N >= kr->kry_dim) return QD_E_ARG;  // assert(N < kry_dim_), :495
  kr->evs_prev.assign((size_t)N, cmake(iv_re, iv_im));
  return QD_OK;
}
