// spmv.hip -- K5 (stored-matrix product), K6 (matrix-free product), and the
// deterministic vector reductions used by eval / norm / Krylov (K10).
//
// Layout.  The stored matrix is the reference's ColMajor arrays (outer =
// source column).  For a Hermitian H, column r conjugated is row r, so
//     y_r = sum_{e in column r} conj(val_e) * x_{inner_e}
// is a row-parallel "pull" product with no atomics and a fixed summation
// order.  Non-Hermitian sums use the columns of the adjoint program instead
// (same formula), built once by qd_linop_from_csr.
#include <mutex>

#include "qd_internal.hpp"
#include "wave.hpp"

typedef double dbl2 __attribute__((ext_vector_type(2)));
#define QD_STREAM_VARIANTS 15
// default CSR-stream variant: spmv_stream3_kernel<8> (software pipelined,
// 512-entry slices), 0.4 % faster than spmv_stream_kernel<512, 8, true> at
// XXZ N=24 in interleaved timing, same bits
#define QD_STREAM_DEFAULT 13

// ---------------------------------------------------------------- K5
// G lanes cooperate on one row: each lane strides the row's entries by G,
// then a butterfly over the G lanes (fixed tree) combines them.  A wave of 64
// lanes covers 64/G consecutive rows, whose entries are contiguous in memory,
// so value/column loads are coalesced.  The matrix streams through with
// non-temporal loads to leave L2/MALL to the gathered x.
//
// Block-to-XCD remap: the hardware deals workgroups round-robin over the 8
// XCDs; logical block = (b % 8) * per + b / 8 gives each XCD a contiguous
// range of rows, so neighbouring rows (which gather neighbouring x entries)
// share one L2.
template <int G>
__global__ void __launch_bounds__(256)
    spmv_pull_kernel(int64_t nrows, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                     const cplx* __restrict__ vals, const cplx* __restrict__ x, cplx* __restrict__ y,
                     const double* __restrict__ xscale, int xscale_idx, const cplx* __restrict__ diag,
                     cplx shift) {
  constexpr int RPB = 256 / G;
  const unsigned nb = gridDim.x;
  const unsigned per = nb >> 3;  // grid is a multiple of 8
  const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const int g = threadIdx.x / G;
  const int l = threadIdx.x % G;
  const int64_t row = (int64_t)lb * RPB + g;
  int32_t beg = 0, end = 0;
  if (row < nrows) {
    beg = outer[row];
    end = outer[row + 1];
  }
  double sx = 0.0, sy = 0.0;
  const dbl2* vp = reinterpret_cast<const dbl2*>(vals);
  for (int32_t e = beg + l; e < end; e += G) {
    dbl2 v = __builtin_nontemporal_load(vp + e);
    int32_t c = __builtin_nontemporal_load(inner + e);
    cplx xv = x[c];
    sx += v.x * xv.x + v.y * xv.y;
    sy += v.x * xv.y - v.y * xv.x;
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    sx += __shfl_xor(sx, o, G);
    sy += __shfl_xor(sy, o, G);
  }
  if (l == 0 && row < nrows) {
    double s = xscale ? xscale[xscale_idx] : 1.0;
    cplx r = cmake(sx * s, sy * s);
    if (diag || shift.x != 0.0 || shift.y != 0.0) {
      cplx xr = x[row];
      cplx d = diag ? cadd(diag[row], shift) : shift;
      r = cadd(r, crmul(s, cmul(d, xr)));
    }
    y[row] = r;
  }
}

// ---------------------------------------------------------------- K5 stream
// CSR-stream form for short rows.  A wave owns 64 consecutive rows.  Phase
// 1: the 64 lanes stream the wave's nonzeros (contiguous in memory) with U
// independent loads in flight per lane, gather x, and park conj(v) * x in
// the wave's LDS slice.  Phase 2: lane l adds the products of row r0 + l in
// entry order.  Rows longer than the slice are handled chunk by chunk.  The
// summation order depends only on the matrix, never on the launch shape of
// other work.
template <int CAPW, int U, bool NT>
__global__ void __launch_bounds__(256)
    spmv_stream_kernel(int64_t nrows, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                       const cplx* __restrict__ vals, const cplx* __restrict__ x, cplx* __restrict__ y,
                       const double* __restrict__ xscale, int xscale_idx, const cplx* __restrict__ diag,
                       cplx shift) {
  __shared__ cplx lds[4][CAPW];
  const unsigned nb = gridDim.x;
  const unsigned per = nb >> 3;
  const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)lb * 4 + w) * 64;
  if (r0 >= nrows) return;  // whole wave leaves together
  const int64_t row = r0 + lane;
  const int32_t nnz_end = outer[nrows];
  int32_t rs = (row < nrows) ? outer[row] : nnz_end;
  int32_t re = __shfl_down(rs, 1, 64);
  if (lane == 63) re = (r0 + 64 < nrows) ? outer[r0 + 64] : nnz_end;
  const int32_t wbeg = __shfl(rs, 0, 64);
  const int32_t wend = __shfl(re, 63, 64);
  cplx* __restrict__ buf = lds[w];
  const dbl2* vp = reinterpret_cast<const dbl2*>(vals);
  double sx = 0.0, sy = 0.0;
  for (int32_t c0 = wbeg; c0 < wend; c0 += CAPW) {
    const int32_t cn = min(wend - c0, (int32_t)CAPW);
    for (int32_t k = lane; k < cn; k += 64 * U) {
      dbl2 v[U];
      int32_t c[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int32_t kk = k + 64 * u;
        if (kk < cn) {
          if (NT) {
            v[u] = __builtin_nontemporal_load(vp + c0 + kk);
            c[u] = __builtin_nontemporal_load(inner + c0 + kk);
          } else {
            v[u] = vp[c0 + kk];
            c[u] = inner[c0 + kk];
          }
        }
      }
      cplx xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k + 64 * u < cn) xv[u] = x[c[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int32_t kk = k + 64 * u;
        if (kk < cn) buf[kk] = cmake(v[u].x * xv[u].x + v[u].y * xv[u].y, v[u].x * xv[u].y - v[u].y * xv[u].x);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int32_t a = max(rs, c0), b = min(re, c0 + cn);
    for (int32_t e = a; e < b; ++e) {
      cplx p = buf[e - c0];
      sx += p.x;
      sy += p.y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (row < nrows) {
    double s = xscale ? xscale[xscale_idx] : 1.0;
    cplx r = cmake(sx * s, sy * s);
    if (diag || shift.x != 0.0 || shift.y != 0.0) {
      cplx xr = x[row];
      cplx d = diag ? cadd(diag[row], shift) : shift;
      r = cadd(r, crmul(s, cmul(d, xr)));
    }
    y[row] = r;
  }
}

// Same partition and summation order as spmv_stream_kernel, with the load
// phase in straight-line code: a block of U x 64 entries that lies wholly in
// the slice loads unconditionally (one wait for the 2U streams, one for the U
// gathers); the tail block clamps its indices to the slice instead of
// branching per load.  Phase 2 reads four products per LDS wait and adds
// them in entry order; y is written non-temporally.
template <int CAPW, int U>
__global__ void __launch_bounds__(256)
    spmv_stream2_kernel(int64_t nrows, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                        const cplx* __restrict__ vals, const cplx* __restrict__ x, cplx* __restrict__ y,
                        const double* __restrict__ xscale, int xscale_idx, const cplx* __restrict__ diag,
                        cplx shift) {
  __shared__ cplx lds[4][CAPW];
  const unsigned nb = gridDim.x;
  const unsigned per = nb >> 3;
  const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)lb * 4 + w) * 64;
  if (r0 >= nrows) return;  // whole wave leaves together
  const int64_t row = r0 + lane;
  const int32_t nnz_end = outer[nrows];
  int32_t rs = (row < nrows) ? outer[row] : nnz_end;
  int32_t re = __shfl_down(rs, 1, 64);
  if (lane == 63) re = (r0 + 64 < nrows) ? outer[r0 + 64] : nnz_end;
  const int32_t wbeg = __shfl(rs, 0, 64);
  const int32_t wend = __shfl(re, 63, 64);
  cplx* __restrict__ buf = lds[w];
  const dbl2* vp = reinterpret_cast<const dbl2*>(vals);
  double sx = 0.0, sy = 0.0;
  for (int32_t c0 = wbeg; c0 < wend; c0 += CAPW) {
    const int32_t cn = min(wend - c0, (int32_t)CAPW);
    const dbl2* __restrict__ vc = vp + c0;
    const int32_t* __restrict__ ic = inner + c0;
    for (int32_t k0 = 0; k0 < cn; k0 += 64 * U) {
      dbl2 v[U];
      int32_t c[U];
      cplx xv[U];
      if (k0 + 64 * U <= cn) {
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(ic + k0 + 64 * u + lane);
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(vc + k0 + 64 * u + lane);
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
        // every gather in flight before the first product
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
          buf[k0 + 64 * u + lane] =
              cmake(v[u].x * xv[u].x + v[u].y * xv[u].y, v[u].x * xv[u].y - v[u].y * xv[u].x);
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(ic + min(k0 + 64 * u + lane, cn - 1));
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(vc + min(k0 + 64 * u + lane, cn - 1));
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t kk = k0 + 64 * u + lane;
          if (kk < cn) buf[kk] = cmake(v[u].x * xv[u].x + v[u].y * xv[u].y, v[u].x * xv[u].y - v[u].y * xv[u].x);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int32_t a = max(rs, c0) - c0, b = min(re, c0 + cn) - c0;
    int32_t e = a;
    for (; e + 4 <= b; e += 4) {
      const cplx p0 = buf[e], p1 = buf[e + 1], p2 = buf[e + 2], p3 = buf[e + 3];
      sx += p0.x;
      sy += p0.y;
      sx += p1.x;
      sy += p1.y;
      sx += p2.x;
      sy += p2.y;
      sx += p3.x;
      sy += p3.y;
    }
    for (; e < b; ++e) {
      const cplx p = buf[e];
      sx += p.x;
      sy += p.y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (row < nrows) {
    double s = xscale ? xscale[xscale_idx] : 1.0;
    cplx r = cmake(sx * s, sy * s);
    if (diag || shift.x != 0.0 || shift.y != 0.0) {
      cplx xr = x[row];
      cplx d = diag ? cadd(diag[row], shift) : shift;
      r = cadd(r, crmul(s, cmul(d, xr)));
    }
    __builtin_nontemporal_store(r.x, &y[row].x);
    __builtin_nontemporal_store(r.y, &y[row].y);
  }
}

// Same partition and summation order as spmv_stream_kernel, software
// pipelined: a slice is exactly one block of 64 x U entries (CAPW = 64 U),
// and the value/index streams of slice s+1 are issued before the wait on the
// x gathers of slice s, so each wave keeps two memory phases in flight
// instead of alternating index load -> gather -> products.  Indices are
// clamped into the slice, so every load stays inside [wbeg, wend).
//
// Launch shape: STREAM3_WPB waves (64 rows each) per block, blocks in row
// order with no XCD remap.  Row ranges differ in their entry counts (XXZ
// N=24: the eight contiguous eighths hold 4.1 M to 4.8 M entries), so giving
// each XCD one contiguous eighth left the XCDs finishing up to 24 us apart
// in a 170 us launch; dealt round-robin, blocks balance the entries over the
// XCDs and the XCDs finish within 4 us.  Small blocks free their slot as
// soon as their own waves end: a 4-wave block held its slot until the
// slowest of its waves finished, which left about a fifth of the wave slots
// idle.  Together: 171 -> 155 us at XXZ N=24, same bits (per-wave start/end
// stamps and interleaved timings on one GPU).
// DOT: the block also sums conj(x_r) y_r over its rows into
// dot_part[blockIdx.x] (a fixed tree: wavefront butterflies, then the
// block's waves in order); the caller sums the partials in a later kernel.
// WRITE = false skips the store of y (an expectation value <x|H|x> without
// the product vector).
constexpr int STREAM3_WPB = 2;
template <int U, bool DOT = false, bool WRITE = true>
__global__ void __launch_bounds__(64 * STREAM3_WPB)
    spmv_stream3_kernel(int64_t nrows, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                        const cplx* __restrict__ vals, const cplx* __restrict__ x, cplx* __restrict__ y,
                        const double* __restrict__ xscale, int xscale_idx, const cplx* __restrict__ diag,
                        cplx shift, cplx* __restrict__ dot_part = nullptr) {
  constexpr int CAPW = 64 * U;
  __shared__ cplx lds[STREAM3_WPB][CAPW];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)blockIdx.x * STREAM3_WPB + w) * 64;
  if (!DOT && r0 >= nrows) return;  // whole wave leaves together
  cplx dsum = cmake(0.0, 0.0);
  if (r0 < nrows) {
  const int64_t row = r0 + lane;
  const int32_t nnz_end = outer[nrows];
  int32_t rs = (row < nrows) ? outer[row] : nnz_end;
  int32_t re = __shfl_down(rs, 1, 64);
  if (lane == 63) re = (r0 + 64 < nrows) ? outer[r0 + 64] : nnz_end;
  const int32_t wbeg = __shfl(rs, 0, 64);
  const int32_t wend = __shfl(re, 63, 64);
  cplx* __restrict__ buf = lds[w];
  const dbl2* vp = reinterpret_cast<const dbl2*>(vals);
  double sx = 0.0, sy = 0.0;
  dbl2 v[U];
  int32_t c[U];
  if (wbeg < wend) {
    const int32_t cn = min(wend - wbeg, (int32_t)CAPW);
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(inner + wbeg + min(64 * u + lane, cn - 1));
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(vp + wbeg + min(64 * u + lane, cn - 1));
  }
  for (int32_t c0 = wbeg; c0 < wend; c0 += CAPW) {
    const int32_t cn = min(wend - c0, (int32_t)CAPW);
    cplx xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
    // the next slice's streams go out behind the gathers (wave-uniform branch)
    const int32_t c1 = c0 + CAPW;
    dbl2 vn[U];
    int32_t cnx[U];
    if (c1 < wend) {
      const int32_t cn1 = min(wend - c1, (int32_t)CAPW);
#pragma unroll
      for (int u = 0; u < U; ++u) cnx[u] = __builtin_nontemporal_load(inner + c1 + min(64 * u + lane, cn1 - 1));
#pragma unroll
      for (int u = 0; u < U; ++u) vn[u] = __builtin_nontemporal_load(vp + c1 + min(64 * u + lane, cn1 - 1));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t kk = 64 * u + lane;
      if (kk < cn) buf[kk] = cmake(v[u].x * xv[u].x + v[u].y * xv[u].y, v[u].x * xv[u].y - v[u].y * xv[u].x);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int32_t a = max(rs, c0) - c0, b = min(re, c0 + cn) - c0;
    int32_t e = a;
    for (; e + 4 <= b; e += 4) {
      const cplx p0 = buf[e], p1 = buf[e + 1], p2 = buf[e + 2], p3 = buf[e + 3];
      sx += p0.x;
      sy += p0.y;
      sx += p1.x;
      sy += p1.y;
      sx += p2.x;
      sy += p2.y;
      sx += p3.x;
      sy += p3.y;
    }
    for (; e < b; ++e) {
      const cplx p = buf[e];
      sx += p.x;
      sy += p.y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (c1 < wend) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        c[u] = cnx[u];
        v[u] = vn[u];
      }
    }
  }
  if (row < nrows) {
    double s = xscale ? xscale[xscale_idx] : 1.0;
    cplx r = cmake(sx * s, sy * s);
    if (diag || shift.x != 0.0 || shift.y != 0.0) {
      cplx xr = x[row];
      cplx d = diag ? cadd(diag[row], shift) : shift;
      r = cadd(r, crmul(s, cmul(d, xr)));
    }
    if (WRITE) {
      __builtin_nontemporal_store(r.x, &y[row].x);
      __builtin_nontemporal_store(r.y, &y[row].y);
    }
    if (DOT) dsum = cdotmul(x[row], r);
  }
  }
  if (DOT) {  // every wave of the block reaches this barrier
    __shared__ cplx wsum[STREAM3_WPB];
    dsum = wave_sum(dsum);
    if (lane == 0) wsum[w] = dsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      cplx t = wsum[0];
#pragma unroll
      for (int q = 1; q < STREAM3_WPB; ++q) t = cadd(t, wsum[q]);
      dot_part[blockIdx.x] = t;
    }
  }
}

// ---------------------------------------------------------------- K6
// Matrix-free pull product from the adjoint program: expanding column r of
// A^dagger = sum conj(c_k) P_k^dagger from state r gives leaves (t, w) with
// w = P_k^dagger[t, r] = conj(P_k[r, t]), so
//     y_r = sum_k c_k sum_leaves conj(w) x_t = sum conj(conj(c_k) w) x_t.
// One thread per row; the term loop is uniform across the wave, so term and
// op records are fetched through the scalar cache.
__global__ void __launch_bounds__(256)
    spmv_matfree_kernel(SectorDesc sec, const uint64_t* __restrict__ states, int64_t n,
                        const DevTerm* __restrict__ aterms, int nterms, const DevOp* __restrict__ aops,
                        const cplx* __restrict__ x, cplx* __restrict__ y, const double* __restrict__ xscale,
                        int xscale_idx, const cplx* __restrict__ diag, cplx shift) {
  const unsigned nb = gridDim.x;
  const unsigned per = nb >> 3;
  const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  int64_t r = (int64_t)lb * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint64_t s = states[r];
  double sx = 0.0, sy = 0.0;
  for (int k = 0; k < nterms; ++k) {
    const DevTerm t = aterms[k];
    const DevOp* tops = aops + t.op_begin;
    int nbitops = 0;
    for (int q = 0; q < t.n_ops; ++q) nbitops += (tops[q].kind == QD_OP_BIT);
    for (int L = 0; L < t.n_leaves; ++L) {
      uint64_t xt;
      cplx w;
      if (!eval_leaf(tops, t.n_ops, nbitops, s, L, &xt, &w)) continue;
      if (!(hypot(w.x, w.y) > QD_LEAF_EPS)) continue;
      uint64_t ti = qd_rank(sec, states, xt);
      if (ti >= (uint64_t)n) continue;
      cplx cw = cmul(t.coeff, w);  // conj(c_k) w
      cplx xv = x[ti];
      sx += cw.x * xv.x + cw.y * xv.y;  // conj(cw) * xv
      sy += cw.x * xv.y - cw.y * xv.x;
    }
  }
  double sc = xscale ? xscale[xscale_idx] : 1.0;
  cplx res = cmake(sx * sc, sy * sc);
  if (diag || shift.x != 0.0 || shift.y != 0.0) {
    cplx d = diag ? cadd(diag[r], shift) : shift;
    res = cadd(res, crmul(sc, cmul(d, x[r])));
  }
  y[r] = res;
}

// ---------------------------------------------------------------- dispatch
static unsigned grid8(int64_t blocks) {
  int64_t g = (blocks + 7) / 8 * 8;
  return (unsigned)(g < 8 ? 8 : g);
}
static unsigned stream3_grid(int64_t n) {
  const int64_t rows = 64 * STREAM3_WPB;
  return (unsigned)std::max<int64_t>(1, (n + rows - 1) / rows);
}

// y = s * (H + diag + shift) x, s = xscale[idx] (device) or 1
int qd_linop_apply(qd_linop* op, const cplx* x, cplx* y, const double* d_xscale, int xscale_idx,
                   const cplx* d_diag, cplx shift) {
  qd_ctx* ctx = op->ctx;
  int64_t n = op->n;
  if (n == 0) return QD_OK;
  if (op->diag0) {
    if (d_diag) return qd_set_error(ctx, QD_E_UNSUPPORTED, "two diagonals on one product");
    d_diag = op->diag0;
  }
  if (op->kind == QD_LINOP_CSR && op->lanes_per_row <= 0) {
    // CSR-stream: 64 rows per wave; variant = -lanes_per_row
    unsigned grid = grid8((n + 255) / 256);
    switch (-op->lanes_per_row) {
#define QD_STREAM_CASE(V, CAPW, U, NT)                                                               \
  case V:                                                                                           \
    spmv_stream_kernel<CAPW, U, NT><<<grid, 256, 0, ctx->stream>>>(n, op->outer, op->inner, op->vals, x, y, \
                                                                   d_xscale, xscale_idx, d_diag, shift);   \
    break;
      QD_STREAM_CASE(0, 512, 8, true)
      QD_STREAM_CASE(1, 512, 4, true)
      QD_STREAM_CASE(2, 256, 4, true)
      QD_STREAM_CASE(3, 256, 8, true)
      QD_STREAM_CASE(4, 1024, 8, true)
      QD_STREAM_CASE(5, 512, 4, false)
      QD_STREAM_CASE(6, 512, 8, false)
      QD_STREAM_CASE(7, 128, 4, true)
      QD_STREAM_CASE(8, 1024, 16, true)
#undef QD_STREAM_CASE
#define QD_STREAM2_CASE(V, CAPW, U)                                                                  \
  case V:                                                                                           \
    spmv_stream2_kernel<CAPW, U><<<grid, 256, 0, ctx->stream>>>(n, op->outer, op->inner, op->vals, x, y, \
                                                                d_xscale, xscale_idx, d_diag, shift);   \
    break;
      QD_STREAM2_CASE(9, 512, 8)
      QD_STREAM2_CASE(10, 256, 4)
      QD_STREAM2_CASE(11, 512, 4)
      QD_STREAM2_CASE(12, 1024, 8)
#undef QD_STREAM2_CASE
#define QD_STREAM3_CASE(V, U)                                                                          \
  case V:                                                                                           \
    spmv_stream3_kernel<U><<<stream3_grid(n), 64 * STREAM3_WPB, 0, ctx->stream>>>(                    \
        n, op->outer, op->inner, op->vals, x, y, d_xscale, xscale_idx, d_diag, shift);                 \
    break;
      QD_STREAM3_CASE(13, 8)
      QD_STREAM3_CASE(14, 4)
#undef QD_STREAM3_CASE
      default:
        return qd_set_error(ctx, QD_E_ARG, "stream variant %d", -op->lanes_per_row);
    }
  } else if (op->kind == QD_LINOP_CSR) {
    switch (op->lanes_per_row) {
#define QD_SPMV_CASE(GG)                                                                                    \
  case GG:                                                                                                  \
    spmv_pull_kernel<GG><<<grid8((n + (256 / GG) - 1) / (256 / GG)), 256, 0, ctx->stream>>>(              \
        n, op->outer, op->inner, op->vals, x, y, d_xscale, xscale_idx, d_diag, shift);                      \
    break;
      QD_SPMV_CASE(2)
      QD_SPMV_CASE(4)
      QD_SPMV_CASE(8)
      QD_SPMV_CASE(16)
      QD_SPMV_CASE(32)
      QD_SPMV_CASE(64)
#undef QD_SPMV_CASE
      default:
        return qd_set_error(ctx, QD_E_ARG, "lanes_per_row %d", op->lanes_per_row);
    }
  } else if (op->kron && op->use_kron && !d_xscale) {
    return qd_kron_apply(ctx, op->kron, x, y, shift, d_diag);
  } else {
    qd_termprog* p = op->prog;
    spmv_matfree_kernel<<<grid8((n + 255) / 256), 256, 0, ctx->stream>>>(
        op->basis->sec, op->basis->d_states, n, p->d_adj_terms, (int)p->h_terms.size(), p->d_adj_ops, x, y,
        d_xscale, xscale_idx, d_diag, shift);
  }
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

static int pick_lanes(int64_t n, int64_t nnz) {
  double avg = n ? (double)nnz / (double)n : 1.0;
  if (avg <= 48.0) return -QD_STREAM_DEFAULT;  // CSR-stream kernel
  int g = 2;
  while (g < 64 && g < avg) g <<= 1;
  return g;
}

extern "C" int qd_linop_from_csr(qd_ctx* ctx, qd_csr* csr, qd_linop** out) {
  if (!ctx || !csr || !out) return qd_set_error(ctx, QD_E_ARG, "qd_linop_from_csr: null argument");
  *out = nullptr;
  qd_csr* use = csr;
  if (csr->hermitian != 1) {
    if (!csr->adj) {
      // the adjoint's CSC: from the adjoint term program, or (a matrix made
      // by the SpMatrix algebra has none) the conjugate transpose
      int rc = (csr->basis && csr->prog) ? qd_csr_build_impl(ctx, csr->basis, csr->prog, 1, &csr->adj)
                                         : qd_csr_adjoint(ctx, csr, &csr->adj);
      if (rc) return rc;
    }
    use = csr->adj;
  }
  qd_linop* op = new qd_linop();
  op->ctx = ctx;
  op->kind = QD_LINOP_CSR;
  op->n = use->n;
  op->outer = use->d_outer;
  op->inner = use->d_inner;
  op->vals = use->d_vals;
  op->nnz = use->nnz;
  op->lanes_per_row = pick_lanes(use->n, use->nnz);
  *out = op;
  return QD_OK;
}

extern "C" int qd_linop_matfree(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog, qd_linop** out) {
  if (!ctx || !basis || !prog || !out) return qd_set_error(ctx, QD_E_ARG, "qd_linop_matfree: null argument");
  *out = nullptr;
  if (prog->nbits > basis->sec.nbits)
    return qd_set_error(ctx, QD_E_ARG, "program acts on %d bits, basis has %d", prog->nbits, basis->sec.nbits);
  if (prog->mdim > 1 || basis->sec.mdim > 1)
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "qd_linop_matfree: a boson mode needs the stored matrix (qd_csr_build)");
  int rc = qd_termprog_ensure_adjoint(prog);
  if (rc) return rc;
  qd_linop* op = new qd_linop();
  op->ctx = ctx;
  op->kind = QD_LINOP_MATFREE;
  op->n = (int64_t)basis->sec.total;
  op->basis = basis;
  op->prog = prog;
  op->kron = qd_kron_try_build(ctx, basis, prog);
  *out = op;
  return QD_OK;
}

// Matrix-free operators: 1 = use the two-block Kronecker kernels when the
// program has that form (default), 0 = always the generic per-row kernel.
// Returns whether the structured form exists through *has_structured.
extern "C" int qd_linop_set_structured(qd_linop* op, int on, int* has_structured) {
  if (!op) return QD_E_ARG;
  op->use_kron = on ? 1 : 0;
  if (has_structured) *has_structured = op->kron ? 1 : 0;
  return QD_OK;
}

extern "C" int qd_linop_destroy(qd_linop* op) {
  if (op && op->kron && !op->view) qd_kron_destroy(op->kron);
  delete op;
  return QD_OK;
}

// A view of `base` that adds a diagonal to every product: (base + diag(d)) x.
// It borrows base's arrays and d_diag (n complex values); both must outlive it.
extern "C" int qd_linop_with_diagonal(qd_ctx* ctx, qd_linop* base, const double* d_diag, qd_linop** out) {
  if (!ctx || !base || !d_diag || !out) return qd_set_error(ctx, QD_E_ARG, "qd_linop_with_diagonal: bad arguments");
  if (base->diag0) return qd_set_error(ctx, QD_E_UNSUPPORTED, "the base operator already carries a diagonal");
  qd_linop* op = new qd_linop(*base);
  op->ctx = ctx;
  op->diag0 = reinterpret_cast<const cplx*>(d_diag);
  op->view = true;
  *out = op;
  return QD_OK;
}

extern "C" int64_t qd_linop_dim(const qd_linop* op) { return op ? op->n : -1; }

// Override the row mapping of a stored-matrix operator: 0, -1, ..., -8
// select a CSR-stream variant (LDS chunk, loads in flight, cache hint),
// 2..64 (powers of two) the vector kernel with that many lanes per row.
extern "C" int qd_linop_set_lanes(qd_linop* op, int lanes) {
  if (!op || op->kind != QD_LINOP_CSR) return QD_E_ARG;
  if (!(lanes <= 0 && lanes >= -QD_STREAM_VARIANTS + 1) && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 &&
      lanes != 32 && lanes != 64)
    return QD_E_ARG;
  op->lanes_per_row = lanes;
  return QD_OK;
}

extern "C" int qd_spmv(qd_ctx* ctx, qd_linop* op, const double* d_x, double* d_y) {
  if (!ctx || !op || (op->n && (!d_x || !d_y))) return qd_set_error(ctx, QD_E_ARG, "qd_spmv: null argument");
  if (d_x == d_y) return qd_set_error(ctx, QD_E_ARG, "qd_spmv: x and y must not alias");
  return qd_linop_apply(op, (const cplx*)d_x, (cplx*)d_y, nullptr, 0, nullptr, cmake(0.0, 0.0));
}

// d_y += alpha * H d_x  (OperatorSum::map_acc, operator_abel.hpp:36-42)
__global__ void axpy_kernel(int64_t n, cplx alpha, const cplx* __restrict__ t, cplx* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx a = cmul(alpha, t[i]);
  cplx v = y[i];
  y[i] = cmake(v.x + a.x, v.y + a.y);
}

extern "C" int qd_spmv_acc(qd_ctx* ctx, qd_linop* op, const double* h_alpha, const double* d_x, double* d_y) {
  if (!ctx || !op || !h_alpha || (op->n && (!d_x || !d_y))) return qd_set_error(ctx, QD_E_ARG, "qd_spmv_acc: null argument");
  int64_t n = op->n;
  if (n == 0) return QD_OK;
  cplx* tmp = nullptr;
  QD_HIP(ctx, hipMallocAsync((void**)&tmp, sizeof(cplx) * n, ctx->stream));
  int rc = qd_linop_apply(op, (const cplx*)d_x, tmp, nullptr, 0, nullptr, cmake(0.0, 0.0));
  if (!rc) {
    axpy_kernel<<<qd_grid(n, 256), 256, 0, ctx->stream>>>(n, cmake(h_alpha[0], h_alpha[1]), tmp, (cplx*)d_y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = qd_set_error(ctx, QD_E_HIP, "axpy launch: %s", hipGetErrorString(e));
  }
  hipFreeAsync(tmp, ctx->stream);
  return rc;
}

// ---------------------------------------------------------------- reductions
// Deterministic dot products: a fixed grid (function of n only) accumulates
// grid-stride partial sums, a wave butterfly and an LDS pass give one value
// per block, and one block sums the block values in index order.  The result
// depends on n alone, never on what else runs on the device.
#define RED_THREADS 256
static int red_blocks(int64_t n) {
  int64_t b = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

__device__ inline cplx block_reduce_c(cplx v, cplx* sh) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  cplx r = cmake(0.0, 0.0);
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = cadd(r, sh[i]);
  return r;
}

// partial[b] = sum over the block's grid-stride slice of conj(a_i) b_i
__global__ void __launch_bounds__(RED_THREADS)
    dot_partial_kernel(int64_t n, const cplx* __restrict__ a, const cplx* __restrict__ b, cplx* __restrict__ part) {
  __shared__ cplx sh[RED_THREADS / 64];
  cplx s = cmake(0.0, 0.0);
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
    cplx av = a[i], bv = b[i];
    s = cadd(s, cdotmul(av, bv));
  }
  cplx r = block_reduce_c(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void __launch_bounds__(RED_THREADS)
    sum_partials_kernel(const cplx* __restrict__ part, int nb, cplx* __restrict__ out) {
  __shared__ cplx sh[RED_THREADS / 64];
  cplx s = cmake(0.0, 0.0);
  for (int i = threadIdx.x; i < nb; i += RED_THREADS) s = cadd(s, part[i]);
  cplx r = block_reduce_c(s, sh);
  if (threadIdx.x == 0) *out = r;
}

// d_out (device, one complex) = a^H b
int qd_dev_dot(qd_ctx* ctx, int64_t n, const cplx* a, const cplx* b, cplx* d_out) {
  int nb = red_blocks(n);
  void* sp = nullptr;
  int rc = qd_ctx_scratch(ctx, sizeof(cplx) * (nb + 8), &sp);
  if (rc) return rc;
  cplx* part = (cplx*)sp;
  dot_partial_kernel<<<nb, RED_THREADS, 0, ctx->stream>>>(n, a, b, part);
  sum_partials_kernel<<<1, RED_THREADS, 0, ctx->stream>>>(part, nb, d_out);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

static int read_one(qd_ctx* ctx, const cplx* d_val, double* h2) {
  double* pin = nullptr;
  int rc = qd_ctx_pinned(ctx, 2, &pin);
  if (rc) return rc;
  QD_HIP(ctx, hipMemcpyAsync(pin, d_val, sizeof(cplx), hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  h2[0] = pin[0];
  h2[1] = pin[1];
  return QD_OK;
}

extern "C" int qd_norm2(qd_ctx* ctx, int64_t n, const double* d_psi, double* h_out) {
  if (!ctx || !h_out || (n && !d_psi)) return qd_set_error(ctx, QD_E_ARG, "qd_norm2: null argument");
  cplx* d_res = nullptr;
  QD_HIP(ctx, hipMallocAsync((void**)&d_res, sizeof(cplx), ctx->stream));
  int rc = qd_dev_dot(ctx, n, (const cplx*)d_psi, (const cplx*)d_psi, d_res);
  double r[2] = {0, 0};
  if (!rc) rc = read_one(ctx, d_res, r);
  hipFreeAsync(d_res, ctx->stream);
  if (rc) return rc;
  *h_out = r[0];
  return QD_OK;
}

extern "C" int qd_expect(qd_ctx* ctx, qd_linop* op, const double* d_psi, double* h_out) {
  if (!ctx || !op || !h_out || (op->n && !d_psi)) return qd_set_error(ctx, QD_E_ARG, "qd_expect: null argument");
  int64_t n = op->n;
  cplx* tmp = nullptr;
  QD_HIP(ctx, hipMallocAsync((void**)&tmp, sizeof(cplx) * (n + 1), ctx->stream));
  int rc = qd_linop_apply(op, (const cplx*)d_psi, tmp, nullptr, 0, nullptr, cmake(0.0, 0.0));
  if (!rc) rc = qd_dev_dot(ctx, n, (const cplx*)d_psi, tmp, tmp + n);
  double r[2] = {0, 0};
  if (!rc) rc = read_one(ctx, tmp + n, r);
  hipFreeAsync(tmp, ctx->stream);
  if (rc) return rc;
  h_out[0] = r[0];
  h_out[1] = r[1];
  return QD_OK;
}

// ------------------------------------------------- product fused with a dot
int64_t qd_linop_dot_blocks(int64_t n) { return stream3_grid(n); }
int qd_linop_dot_rows_per_block() { return 64 * STREAM3_WPB; }

// Stored-matrix operators: y = (H + diag + shift) x in the CSR-stream kernel,
// whose blocks also write conj(x_r) y_r summed over their rows to
// part[0 .. qd_linop_dot_blocks(n)).  Other operators: QD_E_UNSUPPORTED,
// nothing launched.
int qd_linop_apply_dot(qd_linop* op, const cplx* x, cplx* y, cplx* part, bool write) {
  qd_ctx* ctx = op->ctx;
  const int64_t n = op->n;
  if (op->kind != QD_LINOP_CSR || n == 0) return QD_E_UNSUPPORTED;
  const unsigned grid = (unsigned)qd_linop_dot_blocks(n);
  const cplx* diag = op->diag0;
  if (write)
    spmv_stream3_kernel<8, true, true><<<grid, 64 * STREAM3_WPB, 0, ctx->stream>>>(n, op->outer, op->inner, op->vals, x, y, nullptr,
                                                                      0, diag, cmake(0.0, 0.0), part);
  else
    spmv_stream3_kernel<8, true, false><<<grid, 64 * STREAM3_WPB, 0, ctx->stream>>>(n, op->outer, op->inner, op->vals, x, y,
                                                                       nullptr, 0, diag, cmake(0.0, 0.0), part);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

// ---------------------------------------------------------------- seeded state
__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double seeded_uniform(uint64_t seed, uint64_t c) {
  uint64_t u = splitmix64((seed << 32) ^ c);
  return ((double)(u >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}
__global__ void seeded_fill_kernel(int64_t n, uint64_t seed, cplx* __restrict__ psi) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  psi[i] = cmake(seeded_uniform(seed, 2 * (uint64_t)i), seeded_uniform(seed, 2 * (uint64_t)i + 1));
}
__global__ void scale_by_inv_sqrt_kernel(int64_t n, cplx* __restrict__ psi, const cplx* __restrict__ nrm2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = sqrt(nrm2->x);
  cplx v = psi[i];
  psi[i] = cmake(v.x / s, v.y / s);
}

extern "C" int qd_seeded_state(qd_ctx* ctx, int64_t n, uint64_t seed, double* d_psi) {
  if (!ctx || (n && !d_psi)) return qd_set_error(ctx, QD_E_ARG, "qd_seeded_state: null argument");
  if (n == 0) return QD_OK;
  cplx* psi = (cplx*)d_psi;
  seeded_fill_kernel<<<qd_grid(n, 256), 256, 0, ctx->stream>>>(n, seed, psi);
  cplx* d_res = nullptr;
  QD_HIP(ctx, hipMallocAsync((void**)&d_res, sizeof(cplx), ctx->stream));
  int rc = qd_dev_dot(ctx, n, psi, psi, d_res);
  if (!rc) scale_by_inv_sqrt_kernel<<<qd_grid(n, 256), 256, 0, ctx->stream>>>(n, psi, d_res);
  hipFreeAsync(d_res, ctx->stream);
  if (rc) return rc;
  QD_CHECK_LAUNCH(ctx);
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

// ---------------------------------------------------------------- tuner
// Time every row mapping of a stored-matrix operator on (x -> y) with HIP
// events on the context stream and keep the fastest.  The results of all
// mappings agree to rounding (each has a fixed summation order).
extern "C" int qd_linop_autotune(qd_ctx* ctx, qd_linop* op, const double* d_x, double* d_y, int reps,
                                 int* best_out, float* best_us) {
  if (!ctx || !op || !d_x || !d_y || reps < 1) return qd_set_error(ctx, QD_E_ARG, "qd_linop_autotune: bad arguments");
  if (op->kind != QD_LINOP_CSR) {
    if (best_out) *best_out = 0;
    return QD_OK;
  }
  std::vector<int> cands;
  for (int v = 0; v < QD_STREAM_VARIANTS; ++v) cands.push_back(-v);
  for (int g = 4; g <= 64; g <<= 1) cands.push_back(g);
  hipEvent_t e0, e1;
  QD_HIP(ctx, hipEventCreate(&e0));
  QD_HIP(ctx, hipEventCreate(&e1));
  int keep = op->lanes_per_row, best = keep;
  float best_ms = 1e30f;
  for (int c : cands) {
    op->lanes_per_row = c;
    int rc = qd_linop_apply(op, (const cplx*)d_x, (cplx*)d_y, nullptr, 0, nullptr, cmake(0.0, 0.0));
    if (rc) {
      op->lanes_per_row = keep;
      return rc;
    }
    hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps; ++r) qd_linop_apply(op, (const cplx*)d_x, (cplx*)d_y, nullptr, 0, nullptr, cmake(0.0, 0.0));
    hipEventRecord(e1, ctx->stream);
    QD_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  op->lanes_per_row = best;
  if (best_out) *best_out = best;
  if (best_us) *best_us = best_ms * 1000.f / reps;
  return QD_OK;
}

extern "C" int qd_linop_get_lanes(const qd_linop* op) { return op ? op->lanes_per_row : 0; }
