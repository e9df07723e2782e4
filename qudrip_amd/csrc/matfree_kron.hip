// matfree_kron.hip -- K6 fast path: matrix-free product for programs that
// split over a two-block sector (the Hubbard layout of
// programs/hubbard_diss.cpp:97-104: up spins on the low bits, down spins on
// the high bits).
//
// With the sector a mixed radix of block A (low bits, n_A states) and block
// B (high bits, n_B states), a state vector is a dense matrix X[b][a]
// (index b * n_A + a).  When every term is either
//   - a product of hops inside A        (the Jordan-Wigner popcounts of a hop
//                                        pair each pick up the constant
//                                        particle number of B: an even total,
//                                        so the sign is A-local),
//   - any product of ops inside B        (B is the highest block: its JW
//                                        strings only see B bits), or
//   - diagonal (number operators, diagonal bit gates),
// the operator is H = T_A (x) I + I (x) T_B + D and
//   Y[b][a] = sum_a' T_A[a][a'] X[b][a'] + sum_b' T_B[b][b'] X[b'][a] + D(b,a) X[b][a].
// T_A and T_B are small sparse matrices built once with the exact CSR
// builder on the block sectors; D is evaluated on the fly from the diagonal
// terms.  Two kernels:
//   kron_rows_kernel  one workgroup per row b: X[b][:] staged in LDS, the
//                     T_A gathers read LDS, plus D -> writes Y
//   kron_slab_kernel  column slabs of W entries: every row b adds
//                     sum T_B[b][b'] X[b'][slab]; workgroups of one slab run
//                     on one XCD so the slab (n_B * W * 16 B) stays in its L2.
// HBM traffic ~ X twice + Y three times (vs. 32 n algorithmic).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "kron.hpp"

namespace {
typedef double kdbl2 __attribute__((ext_vector_type(2)));
// y is written by the rows pass and read-modified-written by the slab pass,
// once each per product, and is larger than the caches for the sizes where
// the product matters: stream it non-temporally so L2/MALL keep x.
__device__ __forceinline__ cplx y_ld(const cplx* p) {
  const kdbl2 v = __builtin_nontemporal_load(reinterpret_cast<const kdbl2*>(p));
  return cmake(v.x, v.y);
}
__device__ __forceinline__ void y_st(cplx* p, double re, double im) {
  __builtin_nontemporal_store(re, &p->x);
  __builtin_nontemporal_store(im, &p->y);
}
}  // namespace

typedef double dbl2 __attribute__((ext_vector_type(2)));


struct qd_kron {
  uint32_t* d_ellB = nullptr;
  cplx* d_vtabB = nullptr;
  uint32_t* d_ellB4 = nullptr;
  cplx* d_vtabB4 = nullptr;
  uint64_t* d_nmask = nullptr;
  cplx* d_ncoef = nullptr;
  uint32_t* d_ell = nullptr;
  uint32_t* d_ell16 = nullptr;
  uint32_t* d_ell16s = nullptr;
  cplx* d_vtab = nullptr;
  qd_basis* basisA = nullptr;
  qd_basis* basisB = nullptr;
  qd_termprog* progA = nullptr;
  qd_termprog* progB = nullptr;
  qd_csr* csrA = nullptr;
  qd_csr* csrB = nullptr;
  DevTerm* d_dterms = nullptr;
  DevOp* d_dops = nullptr;
  int ndterms = 0;
  KronDesc desc{};
  // launch variants, chosen by timing at the first product: rows * 16 + slab
  // in one word, so a product on another thread reads both or neither
  // (-1: not chosen yet); tune_mu lets one thread time them
  std::atomic<int> variants{-1};
  std::mutex tune_mu;
};

__device__ inline cplx diag_value(const KronDesc& k, uint64_t s) {
  cplx d = cmake(0.0, 0.0);
  for (int t = 0; t < k.nnum; ++t) {
    const uint64_t m = k.nmask[t];
    if ((s & m) == m) d = cadd(d, k.ncoef[t]);
  }
  for (int t = 0; t < k.ndterms; ++t) {
    const DevTerm term = k.dterms[t];
    const DevOp* tops = k.dops + term.op_begin;
    int nbitops = 0;
    for (int q = 0; q < term.n_ops; ++q) nbitops += (tops[q].kind == QD_OP_BIT);
    for (int L = 0; L < term.n_leaves; ++L) {
      uint64_t xt;
      cplx w;
      if (!eval_leaf(tops, term.n_ops, nbitops, s, L, &xt, &w)) continue;
      if (xt != s) continue;
      d = cadd(d, cmul(term.coeff, w));
    }
  }
  return d;
}

// one workgroup per row b; LDS holds X[b][:]
template <bool LDSX>
__global__ void __launch_bounds__(1024)
    kron_rows_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, cplx shift) {
  extern __shared__ cplx xrow[];
  const int64_t b = blockIdx.x;
  const cplx* __restrict__ xb = x + b * k.nA;
  if (LDSX) {
    for (int64_t a = threadIdx.x; a < k.nA; a += blockDim.x) xrow[a] = xb[a];
    __syncthreads();
  }
  const uint64_t sB = k.statesB[b] << k.loB;
  for (int64_t a = threadIdx.x; a < k.nA; a += blockDim.x) {
    double sx = 0.0, sy = 0.0;
    const int32_t e0 = k.Aouter[a], e1 = k.Aouter[a + 1];
    for (int32_t e = e0; e < e1; ++e) {
      cplx v = k.Avals[e];
      cplx xv = LDSX ? xrow[k.Ainner[e]] : xb[k.Ainner[e]];
      sx += v.x * xv.x + v.y * xv.y;  // conj(v) * x
      sy += v.x * xv.y - v.y * xv.x;
    }
    cplx xa = LDSX ? xrow[a] : xb[a];
    cplx d = cadd(diag_value(k, sB | k.statesA[a]), shift);
    if (k.ext_diag) d = cadd(d, k.ext_diag[b * k.nA + a]);
    cplx dx = cmul(d, xa);
    y_st(y + b * k.nA + a, sx + dx.x, sy + dx.y);
  }
}

// Packed variant: RB rows per workgroup share one pass over the packed T_A;
// X rows and the value table live in LDS.
template <int RB>
__global__ void __launch_bounds__(1024)
    kron_rows_packed_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, cplx shift) {
  extern __shared__ cplx lds[];
  cplx* vt = lds;                 // nvtab entries
  cplx* xr = lds + k.nvtab;       // RB rows of nA
  const int64_t b0 = (int64_t)blockIdx.x * RB;
  const int rb = (int)min<int64_t>(RB, k.nB - b0);
  for (int i = threadIdx.x; i < k.nvtab; i += blockDim.x) vt[i] = k.vtab[i];
  for (int r = 0; r < rb; ++r) {
    const cplx* __restrict__ xb = x + (b0 + r) * k.nA;
    for (int64_t a = threadIdx.x; a < k.nA; a += blockDim.x) xr[r * k.nA + a] = xb[a];
  }
  __syncthreads();
  uint64_t sB[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) sB[r] = (r < rb) ? (k.statesB[b0 + r] << k.loB) : 0;
  for (int64_t a = threadIdx.x; a < k.nA; a += blockDim.x) {
    double sx[RB], sy[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) sx[r] = sy[r] = 0.0;
    for (int j = 0; j < k.ell_w; ++j) {
      const uint32_t e = k.ell[(int64_t)j * k.nA + a];
      const cplx v = vt[e & 0xFFFFu];
      const int c = (int)(e >> 16);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const cplx xv = xr[r * k.nA + c];
        sx[r] += v.x * xv.x + v.y * xv.y;  // conj(v) * x
        sy[r] += v.x * xv.y - v.y * xv.x;
      }
    }
    const uint64_t sA = k.statesA[a];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (r < rb) {
        cplx d = cadd(diag_value(k, sB[r] | sA), shift);
        if (k.ext_diag) d = cadd(d, k.ext_diag[(b0 + r) * k.nA + a]);
        cplx dx = cmul(d, xr[r * k.nA + a]);
        y_st(y + (b0 + r) * k.nA + a, sx[r] + dx.x, sy[r] + dx.y);
      }
    }
  }
}

// Y[b][slab] += sum_e conj(TB[b][e]) X[col_e][slab]; W = 32 columns per slab,
// a wave covers two rows (32 lanes each).  Block id -> (XCD, slab, row chunk)
// so that one XCD works through its slabs one after the other.
template <int W>
__global__ void __launch_bounds__(256)
    kron_slab_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, int chunks, int rows_per_chunk) {
  const int xcd = blockIdx.x & 7;
  const int kk = blockIdx.x >> 3;
  const int64_t slab = (int64_t)xcd + 8 * (kk / chunks);
  const int chunk = kk % chunks;
  const int64_t a0 = slab * W;
  if (a0 >= k.nA) return;
  const int lane = threadIdx.x % W;
  const int sub = threadIdx.x / W;  // rows handled by this block in parallel
  const int nsub = blockDim.x / W;
  const int64_t a = a0 + lane;
  const bool col_ok = a < k.nA;
  const int64_t rb = (int64_t)chunk * rows_per_chunk;
  const int64_t re = min<int64_t>(k.nB, rb + rows_per_chunk);
  for (int64_t b = rb + sub; b < re; b += nsub) {
    double sx = 0.0, sy = 0.0;
    const int32_t e0 = k.Bouter[b], e1 = k.Bouter[b + 1];
    for (int32_t e = e0; e < e1; ++e) {
      const cplx v = k.Bvals[e];
      const int32_t bp = k.Binner[e];
      if (col_ok) {
        cplx xv = x[(int64_t)bp * k.nA + a];
        sx += v.x * xv.x + v.y * xv.y;
        sy += v.x * xv.y - v.y * xv.x;
      }
    }
    if (col_ok) {
      const cplx o = y_ld(y + b * k.nA + a);
      y_st(y + b * k.nA + a, o.x + sx, o.y + sy);
    }
  }
}

// Packed slab kernel: T_B entries of a row are read once per 32-lane group
// (broadcast) U at a time, and the U gathers of X row segments are issued
// before any is consumed.
template <int W, int U>
__global__ void __launch_bounds__(256)
    kron_slab_packed_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, int chunks,
                            int rows_per_chunk) {
  __shared__ cplx vt[256];
  const int xcd = blockIdx.x & 7;
  const int kk = blockIdx.x >> 3;
  const int64_t slab = (int64_t)xcd + 8 * (kk / chunks);
  const int chunk = kk % chunks;
  const int64_t a0 = slab * W;
  if (a0 >= k.nA) return;
  for (int i = threadIdx.x; i < k.nvtabB; i += blockDim.x) vt[i] = k.vtabB[i];
  __syncthreads();
  const int lane = threadIdx.x % W;
  const int sub = threadIdx.x / W;
  const int nsub = blockDim.x / W;
  const int64_t a = min<int64_t>(a0 + lane, k.nA - 1);  // clamp: duplicate lanes recompute the last column
  const bool col_ok = a0 + lane < k.nA;
  const int64_t rb = (int64_t)chunk * rows_per_chunk;
  const int64_t re = min<int64_t>(k.nB, rb + rows_per_chunk);
  const int w = k.ellB_w;
  for (int64_t b = rb + sub; b < re; b += nsub) {
    const uint32_t* __restrict__ er = k.ellB + b * w;
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < w; j += U) {
      uint32_t e[U];
      cplx xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) e[u] = (j + u < w) ? er[j + u] : 0u;
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = x[(int64_t)(e[u] >> 16) * k.nA + a];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const cplx v = vt[e[u] & 0xFFFFu];
        sx += v.x * xv[u].x + v.y * xv[u].y;
        sy += v.x * xv[u].y - v.y * xv[u].x;
      }
    }
    if (col_ok) {
      const cplx o = y_ld(y + b * k.nA + a);
      y_st(y + b * k.nA + a, o.x + sx, o.y + sy);
    }
  }
}

// Rows kernel with the loads kept in flight: the RB rows (contiguous in x)
// are staged NU loads per thread at a time, and the packed T_A entries of an
// output are fetched 8 at a time before their products (same summation order
// as kron_rows_packed_kernel: entries in order, pad entries add 0).
template <int RB, int NU>
__global__ void __launch_bounds__(1024)
    kron_rows2_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, cplx shift) {
  extern __shared__ cplx lds[];
  cplx* vt = lds;            // nvtab entries
  cplx* xr = lds + k.nvtab;  // RB rows of nA
  const int64_t b0 = (int64_t)blockIdx.x * RB;
  const int rb = (int)min<int64_t>(RB, k.nB - b0);
  for (int i = threadIdx.x; i < k.nvtab; i += blockDim.x) vt[i] = k.vtab[i];
  const int64_t cnt = (int64_t)rb * k.nA;
  const cplx* __restrict__ xs = x + b0 * k.nA;
  for (int64_t i0 = threadIdx.x; i0 < cnt; i0 += (int64_t)blockDim.x * NU) {
    cplx v[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int64_t i = i0 + (int64_t)u * blockDim.x;
      v[u] = xs[min(i, cnt - 1)];  // clamped: every load in bounds, no stack array
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int64_t i = i0 + (int64_t)u * blockDim.x;
      if (i < cnt) xr[i] = v[u];
    }
  }
  __syncthreads();
  uint64_t sB[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) sB[r] = (r < rb) ? (k.statesB[b0 + r] << k.loB) : 0;
  for (int64_t a = threadIdx.x; a < k.nA; a += blockDim.x) {
    double sx[RB], sy[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) sx[r] = sy[r] = 0.0;
    for (int j0 = 0; j0 < k.ell_w; j0 += 8) {
      uint32_t e[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) e[u] = (j0 + u < k.ell_w) ? k.ell[(int64_t)(j0 + u) * k.nA + a] : 0u;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if ((e[u] & 0xFFFFu) == 0) continue;  // pad entry (value 0): no gather
        const cplx v = vt[e[u] & 0xFFFFu];
        const int c = (int)(e[u] >> 16);
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const cplx xv = xr[r * k.nA + c];
          sx[r] += v.x * xv.x + v.y * xv.y;  // conj(v) * x
          sy[r] += v.x * xv.y - v.y * xv.x;
        }
      }
    }
    const uint64_t sA = k.statesA[a];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (r < rb) {
        cplx d = cadd(diag_value(k, sB[r] | sA), shift);
        if (k.ext_diag) d = cadd(d, k.ext_diag[(b0 + r) * k.nA + a]);
        cplx dx = cmul(d, xr[r * k.nA + a]);
        y_st(y + (b0 + r) * k.nA + a, sx[r] + dx.x, sy[r] + dx.y);
      }
    }
  }
}

// Persistent rows kernel: each block keeps the compressed T_A entries of its
// output columns (a = t + 1024 i, i < SLOTS) in registers for all its rows,
// so T_A is read once per block instead of once per row; the block walks a
// contiguous range of rows with two LDS row buffers: row r+1 is loaded into
// registers while row r is computed, then stored to the other buffer.  Same
// summation order as the packed kernels (entries in order; pad entries,
// which add 0, are skipped).
// D(b, a) from the number-operator masks only (GENDIAG = false) or also the
// general diagonal terms
template <bool GENDIAG>
__device__ __forceinline__ cplx diag_of(const KronDesc& k, uint64_t s) {
  if (GENDIAG) return diag_value(k, s);
  cplx d = cmake(0.0, 0.0);
  for (int t = 0; t < k.nnum; ++t) {
    const uint64_t m = k.nmask[t];
    if ((s & m) == m) d = cadd(d, k.ncoef[t]);
  }
  return d;
}

// Persistent rows kernel: one block per CU walks a contiguous range of rows
// with two LDS row buffers and one row in registers -- while row r is
// computed from one buffer, row r+1 (loaded during row r-1) sits in the
// other and row r+2 is in flight to registers -- so the HBM latency of the
// staging hides behind a whole row of gathers.  Thread t computes the
// outputs a = t + 1024 i (i < S = ceil(nA / 1024)) and keeps the compressed
// T_A words of all S outputs in registers for every row it visits (T_A is
// read from memory once per block): slot 0 is the one in use, and after
// each output the S slots rotate by one, so every register index is a
// compile-time constant while the output loop stays rolled.  Same
// summation order as the packed kernels (entries in order; pad entries,
// which add 0, are skipped).
template <int WP, int S, bool GENDIAG>
__global__ void __launch_bounds__(1024)
    kron_rows3_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, cplx shift, int rows_per_block) {
  extern __shared__ cplx lds[];
  cplx* vt = lds;        // 8 values
  cplx* nc = lds + 8;    // 64 number-product coefficients
  uint64_t* nmA = reinterpret_cast<uint64_t*>(lds + 72);  // their block-A masks
  uint64_t* nmB = nmA + 64;                                // block-B masks
  cplx* buf0 = lds + 136;  // two row buffers of nA
  const int nA = (int)k.nA;
  cplx* buf1 = buf0 + nA;
  const int64_t b0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t b1 = min<int64_t>(k.nB, b0 + rows_per_block);
  if (b0 >= b1) return;  // block-uniform
  const int t = threadIdx.x;
  const int nnum = GENDIAG ? 0 : k.nnum;  // <= 64 when !GENDIAG
  const uint64_t amask = k.loB >= 64 ? ~0ull : ((1ull << k.loB) - 1ull);
  if (t < 8) vt[t] = t < k.nvtab ? k.vtab[t] : cmake(0.0, 0.0);
  if (t < nnum) {
    nc[t] = k.ncoef[t];
    nmA[t] = k.nmask[t] & amask;
    nmB[t] = k.nmask[t] >> k.loB;
  }
  {
    const cplx* __restrict__ xb = x + b0 * nA;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int c = t + 1024 * i;
      if (c < nA) buf0[c] = xb[c];
    }
  }
  __syncthreads();
  // per output: its T_A words, and either its block-A state (general
  // diagonal) or the set F(a) of number products whose A bits it has
  uint32_t e[S][WP];
  uint64_t sa[S];
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int a = t + 1024 * i;
#pragma unroll
    for (int p = 0; p < WP; ++p) e[i][p] = a < nA ? k.ell16[p * nA + a] : 0u;
    const uint64_t st = k.statesA[a < nA ? a : nA - 1];
    if (GENDIAG) {
      sa[i] = st;
    } else {
      uint64_t f = 0;
      for (int q = 0; q < nnum; ++q)
        if ((st & nmA[q]) == nmA[q]) f |= 1ull << q;
      sa[i] = f;
    }
  }
  // row b0 + 1 in registers (plain vector type, so the array stays in
  // registers across the loop): stored to LDS at the top of the first iteration
  kdbl2 pre[S];
  {
    const kdbl2* __restrict__ xn = reinterpret_cast<const kdbl2*>(x + min<int64_t>(b0 + 1, b1 - 1) * nA);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int c = t + 1024 * i;
      pre[i] = xn[c < nA ? c : nA - 1];
    }
  }
  for (int64_t b = b0; b < b1; ++b) {
    const int par = (int)((b - b0) & 1);
    const cplx* cur = par ? buf1 : buf0;
    cplx* nxt = par ? buf0 : buf1;
    // row b+1, loaded during the previous iteration, goes to the free buffer
    // (its last reader, row b-1, finished before the previous barrier); then
    // row b+2 is requested, to arrive while row b is computed
    const uint64_t sbb = k.statesB[b];  // before the prefetch: loads complete in order
    if (b + 1 < b1) {
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const int c = t + 1024 * i;
        if (c < nA) nxt[c] = cmake(pre[i].x, pre[i].y);
      }
    }
    {
      const kdbl2* __restrict__ xn = reinterpret_cast<const kdbl2*>(x + min<int64_t>(b + 2, b1 - 1) * nA);
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const int c = t + 1024 * i;
        pre[i] = xn[c < nA ? c : nA - 1];
      }
    }
    uint64_t g = 0;  // number products whose B bits row b has
    if (!GENDIAG) {
#pragma unroll 8
      for (int q = 0; q < nnum; ++q)
        if ((sbb & nmB[q]) == nmB[q]) g |= 1ull << q;
    }
    const uint64_t sB = sbb << k.loB;
#pragma nounroll
    for (int i = 0; i < S; ++i) {
      const int a = t + 1024 * i;
      if (a < nA) {
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int p = 0; p < WP; ++p) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const uint32_t w = (e[0][p] >> (16 * h)) & 0xFFFFu;
            if (w & 7u) {  // pad words (0) are skipped
              const cplx v = vt[w & 7u];
              const cplx xv = cur[w >> 3];
              sx += v.x * xv.x + v.y * xv.y;  // conj(v) * x
              sy += v.x * xv.y - v.y * xv.x;
            }
          }
        }
        cplx d;
        if (GENDIAG) {
          d = diag_value(k, sB | sa[0]);
        } else {
          // the matching products in mask order: the same sum as diag_value
          d = cmake(0.0, 0.0);
          for (uint64_t m = sa[0] & g; m; m &= m - 1) d = cadd(d, nc[__builtin_ctzll(m)]);
        }
        d = cadd(d, shift);
        if (k.ext_diag) d = cadd(d, k.ext_diag[b * nA + a]);
        const cplx dx = cmul(d, cur[a]);
        y_st(y + b * nA + a, sx + dx.x, sy + dx.y);
      }
      // rotate the slots: the next output's words and state move to slot 0
#pragma unroll
      for (int p = 0; p < WP; ++p) {
        const uint32_t w0 = e[0][p];
#pragma unroll
        for (int q = 0; q + 1 < S; ++q) e[q][p] = e[q + 1][p];
        e[S - 1][p] = w0;
      }
      {
        const uint64_t s0 = sa[0];
#pragma unroll
        for (int q = 0; q + 1 < S; ++q) sa[q] = sa[q + 1];
        sa[S - 1] = s0;
      }
    }
    __syncthreads();
  }
}

// x entry of a row buffer at byte offset `off` (LDS)
__device__ __forceinline__ kdbl2 lds_x(const cplx* buf, uint32_t off) {
  return *reinterpret_cast<const kdbl2*>(reinterpret_cast<const char*>(buf) + off);
}
// d with its sign bit flipped where bit 31 of s is set
__device__ __forceinline__ double flip_sign(double d, uint32_t s) {
  return __longlong_as_double(__double_as_longlong(d) ^ ((long long)s << 32));
}

// Branch-free persistent rows kernel.  Same row walk as kron_rows3_kernel
// (one block per CU over a contiguous range of rows, two LDS row buffers,
// the T_A words of a thread's S outputs a = t + 1024 i in registers), with
// these changes:
//  - every entry slot is evaluated: a pad word (0) points at slot 0 of the
//    row buffer, which holds 0 (the row itself sits at slots 1..nA), so no
//    entry needs a branch.  A pad adds +0: the result differs from skipping
//    it at most in the sign of a zero;
//  - the S * 2 WP entry slots of a thread's row are one unrolled sequence
//    taken in groups of G: the gathers of group g + 1 are issued before the
//    sums of group g (PIPE), so an LDS wait covers G gathers and overlaps
//    G sums;
//  - SIGN: when every value of T_A is +c or -c for one c (hopping with one
//    amplitude), the entry carries the sign in bit 2 instead of a table id,
//    the row sums sum_e (+/-) x[col_e] and multiplies by conj(c) once per
//    output -- one LDS read and two additions per entry instead of two reads
//    and four multiply-adds;
//  - DMA: the next row goes from HBM straight into the free LDS buffer
//    (global_load_lds_dwordx4, 64 entries per wave and instruction) while
//    the current row is computed, instead of through registers.
// The number-product diagonal uses 32-bit sets (at most 32 products).
// Entries are summed in order with fused multiply-adds, so the rounding
// differs from the other rows kernels by a few units in the last place.
template <int WP, int S, bool SIGN, bool GENDIAG, int G, bool PIPE, bool DMA>
__global__ void __launch_bounds__(1024)
    kron_rows4_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, cplx shift, int rows_per_block) {
  extern __shared__ cplx lds[];
  cplx* vt = lds;      // 8 values
  cplx* nc = lds + 8;  // 32 number-product coefficients
  uint64_t* nmB = reinterpret_cast<uint64_t*>(lds + 40);  // their block-B masks
  const int nA = (int)k.nA;
  cplx* buf0 = lds + 56;  // two row buffers of 1 + nA: slot 0 = 0, then the row
  cplx* buf1 = buf0 + nA + 1;
  const int64_t b0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t b1 = min<int64_t>(k.nB, b0 + rows_per_block);
  if (b0 >= b1) return;  // block-uniform
  const int t = threadIdx.x;
  const int nnum = GENDIAG ? 0 : k.nnum;  // <= 32 when !GENDIAG
  const uint64_t amask = k.loB >= 64 ? ~0ull : ((1ull << k.loB) - 1ull);
  if (t < 8) vt[t] = t < k.nvtab ? k.vtab[t] : cmake(0.0, 0.0);
  if (t < nnum) {
    nc[t] = k.ncoef[t];
    nmB[t] = k.nmask[t] >> k.loB;
  }
  if (t == 0) {
    buf0[0] = cmake(0.0, 0.0);
    buf1[0] = cmake(0.0, 0.0);
  }
  // row r of x into slots 1..nA of buf: 64 entries per wave and DMA
  // instruction (lane l writes 16 bytes at the wave's LDS base + 16 l)
  auto dma_row = [&](int64_t r, cplx* buf) {
    const cplx* __restrict__ xr = x + r * nA;
    const int lane = t & 63;
    for (int q = t >> 6; q * 64 < nA; q += 16) {
      const int c = q * 64 + lane;
      if (c < nA)
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(xr + c),
                                         (__attribute__((address_space(3))) void*)(buf + 1 + q * 64), 16, 0, 0);
    }
  };
  if (DMA) {
    dma_row(b0, buf0);
  } else {
    const cplx* __restrict__ xb = x + b0 * nA;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int c = t + 1024 * i;
      if (c < nA) buf0[1 + c] = xb[c];
    }
  }
  uint32_t e[S][WP];
  uint32_t sa[S];
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int a = t + 1024 * i;
#pragma unroll
    for (int p = 0; p < WP; ++p) e[i][p] = a < nA ? k.ell16s[p * nA + a] : 0u;
    const uint64_t st = k.statesA[a < nA ? a : nA - 1];
    if (GENDIAG) {
      sa[i] = (uint32_t)(a < nA ? a : nA - 1);  // the state is read again when needed
    } else {
      uint32_t f = 0;
      for (int q = 0; q < nnum; ++q) {
        const uint64_t m = k.nmask[q] & amask;
        if ((st & m) == m) f |= 1u << q;
      }
      sa[i] = f;
    }
  }
  const cplx ac = k.acoef;
  if (DMA) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's row copies have landed
  __syncthreads();
  kdbl2 pre[DMA ? 1 : S];
  if (!DMA) {
    const kdbl2* __restrict__ xn = reinterpret_cast<const kdbl2*>(x + min<int64_t>(b0 + 1, b1 - 1) * nA);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const int c = t + 1024 * i;
      pre[i] = xn[c < nA ? c : nA - 1];
    }
  }
  constexpr int NSL = 2 * WP;  // entry slots per output
  constexpr int TOT = S * NSL;
  constexpr int NG = (TOT + G - 1) / G;
  for (int64_t b = b0; b < b1; ++b) {
    const int par = (int)((b - b0) & 1);
    const cplx* cur = par ? buf1 : buf0;
    cplx* nxt = par ? buf0 : buf1;
    const uint64_t sbb = k.statesB[b];
    if (DMA) {
      if (b + 1 < b1) dma_row(b + 1, nxt);  // nxt's last reader, row b - 1, ended before the barrier
    } else {
      if (b + 1 < b1) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
          const int c = t + 1024 * i;
          if (c < nA) nxt[1 + c] = cmake(pre[i].x, pre[i].y);
        }
      }
      const kdbl2* __restrict__ xn = reinterpret_cast<const kdbl2*>(x + min<int64_t>(b + 2, b1 - 1) * nA);
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const int c = t + 1024 * i;
        pre[i] = xn[c < nA ? c : nA - 1];
      }
    }
    uint32_t g = 0;  // number products whose B bits row b has
    if (!GENDIAG) {
#pragma unroll 8
      for (int q = 0; q < nnum; ++q)
        if ((sbb & nmB[q]) == nmB[q]) g |= 1u << q;
    }
    const uint64_t sB = sbb << k.loB;
    // decoded afresh for every row: offsets hoisted out of the row loop
    // would hold 2 S WP registers more
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int p = 0; p < WP; ++p) asm volatile("" : "+v"(e[i][p]));
    kdbl2 xr[2][G];
    double sx = 0.0, sy = 0.0;
    // gathers of slot group gg into register set gg & 1
#define QD_ROWS4_ISSUE(gg)                                                   \
  _Pragma("unroll") for (int j = 0; j < G; ++j) {                            \
    const int sl = (gg) * G + j;                                             \
    if (sl < TOT) {                                                          \
      const uint32_t w = e[sl / NSL][(sl % NSL) >> 1];                       \
      const uint32_t hw = (sl & 1) ? (w >> 16) : (w & 0xFFFFu);              \
      xr[(gg) & 1][j] = lds_x(cur, (hw >> 3) << 4);                          \
    }                                                                        \
  }
    QD_ROWS4_ISSUE(0)
#pragma unroll
    for (int gi = 0; gi < NG; ++gi) {
      if (PIPE && gi + 1 < NG) {
        QD_ROWS4_ISSUE(gi + 1)
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int sl = gi * G + j;
        if (sl < TOT) {
          const int i = sl / NSL;
          const uint32_t w = e[i][(sl % NSL) >> 1];
          const uint32_t hw = (sl & 1) ? (w >> 16) : (w & 0xFFFFu);
          const kdbl2 xv = xr[gi & 1][j];
          if (SIGN) {
            const uint32_t s = (hw << 29) & 0x80000000u;  // bit 2: the entry is -c
            sx += flip_sign(xv.x, s);
            sy += flip_sign(xv.y, s);
          } else {
            const cplx v = vt[hw & 7u];
            sx = __builtin_fma(v.x, xv.x, sx);  // conj(v) * x
            sx = __builtin_fma(v.y, xv.y, sx);
            sy = __builtin_fma(v.x, xv.y, sy);
            sy = __builtin_fma(-v.y, xv.x, sy);
          }
          if (sl % NSL == NSL - 1) {  // output i complete
            const int a = t + 1024 * i;
            asm volatile("" : "+v"(sx), "+v"(sy));
            if (SIGN) {
              const double rx = ac.x * sx + ac.y * sy;  // conj(c) * sum
              const double ry = ac.x * sy - ac.y * sx;
              sx = rx;
              sy = ry;
            }
            if (a < nA) {
              cplx d;
              if (GENDIAG) {
                d = diag_value(k, sB | k.statesA[sa[i]]);
              } else {
                d = cmake(0.0, 0.0);
                for (uint32_t m = sa[i] & g; m; m &= m - 1) d = cadd(d, nc[__builtin_ctz(m)]);
              }
              d = cadd(d, shift);
              if (k.ext_diag) d = cadd(d, k.ext_diag[b * nA + a]);
              const cplx dx = cmul(d, cur[1 + a]);
              y_st(y + b * nA + a, sx + dx.x, sy + dx.y);
            }
            sx = 0.0;
            sy = 0.0;
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      if (!PIPE && gi + 1 < NG) {
        QD_ROWS4_ISSUE(gi + 1)
      }
    }
#undef QD_ROWS4_ISSUE
    if (DMA) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the next row has landed
    __syncthreads();
  }
}

// Fused two-block product for small blocks (nA <= 1024, x within the
// MALL): one 256-thread block per row b writes
//   y[b][a] = conj(cA) sum_{T_A entries of a} (+/-) x[b][a']
//           + conj(cB) sum_{T_B entries of b} (+/-) x[b'][a] + (D + shift + ext) x[b][a]
// in one pass.  The row x[b] is staged in LDS for the T_A gathers (the
// branch-free sign form of kron_rows4_kernel); row b's T_B entries are the
// same for the whole block (scalar loads of ellB4: byte offsets of the
// source rows and sign codes), so the T_B part is coalesced loads of the
// source rows, issued before the LDS staging completes.  y is written once,
// and there is one launch per product instead of two.  Needs T_A and T_B in
// sign form and the diagonal from at most 64 number products.
template <int WP, int S>
__global__ void __launch_bounds__(256)
    kron_fused_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, cplx shift) {
  extern __shared__ cplx lds[];
  cplx* nc = lds;                                           // 64 number-product coefficients
  uint64_t* nmB = reinterpret_cast<uint64_t*>(lds + 64);    // their block-B masks (32 cplx)
  cplx* row = lds + 96;                                     // slot 0 = 0, then x[b]
  const int nA = (int)k.nA;
  const int64_t b = blockIdx.x;
  const int t = threadIdx.x;
  const int nnum = k.nnum;
  const uint64_t amask = k.loB >= 64 ? ~0ull : ((1ull << k.loB) - 1ull);
  const cplx* __restrict__ xb = x + b * nA;
  if (t < nnum) {
    nc[t] = k.ncoef[t];
    nmB[t] = k.nmask[t] >> k.loB;
  }
  if (t == 0) row[0] = cmake(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int c = t + 256 * i;
    if (c < nA) row[1 + c] = xb[c];
  }
  // T_B part: row b's entries, uniform over the block
  const uint32_t* __restrict__ er = k.ellB4 + b * k.ellB4_stride;
  const int ne = (int)er[0];
  const char* __restrict__ xc = reinterpret_cast<const char*>(x);
  double bx[S], by[S];
#pragma unroll
  for (int i = 0; i < S; ++i) bx[i] = by[i] = 0.0;
  for (int j = 0; j < ne; j += 4) {
    kdbl2 v[4][S];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t ent = (j + u < ne) ? er[1 + j + u] : 0u;  // padding: row 0, weight 0
#pragma unroll
      for (int i = 0; i < S; ++i) {
        const int c = min(t + 256 * i, nA - 1);
        v[u][i] = *reinterpret_cast<const kdbl2*>(xc + (ent & ~15u) + (size_t)c * sizeof(cplx));
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t ent = (j + u < ne) ? er[1 + j + u] : 0u;
      const uint32_t code = ent & 15u;
      const double sg = code == 1u ? 1.0 : (code == 2u ? -1.0 : 0.0);  // scalar
#pragma unroll
      for (int i = 0; i < S; ++i) {
        bx[i] = __builtin_fma(sg, v[u][i].x, bx[i]);
        by[i] = __builtin_fma(sg, v[u][i].y, by[i]);
      }
    }
  }
  // T_A words and the outputs' number-product sets
  uint32_t e[S][WP];
  uint64_t sa[S];
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int a = t + 256 * i;
#pragma unroll
    for (int p = 0; p < WP; ++p) e[i][p] = a < nA ? k.ell16s[p * nA + a] : 0u;
    const uint64_t st = k.statesA[a < nA ? a : nA - 1];
    uint64_t f = 0;
    for (int q = 0; q < nnum; ++q) {
      const uint64_t m = k.nmask[q] & amask;
      if ((st & m) == m) f |= 1ull << q;
    }
    sa[i] = f;
  }
  __syncthreads();
  const uint64_t sbb = k.statesB[b];
  uint64_t g = 0;  // number products whose B bits row b has
  for (int q = 0; q < nnum; ++q)
    if ((sbb & nmB[q]) == nmB[q]) g |= 1ull << q;
  const cplx ca = k.acoef, cb = k.bcoef;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int a = t + 256 * i;
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int g0 = 0; g0 < 2 * WP; g0 += 4) {
      kdbl2 xv[4];
      uint32_t hv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int sl = g0 + u;
        if (sl < 2 * WP) {
          const uint32_t w = e[i][sl >> 1];
          hv[u] = (sl & 1) ? (w >> 16) : (w & 0xFFFFu);
          xv[u] = lds_x(row, (hv[u] >> 3) << 4);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (g0 + u < 2 * WP) {
          const uint32_t sgn = (hv[u] << 29) & 0x80000000u;  // bit 2: the entry is -cA
          sx += flip_sign(xv[u].x, sgn);
          sy += flip_sign(xv[u].y, sgn);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    asm volatile("" : "+v"(sx), "+v"(sy));
    if (a < nA) {
      double rx = ca.x * sx + ca.y * sy;  // conj(cA) * sum
      double ry = ca.x * sy - ca.y * sx;
      rx += cb.x * bx[i] + cb.y * by[i];  // conj(cB) * sum
      ry += cb.x * by[i] - cb.y * bx[i];
      cplx d = cmake(0.0, 0.0);
      for (uint64_t m = sa[i] & g; m; m &= m - 1) d = cadd(d, nc[__builtin_ctzll(m)]);
      d = cadd(d, shift);
      if (k.ext_diag) d = cadd(d, k.ext_diag[b * nA + a]);
      const cplx dx = cmul(d, row[1 + a]);
      y_st(y + b * nA + a, rx + dx.x, ry + dx.y);
    }
  }
}

// Slab kernel with the block's rows of packed T_B staged in LDS (one
// coalesced read) and U gathers of X row segments in flight per step; same
// summation order as kron_slab_packed_kernel.
template <int W, int U>
__global__ void __launch_bounds__(256)
    kron_slab2_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, int chunks,
                      int rows_per_chunk) {
  extern __shared__ uint32_t se[];  // rows_per_chunk * ellB_w entries
  __shared__ cplx vt[256];
  const int xcd = blockIdx.x & 7;
  const int kk = blockIdx.x >> 3;
  const int64_t slab = (int64_t)xcd + 8 * (kk / chunks);
  const int chunk = kk % chunks;
  const int64_t a0 = slab * W;
  const int64_t rb = (int64_t)chunk * rows_per_chunk;
  const int64_t re = min<int64_t>(k.nB, rb + rows_per_chunk);
  if (a0 >= k.nA || rb >= re) return;  // uniform over the block
  const int w = k.ellB_w;
  for (int i = threadIdx.x; i < k.nvtabB; i += blockDim.x) vt[i] = k.vtabB[i];
  const int64_t ne = (re - rb) * w;
  const uint32_t* __restrict__ src = k.ellB + rb * w;
  for (int64_t i = threadIdx.x; i < ne; i += blockDim.x) se[i] = src[i];
  __syncthreads();
  const int lane = threadIdx.x % W;
  const int sub = threadIdx.x / W;
  const int nsub = blockDim.x / W;
  const int64_t a = min<int64_t>(a0 + lane, k.nA - 1);  // clamp: duplicate lanes recompute the last column
  const bool col_ok = a0 + lane < k.nA;
  for (int64_t b = rb + sub; b < re; b += nsub) {
    const uint32_t* er = se + (b - rb) * w;
    int len = w;  // pad entries (value 0) close the row
    while (len > 0 && (er[len - 1] & 0xFFFFu) == 0) --len;
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < len; j += U) {
      uint32_t e[U];
      cplx xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) e[u] = (j + u < len) ? er[j + u] : 0u;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < len) xv[u] = x[(int64_t)(e[u] >> 16) * k.nA + a];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j + u < len) {
          const cplx v = vt[e[u] & 0xFFFFu];
          sx += v.x * xv[u].x + v.y * xv[u].y;
          sy += v.x * xv[u].y - v.y * xv[u].x;
        }
      }
    }
    if (col_ok) {
      const cplx o = y_ld(y + b * k.nA + a);
      y_st(y + b * k.nA + a, o.x + sx, o.y + sy);
    }
  }
}

// Slab kernel as kron_slab2_kernel, with the read of y -- which misses
// every cache -- issued before the row's gathers instead of after them, and
// the next row's y requested while this row is gathered, so neither waits
// for HBM at the end of a row.  Same summation order.
template <int W, int U>
__global__ void __launch_bounds__(256)
    kron_slab3_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, int chunks,
                      int rows_per_chunk) {
  extern __shared__ uint32_t se[];  // rows_per_chunk * ellB_w entries
  __shared__ cplx vt[256];
  const int xcd = blockIdx.x & 7;
  const int kk = blockIdx.x >> 3;
  const int64_t slab = (int64_t)xcd + 8 * (kk / chunks);
  const int chunk = kk % chunks;
  const int64_t a0 = slab * W;
  const int64_t rb = (int64_t)chunk * rows_per_chunk;
  const int64_t re = min<int64_t>(k.nB, rb + rows_per_chunk);
  if (a0 >= k.nA || rb >= re) return;  // uniform over the block
  const int w = k.ellB_w;
  for (int i = threadIdx.x; i < k.nvtabB; i += blockDim.x) vt[i] = k.vtabB[i];
  const int64_t ne = (re - rb) * w;
  const uint32_t* __restrict__ src = k.ellB + rb * w;
  for (int64_t i = threadIdx.x; i < ne; i += blockDim.x) se[i] = src[i];
  __syncthreads();
  const int lane = threadIdx.x % W;
  const int sub = threadIdx.x / W;
  const int nsub = blockDim.x / W;
  const int64_t a = min<int64_t>(a0 + lane, k.nA - 1);  // clamp: duplicate lanes recompute the last column
  const bool col_ok = a0 + lane < k.nA;
  int64_t b = rb + sub;
  cplx o = b < re ? y_ld(y + b * k.nA + a) : cmake(0.0, 0.0);
  for (; b < re; b += nsub) {
    const int64_t bn = b + nsub;
    const cplx on = bn < re ? y_ld(y + bn * k.nA + a) : cmake(0.0, 0.0);  // the next row's y, in flight
    const uint32_t* er = se + (b - rb) * w;
    int len = w;  // pad entries (value 0) close the row
    while (len > 0 && (er[len - 1] & 0xFFFFu) == 0) --len;
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < len; j += U) {
      uint32_t e[U];
      cplx xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) e[u] = (j + u < len) ? er[j + u] : 0u;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < len) xv[u] = x[(int64_t)(e[u] >> 16) * k.nA + a];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j + u < len) {
          const cplx v = vt[e[u] & 0xFFFFu];
          sx += v.x * xv[u].x + v.y * xv[u].y;
          sy += v.x * xv[u].y - v.y * xv[u].x;
        }
      }
    }
    if (col_ok) y_st(y + b * k.nA + a, o.x + sx, o.y + sy);
    o = on;
  }
}

// Slab pass with one wave per (row b, 64-column slab): the entries of T_B's
// row b are the same for every lane, so they sit in scalar registers (one
// scalar load of the whole row) and every entry is one coalesced 1 KB load
// of the x row segment, addressed as the entry's byte offset (scalar) plus
// the lane's column.  Entries go in groups of U, a group at a time (a
// uniform branch per group, none per entry; pad entries of a group point at
// row 0 with coefficient 0).  SIGN: every T_B value is +c or -c, the entry
// code (+1, -1, 0) is a scalar multiplier and conj(c) scales the row sum;
// otherwise the code is a value id (0 pads).  y of the row is requested
// before its gathers.
template <bool SIGN, int U, int NE>
__global__ void __launch_bounds__(256)
    kron_slab4_kernel(KronDesc k, const cplx* __restrict__ x, cplx* __restrict__ y, int chunks,
                      int rows_per_chunk) {
  const int xcd = blockIdx.x & 7;
  const int kk = blockIdx.x >> 3;
  const int64_t slab = (int64_t)xcd + 8 * (kk / chunks);
  const int chunk = kk % chunks;
  const int64_t a0 = slab * 64;
  const int64_t rb = (int64_t)chunk * rows_per_chunk;
  const int64_t re = min<int64_t>(k.nB, rb + rows_per_chunk);
  if (a0 >= k.nA || rb >= re) return;  // uniform over the block
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t a = min<int64_t>(a0 + lane, k.nA - 1);  // clamp: duplicate lanes recompute the last column
  const bool col_ok = a0 + lane < k.nA;
  const char* __restrict__ xa = reinterpret_cast<const char*>(x + a);
  const int stride = k.ellB4_stride;  // NE + 1
  const cplx bc = k.bcoef;
  for (int64_t b = rb + wv; b < re; b += 4) {
    const uint32_t* __restrict__ er = k.ellB4 + b * stride;
    uint32_t ent[NE + 1];
#pragma unroll
    for (int j = 0; j <= NE; ++j) ent[j] = er[j];
    const int ne = (int)ent[0];
    const cplx o = y_ld(y + b * k.nA + a);
    double px = 0.0, py = 0.0;
#pragma unroll
    for (int g = 0; g < NE; g += U) {
      if (g < ne) {
        kdbl2 xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = *reinterpret_cast<const kdbl2*>(xa + (ent[1 + g + u] & ~15u));
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t code = ent[1 + g + u] & 15u;
          if (SIGN) {
            const double sg = code == 1u ? 1.0 : (code == 2u ? -1.0 : 0.0);
            px = __builtin_fma(sg, xv[u].x, px);
            py = __builtin_fma(sg, xv[u].y, py);
          } else {
            const cplx v = k.vtabB4[code];
            px = __builtin_fma(v.x, xv[u].x, px);  // conj(v) * x
            px = __builtin_fma(v.y, xv[u].y, px);
            py = __builtin_fma(v.x, xv[u].y, py);
            py = __builtin_fma(-v.y, xv[u].x, py);
          }
        }
      }
    }
    double rx = px, ry = py;
    if (SIGN) {
      rx = bc.x * px + bc.y * py;  // conj(c) * sum
      ry = bc.x * py - bc.y * px;
    }
    if (col_ok) y_st(y + b * k.nA + a, o.x + rx, o.y + ry);
  }
}

// ------------------------------------------------------------------ host
// pack a pull-form block matrix: entries (col << 16 | value id); row-major
// when rowmajor, else entry-major (ELL column j contiguous over rows)
static bool pack_block(qd_ctx* ctx, const qd_csr* c, int64_t n, bool rowmajor, int max_tab, uint32_t** d_ell,
                       cplx** d_tab, int* width, int* ntab) {
  if (n >= 65536) return false;
  int64_t nnz = 0;
  qd_csr_info(c, nullptr, &nnz, nullptr);
  std::vector<int32_t> ao(n + 1), ai(nnz);
  std::vector<cplx> av(nnz);
  if (qd_csr_download(const_cast<qd_csr*>(c), ao.data(), ai.data(), (double*)av.data()) != QD_OK) return false;
  std::vector<cplx> tab(1, cmake(0.0, 0.0));
  std::vector<uint32_t> vid(nnz);
  for (int64_t e = 0; e < nnz; ++e) {
    size_t q = 0;
    for (; q < tab.size(); ++q)
      if (std::memcmp(&tab[q], &av[e], sizeof(cplx)) == 0) break;
    if (q == tab.size()) {
      if ((int)tab.size() >= max_tab) return false;
      tab.push_back(av[e]);
    }
    vid[e] = (uint32_t)q;
  }
  int w = 1;
  for (int64_t r = 0; r < n; ++r) w = std::max(w, ao[r + 1] - ao[r]);
  std::vector<uint32_t> ell((size_t)w * n);
  for (int64_t r = 0; r < n; ++r)
    for (int j = 0; j < w; ++j) {
      int32_t e = ao[r] + j;
      uint32_t v = (e < ao[r + 1]) ? (((uint32_t)ai[e] << 16) | vid[e]) : ((uint32_t)r << 16);
      ell[rowmajor ? (size_t)r * w + j : (size_t)j * n + r] = v;
    }
  if (hipMalloc(d_ell, sizeof(uint32_t) * ell.size()) != hipSuccess) return false;
  if (hipMalloc(d_tab, sizeof(cplx) * tab.size()) != hipSuccess) {
    hipFree(*d_ell);
    *d_ell = nullptr;
    return false;
  }
  hipMemcpyAsync(*d_ell, ell.data(), sizeof(uint32_t) * ell.size(), hipMemcpyHostToDevice, ctx->stream);
  hipMemcpyAsync(*d_tab, tab.data(), sizeof(cplx) * tab.size(), hipMemcpyHostToDevice, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return false;
  *width = w;
  *ntab = (int)tab.size();
  return true;
}

// pack a pull-form block matrix for kron_slab4_kernel (rows of uniform
// entries): per row the entry count, then the byte offsets of the source
// rows of x (row length `rowbytes`) | a code: +1 / -1 (1, 2) in the sign
// form, a value id (1..15) otherwise; pads (code 0, offset 0) fill the row to
// a multiple of 8 entries
static bool pack_block_rows(qd_ctx* ctx, const qd_csr* c, int64_t n, uint64_t rowbytes, KronDesc* d,
                            uint32_t** d_ell, cplx** d_tab) {
  if (n >= 65536 || rowbytes % 16 != 0 || rowbytes * (uint64_t)n >= (1ull << 32)) return false;
  int64_t nnz = 0;
  qd_csr_info(c, nullptr, &nnz, nullptr);
  std::vector<int32_t> ao(n + 1), ai(nnz);
  std::vector<cplx> av(nnz);
  if (qd_csr_download(const_cast<qd_csr*>(c), ao.data(), ai.data(), (double*)av.data()) != QD_OK) return false;
  int w = 0;
  for (int64_t r = 0; r < n; ++r) w = std::max(w, ao[r + 1] - ao[r]);
  if (w > 32) return false;
  bool sign = nnz > 0;
  const cplx c0 = nnz > 0 ? av[0] : cmake(0.0, 0.0);
  for (int64_t e = 0; e < nnz && sign; ++e)
    sign = (av[e].x == c0.x && av[e].y == c0.y) || (av[e].x == -c0.x && av[e].y == -c0.y);
  std::vector<cplx> tab(1, cmake(0.0, 0.0));  // id 0 pads
  std::vector<uint32_t> code(nnz, 0u);
  for (int64_t e = 0; e < nnz; ++e) {
    if (sign) {
      code[e] = (av[e].x == c0.x && av[e].y == c0.y) ? 1u : 2u;
      continue;
    }
    size_t q = 1;
    for (; q < tab.size(); ++q)
      if (std::memcmp(&tab[q], &av[e], sizeof(cplx)) == 0) break;
    if (q == tab.size()) {
      if (tab.size() >= 16) return false;
      tab.push_back(av[e]);
    }
    code[e] = (uint32_t)q;
  }
  const int stride = 1 + std::max(8, (w + 7) / 8 * 8);
  std::vector<uint32_t> ell((size_t)stride * n, 0u);
  for (int64_t r = 0; r < n; ++r) {
    uint32_t* row = ell.data() + (size_t)r * stride;
    row[0] = (uint32_t)(ao[r + 1] - ao[r]);
    for (int32_t e = ao[r]; e < ao[r + 1]; ++e)
      row[1 + (e - ao[r])] = (uint32_t)((uint64_t)ai[e] * rowbytes) | code[e];
  }
  if (hipMalloc(d_ell, sizeof(uint32_t) * ell.size()) != hipSuccess) return false;
  if (hipMalloc(d_tab, sizeof(cplx) * tab.size()) != hipSuccess) {
    hipFree(*d_ell);
    *d_ell = nullptr;
    return false;
  }
  hipMemcpyAsync(*d_ell, ell.data(), sizeof(uint32_t) * ell.size(), hipMemcpyHostToDevice, ctx->stream);
  hipMemcpyAsync(*d_tab, tab.data(), sizeof(cplx) * tab.size(), hipMemcpyHostToDevice, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return false;
  d->ellB4 = *d_ell;
  d->ellB4_stride = stride;
  d->bsign = sign ? 1 : 0;
  d->bcoef = c0;
  d->vtabB4 = *d_tab;
  return true;
}

static bool is_diag_op(const DevOp& o) {
  if (o.kind == QD_OP_HOP) return o.a == o.b;
  return o.U[1].x == 0.0 && o.U[1].y == 0.0 && o.U[2].x == 0.0 && o.U[2].y == 0.0;
}

// Try to build the Kronecker form; returns nullptr (no error) if the program
// or sector does not have it.
qd_kron* qd_kron_try_build(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog) {
  const SectorDesc& sec = basis->sec;
  if (!sec.cns || sec.nblocks != 2 || sec.cnt[0] < 0 || sec.cnt[1] < 0) return nullptr;
  const int loA = sec.lo[0], lenA = sec.len[0], loB = sec.lo[1], lenB = sec.len[1];
  if (loA != 0 || loB != lenA || loB + lenB != sec.nbits) return nullptr;
  std::vector<DevTerm> tA, tB, tD;
  std::vector<DevOp> oA, oB, oD;
  auto inA = [&](int s) { return s >= loA && s < loA + lenA; };
  auto inB = [&](int s) { return s >= loB && s < loB + lenB; };
  for (const DevTerm& t : prog->h_terms) {
    bool diag = true, allA = true, allB = true, hopsOnly = true;
    for (int q = 0; q < t.n_ops; ++q) {
      const DevOp& o = prog->h_ops[t.op_begin + q];
      diag = diag && is_diag_op(o);
      allA = allA && inA(o.a) && inA(o.b);
      allB = allB && inB(o.a) && inB(o.b);
      hopsOnly = hopsOnly && (o.kind == QD_OP_HOP);
    }
    std::vector<DevTerm>* dt;
    std::vector<DevOp>* dopv;
    int shift = 0;
    if (diag) {
      dt = &tD;
      dopv = &oD;
    } else if (allA && hopsOnly) {
      dt = &tA;
      dopv = &oA;
    } else if (allB) {
      dt = &tB;
      dopv = &oB;
      shift = loB;
    } else {
      return nullptr;
    }
    DevTerm c = t;
    c.op_begin = (int)dopv->size();
    for (int q = 0; q < t.n_ops; ++q) {
      DevOp o = prog->h_ops[t.op_begin + q];
      o.a -= shift;
      o.b -= shift;
      dopv->push_back(o);
    }
    dt->push_back(c);
  }
  // sub-programs through the public constructors (validation included)
  auto make_prog = [&](int nbits, const std::vector<DevTerm>& ts, const std::vector<DevOp>& os) -> qd_termprog* {
    std::vector<qd_term> qt(ts.size());
    std::vector<qd_op> qo(os.size());
    for (size_t i = 0; i < ts.size(); ++i) {
      qt[i].coeff[0] = ts[i].coeff.x;
      qt[i].coeff[1] = ts[i].coeff.y;
      qt[i].op_begin = ts[i].op_begin;
      qt[i].n_ops = ts[i].n_ops;
    }
    for (size_t i = 0; i < os.size(); ++i) {
      qo[i].kind = os[i].kind;
      qo[i].site_a = os[i].a;
      qo[i].site_b = os[i].b;
      qo[i].string = os[i].string;
      for (int u = 0; u < 4; ++u) {
        qo[i].U[2 * u] = os[i].U[u].x;
        qo[i].U[2 * u + 1] = os[i].U[u].y;
      }
    }
    qd_termprog* p = nullptr;
    if (qd_termprog_create(ctx, nbits, qt.data(), (int)qt.size(), qo.data(), (int)qo.size(), &p) != QD_OK) return nullptr;
    return p;
  };
  qd_kron* k = new qd_kron();
  auto fail = [&]() -> qd_kron* {
    if (k->csrA) qd_csr_destroy(k->csrA);
    if (k->csrB) qd_csr_destroy(k->csrB);
    if (k->progA) qd_termprog_destroy(k->progA);
    if (k->progB) qd_termprog_destroy(k->progB);
    if (k->basisA) qd_basis_destroy(k->basisA);
    if (k->basisB) qd_basis_destroy(k->basisB);
    hipFree(k->d_dterms);
    hipFree(k->d_dops);
    delete k;
    return nullptr;
  };
  qd_constraint cA{(lenA >= 64) ? ~0ull : ((1ull << lenA) - 1ull), sec.cnt[0]};
  qd_constraint cB{(lenB >= 64) ? ~0ull : ((1ull << lenB) - 1ull), sec.cnt[1]};
  if (qd_basis_create(ctx, lenA, &cA, 1, &k->basisA) != QD_OK || qd_basis_create(ctx, lenB, &cB, 1, &k->basisB) != QD_OK)
    return fail();
  // an empty block program is a zero matrix: give it one zero-coefficient number term
  DevOp zop{};
  zop.kind = QD_OP_HOP;
  zop.string = 1;
  if (tA.empty()) {
    tA.push_back(DevTerm{cmake(0.0, 0.0), 0, 1, 0, 1});
    oA.push_back(zop);
  }
  if (tB.empty()) {
    tB.push_back(DevTerm{cmake(0.0, 0.0), 0, 1, 0, 1});
    oB.push_back(zop);
  }
  k->progA = make_prog(lenA, tA, oA);
  k->progB = make_prog(lenB, tB, oB);
  if (!k->progA || !k->progB) return fail();
  if (qd_csr_build(ctx, k->basisA, k->progA, &k->csrA) != QD_OK || qd_csr_build(ctx, k->basisB, k->progB, &k->csrB) != QD_OK)
    return fail();
  // pull form needs rows of each block operator: for a non-Hermitian block use
  // the CSC of its adjoint program
  const qd_csr* useA = k->csrA;
  const qd_csr* useB = k->csrB;
  if (k->csrA->hermitian != 1) {
    if (qd_csr_build_impl(ctx, k->basisA, k->progA, 1, &k->csrA->adj) != QD_OK) return fail();
    useA = k->csrA->adj;
  }
  if (k->csrB->hermitian != 1) {
    if (qd_csr_build_impl(ctx, k->basisB, k->progB, 1, &k->csrB->adj) != QD_OK) return fail();
    useB = k->csrB->adj;
  }
  // number-operator products -> (mask, coefficient)
  std::vector<uint64_t> nmask;
  std::vector<cplx> ncoef;
  {
    std::vector<DevTerm> rest;
    std::vector<DevOp> rest_ops;
    for (const DevTerm& t : tD) {
      bool num = true;
      uint64_t m = 0;
      for (int q = 0; q < t.n_ops; ++q) {
        const DevOp& o = oD[t.op_begin + q];
        if (o.kind != QD_OP_HOP || o.a != o.b) num = false;
        else m |= (1ull << o.a);
      }
      if (num) {
        nmask.push_back(m);
        ncoef.push_back(t.coeff);
      } else {
        DevTerm c = t;
        c.op_begin = (int)rest_ops.size();
        for (int q = 0; q < t.n_ops; ++q) rest_ops.push_back(oD[t.op_begin + q]);
        rest.push_back(c);
      }
    }
    tD.swap(rest);
    oD.swap(rest_ops);
  }
  if (hipMalloc(&k->d_nmask, sizeof(uint64_t) * std::max<size_t>(1, nmask.size())) != hipSuccess ||
      hipMalloc(&k->d_ncoef, sizeof(cplx) * std::max<size_t>(1, ncoef.size())) != hipSuccess)
    return fail();
  if (!nmask.empty()) {
    hipMemcpyAsync(k->d_nmask, nmask.data(), sizeof(uint64_t) * nmask.size(), hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(k->d_ncoef, ncoef.data(), sizeof(cplx) * ncoef.size(), hipMemcpyHostToDevice, ctx->stream);
  }
  k->ndterms = (int)tD.size();
  if (hipMalloc(&k->d_dterms, sizeof(DevTerm) * std::max<size_t>(1, tD.size())) != hipSuccess ||
      hipMalloc(&k->d_dops, sizeof(DevOp) * std::max<size_t>(1, oD.size())) != hipSuccess)
    return fail();
  if (!tD.empty()) {
    // recompute leaf counts for the diagonal sub-program layout
    for (auto& t : tD) {
      int nl = 1;
      for (int q = 0; q < t.n_ops; ++q) nl *= (oD[t.op_begin + q].kind == QD_OP_BIT) ? 2 : 1;
      t.n_leaves = nl;
    }
    hipMemcpyAsync(k->d_dterms, tD.data(), sizeof(DevTerm) * tD.size(), hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(k->d_dops, oD.data(), sizeof(DevOp) * oD.size(), hipMemcpyHostToDevice, ctx->stream);
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail();
  KronDesc& d = k->desc;
  d.nA = (int64_t)k->basisA->sec.total;
  d.nB = (int64_t)k->basisB->sec.total;
  d.loB = loB;
  d.Aouter = useA->d_outer;
  d.Ainner = useA->d_inner;
  d.Avals = useA->d_vals;
  d.Bouter = useB->d_outer;
  d.Binner = useB->d_inner;
  d.Bvals = useB->d_vals;
  d.statesA = k->basisA->d_states;
  d.statesB = k->basisB->d_states;
  d.dterms = k->d_dterms;
  d.dops = k->d_dops;
  d.ndterms = k->ndterms;
  d.nmask = k->d_nmask;
  d.ncoef = k->d_ncoef;
  d.nnum = (int)nmask.size();
  // packed T_A when columns fit 16 bits and the block has < 65536 distinct values
  d.ell = nullptr;
  if (d.nA < 65536) {
    int64_t nnz = 0;
    qd_csr_info(useA, nullptr, &nnz, nullptr);
    std::vector<int32_t> ao(d.nA + 1), ai(nnz);
    std::vector<cplx> av(nnz);
    if (qd_csr_download(const_cast<qd_csr*>(useA), ao.data(), ai.data(), (double*)av.data()) == QD_OK) {
      std::vector<cplx> tab(1, cmake(0.0, 0.0));
      std::vector<uint32_t> vid(nnz);
      bool ok = true;
      for (int64_t e = 0; e < nnz && ok; ++e) {
        size_t q = 0;
        for (; q < tab.size(); ++q)
          if (std::memcmp(&tab[q], &av[e], sizeof(cplx)) == 0) break;
        if (q == tab.size()) {
          if (tab.size() >= 4096) ok = false;
          tab.push_back(av[e]);
        }
        vid[e] = (uint32_t)q;
      }
      int w = 0;
      for (int64_t a = 0; a < d.nA; ++a) w = std::max(w, ao[a + 1] - ao[a]);
      if (ok) {
        std::vector<uint32_t> ell((size_t)w * d.nA);
        for (int64_t a = 0; a < d.nA; ++a)
          for (int j = 0; j < w; ++j) {
            int32_t e = ao[a] + j;
            ell[(size_t)j * d.nA + a] = (e < ao[a + 1]) ? (((uint32_t)ai[e] << 16) | vid[e]) : ((uint32_t)a << 16);
          }
        if (hipMalloc(&k->d_ell, sizeof(uint32_t) * std::max<size_t>(1, ell.size())) == hipSuccess &&
            hipMalloc(&k->d_vtab, sizeof(cplx) * tab.size()) == hipSuccess) {
          hipMemcpyAsync(k->d_ell, ell.data(), sizeof(uint32_t) * ell.size(), hipMemcpyHostToDevice, ctx->stream);
          hipMemcpyAsync(k->d_vtab, tab.data(), sizeof(cplx) * tab.size(), hipMemcpyHostToDevice, ctx->stream);
          if (hipStreamSynchronize(ctx->stream) == hipSuccess) {
            d.ell = k->d_ell;
            d.ell_w = w;
            d.vtab = k->d_vtab;
            d.nvtab = (int)tab.size();
          }
        }
        // the register-resident form of the persistent rows kernel
        d.ell16 = nullptr;
        if (d.ell && d.nA < 8192 && tab.size() <= 8 && w <= 2 * KRON_WP) {
          std::vector<uint32_t> e16((size_t)KRON_WP * d.nA, 0u);
          for (int64_t a = 0; a < d.nA; ++a)
            for (int j = 0; j < w; ++j) {
              const int32_t e = ao[a] + j;
              if (e >= ao[a + 1]) break;
              const uint32_t v = ((uint32_t)ai[e] << 3) | vid[e];
              e16[(size_t)(j / 2) * d.nA + a] |= v << (16 * (j & 1));
            }
          if (hipMalloc(&k->d_ell16, sizeof(uint32_t) * e16.size()) == hipSuccess) {
            hipMemcpyAsync(k->d_ell16, e16.data(), sizeof(uint32_t) * e16.size(), hipMemcpyHostToDevice, ctx->stream);
            if (hipStreamSynchronize(ctx->stream) == hipSuccess) d.ell16 = k->d_ell16;
          }
        }
        // the branch-free form: columns shifted by one (0 pads), and the sign
        // form when every nonzero table value is +c or -c
        d.ell16s = nullptr;
        d.asign = 0;
        d.acoef = cmake(0.0, 0.0);
        if (d.ell16 && d.nA + 1 < 8192) {
          bool sign = tab.size() >= 2;
          const cplx c = sign ? tab[1] : cmake(0.0, 0.0);
          std::vector<uint32_t> low(tab.size(), 0u);
          for (size_t q = 1; q < tab.size() && sign; ++q) {
            if (tab[q].x == c.x && tab[q].y == c.y)
              low[q] = 0u;
            else if (tab[q].x == -c.x && tab[q].y == -c.y)
              low[q] = 4u;
            else
              sign = false;
          }
          std::vector<uint32_t> e16s((size_t)KRON_WP * d.nA, 0u);
          for (int64_t a = 0; a < d.nA; ++a)
            for (int j = 0; j < w; ++j) {
              const int32_t e = ao[a] + j;
              if (e >= ao[a + 1]) break;
              const uint32_t v = (((uint32_t)ai[e] + 1u) << 3) | (sign ? low[vid[e]] : vid[e]);
              e16s[(size_t)(j / 2) * d.nA + a] |= v << (16 * (j & 1));
            }
          if (hipMalloc(&k->d_ell16s, sizeof(uint32_t) * e16s.size()) == hipSuccess) {
            hipMemcpyAsync(k->d_ell16s, e16s.data(), sizeof(uint32_t) * e16s.size(), hipMemcpyHostToDevice,
                           ctx->stream);
            if (hipStreamSynchronize(ctx->stream) == hipSuccess) {
              d.ell16s = k->d_ell16s;
              d.asign = sign ? 1 : 0;
              d.acoef = c;
            }
          }
        }
      }
    }
  }
  d.ellB = nullptr;
  if (pack_block(ctx, useB, d.nB, true, 256, &k->d_ellB, &k->d_vtabB, &d.ellB_w, &d.nvtabB)) {
    d.ellB = k->d_ellB;
    d.vtabB = k->d_vtabB;
  }
  d.ellB4 = nullptr;
  pack_block_rows(ctx, useB, d.nB, (uint64_t)d.nA * sizeof(cplx), &d, &k->d_ellB4, &k->d_vtabB4);
  return k;
}

const KronDesc* qd_kron_desc(const qd_kron* k) { return k ? &k->desc : nullptr; }


void qd_kron_destroy(qd_kron* k) {
  if (!k) return;
  hipFree(k->d_ell);
  hipFree(k->d_ell16);
  hipFree(k->d_ell16s);
  hipFree(k->d_vtab);
  hipFree(k->d_nmask);
  hipFree(k->d_ncoef);
  hipFree(k->d_ellB);
  hipFree(k->d_vtabB);
  hipFree(k->d_ellB4);
  hipFree(k->d_vtabB4);
  qd_csr_destroy(k->csrA);
  qd_csr_destroy(k->csrB);
  qd_termprog_destroy(k->progA);
  qd_termprog_destroy(k->progB);
  qd_basis_destroy(k->basisA);
  qd_basis_destroy(k->basisB);
  hipFree(k->d_dterms);
  hipFree(k->d_dops);
  delete k;
}

// launch shapes of the two passes
namespace {
constexpr int KRON_SLAB_W = 32;

struct SlabShape {
  int chunks, rows_per_chunk;
  unsigned grid;
};

// One column slab is split into `chunks` blocks of rows.  With target 128 an
// XCD that holds more than 128 blocks at once works on two or three slabs
// together; `target` = the blocks one XCD holds keeps one slab (W columns of
// x over all nB rows) live per XCD, so it stays in that XCD's L2.
SlabShape slab_shape(const KronDesc& d, int W, int target = 128) {
  SlabShape s;
  const int64_t nslabs = (d.nA + W - 1) / W;
  const int64_t c = std::min<int64_t>(target, std::max<int64_t>(1, d.nB / 8));
  s.rows_per_chunk = (int)((d.nB + c - 1) / c);
  s.chunks = (int)((d.nB + s.rows_per_chunk - 1) / s.rows_per_chunk);  // no empty chunks
  const int64_t slab_groups = (nslabs + 7) / 8;  // slabs per XCD
  s.grid = (unsigned)(slab_groups * s.chunks * 8);
  return s;
}

// blocks of a 256-thread kernel with `lds` dynamic LDS that one XCD holds at
// once; the occupancy query runs once per (kernel, LDS size) and process
int xcd_resident(const void* fn, size_t lds, int ncu, int threads = 256) {
  static std::mutex mu;
  static std::map<std::pair<const void*, size_t>, int> memo;
  int occ = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = memo.find({fn, lds});
    if (it != memo.end()) occ = it->second;
  }
  if (occ == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, lds) != hipSuccess || occ < 1) occ = 1;
    std::lock_guard<std::mutex> lk(mu);
    memo[{fn, lds}] = occ;
  }
  return occ * std::max(1, ncu / 8);
}

int kron_optin(int device) {
  // opt-in LDS limit, and the kernels' attributes raised to it, once per
  // process (thread-safe static initialisation; one device per process)
  static const int optin = [device]() {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, device) != hipSuccess) v = 65536;
    if (v > 65536) {
      hipFuncSetAttribute((const void*)kron_rows_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
      hipFuncSetAttribute((const void*)kron_rows_packed_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
      hipFuncSetAttribute((const void*)kron_rows2_kernel<1, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
      hipFuncSetAttribute((const void*)kron_rows2_kernel<2, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
      hipFuncSetAttribute((const void*)kron_rows2_kernel<4, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    }
    return v;
  }();
  return optin;
}

// persistent rows kernel: SLOTS = ceil(nA / 1024) <= 5, WP words of
// compressed entries per output, two row buffers in LDS; grid = the blocks
// the device holds at once (at most one per row)
template <int WP, int S, bool GD>
bool launch_rows3(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  const size_t lds = sizeof(cplx) * (136 + 2 * (size_t)d.nA);
  if (lds > (size_t)optin) return false;
  const void* fn = (const void*)kron_rows3_kernel<WP, S, GD>;
  static std::once_flag once;
  std::call_once(once, [fn, optin]() { hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, optin); });
  const int per_xcd = xcd_resident(fn, lds, ncu, 1024);
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(d.nB, (int64_t)per_xcd * 8));
  const int rpb = (int)((d.nB + grid - 1) / grid);
  const unsigned g = (unsigned)((d.nB + rpb - 1) / rpb);
  kron_rows3_kernel<WP, S, GD><<<g, 1024, lds, st>>>(d, x, y, shift, rpb);
  return true;
}

template <int WP, int S>
bool launch_rows3_d(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  // the number products go through per-output bit sets when there are at
  // most 64 and no other diagonal terms
  return (d.ndterms > 0 || d.nnum > 64) ? launch_rows3<WP, S, true>(d, optin, ncu, x, y, shift, st)
                                        : launch_rows3<WP, S, false>(d, optin, ncu, x, y, shift, st);
}

template <int S>
bool launch_rows3_w(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  const int wp = (d.ell_w + 1) / 2;  // words holding the widest row
  if (wp <= 8) return launch_rows3_d<8, S>(d, optin, ncu, x, y, shift, st);
  if (wp <= 10) return launch_rows3_d<10, S>(d, optin, ncu, x, y, shift, st);
  return launch_rows3_d<12, S>(d, optin, ncu, x, y, shift, st);
}

// branch-free persistent rows kernel (kron_rows4_kernel), G entries per
// LDS wait (0: compiler order)
template <int WP, int S, bool SG, bool GD, int G, bool PI, bool DM>
bool launch_rows4(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  const size_t lds = sizeof(cplx) * (56 + 2 * (size_t)(d.nA + 1));
  if (lds > (size_t)optin) return false;
  const void* fn = (const void*)kron_rows4_kernel<WP, S, SG, GD, G, PI, DM>;
  static std::once_flag once;
  std::call_once(once, [fn, optin]() { hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, optin); });
  const int per_xcd = xcd_resident(fn, lds, ncu, 1024);
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(d.nB, (int64_t)per_xcd * 8));
  const int rpb = (int)((d.nB + grid - 1) / grid);
  const unsigned g = (unsigned)((d.nB + rpb - 1) / rpb);
  kron_rows4_kernel<WP, S, SG, GD, G, PI, DM><<<g, 1024, lds, st>>>(d, x, y, shift, rpb);
  return true;
}

template <int WP, int S, int G, bool PI, bool DM>
bool launch_rows4_d(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  const bool gd = d.ndterms > 0 || d.nnum > 32;
  if (d.asign) return gd ? launch_rows4<WP, S, true, true, G, PI, DM>(d, optin, ncu, x, y, shift, st)
                         : launch_rows4<WP, S, true, false, G, PI, DM>(d, optin, ncu, x, y, shift, st);
  return gd ? launch_rows4<WP, S, false, true, G, PI, DM>(d, optin, ncu, x, y, shift, st)
            : launch_rows4<WP, S, false, false, G, PI, DM>(d, optin, ncu, x, y, shift, st);
}

template <int S, int G, bool PI, bool DM>
bool launch_rows4_w(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  const int wp = (d.ell_w + 1) / 2;
  if (wp <= 8) return launch_rows4_d<8, S, G, PI, DM>(d, optin, ncu, x, y, shift, st);
  if (wp <= 10) return launch_rows4_d<10, S, G, PI, DM>(d, optin, ncu, x, y, shift, st);
  return launch_rows4_d<12, S, G, PI, DM>(d, optin, ncu, x, y, shift, st);
}

template <int G, bool PI, bool DM>
bool launch_rows4_s(const KronDesc& d, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  if (!d.ell16s) return false;
  switch ((d.nA + 1023) / 1024) {
    case 1: return launch_rows4_w<1, G, PI, DM>(d, optin, ncu, x, y, shift, st);
    case 2: return launch_rows4_w<2, G, PI, DM>(d, optin, ncu, x, y, shift, st);
    case 3: return launch_rows4_w<3, G, PI, DM>(d, optin, ncu, x, y, shift, st);
    case 4: return launch_rows4_w<4, G, PI, DM>(d, optin, ncu, x, y, shift, st);
    case 5: return launch_rows4_w<5, G, PI, DM>(d, optin, ncu, x, y, shift, st);
    default: return false;
  }
}

// the fused one-pass product (kron_fused_kernel), or false when the form
// does not qualify (block sizes, sign forms, diagonal terms)
template <int WP, int S>
bool launch_fused_k(const KronDesc& d, int optin, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  const size_t lds = sizeof(cplx) * (96 + 1 + (size_t)d.nA);
  if (lds > (size_t)optin) return false;
  kron_fused_kernel<WP, S><<<(unsigned)d.nB, 256, lds, st>>>(d, x, y, shift);
  return true;
}

bool launch_fused(const KronDesc& d, int optin, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  if (!d.ell16s || !d.asign || !d.ellB4 || !d.bsign || d.ndterms > 0 || d.nnum > 64 || d.nA > 1024) return false;
  if ((double)d.nA * (double)d.nB * sizeof(cplx) > 128e6) return false;  // x well within the MALL
  const int wp = (d.ell_w + 1) / 2;
#define QD_FUSED_S(S)                                                            \
  if ((d.nA + 255) / 256 == S) {                                                 \
    if (wp <= 8) return launch_fused_k<8, S>(d, optin, x, y, shift, st);         \
    if (wp <= 10) return launch_fused_k<10, S>(d, optin, x, y, shift, st);       \
    return launch_fused_k<12, S>(d, optin, x, y, shift, st);                     \
  }
  QD_FUSED_S(1)
  QD_FUSED_S(2)
  QD_FUSED_S(3)
  QD_FUSED_S(4)
#undef QD_FUSED_S
  return false;
}
constexpr int KRON_FUSED = 15;  // variant word 15 * 16 + 15: the fused product

// rows pass (writes y): variant 0 = kron_rows_packed<2>, 1 = rows2<2,8>,
// 2 = rows2<1,8>, 3 = rows2<4,4>, 4 = the persistent rows3 (T_A in
// registers; needs the compressed table), 5-8 = the branch-free rows4 (needs
// its table): 5 groups of 4 entries, rows staged through registers; 6 and 7
// groups of 4 and 6, pipelined, rows copied straight to LDS; 8 groups of 8,
// rows copied to LDS; a variant whose LDS does not fit returns false.
// Without the packed table: kron_rows_kernel.
constexpr int KRON_ROWS_VARIANTS = 9;
bool launch_rows(const KronDesc& d, int v, int optin, int ncu, const cplx* x, cplx* y, cplx shift, hipStream_t st) {
  if (v < 0 || v >= KRON_ROWS_VARIANTS) return false;
  if (v == 5) return launch_rows4_s<4, false, false>(d, optin, ncu, x, y, shift, st);
  if (v == 6) return launch_rows4_s<4, true, true>(d, optin, ncu, x, y, shift, st);
  if (v == 7) return launch_rows4_s<6, true, true>(d, optin, ncu, x, y, shift, st);
  if (v == 8) return launch_rows4_s<8, false, true>(d, optin, ncu, x, y, shift, st);
  if (v == 4) {
    if (!d.ell16) return false;
    switch ((d.nA + 1023) / 1024) {  // row loads per thread when staging
      case 1: return launch_rows3_w<1>(d, optin, ncu, x, y, shift, st);
      case 2: return launch_rows3_w<2>(d, optin, ncu, x, y, shift, st);
      case 3: return launch_rows3_w<3>(d, optin, ncu, x, y, shift, st);
      case 4: return launch_rows3_w<4>(d, optin, ncu, x, y, shift, st);
      case 5: return launch_rows3_w<5>(d, optin, ncu, x, y, shift, st);
      default: return false;  // more slots than the registers of a 1024-thread block hold
    }
  }
  const size_t lds1 = sizeof(cplx) * (size_t)d.nA;
  if (!d.ell) {
    if (v != 0) return false;
    if (lds1 <= (size_t)optin)
      kron_rows_kernel<true><<<(unsigned)d.nB, 1024, lds1, st>>>(d, x, y, shift);
    else
      kron_rows_kernel<false><<<(unsigned)d.nB, 1024, 0, st>>>(d, x, y, shift);
    return true;
  }
  const int rb = v == 2 ? 1 : v == 3 ? 4 : 2;
  const size_t lds = sizeof(cplx) * ((size_t)d.nvtab + (size_t)rb * (size_t)d.nA);
  if (lds > (size_t)optin) {
    if (v != 0) return false;
    if (lds1 <= (size_t)optin)
      kron_rows_kernel<true><<<(unsigned)d.nB, 1024, lds1, st>>>(d, x, y, shift);
    else
      kron_rows_kernel<false><<<(unsigned)d.nB, 1024, 0, st>>>(d, x, y, shift);
    return true;
  }
  const unsigned grid = (unsigned)((d.nB + rb - 1) / rb);
  switch (v) {
    case 0: kron_rows_packed_kernel<2><<<grid, 1024, lds, st>>>(d, x, y, shift); break;
    case 1: kron_rows2_kernel<2, 8><<<grid, 1024, lds, st>>>(d, x, y, shift); break;
    case 2: kron_rows2_kernel<1, 8><<<grid, 1024, lds, st>>>(d, x, y, shift); break;
    default: kron_rows2_kernel<4, 4><<<grid, 1024, lds, st>>>(d, x, y, shift); break;
  }
  return true;
}

// wave-per-row slab kernel: 64-column slabs, one slab per XCD at a time
template <int U, int NE>
bool launch_slab4_n(const KronDesc& d, int ncu, const cplx* x, cplx* y, hipStream_t st) {
  const void* fn =
      d.bsign ? (const void*)kron_slab4_kernel<true, U, NE> : (const void*)kron_slab4_kernel<false, U, NE>;
  const SlabShape sh = slab_shape(d, 64, xcd_resident(fn, 0, ncu));
  if (d.bsign)
    kron_slab4_kernel<true, U, NE><<<sh.grid, 256, 0, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk);
  else
    kron_slab4_kernel<false, U, NE><<<sh.grid, 256, 0, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk);
  return true;
}

template <int U>
bool launch_slab4(const KronDesc& d, int ncu, const cplx* x, cplx* y, hipStream_t st) {
  if (!d.ellB4) return false;
  const int ne = d.ellB4_stride - 1;  // entry slots per row (a multiple of 8)
  if (ne <= 8) return launch_slab4_n<U, 8>(d, ncu, x, y, st);
  if (ne <= 16) return launch_slab4_n<U, 16>(d, ncu, x, y, st);
  if (ne <= 24) return launch_slab4_n<U, 24>(d, ncu, x, y, st);
  if (ne <= 32) return launch_slab4_n<U, 32>(d, ncu, x, y, st);
  return false;
}

// slab pass (adds into y): variant 0 = kron_slab_packed<32,4>, 1 =
// slab2<32,8>, 2 = slab2<32,16>, each at 128 chunks per slab; 3 = slab2<32,8>,
// 4 = slab2<32,16>, 5 = slab3<32,8> and 6 = slab3<32,4> with one slab per
// XCD at a time; 7 = slab4<8> and 8 = slab4<4> (one wave per row of a
// 64-column slab; needs its table).  Without the packed table:
// kron_slab_kernel.
constexpr int KRON_SLAB_VARIANTS = 9;

bool launch_slab(const KronDesc& d, int v, int ncu, const cplx* x, cplx* y, hipStream_t st) {
  if (v == 7) return launch_slab4<8>(d, ncu, x, y, st);
  if (v == 8) return launch_slab4<4>(d, ncu, x, y, st);
  if (!d.ellB) {
    if (v != 0) return false;
    const SlabShape sh = slab_shape(d, KRON_SLAB_W);
    kron_slab_kernel<KRON_SLAB_W><<<sh.grid, 256, 0, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk);
    return true;
  }
  if (v < 0 || v >= KRON_SLAB_VARIANTS) return false;
  const void* fn = v == 0                ? (const void*)kron_slab_packed_kernel<KRON_SLAB_W, 4>
                   : (v == 1 || v == 3) ? (const void*)kron_slab2_kernel<KRON_SLAB_W, 8>
                   : v == 5             ? (const void*)kron_slab3_kernel<KRON_SLAB_W, 8>
                   : v == 6             ? (const void*)kron_slab3_kernel<KRON_SLAB_W, 4>
                                        : (const void*)kron_slab2_kernel<KRON_SLAB_W, 16>;
  SlabShape sh = slab_shape(d, KRON_SLAB_W);
  size_t lds = sizeof(uint32_t) * (size_t)sh.rows_per_chunk * (size_t)d.ellB_w;
  if (v >= 3) {  // the LDS per block shrinks with the chunk, which can raise residency: take the fixed point
    for (int it = 0; it < 3; ++it) {
      sh = slab_shape(d, KRON_SLAB_W, xcd_resident(fn, lds, ncu));
      lds = sizeof(uint32_t) * (size_t)sh.rows_per_chunk * (size_t)d.ellB_w;
    }
  }
  if (v != 0 && lds > 32768) return false;
  switch (v) {
    case 0: kron_slab_packed_kernel<KRON_SLAB_W, 4><<<sh.grid, 256, 0, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk); break;
    case 1:
    case 3: kron_slab2_kernel<KRON_SLAB_W, 8><<<sh.grid, 256, lds, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk); break;
    case 5: kron_slab3_kernel<KRON_SLAB_W, 8><<<sh.grid, 256, lds, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk); break;
    case 6: kron_slab3_kernel<KRON_SLAB_W, 4><<<sh.grid, 256, lds, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk); break;
    default: kron_slab2_kernel<KRON_SLAB_W, 16><<<sh.grid, 256, lds, st>>>(d, x, y, sh.chunks, sh.rows_per_chunk); break;
  }
  return true;
}

// Variant choice per block shape, shared by every product of the process
// (the first product of a shape times each variant: 3 launches after one
// warm-up, best average).
// The key holds everything that decides which variants can run on a form
// (the packed tables built, their sign forms, the diagonal terms), so a
// cached pair is always one the form supports.
struct KronKey {
  int64_t nA, nB;
  int w, wB, nvtab, nvtabB, feat;
  bool operator<(const KronKey& o) const {
    return std::tie(nA, nB, w, wB, nvtab, nvtabB, feat) < std::tie(o.nA, o.nB, o.w, o.wB, o.nvtab, o.nvtabB, o.feat);
  }
};
int kron_features(const KronDesc& d) {
  return (d.ell16 ? 1 : 0) | (d.ell16s ? 2 : 0) | (d.asign ? 4 : 0) | (d.ellB4 ? 8 : 0) | (d.bsign ? 16 : 0) |
         (d.ndterms > 0 ? 32 : 0) | (d.nnum > 32 ? 64 : 0) | (d.nnum > 64 ? 128 : 0);
}
std::mutex g_tune_mu;
std::map<KronKey, std::pair<int, int>> g_tuned;

template <typename F>
float time_launches(hipStream_t st, F&& f, int reps) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 1e30f;
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return 1e30f;
  }
  f();
  hipEventRecord(e0, st);
  for (int r = 0; r < reps; ++r) f();
  hipEventRecord(e1, st);
  float ms = 1e30f;
  if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 1e30f;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return ms;
}

int kron_tune_locked(qd_ctx* ctx, qd_kron* k, const cplx* x, cplx* y) {
  const KronDesc& d = k->desc;
  // QD_KRON_VARIANTS="rows,slab" fixes both variants instead of timing them
  // (tests run every variant through it; "15,15" is the fused product); an
  // unfit variant then makes the product return QD_E_UNSUPPORTED.
  if (const char* f = std::getenv("QD_KRON_VARIANTS")) {
    int r = -1, s = -1;
    if (std::sscanf(f, "%d,%d", &r, &s) == 2 && r >= 0 && s >= 0 && r < 16 && s < 16) {
      k->variants.store(r * 16 + s);
      return QD_OK;
    }
  }
  const KronKey key{d.nA, d.nB, d.ell ? d.ell_w : -1, d.ellB ? d.ellB_w : -1, d.nvtab, d.ellB ? d.nvtabB : -1,
                    kron_features(d)};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      k->variants.store(it->second.first * 16 + it->second.second);
      return QD_OK;
    }
  }
  const int optin = kron_optin(ctx->device);
  hipStream_t st = ctx->stream;
  int best_r = 0, best_s = 0;
  float tr = 1e30f, ts = 1e30f;
  for (int v = 0; v < KRON_ROWS_VARIANTS; ++v) {
    bool ok = true;
    float ms = time_launches(st, [&] { ok = ok && launch_rows(d, v, optin, ctx->num_cus, x, y, cmake(0.0, 0.0), st); }, 3);
    if (ok && std::getenv("QD_TUNE_LOG")) std::fprintf(stderr, "[qd]   rows variant %d: %.1f us\n", v, 1000.0 * ms / 3);
    if (ok && ms < tr) {
      tr = ms;
      best_r = v;
    }
  }
  for (int v = 0; v < KRON_SLAB_VARIANTS; ++v) {
    bool ok = true;
    float ms = time_launches(st, [&] { ok = ok && launch_slab(d, v, ctx->num_cus, x, y, st); }, 3);
    if (ok && std::getenv("QD_TUNE_LOG")) std::fprintf(stderr, "[qd]   slab variant %d: %.1f us\n", v, 1000.0 * ms / 3);
    if (ok && ms < ts) {
      ts = ms;
      best_s = v;
    }
  }
  {  // the fused product, against the best pair of passes
    bool ok = true;
    const float ms = time_launches(st, [&] { ok = ok && launch_fused(d, optin, x, y, cmake(0.0, 0.0), st); }, 3);
    if (ok && std::getenv("QD_TUNE_LOG")) std::fprintf(stderr, "[qd]   fused: %.1f us\n", 1000.0 * ms / 3);
    if (ok && ms < tr + ts) {
      best_r = best_s = KRON_FUSED;
      tr = ms;
      ts = 0.0f;
    }
  }
  QD_CHECK_LAUNCH(ctx);
  if (std::getenv("QD_TUNE_LOG"))
    std::fprintf(stderr, "[qd] Kronecker nA=%lld nB=%lld w=%d wB=%d: rows variant %d (%.1f us), slab variant %d (%.1f us)\n",
                 (long long)d.nA, (long long)d.nB, d.ell ? d.ell_w : -1, d.ellB ? d.ellB_w : -1, best_r,
                 1000.0 * tr / 3, best_s, 1000.0 * ts / 3);
  k->variants.store(best_r * 16 + best_s);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key] = {best_r, best_s};
  return QD_OK;
}

// one thread times the variants of a form; others that arrive meanwhile wait
// for its choice (x, y: n-vector scratch of the calling thread)
int kron_tune(qd_ctx* ctx, qd_kron* k, const cplx* x, cplx* y) {
  std::lock_guard<std::mutex> lk(k->tune_mu);
  if (k->variants.load() >= 0) return QD_OK;
  return kron_tune_locked(ctx, k, x, y);
}
}  // namespace

int qd_kron_apply(qd_ctx* ctx, qd_kron* k, const cplx* x, cplx* y, cplx shift, const cplx* ext_diag) {
  if (k->desc.nA == 0 || k->desc.nB == 0) return QD_OK;
  int var = k->variants.load();
  if (var < 0) {
    int rc = kron_tune(ctx, k, x, y);  // leaves y overwritten: the product below rewrites it
    if (rc) return rc;
    var = k->variants.load();
  }
  const int rows_variant = var / 16, slab_variant = var % 16;
  KronDesc d = k->desc;
  d.ext_diag = ext_diag;
  const int optin = kron_optin(ctx->device);
  if (rows_variant == KRON_FUSED && slab_variant == KRON_FUSED) {
    if (!launch_fused(d, optin, x, y, shift, ctx->stream))
      return qd_set_error(ctx, QD_E_UNSUPPORTED, "Kronecker fused product");
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  }
  if (!launch_rows(d, rows_variant, optin, ctx->num_cus, x, y, shift, ctx->stream))
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "Kronecker rows variant %d", rows_variant);
  QD_CHECK_LAUNCH(ctx);
  if (!launch_slab(d, slab_variant, ctx->num_cus, x, y, ctx->stream))
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "Kronecker slab variant %d", slab_variant);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

int qd_kron_prepare(qd_ctx* ctx, qd_kron* k, const cplx* x, cplx* y) {
  if (!k || k->variants.load() >= 0 || k->desc.nA == 0 || k->desc.nB == 0) return QD_OK;
  return kron_tune(ctx, k, x, y);
}
