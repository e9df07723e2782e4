// mrhs_kron.hpp -- multi-RHS Kronecker product for the batched Arnoldi step
// (mrhs.hpp) of one two-block sector: H = T_A (x) I + I (x) T_B plus a
// per-slot diagonal, applied to m slots stored slot-minor (entry (s, t) at
// s * lds + t, s = b nA + a).  Same two passes as the single-vector product
// of matfree_kron.hip, in its summation order: a rows pass
//   x = f_t src[s, t],  Vout[s, t] = x,
//   Wout[s, t] = f_t sum_a' conj(T_A[a][a']) src[b nA + a', t] + (alpha_t D[s] + beta_t) x
// and a slab pass
//   Wout[s, t] += f_t sum_b' conj(T_B[b][b']) src[b' nA + a, t].
// MP lanes per state (slot t on lane t, MP = m rounded up to a power of two),
// 256 / MP states per block: every gather reads m consecutive slots (16 m
// bytes) of one state.  The rows pass reads within one row b (nA m 16 B); the
// slab pass gives each XCD column slabs of 256 / MP states whose nB rows
// (nB 256 16 B at most) stay in its L2 while its blocks sweep the rows.
#pragma once

#include "kron.hpp"
#include "mrhs.hpp"

namespace {

template <int MP>
__global__ void __launch_bounds__(256)
    g_kron_rows_kernel(KronDesc k, const GSec* __restrict__ gs, int from_x, const cplx* __restrict__ Wsrc,
                       const double* __restrict__ f, const cplx* __restrict__ alpha, const cplx* __restrict__ beta,
                       cplx* __restrict__ Vout, cplx* __restrict__ Wout) {
  constexpr int APB = 256 / MP;
  const GSec& S = gs[0];
  const int m = S.m;
  const int t = (int)threadIdx.x % MP;
  const int64_t a = (int64_t)blockIdx.x * APB + threadIdx.x / MP;
  const int64_t b = blockIdx.y;
  if (a >= k.nA || t >= m) return;
  const cplx* __restrict__ src = from_x ? S.X : Wsrc + S.vo;
  const int64_t lds = from_x ? S.ldx : m;
  const int g = S.so + t;
  const double ft = f ? f[g] : 1.0;
  const cplx* __restrict__ rowb = src + (b * k.nA) * lds + t;
  double sx = 0.0, sy = 0.0;
  for (int j0 = 0; j0 < k.ell_w; j0 += 4) {
    uint32_t e[4];
    cplx xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) e[u] = (j0 + u < k.ell_w) ? k.ell[(int64_t)(j0 + u) * k.nA + a] : 0u;
#pragma unroll
    for (int u = 0; u < 4; ++u) xv[u] = rowb[(int64_t)(e[u] >> 16) * lds];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const cplx v = k.vtab[e[u] & 0xFFFFu];
      sx += v.x * xv[u].x + v.y * xv[u].y;  // conj(v) * x
      sy += v.x * xv[u].y - v.y * xv[u].x;
    }
  }
  const int64_t s = b * k.nA + a;
  const cplx x0 = src[s * lds + t];
  const cplx x = cmake(x0.x * ft, x0.y * ft);
  const cplx al = alpha[g], be = beta[g];
  const double D = S.D[s];
  const cplx dx = cmul(cmake(al.x * D + be.x, al.y * D + be.y), x);
  if (Vout) Vout[S.vo + s * m + t] = x;
  Wout[S.vo + s * m + t] = cmake(sx * ft + dx.x, sy * ft + dx.y);
}

// block -> (XCD, slab, row chunk) as in kron_slab_packed_kernel
template <int MP>
__global__ void __launch_bounds__(256)
    g_kron_slab_kernel(KronDesc k, const GSec* __restrict__ gs, int from_x, const cplx* __restrict__ Wsrc,
                       const double* __restrict__ f, cplx* __restrict__ Wout, int chunks, int rows_per_chunk) {
  constexpr int APB = 256 / MP;
  const int xcd = blockIdx.x & 7;
  const int kk = blockIdx.x >> 3;
  const int64_t slab = (int64_t)xcd + 8 * (kk / chunks);
  const int chunk = kk % chunks;
  const GSec& S = gs[0];
  const int m = S.m;
  const int t = (int)threadIdx.x % MP;
  const int64_t a = slab * APB + threadIdx.x / MP;
  if (a >= k.nA || t >= m) return;
  const int64_t rb = (int64_t)chunk * rows_per_chunk;
  const int64_t re = min<int64_t>(k.nB, rb + rows_per_chunk);
  const cplx* __restrict__ src = from_x ? S.X : Wsrc + S.vo;
  const int64_t lds = from_x ? S.ldx : m;
  const double ft = f ? f[S.so + t] : 1.0;
  const cplx* __restrict__ cola = src + a * lds + t;
  const int w = k.ellB_w;
  for (int64_t b = rb; b < re; ++b) {
    const uint32_t* __restrict__ er = k.ellB + b * w;
    double sx = 0.0, sy = 0.0;
    for (int j0 = 0; j0 < w; j0 += 4) {
      uint32_t e[4];
      cplx xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) e[u] = (j0 + u < w) ? er[j0 + u] : 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = cola[(int64_t)(e[u] >> 16) * k.nA * lds];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const cplx v = k.vtabB[e[u] & 0xFFFFu];
        sx += v.x * xv[u].x + v.y * xv[u].y;
        sy += v.x * xv[u].y - v.y * xv[u].x;
      }
    }
    cplx* __restrict__ o = Wout + S.vo + (b * k.nA + a) * m + t;
    const cplx ov = *o;
    *o = cmake(ov.x + sx * ft, ov.y + sy * ft);
  }
}

// The product of mrhs_arnoldi_with for one sector in two-block form (the
// packed T_A and T_B tables must exist, m <= 64).
struct KronProduct {
  KronDesc d;
  const GSec* gs;
  int m;
  template <int MP>
  void launch(hipStream_t st, int from_x, const cplx* Wsrc, const double* f, const cplx* alpha, const cplx* beta,
              cplx* Vout, cplx* Wout) const {
    constexpr int APB = 256 / MP;
    const dim3 rgrid((unsigned)((d.nA + APB - 1) / APB), (unsigned)d.nB);
    g_kron_rows_kernel<MP><<<rgrid, 256, 0, st>>>(d, gs, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    const int64_t nslabs = (d.nA + APB - 1) / APB;
    const int chunks = (int)std::min<int64_t>(128, std::max<int64_t>(1, d.nB / 8));
    const int rpc = (int)((d.nB + chunks - 1) / chunks);
    const unsigned sgrid = (unsigned)((nslabs + 7) / 8 * chunks * 8);
    g_kron_slab_kernel<MP><<<sgrid, 256, 0, st>>>(d, gs, from_x, Wsrc, f, Wout, chunks, rpc);
  }
  void operator()(hipStream_t st, int from_x, const cplx* Wsrc, const double* f, const cplx* alpha,
                  const cplx* beta, cplx* Vout, cplx* Wout) const {
    if (m <= 1) launch<1>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    else if (m <= 2) launch<2>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    else if (m <= 4) launch<4>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    else if (m <= 8) launch<8>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    else if (m <= 16) launch<16>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    else if (m <= 32) launch<32>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
    else launch<64>(st, from_x, Wsrc, f, alpha, beta, Vout, Wout);
  }
};

}  // namespace
