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
// (row j+1).  The reference's Lanczos (expandBasis<1>, :395-425) keeps only
// the last two vectors; full reorthogonalisation changes no exact-arithmetic
// quantity (same Krylov space, same tridiagonal) and keeps the Ritz values
// free of ghost copies.  Every reduction has a fixed tree (block partials
// summed in index order), so a solve gives the same bits whatever else runs.
//
// Breakdown (beta < 1e-10, :387-390, :421) sets a device flag: the remaining
// kernels of the expansion see it and return, so the whole expansion is
// enqueued without a host round trip and read back once at the end.
#include <algorithm>
#include <cmath>

#include "qd_internal.hpp"

int qd_linop_apply(qd_linop* op, const cplx* x, cplx* y, const double* d_xscale, int xscale_idx,
                   const cplx* d_diag, cplx shift);

namespace {

constexpr int KT = 256;  // threads per block for the tall-skinny kernels

__device__ inline cplx wave_sum(cplx s) {
  for (int o = 32; o > 0; o >>= 1) {
    s.x += __shfl_xor(s.x, o, 64);
    s.y += __shfl_xor(s.y, o, 64);
  }
  return s;
}

// part[b * ncols + c] = sum over rows of block b of conj(V[r, c]) * w[r]
template <int RPT>
__global__ void __launch_bounds__(KT)
    gemv_t_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ w,
                  cplx* __restrict__ part, const int* __restrict__ flag) {
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
  for (int c0 = 0; c0 < ncols; c0 += 64) {
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
__global__ void __launch_bounds__(KT)
    reduce_cols_kernel(const cplx* __restrict__ part, int nb, int ncols, cplx* __restrict__ h,
                       cplx* __restrict__ kcol, int accumulate, const int* __restrict__ flag) {
  if (flag && *flag) return;
  __shared__ cplx sh[KT / 64];
  const int c = blockIdx.x;
  cplx s = cmake(0.0, 0.0);
  for (int b = threadIdx.x; b < nb; b += KT) s = cadd(s, part[(int64_t)b * ncols + c]);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    cplx t = sh[0];
    for (int q = 1; q < KT / 64; ++q) t = cadd(t, sh[q]);
    h[c] = t;
    if (kcol) kcol[c] = accumulate ? cadd(kcol[c], t) : t;
  }
}

// w[r] -= sum_c V[r, c] h[c]; optional block partial of |w|^2
template <int RPT>
__global__ void __launch_bounds__(KT)
    gemv_n_kernel(int64_t n, const cplx* __restrict__ V, int64_t ld, int ncols, const cplx* __restrict__ h,
                  cplx* __restrict__ w, double* __restrict__ norm_part, const int* __restrict__ flag) {
  if (flag && *flag) return;
  __shared__ double sh[KT / 64];
  const int64_t base = (int64_t)blockIdx.x * KT * RPT;
  cplx acc[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    int64_t r = base + (int64_t)q * KT + threadIdx.x;
    acc[q] = (r < n) ? w[r] : cmake(0.0, 0.0);
  }
  for (int c = 0; c < ncols; ++c) {
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
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
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
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
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
        cplx v = Vc[r];
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
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
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
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0];
    for (int q = 1; q < KT / 64; ++q) t += sh[q];
    *out = t;
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
};

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
         sizeof(double) * (size_t)(nb + 1) + 256 + sizeof(cplx) * (size_t)nb * (K + 1);
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
  if (hipMalloc(&kr->d_V, vbytes) != hipSuccess) {
    delete kr;
    return qd_set_error(ctx, QD_E_OOM, "Krylov basis: %zu bytes", vbytes);
  }
  if (hipMalloc(&kr->d_scal, kr_scal_bytes(n, kry_dim)) != hipSuccess) {
    hipFree(kr->d_V);
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
  hipFree(kr->d_V);
  hipFree(kr->d_scal);
  delete kr;
  return QD_OK;
}

extern "C" int qd_krylov_reset(qd_krylov* kr) {
  if (!kr) return QD_E_ARG;
  kr->act_dim = 0;
  kr->breakdown = 0;
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
  QD_CHECK_LAUNCH(ctx);
  kr->act_dim = 0;
  kr->breakdown = 0;
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

// extend the basis from act_dim by add columns; reads KR back.
static int kr_expand(qd_krylov* kr, qd_linop* op, int add) {
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
  int hflag[2] = {0, 0};
  QD_HIP(ctx, hipMemcpyAsync(hflag, d.flag, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipMemcpyAsync(cols.data(), d.KR + (size_t)start * (K + 2), sizeof(cplx) * cols.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int end = start + add;
  if (hflag[0]) {
    end = hflag[1] + 1;  // column bd_step is complete; the space is invariant
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
  qd_ctx* ctx = kr->ctx;
  if (kr->act_dim >= kr->kry_dim && !kr->evals.empty()) {  // :567
    if (h_E0) *h_E0 = kr->evals[0];
    return QD_OK;
  }
  int rc;
  if (kr->act_dim == 0) {
    rc = kr_init(kr, (const cplx*)d_psi);
    if (rc) return rc;
  }
  int add = add_dim;
  if (add == -1 || kr->kry_dim < kr->act_dim + add) add = kr->kry_dim - kr->act_dim;  // :523-524
  rc = kr_expand(kr, op, add);
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

extern "C" int qd_krylov_get_info(qd_krylov* kr, qd_krylov_info* info) {
  if (!kr || !info) return QD_E_ARG;
  info->act_dim = kr->act_dim;
  info->breakdown = kr->breakdown;
  info->error = kr->err;
  info->beta_last = kr->beta_last;
  info->n_spmv = kr->n_spmv;
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
