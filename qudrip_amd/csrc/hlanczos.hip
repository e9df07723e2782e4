// hlanczos.hip -- Hermitian::LanczosSolver (sparse_algorithm.hpp:236-313):
// the reference's two-vector Lanczos, without reorthogonalisation, on the
// device.  Two passes of the three-term recurrence per solve:
//   pass 1 (getKrylovReprAd, :67-121) builds the tridiagonal Krylov matrix:
//     q_0 = psi / |psi|, a_0 = q_0^H H q_0, r = H q_0 - a_0 q_0, then while
//     |r|^2 > eps (1e-10, on the squared norm): b = |r|, q = r / b,
//     a = q^H H q (complex), r = H q - a q - b q_prev; it stops once
//     kry_dim + 1 coefficients a are held.  error() = |r| of the last r.
//   the tridiagonal (Re a on the diagonal, b off it) is solved on the host
//     in FP64; getGroundState (:285-300) takes the lowest eigenvector as
//     coefficients c, evolve (:257-283) c = U diag(e^{-i lam h}) U^T e_0;
//   pass 2 (getKrylovGS, :166-233) reruns the recurrence from psi with the
//     real part of a and accumulates out = sum_i c_i q_i.
// Two raw vectors hold the recurrence (q = s * raw with a device-side scale
// s, so no normalisation pass), a third the product.  Per step three
// launches: the product fused with the block sums of <raw|H raw> (stored
// matrices: the CSR-stream kernel writes them; other operators: the product
// then a dot pass), one one-block kernel that closes the previous step's
// |r|^2 (next scale, b, stopping rule) and this step's a, and the update
// pass (3 read, 1 written; pass 2 also accumulates out) with the block sums
// of |r|^2.  Kernel boundaries order the sums; every kernel returns at once
// after the recurrence has stopped, so a whole pass is launched without a
// host round trip.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "qd_internal.hpp"
#include "wave.hpp"

int qd_linop_apply(qd_linop* op, const cplx* x, cplx* y, const double* d_xscale, int xscale_idx,
                   const cplx* d_diag, cplx shift);

namespace {
typedef double hdbl2 __attribute__((ext_vector_type(2)));
// Vector accesses of the recurrence.  NT: non-temporal, chosen when a
// vector is larger than the 256 MB MALL, where none of them survives to the
// next kernel anyway and ordinary accesses would only evict the product's
// operands; below that size the vectors are reused from MALL (measured: XXZ
// N = 24 slows by 4 % with NT, Hubbard 4x4 gains 3 %).
constexpr int64_t kNtBytes = (int64_t)256 << 20;
template <bool NT>
__device__ __forceinline__ cplx hld(const cplx* p) {
  if (!NT) return *p;
  const hdbl2 v = __builtin_nontemporal_load(reinterpret_cast<const hdbl2*>(p));
  return cmake(v.x, v.y);
}
template <bool NT>
__device__ __forceinline__ void hst(cplx* p, cplx v) {
  if (!NT) {
    *p = v;
    return;
  }
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}

constexpr int HT = 256;  // threads per block
constexpr int HR = 4;    // rows per thread per tile

// device scalars of one pass
struct HlzDev {
  double s[2];  // q = s[k] * raw vector k
  double nrm2;  // |r|^2 of the newest raw vector
  double err;   // sqrt(nrm2) after the last step of pass 1
  double start; // |psi|^2
  cplx a;       // a of the current step (pass 2: its real part)
  int stop;     // the recurrence has ended
  int dim;      // count of a held (pass 1)
};

__device__ inline cplx block_sum(cplx v, cplx* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  cplx t = sh[0];
  for (int q = 1; q < HT / 64; ++q) t = cadd(t, sh[q]);
  return t;
}

// raw0 = psi, part[b] = sum |psi|^2 over the block's rows
__global__ void __launch_bounds__(HT) hlz_start_kernel(int64_t n, const cplx* __restrict__ psi,
                                                       cplx* __restrict__ raw0, cplx* __restrict__ part) {
  __shared__ cplx sh[HT / 64];
  double acc = 0.0;
  for (int64_t base = (int64_t)blockIdx.x * HT * HR; base < n; base += (int64_t)gridDim.x * HT * HR) {
#pragma unroll
    for (int q = 0; q < HR; ++q) {
      const int64_t r = base + (int64_t)q * HT + threadIdx.x;
      if (r < n) {
        const cplx v = psi[r];
        raw0[r] = v;
        acc += v.x * v.x + v.y * v.y;
      }
    }
  }
  const cplx t = block_sum(cmake(acc, 0.0), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// part[b] = sum conj(x) y over the block's rows
template <bool NT>
__global__ void __launch_bounds__(HT) hlz_dot_kernel(int64_t n, const cplx* __restrict__ x,
                                                     const cplx* __restrict__ y, cplx* __restrict__ part,
                                                     const HlzDev* __restrict__ dev) {
  if (dev->stop) return;
  __shared__ cplx sh[HT / 64];
  cplx acc = cmake(0.0, 0.0);
  for (int64_t base = (int64_t)blockIdx.x * HT * HR; base < n; base += (int64_t)gridDim.x * HT * HR) {
#pragma unroll
    for (int q = 0; q < HR; ++q) {
      const int64_t r = base + (int64_t)q * HT + threadIdx.x;
      if (r < n) acc = cadd(acc, cdotmul(hld<NT>(x + r), hld<NT>(y + r)));
    }
  }
  const cplx t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// One-block scalar step after hlz_start_kernel: s[0] = 1 / |psi|.
__global__ void __launch_bounds__(HT) hlz_start_scalars(const cplx* __restrict__ part, int nb,
                                                        HlzDev* __restrict__ dev) {
  __shared__ cplx sh[HT / 64];
  cplx acc = cmake(0.0, 0.0);
  for (int b = threadIdx.x; b < nb; b += HT) acc = cadd(acc, part[b]);
  const cplx t = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    dev->s[0] = 1.0 / sqrt(t.x);
    dev->s[1] = 0.0;
    dev->nrm2 = t.x;
    dev->start = t.x;
    dev->stop = 0;
    dev->dim = 0;
    dev->err = 0.0;
  }
}

// One-block kernel between the product and the update of step `step`
// (FT threads, fixed summation order):
//  - nb_norm > 0: |r|^2 = sum pnorm (the previous step's update) gives the
//    scale of raw vector cx; pass 1 also records b = |r| of step - 1,
//    error() and the stopping rule of getKrylovReprAd: stop when
//    |r|^2 <= eps, or, from step 1 on, when the count of a exceeds kry_dim;
//  - nb_dot > 0 (and not stopped): a = s_x^2 sum pdot (the product was
//    applied to the raw vector); pass 1 records it, pass 2 keeps its real
//    part.
constexpr int FT = 1024;
__device__ inline cplx block_sum_ft(cplx v, cplx* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  cplx t = sh[0];
  for (int q = 1; q < FT / 64; ++q) t = cadd(t, sh[q]);
  __syncthreads();
  return t;
}
// this thread's share of part[0 .. m): entries t, t + FT, ... in four
// interleaved accumulators (four loads in flight), combined in a fixed order
__device__ inline cplx strided_sum(const cplx* __restrict__ part, int m) {
  cplx a0 = cmake(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
  int b = threadIdx.x;
  for (; b + 3 * FT < m; b += 4 * FT) {
    a0 = cadd(a0, part[b]);
    a1 = cadd(a1, part[b + FT]);
    a2 = cadd(a2, part[b + 2 * FT]);
    a3 = cadd(a3, part[b + 3 * FT]);
  }
  for (; b < m; b += FT) a0 = cadd(a0, part[b]);
  return cadd(cadd(a0, a1), cadd(a2, a3));
}
__global__ void __launch_bounds__(FT)
    hlz_finish_kernel(const cplx* __restrict__ pnorm, int nb_norm, const cplx* __restrict__ pdot, int nb_dot,
                      HlzDev* __restrict__ dev, cplx* __restrict__ alpha, double* __restrict__ beta, int step, int cx,
                      int kry_dim, int pass2) {
  if (dev->stop) return;
  __shared__ cplx sh[FT / 64];
  __shared__ int stopped;
  if (nb_norm > 0) {
    const cplx t = block_sum_ft(strided_sum(pnorm, nb_norm), sh);
    if (threadIdx.x == 0) {
      const double n2 = t.x;
      const int prev = step - 1;
      dev->nrm2 = n2;
      dev->s[cx] = 1.0 / sqrt(n2);
      int st = 0;
      if (!pass2) {
        beta[prev] = sqrt(n2);
        dev->err = sqrt(n2);
        if (!(n2 > QD_LEAF_EPS) || (prev >= 1 && prev + 1 > kry_dim)) st = 1;
      }
      dev->stop = st;
      stopped = st;
    }
    __syncthreads();
    if (stopped) return;
  }
  if (nb_dot > 0) {
    const cplx t = block_sum_ft(strided_sum(pdot, nb_dot), sh);
    if (threadIdx.x == 0) {
      const double s2 = dev->s[cx] * dev->s[cx];
      const cplx a = crmul(s2, t);
      if (pass2) {
        dev->a = cmake(a.x, 0.0);
      } else {
        dev->a = a;
        alpha[step] = a;
        dev->dim = step + 1;
      }
    }
  }
}

// r = s_x y - a s_x x - b s_p p with b = 1 / s_x (= |r_prev|), written over
// p's raw vector; part[b] = sum |r|^2.  FIRST: no p.  ACC: out (+)= c s_x x
// (ACC_SET: the first term, out not read).
template <bool FIRST, bool ACC, bool ACC_SET, bool NT>
__global__ void __launch_bounds__(HT)
    hlz_update_kernel(int64_t n, const cplx* __restrict__ y, const cplx* __restrict__ x, cplx* __restrict__ p,
                      cplx* __restrict__ out, const cplx* __restrict__ coef, int step, int cx,
                      cplx* __restrict__ part, const HlzDev* __restrict__ dev) {
  if (dev->stop) return;
  __shared__ cplx sh[HT / 64];
  const double sx = dev->s[cx], sp = dev->s[cx ^ 1];
  const double b = FIRST ? 0.0 : 1.0 / sx;
  const cplx a = dev->a;
  const cplx as = crmul(sx, a);
  const double bs = b * sp;
  const cplx c = ACC ? crmul(sx, coef[step]) : cmake(0.0, 0.0);
  double acc = 0.0;
  for (int64_t base = (int64_t)blockIdx.x * HT * HR; base < n; base += (int64_t)gridDim.x * HT * HR) {
#pragma unroll
    for (int q = 0; q < HR; ++q) {
      const int64_t r = base + (int64_t)q * HT + threadIdx.x;
      if (r < n) {
        const cplx xv = hld<NT>(x + r), yv = hld<NT>(y + r);
        cplx v = csub(crmul(sx, yv), cmul(as, xv));
        if (!FIRST) v = csub(v, crmul(bs, hld<NT>(p + r)));
        hst<NT>(p + r, v);
        acc += v.x * v.x + v.y * v.y;
        if (ACC) hst<NT>(out + r, ACC_SET ? cmul(c, xv) : cadd(hld<NT>(out + r), cmul(c, xv)));
      }
    }
  }
  const cplx t = block_sum(cmake(acc, 0.0), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out (+)= c s_x x: the last term of pass 2, whose next vector is not needed
// (set: the only term, out not read)
__global__ void __launch_bounds__(HT) hlz_acc_kernel(int64_t n, const cplx* __restrict__ x, cplx* __restrict__ out,
                                                     const cplx* __restrict__ coef, int step, int cx, int set,
                                                     const HlzDev* __restrict__ dev) {
  const cplx c = crmul(dev->s[cx], coef[step]);
  for (int64_t r = (int64_t)blockIdx.x * HT + threadIdx.x; r < n; r += (int64_t)gridDim.x * HT)
    out[r] = set ? cmul(c, x[r]) : cadd(out[r], cmul(c, x[r]));
}

}  // namespace

struct qd_hlanczos {
  qd_ctx* ctx = nullptr;
  int64_t n = 0;
  int kry_dim = 0;
  int nb = 1;               // blocks of the vector kernels (partials)
  cplx* d_raw[2] = {nullptr, nullptr};
  cplx* d_y = nullptr;
  char* d_small = nullptr;  // HlzDev | alpha[kry_dim+2] | beta | coef | partials
  HlzDev* d_dev = nullptr;
  cplx* d_alpha = nullptr;
  double* d_beta = nullptr;
  cplx* d_coef = nullptr;
  cplx* d_part = nullptr;    // block sums of the start and update passes (nb)
  cplx* d_pdot = nullptr;    // block sums of <raw|H raw> (dot_blocks)
  int dot_blocks = 0;        // blocks of the fused product (stored matrices)
  // host copies of the last solve
  int dim = 0;
  double err = 0.0;
  int64_t n_spmv = 0;
  std::vector<cplx> alpha;
  std::vector<double> beta;
  std::vector<double> evals;
};

namespace {

int hlz_fail(qd_hlanczos* lz, int code, const char* what) { return qd_set_error(lz->ctx, code, "%s", what); }

// access form of the vector kernels: non-temporal above the MALL size;
// QD_HLZ_NT=0 / 1 overrides the size rule (the two forms give the same bits)
bool use_nt(int64_t n) {
  if (const char* e = std::getenv("QD_HLZ_NT")) return e[0] == '1';
  return n * (int64_t)sizeof(cplx) > kNtBytes;
}

// y = H raw_x and the block sums of <raw_x|y> in d_pdot: one fused launch
// for stored matrices, else the product and a dot pass.  Returns the number
// of block sums through *npart.
int step_product(qd_hlanczos* lz, qd_linop* op, int cx, int* npart) {
  qd_ctx* ctx = lz->ctx;
  const int64_t n = lz->n;
  int rc = qd_linop_apply_dot(op, lz->d_raw[cx], lz->d_y, lz->d_pdot, true);
  if (rc == QD_OK) {
    *npart = lz->dot_blocks;
    return QD_OK;
  }
  if (rc != QD_E_UNSUPPORTED) return rc;
  rc = qd_linop_apply(op, lz->d_raw[cx], lz->d_y, nullptr, 0, nullptr, cmake(0.0, 0.0));
  if (rc) return rc;
  if (use_nt(n))
    hlz_dot_kernel<true><<<lz->nb, HT, 0, ctx->stream>>>(n, lz->d_raw[cx], lz->d_y, lz->d_pdot, lz->d_dev);
  else
    hlz_dot_kernel<false><<<lz->nb, HT, 0, ctx->stream>>>(n, lz->d_raw[cx], lz->d_y, lz->d_pdot, lz->d_dev);
  QD_CHECK_LAUNCH(ctx);
  *npart = lz->nb;
  return QD_OK;
}

void launch_finish(qd_hlanczos* lz, int nb_norm, int nb_dot, int step, int cx, int pass2) {
  hlz_finish_kernel<<<1, FT, 0, lz->ctx->stream>>>(lz->d_part, nb_norm, lz->d_pdot, nb_dot, lz->d_dev, lz->d_alpha,
                                                   lz->d_beta, step, cx, lz->kry_dim, pass2);
}

template <bool FIRST, bool ACC, bool ACC_SET>
void launch_update(qd_hlanczos* lz, int step, int cx, cplx* out) {
  const int64_t n = lz->n;
  hipStream_t st = lz->ctx->stream;
  cplx* y = lz->d_y;
  const cplx* x = lz->d_raw[cx];
  cplx* p = lz->d_raw[cx ^ 1];
  if (use_nt(n))
    hlz_update_kernel<FIRST, ACC, ACC_SET, true><<<lz->nb, HT, 0, st>>>(n, y, x, p, out, lz->d_coef, step, cx,
                                                                         lz->d_part, lz->d_dev);
  else
    hlz_update_kernel<FIRST, ACC, ACC_SET, false><<<lz->nb, HT, 0, st>>>(n, y, x, p, out, lz->d_coef, step, cx,
                                                                          lz->d_part, lz->d_dev);
}

// pass 1: the tridiagonal coefficients of psi's Krylov space
int hlz_pass1(qd_hlanczos* lz, qd_linop* op, const cplx* psi) {
  qd_ctx* ctx = lz->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = lz->n;
  const int nb = lz->nb;
  hlz_start_kernel<<<nb, HT, 0, st>>>(n, psi, lz->d_raw[0], lz->d_part);
  hlz_start_scalars<<<1, HT, 0, st>>>(lz->d_part, nb, lz->d_dev);
  QD_CHECK_LAUNCH(ctx);
  for (int i = 0; i <= lz->kry_dim; ++i) {
    const int cx = i & 1;
    int npart = 0;
    int rc = step_product(lz, op, cx, &npart);
    if (rc) return rc;
    launch_finish(lz, i == 0 ? 0 : nb, npart, i, cx, 0);
    if (i == 0)
      launch_update<true, false, false>(lz, i, cx, nullptr);
    else
      launch_update<false, false, false>(lz, i, cx, nullptr);
    QD_CHECK_LAUNCH(ctx);
  }
  // the last step's |r|^2: b, error() and the stop flag
  launch_finish(lz, nb, 0, lz->kry_dim + 1, (lz->kry_dim + 1) & 1, 0);
  QD_CHECK_LAUNCH(ctx);
  HlzDev h{};
  QD_HIP(ctx, hipMemcpyAsync(&h, lz->d_dev, sizeof(h), hipMemcpyDeviceToHost, st));
  QD_HIP(ctx, hipStreamSynchronize(st));
  if (!(h.start > 0.0) || !std::isfinite(h.start))
    return hlz_fail(lz, QD_E_ARG, "Hermitian Lanczos: the start vector is zero or not finite");
  lz->dim = h.dim;
  lz->err = h.err;
  lz->n_spmv += h.dim;
  lz->alpha.assign((size_t)h.dim, cmake(0.0, 0.0));
  lz->beta.assign((size_t)h.dim, 0.0);
  QD_HIP(ctx, hipMemcpy(lz->alpha.data(), lz->d_alpha, sizeof(cplx) * h.dim, hipMemcpyDeviceToHost));
  QD_HIP(ctx, hipMemcpy(lz->beta.data(), lz->d_beta, sizeof(double) * h.dim, hipMemcpyDeviceToHost));
  return QD_OK;
}

// the tridiagonal's eigenpairs (ascending), z column-major
int hlz_eig(qd_hlanczos* lz, std::vector<double>& w, std::vector<double>& z) {
  const int m = lz->dim;
  std::vector<double> d(m), e(std::max(m - 1, 1), 0.0);
  for (int i = 0; i < m; ++i) d[i] = lz->alpha[i].x;  // the eigen-solver reads the real diagonal
  for (int i = 0; i + 1 < m; ++i) e[i] = lz->beta[i];
  w.assign(m, 0.0);
  z.assign((size_t)m * m, 0.0);
  if (qd_tridiag_eig(m, d.data(), e.data(), w.data(), z.data(), 1) != 0)
    return qd_set_error(lz->ctx, QD_E_EIGEN, "failed to solve the Krylov matrix");
  lz->evals = w;
  return QD_OK;
}

// pass 2: out = sum_i coef[i] q_i over the dim vectors of pass 1
int hlz_pass2(qd_hlanczos* lz, qd_linop* op, const cplx* psi, cplx* out, const std::vector<cplx>& coef) {
  qd_ctx* ctx = lz->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = lz->n;
  const int nb = lz->nb;
  const int m = lz->dim;
  QD_HIP(ctx, hipMemcpyAsync(lz->d_coef, coef.data(), sizeof(cplx) * m, hipMemcpyHostToDevice, st));
  hlz_start_kernel<<<nb, HT, 0, st>>>(n, psi, lz->d_raw[0], lz->d_part);
  hlz_start_scalars<<<1, HT, 0, st>>>(lz->d_part, nb, lz->d_dev);
  for (int i = 0; i + 1 < m; ++i) {
    const int cx = i & 1;
    int npart = 0;
    int rc = step_product(lz, op, cx, &npart);
    if (rc) return rc;
    ++lz->n_spmv;
    launch_finish(lz, i == 0 ? 0 : nb, npart, i, cx, 1);
    if (i == 0)
      launch_update<true, true, true>(lz, i, cx, out);
    else
      launch_update<false, true, false>(lz, i, cx, out);
    QD_CHECK_LAUNCH(ctx);
  }
  // the scale of the last vector
  if (m > 1) launch_finish(lz, nb, 0, m - 1, (m - 1) & 1, 1);
  const int cl = (m - 1) & 1;
  hlz_acc_kernel<<<nb, HT, 0, st>>>(n, lz->d_raw[cl], out, lz->d_coef, m - 1, cl, m == 1, lz->d_dev);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

}  // namespace

extern "C" int qd_hlanczos_create(qd_ctx* ctx, int64_t n, int kry_dim, qd_hlanczos** out) {
  if (!ctx || !out || n < 1 || kry_dim < 1 || kry_dim > 100000)
    return qd_set_error(ctx, QD_E_ARG, "qd_hlanczos_create: bad arguments");
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_hlanczos* lz = new qd_hlanczos();
  lz->ctx = ctx;
  lz->n = n;
  lz->kry_dim = kry_dim;
  const int64_t tiles = (n + (int64_t)HT * HR - 1) / ((int64_t)HT * HR);
  lz->nb = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)ctx->num_cus * 8));
  const size_t vb = sizeof(cplx) * (size_t)n;
  const size_t k2 = (size_t)kry_dim + 2;
  // block sums: the start / update passes (nb), and the dot of each step
  // (the fused product's blocks, or nb from the dot pass)
  lz->dot_blocks = (int)qd_linop_dot_blocks(n);
  const int64_t npdot = std::max<int64_t>(lz->nb, lz->dot_blocks);
  const size_t small = 256 + sizeof(cplx) * k2 * 2 + sizeof(double) * k2 + sizeof(cplx) * (size_t)lz->nb +
                       sizeof(cplx) * (size_t)npdot + 2 * 256;
  if (qd_dev_alloc(ctx, vb, (void**)&lz->d_raw[0]) != QD_OK || qd_dev_alloc(ctx, vb, (void**)&lz->d_raw[1]) != QD_OK ||
      qd_dev_alloc(ctx, vb, (void**)&lz->d_y) != QD_OK || qd_dev_alloc(ctx, small, (void**)&lz->d_small) != QD_OK) {
    qd_dev_free(ctx, lz->d_raw[0]);
    qd_dev_free(ctx, lz->d_raw[1]);
    qd_dev_free(ctx, lz->d_y);
    qd_dev_free(ctx, lz->d_small);
    delete lz;
    return qd_set_error(ctx, QD_E_OOM, "Hermitian Lanczos: 3 vectors of %zu bytes", vb);
  }
  char* p = lz->d_small;
  lz->d_dev = (HlzDev*)p;
  p += 256;
  lz->d_alpha = (cplx*)p;
  p += sizeof(cplx) * k2;
  lz->d_coef = (cplx*)p;
  p += sizeof(cplx) * k2;
  lz->d_beta = (double*)p;
  p += sizeof(double) * k2;
  p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
  lz->d_part = (cplx*)p;
  p += sizeof(cplx) * (size_t)lz->nb;
  p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
  lz->d_pdot = (cplx*)p;
  *out = lz;
  return QD_OK;
}

extern "C" int qd_hlanczos_destroy(qd_hlanczos* lz) {
  if (!lz) return QD_OK;
  hipSetDevice(lz->ctx->device);
  hipStreamSynchronize(lz->ctx->stream);
  qd_dev_free(lz->ctx, lz->d_raw[0]);
  qd_dev_free(lz->ctx, lz->d_raw[1]);
  qd_dev_free(lz->ctx, lz->d_y);
  qd_dev_free(lz->ctx, lz->d_small);
  delete lz;
  return QD_OK;
}

static int hlz_check_op(qd_hlanczos* lz, qd_linop* op) {
  if (!op) return qd_set_error(lz->ctx, QD_E_ARG, "Hermitian Lanczos: no operator");
  if (op->n != lz->n)
    return qd_set_error(lz->ctx, QD_E_ARG, "Hermitian Lanczos: operator dimension %lld, solver %lld",
                        (long long)op->n, (long long)lz->n);
  return QD_OK;
}

extern "C" int qd_hlanczos_ground_state(qd_hlanczos* lz, qd_linop* op, double* d_psi, double* h_E0) {
  if (!lz || !d_psi) return QD_E_ARG;
  hipSetDevice(lz->ctx->device);
  int rc = hlz_check_op(lz, op);
  if (rc) return rc;
  cplx* psi = (cplx*)d_psi;
  if ((rc = hlz_pass1(lz, op, psi))) return rc;
  std::vector<double> w, z;
  if ((rc = hlz_eig(lz, w, z))) return rc;
  std::vector<cplx> c((size_t)lz->dim);
  for (int i = 0; i < lz->dim; ++i) c[i] = cmake(z[i], 0.0);  // column 0: the lowest eigenvalue
  if ((rc = hlz_pass2(lz, op, psi, psi, c))) return rc;
  if (h_E0) *h_E0 = w[0];
  QD_HIP(lz->ctx, hipStreamSynchronize(lz->ctx->stream));
  return QD_OK;
}

extern "C" int qd_hlanczos_evolve(qd_hlanczos* lz, qd_linop* op, const double* d_psi_in, double* d_psi_out,
                                  double h) {
  if (!lz || !d_psi_in || !d_psi_out) return QD_E_ARG;
  hipSetDevice(lz->ctx->device);
  int rc = hlz_check_op(lz, op);
  if (rc) return rc;
  if ((rc = hlz_pass1(lz, op, (const cplx*)d_psi_in))) return rc;
  std::vector<double> w, z;
  if ((rc = hlz_eig(lz, w, z))) return rc;
  const int m = lz->dim;
  // c = U diag(exp(-i lam h)) U^T e_0 (U real orthogonal)
  std::vector<cplx> c((size_t)m, cmake(0.0, 0.0));
  for (int j = 0; j < m; ++j) {
    const double u0 = z[(size_t)j * m];  // U[0][j]
    const cplx ph = cmake(std::cos(w[j] * h) * u0, -std::sin(w[j] * h) * u0);
    for (int i = 0; i < m; ++i) c[i] = cadd(c[i], crmul(z[(size_t)j * m + i], ph));
  }
  if ((rc = hlz_pass2(lz, op, (const cplx*)d_psi_in, (cplx*)d_psi_out, c))) return rc;
  QD_HIP(lz->ctx, hipStreamSynchronize(lz->ctx->stream));
  return QD_OK;
}

extern "C" int qd_hlanczos_get_info(qd_hlanczos* lz, qd_hlanczos_info* info) {
  if (!lz || !info) return QD_E_ARG;
  info->dim = lz->dim;
  info->error = lz->err;
  info->n_spmv = lz->n_spmv;
  return QD_OK;
}

extern "C" int qd_hlanczos_eigenvalues(qd_hlanczos* lz, double* h_evals, int cap, int* count) {
  if (!lz || (cap > 0 && !h_evals)) return QD_E_ARG;
  const int m = std::min<int>(cap, (int)lz->evals.size());
  for (int i = 0; i < m; ++i) h_evals[i] = lz->evals[i];
  if (count) *count = m;
  return QD_OK;
}

extern "C" int qd_hlanczos_coefficients(qd_hlanczos* lz, double* h_alpha, double* h_beta, int cap, int* count) {
  if (!lz || (cap > 0 && (!h_alpha || !h_beta))) return QD_E_ARG;
  const int m = std::min<int>(cap, lz->dim);
  for (int i = 0; i < m; ++i) {
    h_alpha[2 * i] = lz->alpha[i].x;
    h_alpha[2 * i + 1] = lz->alpha[i].y;
    h_beta[i] = lz->beta[i];
  }
  if (count) *count = m;
  return QD_OK;
}
