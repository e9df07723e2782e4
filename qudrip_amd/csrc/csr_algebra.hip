// csr_algebra.hip -- arithmetic between stored matrices (the reference's
// SpMatrix = Eigen::SparseMatrix<complex<double>> ColMajor, matrix.hpp:11),
// on the device and with Eigen's value rules, so that a caller can combine
// the matrices `>>` returns the way programs/hubbard_diss.cpp does:
//
//   SpMatrix Id(n, n); Id.setIdentity();              hubbard_diss.cpp:110,117
//   Tp += thop * hop21u >> hubbard;                     :135-136
//   qj_tot += gamma * SpMatrix(ni.adjoint()) * ni;      :156
//   Htot = Tdr * Tp + std::conj(Tdr) * Tm + (ndr + U0) * docc;   :443
//   H -= 0.5 * II * (qj_tot - wt * Id);                 :514-517
//
// Every result is a new ColMajor matrix (outer = column, inner rows
// ascending), built in two passes (count, then fill at scanned offsets).
// The value of each entry follows Eigen's element rules for these
// expressions (SparseCwiseBinaryOp.h, ConservativeSparseSparseProduct.h):
//   A + B, A - B   union of the two patterns; an entry of one side only is
//                  op(a, 0) or op(0, b), so -0 becomes +0 and 0 - b = -b;
//                  explicit zeros are kept (no pruning);
//   c * A          c * a per entry (std::complex product; a real c scales
//                  both parts: (c re, c im));
//   A.adjoint()    conjugate transpose, each new column's rows ascending;
//   A * B          column j of the product visits B's column j in order and,
//                  for each entry (k, y), A's column k in order; the first
//                  visit of row i stores x * y, later visits add x * y.
// Evaluating an expression tree node by node gives the same numbers as
// Eigen's lazy evaluation of the tree: each entry sees the same operations
// on the same operands.  This file is compiled without FMA contraction.
#include "qd_internal.hpp"

namespace {

constexpr int ALG_THREADS = 256;

__global__ void alg_offs_to_outer(const int64_t* __restrict__ offs, int64_t n, int32_t* __restrict__ outer) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) outer[i] = (int32_t)offs[i];
}

// ------------------------------------------------------------------ A +- B
// one thread per column: merge of two ascending row lists
template <bool FILL>
__global__ void add_kernel(int64_t n, const int32_t* __restrict__ ao, const int32_t* __restrict__ ai,
                           const cplx* __restrict__ av, const int32_t* __restrict__ bo,
                           const int32_t* __restrict__ bi, const cplx* __restrict__ bv, double sign,
                           int32_t* __restrict__ counts, const int64_t* __restrict__ offs,
                           int32_t* __restrict__ ci, cplx* __restrict__ cv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int32_t p = ao[j], pe = ao[j + 1], q = bo[j], qe = bo[j + 1];
  int64_t w = FILL ? offs[j] : 0;
  int32_t cnt = 0;
  const cplx zero = cmake(0.0, 0.0);
  while (p < pe || q < qe) {
    int32_t r;
    cplx a = zero, b = zero;
    if (p < pe && q < qe && ai[p] == bi[q]) {
      r = ai[p];
      a = av[p++];
      b = bv[q++];
    } else if (p < pe && (q >= qe || ai[p] < bi[q])) {
      r = ai[p];
      a = av[p++];
    } else {
      r = bi[q];
      b = bv[q++];
    }
    if (FILL) {
      // op(a, b) with 0 for the missing side; sign = -1 is Eigen's
      // scalar_difference_op (a - b), not a + (-1) b
      ci[w] = r;
      cv[w] = sign > 0.0 ? cadd(a, b) : csub(a, b);
      ++w;
    }
    ++cnt;
  }
  if (!FILL) counts[j] = cnt;
}

// ------------------------------------------------------------------ c * A
__global__ void scale_kernel(int64_t nnz, cplx c, int real_scalar, const cplx* __restrict__ av,
                             cplx* __restrict__ cv) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  cv[e] = real_scalar ? crmul(c.x, av[e]) : cmul(c, av[e]);
}

__global__ void copy_i32_kernel(int64_t m, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) b[e] = a[e];
}

// -------------------------------------------------------------- identity
__global__ void identity_kernel(int64_t n, int32_t* __restrict__ outer, int32_t* __restrict__ inner,
                                cplx* __restrict__ vals) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j <= n) outer[j] = (int32_t)j;
  if (j < n) {
    inner[j] = (int32_t)j;
    vals[j] = cmake(1.0, 0.0);
  }
}

__global__ void zero_i32_kernel(int64_t m, int32_t* __restrict__ a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) a[e] = 0;
}

// ------------------------------------------------------------- adjoint
__global__ void row_count_kernel(int64_t nnz, const int32_t* __restrict__ ai, int32_t* __restrict__ counts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) atomicAdd(&counts[ai[e]], 1);
}

// entry (i, j, v) of A goes to new column i as (j, conj v); the slot inside
// the column comes from an atomic cursor, the column is sorted afterwards
__global__ void adj_scatter_kernel(int64_t n, const int32_t* __restrict__ ao, const int32_t* __restrict__ ai,
                                   const cplx* __restrict__ av, const int64_t* __restrict__ offs,
                                   int32_t* __restrict__ cursor, int32_t* __restrict__ ci, cplx* __restrict__ cv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int32_t e = ao[j]; e < ao[j + 1]; ++e) {
    const int32_t i = ai[e];
    const int64_t w = offs[i] + atomicAdd(&cursor[i], 1);
    ci[w] = (int32_t)j;
    cv[w] = cconj(av[e]);
  }
}

// sorts each column's (row, value) pairs by row: insertion sort for short
// columns, Shell sort (Ciura gaps) for long ones; rows within a column are
// distinct, so the order is unique
__global__ void col_sort_kernel(int64_t n, const int64_t* __restrict__ offs, int32_t* __restrict__ ci,
                                cplx* __restrict__ cv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t lo = offs[j], len = offs[j + 1] - offs[j];
  int32_t* r = ci + lo;
  cplx* v = cv + lo;
  const int gaps[9] = {701, 301, 132, 57, 23, 10, 4, 1, 0};
  for (int g = 0; gaps[g] > 0; ++g) {
    const int64_t gap = gaps[g];
    if (gap >= len && gap > 1) continue;
    for (int64_t a = gap; a < len; ++a) {
      const int32_t key = r[a];
      const cplx val = v[a];
      int64_t b = a;
      while (b >= gap && r[b - gap] > key) {
        r[b] = r[b - gap];
        v[b] = v[b - gap];
        b -= gap;
      }
      r[b] = key;
      v[b] = val;
    }
  }
}

// ------------------------------------------------------------- A * B
// One wave per column j of the product.  The column's candidate list --
// (row i, x * y) for every entry (k, y) of B's column j in order and every
// entry (i, x) of A's column k in order -- is laid out in LDS in that visit
// order.  The first visit of each row then sums that row's values in visit
// order and writes the sum at the row's rank among the column's rows.
constexpr int MM_WAVES = 2;
constexpr int MM_CAP = 1024;  // candidates per column (QD_E_CAPACITY beyond)
template <bool FILL>
__global__ void __launch_bounds__(64 * MM_WAVES)
    spgemm_kernel(int64_t n, const int32_t* __restrict__ ao, const int32_t* __restrict__ ai,
                  const cplx* __restrict__ av, const int32_t* __restrict__ bo, const int32_t* __restrict__ bi,
                  const cplx* __restrict__ bv, int32_t* __restrict__ counts, const int64_t* __restrict__ offs,
                  int32_t* __restrict__ ci, cplx* __restrict__ cv, int* __restrict__ overflow) {
  __shared__ int32_t s_row[MM_WAVES][MM_CAP];
  __shared__ cplx s_val[MM_WAVES][FILL ? MM_CAP : 1];
  __shared__ uint8_t s_first[MM_WAVES][MM_CAP];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * MM_WAVES + wv;
  const bool active = j < n;  // every wave reaches the block barriers
  int32_t* rows = s_row[wv];
  cplx* vals = s_val[wv];
  uint8_t* first = s_first[wv];
  // candidate offsets: exclusive prefix over B's entries of |A col k|, in
  // chunks of 64 entries (one per lane)
  int total = 0;
  const int32_t qb = active ? bo[j] : 0, qe = active ? bo[j + 1] : 0;
  for (int32_t q0 = qb; q0 < qe; q0 += 64) {
    const int32_t q = q0 + lane;
    int len = 0;
    int32_t k = 0;
    if (q < qe) {
      k = bi[q];
      len = ao[k + 1] - ao[k];
    }
    int incl = len;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    const int base = total + incl - len;
    if (q < qe && base + len <= MM_CAP) {
      const cplx y = bv[q];
      for (int t = 0; t < len; ++t) {
        const int32_t e = ao[k] + t;
        rows[base + t] = ai[e];
        if (FILL) vals[base + t] = cmul(av[e], y);  // x * y
      }
    }
    total += __shfl(incl, 63, 64);
  }
  const bool over = total > MM_CAP;
  if (active && over && lane == 0) atomicOr(overflow, 1);
  const int tot = over ? 0 : total;
  __syncthreads();
  // first visit of each row
  int nfirst = 0;
  for (int c = lane; c < tot; c += 64) {
    const int32_t r = rows[c];
    int f = 1;
    for (int d = 0; d < c; ++d)
      if (rows[d] == r) {
        f = 0;
        break;
      }
    first[c] = (uint8_t)f;
    nfirst += f;
  }
  for (int o = 32; o > 0; o >>= 1) nfirst += __shfl_xor(nfirst, o, 64);
  if (!FILL) {
    if (active && lane == 0) counts[j] = nfirst;
    return;
  }
  __syncthreads();
  if (!active) return;
  const int64_t w0 = offs[j];
  for (int c = lane; c < tot; c += 64) {
    if (!first[c]) continue;
    const int32_t r = rows[c];
    cplx acc = vals[c];
    int rank = 0;
    for (int d = 0; d < tot; ++d) {
      if (d > c && rows[d] == r) acc = cadd(acc, vals[d]);
      if (first[d] && rows[d] < r) ++rank;
    }
    ci[w0 + rank] = r;
    cv[w0 + rank] = acc;
  }
}

// exact Hermiticity: every (r, c, v) has (c, r, conj v) bit for bit, or v is
// exactly zero with no partner
__global__ void herm_exact_kernel(int64_t n, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                                  const cplx* __restrict__ vals, int* __restrict__ flag) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int32_t e = outer[c]; e < outer[c + 1]; ++e) {
    const int32_t r = inner[e];
    const cplx v = vals[e];
    int32_t lo = outer[r], hi = outer[r + 1];
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (inner[mid] < (int32_t)c)
        lo = mid + 1;
      else
        hi = mid;
    }
    const bool found = lo < outer[r + 1] && inner[lo] == (int32_t)c;
    const cplx w = found ? vals[lo] : cmake(0.0, 0.0);
    if (!(v.x == w.x && v.y == -w.y)) *flag = 1;
  }
}

unsigned gridn(int64_t m) { return qd_grid(m, ALG_THREADS); }

qd_csr* new_csr(qd_ctx* ctx, int64_t n) {
  qd_csr* c = new qd_csr();
  c->ctx = ctx;
  c->n = n;
  return c;
}

// outer from per-column counts; allocates inner / vals of the total
int alloc_from_counts(qd_ctx* ctx, qd_csr* c, const int32_t* counts, int64_t** offs_out) {
  const int64_t n = c->n;
  int64_t* offs = nullptr;
  if (hipMalloc(&offs, sizeof(int64_t) * (n + 1)) != hipSuccess) return qd_set_error(ctx, QD_E_OOM, "offsets");
  *offs_out = offs;
  int rc = qd_scan_counts(ctx, counts, n, offs);
  if (rc) return rc;
  int64_t nnz = 0;
  QD_HIP(ctx, hipMemcpyAsync(&nnz, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nnz >= (int64_t)INT32_MAX)
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "nnz %lld exceeds int32 storage indices", (long long)nnz);
  c->nnz = nnz;
  if (hipMalloc(&c->d_outer, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&c->d_inner, sizeof(int32_t) * (nnz ? nnz : 1)) != hipSuccess ||
      hipMalloc(&c->d_vals, sizeof(cplx) * (nnz ? nnz : 1)) != hipSuccess)
    return qd_set_error(ctx, QD_E_OOM, "matrix arrays (nnz %lld)", (long long)nnz);
  alg_offs_to_outer<<<gridn(n + 1), ALG_THREADS, 0, ctx->stream>>>(offs, n, c->d_outer);
  return QD_OK;
}

int finish(qd_ctx* ctx, qd_csr* c, int rc, qd_csr** out, int32_t* counts, int64_t* offs, int herm_hint) {
  if (rc == QD_OK) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = qd_set_error(ctx, QD_E_HIP, "matrix algebra launch: %s", hipGetErrorString(e));
  }
  if (rc == QD_OK) {
    if (herm_hint >= 0) {
      c->hermitian = herm_hint;
    } else {
      int* dflag = nullptr;
      int h = 0;
      if (hipMalloc(&dflag, sizeof(int)) != hipSuccess) {
        rc = qd_set_error(ctx, QD_E_OOM, "hermitian flag");
      } else {
        hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream);
        if (c->n > 0)
          herm_exact_kernel<<<gridn(c->n), ALG_THREADS, 0, ctx->stream>>>(c->n, c->d_outer, c->d_inner, c->d_vals,
                                                                          dflag);
        hipMemcpyAsync(&h, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = qd_set_error(ctx, QD_E_HIP, "hermitian check");
        hipFree(dflag);
        c->hermitian = h ? 0 : 1;
      }
    }
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(counts);
  hipFree(offs);
  if (rc != QD_OK) {
    qd_csr_destroy(c);
    return rc;
  }
  *out = c;
  return QD_OK;
}

bool same_ctx_dims(qd_ctx* ctx, const qd_csr* a, const qd_csr* b) { return a && b && a->n == b->n && ctx; }

}  // namespace

extern "C" int qd_csr_zero(qd_ctx* ctx, int64_t n, qd_csr** out) {
  if (!ctx || !out || n < 0 || n >= (int64_t)INT32_MAX) return qd_set_error(ctx, QD_E_ARG, "qd_csr_zero: bad arguments");
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_csr* c = new_csr(ctx, n);
  int rc = QD_OK;
  if (hipMalloc(&c->d_outer, sizeof(int32_t) * (n + 1)) != hipSuccess || hipMalloc(&c->d_inner, 4) != hipSuccess ||
      hipMalloc(&c->d_vals, sizeof(cplx)) != hipSuccess)
    rc = qd_set_error(ctx, QD_E_OOM, "zero matrix");
  else
    zero_i32_kernel<<<gridn(n + 1), ALG_THREADS, 0, ctx->stream>>>(n + 1, c->d_outer);
  return finish(ctx, c, rc, out, nullptr, nullptr, 1);
}

extern "C" int qd_csr_identity(qd_ctx* ctx, int64_t n, qd_csr** out) {
  if (!ctx || !out || n < 0 || n >= (int64_t)INT32_MAX)
    return qd_set_error(ctx, QD_E_ARG, "qd_csr_identity: bad arguments");
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_csr* c = new_csr(ctx, n);
  c->nnz = n;
  int rc = QD_OK;
  if (hipMalloc(&c->d_outer, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&c->d_inner, sizeof(int32_t) * (n ? n : 1)) != hipSuccess ||
      hipMalloc(&c->d_vals, sizeof(cplx) * (n ? n : 1)) != hipSuccess)
    rc = qd_set_error(ctx, QD_E_OOM, "identity matrix");
  else
    identity_kernel<<<gridn(n + 1), ALG_THREADS, 0, ctx->stream>>>(n, c->d_outer, c->d_inner, c->d_vals);
  return finish(ctx, c, rc, out, nullptr, nullptr, 1);
}

extern "C" int qd_csr_scale(qd_ctx* ctx, const qd_csr* a, double c_re, double c_im, int real_scalar, qd_csr** out) {
  if (!ctx || !a || !out) return qd_set_error(ctx, QD_E_ARG, "qd_csr_scale: null argument");
  if (real_scalar && c_im != 0.0) return qd_set_error(ctx, QD_E_ARG, "qd_csr_scale: a real scalar has no imaginary part");
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_csr* c = new_csr(ctx, a->n);
  c->nnz = a->nnz;
  int rc = QD_OK;
  const int64_t m = a->nnz;
  if (hipMalloc(&c->d_outer, sizeof(int32_t) * (a->n + 1)) != hipSuccess ||
      hipMalloc(&c->d_inner, sizeof(int32_t) * (m ? m : 1)) != hipSuccess ||
      hipMalloc(&c->d_vals, sizeof(cplx) * (m ? m : 1)) != hipSuccess) {
    rc = qd_set_error(ctx, QD_E_OOM, "scaled matrix");
  } else {
    copy_i32_kernel<<<gridn(a->n + 1), ALG_THREADS, 0, ctx->stream>>>(a->n + 1, a->d_outer, c->d_outer);
    if (m) {
      copy_i32_kernel<<<gridn(m), ALG_THREADS, 0, ctx->stream>>>(m, a->d_inner, c->d_inner);
      scale_kernel<<<gridn(m), ALG_THREADS, 0, ctx->stream>>>(m, cmake(c_re, c_im), real_scalar, a->d_vals, c->d_vals);
    }
  }
  // a real multiple of a Hermitian matrix is Hermitian; otherwise check
  const int hint = (real_scalar && a->hermitian == 1) ? 1 : -1;
  return finish(ctx, c, rc, out, nullptr, nullptr, hint);
}

extern "C" int qd_csr_add(qd_ctx* ctx, const qd_csr* a, const qd_csr* b, int sign, qd_csr** out) {
  if (!out || (sign != 1 && sign != -1) || !same_ctx_dims(ctx, a, b))
    return qd_set_error(ctx, QD_E_ARG, "qd_csr_add: null argument, sign not +-1 or dimensions differ");
  *out = nullptr;
  hipSetDevice(ctx->device);
  const int64_t n = a->n;
  qd_csr* c = new_csr(ctx, n);
  int32_t* counts = nullptr;
  int64_t* offs = nullptr;
  if (hipMalloc(&counts, sizeof(int32_t) * (n ? n : 1)) != hipSuccess)
    return finish(ctx, c, qd_set_error(ctx, QD_E_OOM, "column counts"), out, nullptr, nullptr, -1);
  const double sg = sign > 0 ? 1.0 : -1.0;
  if (n)
    add_kernel<false><<<gridn(n), ALG_THREADS, 0, ctx->stream>>>(n, a->d_outer, a->d_inner, a->d_vals, b->d_outer,
                                                                  b->d_inner, b->d_vals, sg, counts, nullptr,
                                                                  nullptr, nullptr);
  int rc = alloc_from_counts(ctx, c, counts, &offs);
  if (rc == QD_OK && n)
    add_kernel<true><<<gridn(n), ALG_THREADS, 0, ctx->stream>>>(n, a->d_outer, a->d_inner, a->d_vals, b->d_outer,
                                                                 b->d_inner, b->d_vals, sg, nullptr, offs,
                                                                 c->d_inner, c->d_vals);
  return finish(ctx, c, rc, out, counts, offs, (a->hermitian == 1 && b->hermitian == 1) ? 1 : -1);
}

extern "C" int qd_csr_adjoint(qd_ctx* ctx, const qd_csr* a, qd_csr** out) {
  if (!ctx || !a || !out) return qd_set_error(ctx, QD_E_ARG, "qd_csr_adjoint: null argument");
  *out = nullptr;
  hipSetDevice(ctx->device);
  const int64_t n = a->n;
  qd_csr* c = new_csr(ctx, n);
  int32_t* counts = nullptr;
  int64_t* offs = nullptr;
  if (hipMalloc(&counts, sizeof(int32_t) * (n ? n : 1)) != hipSuccess)
    return finish(ctx, c, qd_set_error(ctx, QD_E_OOM, "row counts"), out, nullptr, nullptr, -1);
  hipMemsetAsync(counts, 0, sizeof(int32_t) * (n ? n : 1), ctx->stream);
  if (a->nnz) row_count_kernel<<<gridn(a->nnz), ALG_THREADS, 0, ctx->stream>>>(a->nnz, a->d_inner, counts);
  int rc = alloc_from_counts(ctx, c, counts, &offs);
  if (rc == QD_OK && n) {
    hipMemsetAsync(counts, 0, sizeof(int32_t) * n, ctx->stream);  // reused as the fill cursors
    adj_scatter_kernel<<<gridn(n), ALG_THREADS, 0, ctx->stream>>>(n, a->d_outer, a->d_inner, a->d_vals, offs,
                                                                  counts, c->d_inner, c->d_vals);
    col_sort_kernel<<<gridn(n), ALG_THREADS, 0, ctx->stream>>>(n, offs, c->d_inner, c->d_vals);
  }
  return finish(ctx, c, rc, out, counts, offs, a->hermitian == 1 ? 1 : (a->hermitian == 0 ? 0 : -1));
}

extern "C" int qd_csr_multiply(qd_ctx* ctx, const qd_csr* a, const qd_csr* b, qd_csr** out) {
  if (!out || !same_ctx_dims(ctx, a, b)) return qd_set_error(ctx, QD_E_ARG, "qd_csr_multiply: null argument or dimensions differ");
  *out = nullptr;
  hipSetDevice(ctx->device);
  const int64_t n = a->n;
  qd_csr* c = new_csr(ctx, n);
  int32_t* counts = nullptr;
  int64_t* offs = nullptr;
  int* dflag = nullptr;
  if (hipMalloc(&counts, sizeof(int32_t) * (n ? n : 1)) != hipSuccess || hipMalloc(&dflag, sizeof(int)) != hipSuccess) {
    hipFree(dflag);
    return finish(ctx, c, qd_set_error(ctx, QD_E_OOM, "product counts"), out, counts, nullptr, -1);
  }
  hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream);
  const unsigned grid = (unsigned)std::max<int64_t>(1, (n + MM_WAVES - 1) / MM_WAVES);
  if (n)
    spgemm_kernel<false><<<grid, 64 * MM_WAVES, 0, ctx->stream>>>(n, a->d_outer, a->d_inner, a->d_vals, b->d_outer,
                                                                   b->d_inner, b->d_vals, counts, nullptr, nullptr,
                                                                   nullptr, dflag);
  int hflag = 0;
  hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  int rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? QD_OK : qd_set_error(ctx, QD_E_HIP, "product count pass");
  if (rc == QD_OK && hflag)
    rc = qd_set_error(ctx, QD_E_CAPACITY, "a product column visits more than %d entries", MM_CAP);
  if (rc == QD_OK) rc = alloc_from_counts(ctx, c, counts, &offs);
  if (rc == QD_OK && n)
    spgemm_kernel<true><<<grid, 64 * MM_WAVES, 0, ctx->stream>>>(n, a->d_outer, a->d_inner, a->d_vals, b->d_outer,
                                                                  b->d_inner, b->d_vals, nullptr, offs, c->d_inner,
                                                                  c->d_vals, dflag);
  hipStreamSynchronize(ctx->stream);
  hipFree(dflag);
  return finish(ctx, c, rc, out, counts, offs, -1);
}
