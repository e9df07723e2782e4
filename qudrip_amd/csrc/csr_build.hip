// csr_build.hip -- K3/K4: two-pass count/fill build of `ham >> index`.
//
// COMPILED WITH -ffp-contract=off: every complex product below is two
// multiplies and one add/sub per component, like the reference's host code.
//
// One thread owns one source state (one column of the reference's ColMajor
// matrix, operator.hpp:250-262).  It walks the terms in program order and,
// per term, the leaves of the operator tree in the reference's leaf order
// (operator.hpp:135-206: ops right-to-left, first-applied op most
// significant).  Each surviving leaf (alive, |v| > 1e-10, target inside the
// sector) is folded into a per-thread list kept sorted by target row:
//   - leaves of one term with equal targets are summed in leaf order first
//     (Eigen setFromTriplets, operator.hpp:264),
//   - then acc = acc + c_k * v_k, a new entry starting from (+0, +0)
//     (operator_abel.hpp:28-31 through Eigen's sparse sum).
// Pass 1 counts the distinct targets per column; an exclusive scan gives the
// column pointers; pass 2 writes rows ascending and values.  When every
// column has at most 32 distinct targets (the common case) one expansion is
// enough: the fill pass writes each column's sorted list to a staging area
// at a fixed stride of 32 together with its count, the scan gives the column
// pointers, and a compaction copies the lists into place.  Same per-column
// arithmetic, so the same bits; a longer column makes the staged pass flag an
// overflow and the two-pass build runs instead.
#include "qd_internal.hpp"

// sorted per-thread list: find or insert key; returns position or -1 when full
template <int CAP, bool FILL>
__device__ inline int list_add(int32_t* key, cplx* acc, int& ne, int32_t k, cplx prod) {
  int lo = 0, hi = ne;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (key[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < ne && key[lo] == k) {
    if (FILL) acc[lo] = cadd(acc[lo], prod);
    return lo;
  }
  if (ne >= CAP) return -1;
  for (int q = ne; q > lo; --q) {
    key[q] = key[q - 1];
    if (FILL) acc[q] = acc[q - 1];
  }
  key[lo] = k;
  if (FILL) acc[lo] = cadd(cmake(0.0, 0.0), prod);
  ++ne;
  return lo;
}

#define MAXDUP 32

// mode gates of a composite index (MODES kernels only)
struct ModeArgs {
  int mdim;
  const cplx* mats;
};

template <int CAP, bool FILL, bool MODES, bool STAGE = false>
__global__ void __launch_bounds__(128)
    csr_rows_kernel(SectorDesc sec, const uint64_t* __restrict__ states, int64_t n,
                    const DevTerm* __restrict__ terms, int nterms, const DevOp* __restrict__ ops,
                    int32_t* __restrict__ counts, const int64_t* __restrict__ offs, int32_t* __restrict__ inner,
                    cplx* __restrict__ vals, int* __restrict__ flag, ModeArgs ma) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t s = states[i];
  int32_t key[CAP];
  cplx acc[FILL ? CAP : 1];
  int ne = 0;
  bool overflow = false;
  for (int k = 0; k < nterms && !overflow; ++k) {
    const DevTerm t = terms[k];
    const DevOp* tops = ops + t.op_begin;
    int nbitops = 0;
    if (!MODES)
      for (int q = 0; q < t.n_ops; ++q) nbitops += (tops[q].kind == QD_OP_BIT);
    auto leaf = [&](int L, uint64_t* x, cplx* v) {
      return MODES ? eval_leaf_modes(tops, t.n_ops, s, L, ma.mdim, ma.mats, x, v)
                   : eval_leaf(tops, t.n_ops, nbitops, s, L, x, v);
    };
    if (!t.may_dup) {
      for (int L = 0; L < t.n_leaves; ++L) {
        uint64_t x;
        cplx v;
        if (!leaf(L, &x, &v)) continue;
        if (!(hypot(v.x, v.y) > QD_LEAF_EPS)) continue;
        uint64_t r = qd_rank(sec, states, x);
        if (r >= (uint64_t)n) continue;
        cplx prod = FILL ? cmul(t.coeff, v) : v;
        if (list_add<CAP, FILL>(key, acc, ne, (int32_t)r, prod) < 0) {
          overflow = true;
          break;
        }
      }
    } else {
      int32_t tk[MAXDUP];
      cplx tv[MAXDUP];
      int nt = 0;
      for (int L = 0; L < t.n_leaves; ++L) {
        uint64_t x;
        cplx v;
        if (!leaf(L, &x, &v)) continue;
        if (!(hypot(v.x, v.y) > QD_LEAF_EPS)) continue;
        uint64_t r = qd_rank(sec, states, x);
        if (r >= (uint64_t)n) continue;
        int f = -1;
        for (int q = 0; q < nt; ++q)
          if (tk[q] == (int32_t)r) {
            f = q;
            break;
          }
        if (f >= 0) {
          tv[f] = cadd(tv[f], v);
        } else {
          if (nt >= MAXDUP) {
            overflow = true;
            break;
          }
          tk[nt] = (int32_t)r;
          tv[nt] = v;
          ++nt;
        }
      }
      for (int q = 0; q < nt && !overflow; ++q) {
        cplx prod = FILL ? cmul(t.coeff, tv[q]) : tv[q];
        if (list_add<CAP, FILL>(key, acc, ne, tk[q], prod) < 0) overflow = true;
      }
    }
  }
  if (overflow) {
    atomicOr(flag, 1);
    if (!FILL || STAGE) counts[i] = 0;
    return;
  }
  if (!FILL) {
    counts[i] = ne;
  } else {
    if (STAGE) counts[i] = ne;
    const int64_t base = STAGE ? i * CAP : offs[i];  // staging: fixed stride CAP per column
    for (int e = 0; e < ne; ++e) {
      inner[base + e] = key[e];
      vals[base + e] = acc[e];
    }
  }
}

// staged lists -> their final places (one thread per staging slot; slots
// past a column's count are skipped)
template <int CAP>
__global__ void stage_compact_kernel(int64_t n, const int32_t* __restrict__ counts, const int64_t* __restrict__ offs,
                                     const int32_t* __restrict__ s_inner, const cplx* __restrict__ s_vals,
                                     int32_t* __restrict__ inner, cplx* __restrict__ vals) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * CAP) return;
  const int64_t col = e / CAP;
  const int k = (int)(e % CAP);
  if (k >= counts[col]) return;
  inner[offs[col] + k] = s_inner[e];
  vals[offs[col] + k] = s_vals[e];
}

__global__ void max_count_kernel(const int32_t* __restrict__ c, int64_t n, int* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int m = 0;
  for (; i < n; i += (int64_t)gridDim.x * blockDim.x) m = max(m, c[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void offs_to_outer(const int64_t* __restrict__ offs, int64_t n, int32_t* __restrict__ outer) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) outer[i] = (int32_t)offs[i];
}

// Hermiticity: every stored (r, c, v) must have (c, r, v') with
// |v - conj(v')| <= 1e-12 max(|v|, |v'|), a missing partner counting as 0.
__global__ void herm_check_kernel(int64_t n, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                                  const cplx* __restrict__ vals, int* __restrict__ flag) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int32_t e = outer[c]; e < outer[c + 1]; ++e) {
    int32_t r = inner[e];
    cplx v = vals[e];
    int32_t lo = outer[r], hi = outer[r + 1];
    while (lo < hi) {
      int32_t mid = (lo + hi) >> 1;
      if (inner[mid] < (int32_t)c)
        lo = mid + 1;
      else
        hi = mid;
    }
    cplx w = cmake(0.0, 0.0);
    if (lo < outer[r + 1] && inner[lo] == (int32_t)c) w = vals[lo];
    double dx = v.x - w.x, dy = v.y + w.y;
    double d = sqrt(dx * dx + dy * dy);
    double m = fmax(sqrt(v.x * v.x + v.y * v.y), sqrt(w.x * w.x + w.y * w.y));
    if (d > 1e-12 * m) {
      atomicOr(flag, 1);
      return;
    }
  }
}

template <int CAP, bool MODES>
static int run_pass(qd_ctx* ctx, qd_basis* basis, int64_t n, const DevTerm* dterms, int nterms, const DevOp* dops,
                    bool fill, int32_t* counts, const int64_t* offs, int32_t* inner, cplx* vals, int* dflag,
                    ModeArgs ma) {
  const int B = 128;
  if (fill)
    csr_rows_kernel<CAP, true, MODES><<<qd_grid(n, B), B, 0, ctx->stream>>>(
        basis->sec, basis->d_states, n, dterms, nterms, dops, counts, offs, inner, vals, dflag, ma);
  else
    csr_rows_kernel<CAP, false, MODES><<<qd_grid(n, B), B, 0, ctx->stream>>>(
        basis->sec, basis->d_states, n, dterms, nterms, dops, counts, offs, inner, vals, dflag, ma);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

template <bool MODES>
static int pass_cap(int cap, qd_ctx* ctx, qd_basis* basis, int64_t n, const DevTerm* dterms, int nterms,
                    const DevOp* dops, bool fill, int32_t* counts, const int64_t* offs, int32_t* inner, cplx* vals,
                    int* dflag, ModeArgs ma) {
  if (cap <= 32)
    return run_pass<32, MODES>(ctx, basis, n, dterms, nterms, dops, fill, counts, offs, inner, vals, dflag, ma);
  if (cap <= 64)
    return run_pass<64, MODES>(ctx, basis, n, dterms, nterms, dops, fill, counts, offs, inner, vals, dflag, ma);
  if (cap <= 256)
    return run_pass<256, MODES>(ctx, basis, n, dterms, nterms, dops, fill, counts, offs, inner, vals, dflag, ma);
  return run_pass<1024, MODES>(ctx, basis, n, dterms, nterms, dops, fill, counts, offs, inner, vals, dflag, ma);
}

static int pass_dispatch(int cap, qd_ctx* ctx, qd_basis* basis, int64_t n, const DevTerm* dterms, int nterms,
                         const DevOp* dops, bool fill, int32_t* counts, const int64_t* offs, int32_t* inner,
                         cplx* vals, int* dflag, ModeArgs ma) {
  if (ma.mdim > 1)
    return pass_cap<true>(cap, ctx, basis, n, dterms, nterms, dops, fill, counts, offs, inner, vals, dflag, ma);
  return pass_cap<false>(cap, ctx, basis, n, dterms, nterms, dops, fill, counts, offs, inner, vals, dflag, ma);
}

constexpr int STAGE_CAP = 32;
// staging is used while it stays under this many bytes (the two-pass build
// needs no staging)
constexpr size_t STAGE_MAX_BYTES = (size_t)16 << 30;

static int run_stage(qd_ctx* ctx, qd_basis* basis, int64_t n, const DevTerm* dterms, int nterms, const DevOp* dops,
                     int32_t* counts, int32_t* s_inner, cplx* s_vals, int* dflag, ModeArgs ma) {
  const int B = 128;
  if (ma.mdim > 1)
    csr_rows_kernel<STAGE_CAP, true, true, true><<<qd_grid(n, B), B, 0, ctx->stream>>>(
        basis->sec, basis->d_states, n, dterms, nterms, dops, counts, nullptr, s_inner, s_vals, dflag, ma);
  else
    csr_rows_kernel<STAGE_CAP, true, false, true><<<qd_grid(n, B), B, 0, ctx->stream>>>(
        basis->sec, basis->d_states, n, dterms, nterms, dops, counts, nullptr, s_inner, s_vals, dflag, ma);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

int qd_csr_build_impl(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog, int adjoint, qd_csr** out) {
  *out = nullptr;
  int64_t n = (int64_t)basis->sec.total;
  if (n >= (int64_t)INT32_MAX)
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "sector of %lld states exceeds int32 CSR indices", (long long)n);
  if (prog->nbits > basis->sec.nbits)
    return qd_set_error(ctx, QD_E_ARG, "program acts on %d bits, basis has %d", prog->nbits, basis->sec.nbits);
  // a program without mode gates acts as the identity on a basis's mode
  const int bmdim = basis->sec.mdim > 1 ? (int)basis->sec.mdim : 1;
  if (prog->mdim > 1 && bmdim != prog->mdim)
    return qd_set_error(ctx, QD_E_ARG, "program mode dimension %d, basis %d", prog->mdim, bmdim);
  const ModeArgs ma{bmdim, prog->d_mats};
  const DevTerm* dterms = prog->d_terms;
  const DevOp* dops = prog->d_ops;
  if (adjoint) {
    int rc = qd_termprog_ensure_adjoint(prog);
    if (rc) return rc;
    dterms = prog->d_adj_terms;
    dops = prog->d_adj_ops;
  }
  int nterms = (int)prog->h_terms.size();
  hipSetDevice(ctx->device);
  qd_csr* c = new qd_csr();
  c->ctx = ctx;
  c->n = n;
  c->basis = basis;
  c->prog = prog;
  int32_t* counts = nullptr;
  int64_t* offs = nullptr;
  int* dflag = nullptr;
  int32_t* s_inner = nullptr;
  cplx* s_vals = nullptr;
  int rc = QD_OK;
  int hflag[2] = {0, 0};
  int64_t nnz = 0;
  int cap = 32;
  auto fail = [&](int code, const char* what) {
    hipStreamSynchronize(ctx->stream);
    hipFree(s_inner);
    hipFree(s_vals);
    hipFree(counts);
    hipFree(offs);
    hipFree(dflag);
    hipFree(c->d_outer);
    hipFree(c->d_inner);
    hipFree(c->d_vals);
    delete c;
    if (code == QD_E_OOM) return qd_set_error(ctx, code, "csr build: allocation failed (%s)", what);
    return code;
  };
  if (hipMalloc(&counts, sizeof(int32_t) * (n ? n : 1)) != hipSuccess ||
      hipMalloc(&offs, sizeof(int64_t) * (n + 1)) != hipSuccess || hipMalloc(&dflag, 2 * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->d_outer, sizeof(int32_t) * (n + 1)) != hipSuccess)
    return fail(QD_E_OOM, "counts");
  // one expansion with staged lists, when the staging fits
  const size_t stage_bytes = (size_t)n * STAGE_CAP * (sizeof(int32_t) + sizeof(cplx));
  if (n > 0 && stage_bytes <= STAGE_MAX_BYTES && hipMalloc(&s_inner, sizeof(int32_t) * (size_t)n * STAGE_CAP) == hipSuccess &&
      hipMalloc(&s_vals, sizeof(cplx) * (size_t)n * STAGE_CAP) == hipSuccess) {
    hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream);
    rc = run_stage(ctx, basis, n, dterms, nterms, dops, counts, s_inner, s_vals, dflag, ma);
    if (rc) return fail(rc, "staged fill");
    hipMemcpyAsync(hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(qd_set_error(ctx, QD_E_HIP, "staged fill"), "");
    if (!hflag[0]) {
      rc = qd_scan_counts(ctx, counts, n, offs);
      if (rc) return fail(rc, "scan");
      hipMemcpyAsync(&nnz, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream);
      if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(qd_set_error(ctx, QD_E_HIP, "staged scan"), "");
      if (nnz >= (int64_t)INT32_MAX)
        return fail(qd_set_error(ctx, QD_E_UNSUPPORTED, "nnz %lld exceeds int32 storage indices", (long long)nnz), "");
      c->nnz = nnz;
      offs_to_outer<<<qd_grid(n + 1, 256), 256, 0, ctx->stream>>>(offs, n, c->d_outer);
      if (hipMalloc(&c->d_inner, sizeof(int32_t) * (nnz ? nnz : 1)) != hipSuccess ||
          hipMalloc(&c->d_vals, sizeof(cplx) * (nnz ? nnz : 1)) != hipSuccess)
        return fail(QD_E_OOM, "values");
      stage_compact_kernel<STAGE_CAP><<<qd_grid(n * STAGE_CAP, 256), 256, 0, ctx->stream>>>(
          n, counts, offs, s_inner, s_vals, c->d_inner, c->d_vals);
      QD_CHECK_LAUNCH(ctx);
      hipStreamSynchronize(ctx->stream);
      hipFree(s_inner);
      hipFree(s_vals);
      hipFree(counts);
      hipFree(offs);
      hipFree(dflag);
      *out = c;
      return QD_OK;
    }
    cap = 2 * STAGE_CAP;  // some column holds more than 32 targets: two passes from 64
  } else if (n > 0 && stage_bytes <= STAGE_MAX_BYTES) {
    (void)hipGetLastError();  // a failed staging allocation is not an error: two passes instead
  }
  hipFree(s_inner);
  hipFree(s_vals);
  s_inner = nullptr;
  s_vals = nullptr;
  // pass 1 with growing capacity
  for (;;) {
    hipMemsetAsync(dflag, 0, 2 * sizeof(int), ctx->stream);
    if (n > 0) rc = pass_dispatch(cap, ctx, basis, n, dterms, nterms, dops, false, counts, nullptr, nullptr, nullptr, dflag, ma);
    if (rc) return fail(rc, "count");
    hipMemcpyAsync(hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    if (!hflag[0]) break;
    if (cap >= 1024) return fail(qd_set_error(ctx, QD_E_CAPACITY, "a column has more than 1024 distinct rows"), "");
    cap = (cap <= 32) ? 64 : (cap <= 64 ? 256 : 1024);
  }
  // exact per-column maximum decides the fill instantiation
  if (n > 0) {
    hipMemsetAsync(dflag + 1, 0, sizeof(int), ctx->stream);
    max_count_kernel<<<256, 256, 0, ctx->stream>>>(counts, n, dflag + 1);
    hipMemcpyAsync(hflag + 1, dflag + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  }
  rc = qd_scan_counts(ctx, counts, n, offs);
  if (rc) return fail(rc, "scan");
  hipMemcpyAsync(&nnz, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(qd_set_error(ctx, QD_E_HIP, "count pass"), "");
  if (nnz >= (int64_t)INT32_MAX)
    return fail(qd_set_error(ctx, QD_E_UNSUPPORTED, "nnz %lld exceeds int32 storage indices", (long long)nnz), "");
  c->nnz = nnz;
  offs_to_outer<<<qd_grid(n + 1, 256), 256, 0, ctx->stream>>>(offs, n, c->d_outer);
  if (hipMalloc(&c->d_inner, sizeof(int32_t) * (nnz ? nnz : 1)) != hipSuccess ||
      hipMalloc(&c->d_vals, sizeof(cplx) * (nnz ? nnz : 1)) != hipSuccess)
    return fail(QD_E_OOM, "values");
  if (n > 0) {
    hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream);
    rc = pass_dispatch(hflag[1] > cap ? 1024 : (hflag[1] > 0 ? hflag[1] : 1), ctx, basis, n, dterms, nterms, dops,
                       true, counts, offs, c->d_inner, c->d_vals, dflag, ma);
    if (rc) return fail(rc, "fill");
    hipMemcpyAsync(hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    if (hflag[0]) return fail(qd_set_error(ctx, QD_E_CAPACITY, "fill pass overflow"), "");
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(counts);
  hipFree(offs);
  hipFree(dflag);
  *out = c;
  return QD_OK;
}

static int check_hermitian(qd_csr* c) {
  qd_ctx* ctx = c->ctx;
  int* dflag = nullptr;
  if (hipMalloc(&dflag, sizeof(int)) != hipSuccess) return qd_set_error(ctx, QD_E_OOM, "hermitian flag");
  hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream);
  if (c->n > 0)
    herm_check_kernel<<<qd_grid(c->n, 256), 256, 0, ctx->stream>>>(c->n, c->d_outer, c->d_inner, c->d_vals, dflag);
  int h = 0;
  hipMemcpyAsync(&h, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  hipFree(dflag);
  if (e != hipSuccess) return qd_set_error(ctx, QD_E_HIP, "hermitian check: %s", hipGetErrorString(e));
  c->hermitian = h ? 0 : 1;
  return QD_OK;
}

extern "C" int qd_csr_build(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog, qd_csr** out) {
  if (!ctx || !basis || !prog || !out) return qd_set_error(ctx, QD_E_ARG, "qd_csr_build: null argument");
  int rc = qd_csr_build_impl(ctx, basis, prog, 0, out);
  if (rc) return rc;
  rc = check_hermitian(*out);
  if (rc) {
    qd_csr_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

extern "C" int qd_csr_destroy(qd_csr* c) {
  if (!c) return QD_OK;
  hipSetDevice(c->ctx->device);
  hipStreamSynchronize(c->ctx->stream);
  hipFree(c->d_outer);
  hipFree(c->d_inner);
  hipFree(c->d_vals);
  if (c->adj) qd_csr_destroy(c->adj);
  delete c;
  return QD_OK;
}

extern "C" int qd_csr_info(const qd_csr* c, int64_t* n, int64_t* nnz, int* hermitian) {
  if (!c) return QD_E_ARG;
  if (n) *n = c->n;
  if (nnz) *nnz = c->nnz;
  if (hermitian) *hermitian = c->hermitian;
  return QD_OK;
}

extern "C" int qd_csr_download(qd_csr* c, int32_t* h_outer, int32_t* h_inner, double* h_vals) {
  if (!c) return QD_E_ARG;
  qd_ctx* ctx = c->ctx;
  if (h_outer) QD_HIP(ctx, hipMemcpyAsync(h_outer, c->d_outer, sizeof(int32_t) * (c->n + 1), hipMemcpyDeviceToHost, ctx->stream));
  if (h_inner && c->nnz)
    QD_HIP(ctx, hipMemcpyAsync(h_inner, c->d_inner, sizeof(int32_t) * c->nnz, hipMemcpyDeviceToHost, ctx->stream));
  if (h_vals && c->nnz)
    QD_HIP(ctx, hipMemcpyAsync(h_vals, c->d_vals, sizeof(cplx) * c->nnz, hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

// device arrays of the stored matrix (for callers that run their own kernels)
extern "C" int qd_csr_device_arrays(qd_csr* c, const int32_t** outer, const int32_t** inner, const double** vals) {
  if (!c) return QD_E_ARG;
  if (outer) *outer = c->d_outer;
  if (inner) *inner = c->d_inner;
  if (vals) *vals = (const double*)c->d_vals;
  return QD_OK;
}
