// noalloc.hip -- an expectation value that allocates nothing.
//
// No reference counterpart: for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  qd_expect_noalloc computes <psi|H|psi> of a stored-matrix
// operator with the CSR-stream product fused with the dot (the product
// vector is never stored): each block of 128 rows writes its sum to a
// file-scope device buffer of fixed size (4 MiB), then one block adds the
// sums in a fixed order and writes the result.  Operators of at most
// 2^18 * 128 = 33,554,432 rows are accepted.  Every call on one GPU shares that buffer:
// calls that overlap on one GPU (two streams, two threads) race and must be
// serialised by the caller.  Both kernels run asynchronously on the
// context's stream and write no memset node, so the call may be captured.
#include <mutex>

#include "qd_internal.hpp"
#include "wave.hpp"

namespace {
constexpr int64_t kNoallocBlocks = (int64_t)1 << 18;  // one partial per block of the fused product
__device__ __align__(256) cplx scratch[kNoallocBlocks];

// out = sum part[0 .. nb) in a fixed order (one block)
constexpr int NA_T = 1024;
__global__ void __launch_bounds__(NA_T) noalloc_sum_kernel(const cplx* __restrict__ part, int nb,
                                                           double* __restrict__ out) {
  __shared__ cplx sh[NA_T / 64];
  cplx acc = cmake(0.0, 0.0);
  for (int b = threadIdx.x; b < nb; b += NA_T) acc = cadd(acc, part[b]);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    cplx t = sh[0];
    for (int q = 1; q < NA_T / 64; ++q) t = cadd(t, sh[q]);
    out[0] = t.x;
    out[1] = t.y;
  }
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
cplx* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(qd_ctx* ctx, cplx** out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return qd_set_error(ctx, QD_E_ARG, "device %d", ctx->device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return qd_set_error(ctx, QD_E_HIP, "hipGetDevice");
    if (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess)
      return qd_set_error(ctx, QD_E_HIP, "hipSetDevice(%d)", ctx->device);
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != ctx->device) hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return qd_set_error(ctx, QD_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = static_cast<cplx*>(p);
  }
  *out = g_scratch_ptr[ctx->device];
  return QD_OK;
}
}  // namespace

extern "C" int qd_expect_noalloc(qd_ctx* ctx, qd_linop* op, const double* d_psi, double* d_out) {
  if (!ctx || !op || !d_out || (op->n && !d_psi)) return qd_set_error(ctx, QD_E_ARG, "qd_expect_noalloc: null argument");
  if (op->kind != QD_LINOP_CSR)
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "qd_expect_noalloc: stored-matrix operators only (qd_linop_from_csr)");
  const int64_t n = op->n;
  const int64_t nb = qd_linop_dot_blocks(n);
  if (nb > kNoallocBlocks)
    return qd_set_error(ctx, QD_E_CAPACITY, "qd_expect_noalloc: dimension %lld exceeds the scratch limit of %lld rows",
                        (long long)n, (long long)(kNoallocBlocks * qd_linop_dot_rows_per_block()));
  cplx* part = nullptr;
  int rc = noalloc_scratch(ctx, &part);
  if (rc) return rc;
  if (n > 0) {
    rc = qd_linop_apply_dot(op, (const cplx*)d_psi, nullptr, part, false);
    if (rc) return rc;
  }
  noalloc_sum_kernel<<<1, NA_T, 0, ctx->stream>>>(part, (int)(n > 0 ? nb : 0), d_out);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

