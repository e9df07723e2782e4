// ctx.hip -- context, error reporting, device memory helpers, prefix scan.
#include <cstdarg>
#include <cstdio>

#include "qd_internal.hpp"

int qd_set_error(qd_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

extern "C" int qd_abi_version(void) { return 1; }

extern "C" int qd_ctx_create(int device, qd_ctx** out) {
  if (!out) return QD_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return QD_E_HIP;
  if (device < 0 || device >= ndev) return QD_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return QD_E_HIP;
  qd_ctx* c = new qd_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return QD_E_HIP;
  }
  c->own_stream = true;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  *out = c;
  return QD_OK;
}

// Bind a context to an existing stream (e.g. the caller's framework stream),
// so that work on both is ordered without extra synchronisation.
extern "C" int qd_ctx_create_on_stream(int device, void* stream, qd_ctx** out) {
  int rc = qd_ctx_create(device, out);
  if (rc != QD_OK) return rc;
  qd_ctx* c = *out;
  hipStreamDestroy(c->stream);
  c->stream = (hipStream_t)stream;
  c->own_stream = false;
  return QD_OK;
}

extern "C" int qd_ctx_destroy(qd_ctx* ctx) {
  if (!ctx) return QD_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) hipFree(ctx->scratch);
  if (ctx->pinned) hipHostFree(ctx->pinned);
  if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return QD_OK;
}

extern "C" const char* qd_last_error(const qd_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int qd_dev_alloc(qd_ctx* ctx, size_t bytes, void** p) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  qd_devpool* P = ctx->pool;
  if (!P) return hipMalloc(p, bytes) == hipSuccess ? QD_OK : QD_E_OOM;
  std::vector<void*> release;
  {
    std::lock_guard<std::mutex> lk(P->mu);
    auto it = P->free_blocks.lower_bound(bytes);
    if (it != P->free_blocks.end() && it->first <= bytes + bytes / 8) {
      *p = it->second;
      P->cached -= it->first;
      P->in_use += it->first;
      P->live[*p] = it->first;
      P->free_blocks.erase(it);
      return QD_OK;
    }
    // make room under the cap by dropping cached blocks, largest first
    while (!P->free_blocks.empty() && P->in_use + P->cached + bytes > P->cap) {
      auto last = std::prev(P->free_blocks.end());
      P->cached -= last->first;
      release.push_back(last->second);
      P->free_blocks.erase(last);
    }
  }
  for (void* q : release) hipFree(q);
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    return QD_E_OOM;
  }
  std::lock_guard<std::mutex> lk(P->mu);
  P->in_use += bytes;
  P->live[*p] = bytes;
  return QD_OK;
}

void qd_dev_free(qd_ctx* ctx, void* p) {
  if (!p) return;
  qd_devpool* P = ctx->pool;
  if (P) {
    std::lock_guard<std::mutex> lk(P->mu);
    auto it = P->live.find(p);
    if (it != P->live.end()) {
      const size_t sz = it->second;
      P->live.erase(it);
      P->in_use -= sz;
      P->free_blocks.emplace(sz, p);
      P->cached += sz;
      return;
    }
  }
  hipFree(p);
}

void qd_devpool_release_all(qd_devpool* P) {
  if (!P) return;
  std::lock_guard<std::mutex> lk(P->mu);
  for (auto& kv : P->free_blocks) hipFree(kv.second);
  P->free_blocks.clear();
  P->cached = 0;
}

extern "C" int qd_ctx_sync(qd_ctx* ctx) {
  if (!ctx) return QD_E_ARG;
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

extern "C" void* qd_ctx_stream(qd_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int qd_malloc(qd_ctx* ctx, size_t bytes, void** d_ptr) {
  if (!ctx || !d_ptr) return QD_E_ARG;
  *d_ptr = nullptr;
  if (bytes == 0) bytes = 256;
  hipSetDevice(ctx->device);
  hipError_t e = hipMalloc(d_ptr, bytes);
  if (e != hipSuccess) return qd_set_error(ctx, QD_E_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return QD_OK;
}

extern "C" int qd_free(qd_ctx* ctx, void* d_ptr) {
  if (!ctx) return QD_E_ARG;
  if (!d_ptr) return QD_OK;
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  QD_HIP(ctx, hipFree(d_ptr));
  return QD_OK;
}

extern "C" int qd_memcpy_h2d(qd_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  if (!ctx || (!d_dst && bytes) || (!h_src && bytes)) return QD_E_ARG;
  QD_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

extern "C" int qd_memcpy_d2h(qd_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
  if (!ctx || (!d_src && bytes) || (!h_dst && bytes)) return QD_E_ARG;
  QD_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

int qd_ctx_scratch(qd_ctx* ctx, size_t bytes, void** p) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
      hipFree(ctx->scratch);
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    size_t nb = bytes < (1u << 20) ? (1u << 20) : bytes;
    hipError_t e = hipMalloc(&ctx->scratch, nb);
    if (e != hipSuccess) return qd_set_error(ctx, QD_E_OOM, "scratch hipMalloc(%zu): %s", nb, hipGetErrorString(e));
    ctx->scratch_bytes = nb;
  }
  *p = ctx->scratch;
  return QD_OK;
}

int qd_ctx_pinned(qd_ctx* ctx, size_t count, double** p) {
  if (count > ctx->pinned_count) {
    if (ctx->pinned) {
      QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
      hipHostFree(ctx->pinned);
      ctx->pinned = nullptr;
      ctx->pinned_count = 0;
    }
    size_t nc = count < 4096 ? 4096 : count;
    QD_HIP(ctx, hipHostMalloc((void**)&ctx->pinned, nc * sizeof(double), hipHostMallocDefault));
    ctx->pinned_count = nc;
  }
  *p = ctx->pinned;
  return QD_OK;
}

// ------------------------------------------------------------------ scan
// Exclusive scan of int32 counts into int64 offsets[n+1]: per-block sums,
// a single-block scan of the block sums, then a per-block local scan.
#define SCAN_BLOCK 1024
#define SCAN_ITEMS 4  // items per thread

__global__ void scan_block_sums(const int32_t* __restrict__ counts, int64_t n, int64_t* __restrict__ bsum) {
  __shared__ int64_t red[SCAN_BLOCK / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK * SCAN_ITEMS;
  int64_t s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)k * SCAN_BLOCK + threadIdx.x;
    if (i < n) s += counts[i];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += red[w];
    bsum[blockIdx.x] = t;
  }
}

// single block: exclusive scan of nb block sums in place, total to out_total
__global__ void scan_bsums(int64_t* bsum, int64_t nb, int64_t* out_total) {
  __shared__ int64_t carry;
  __shared__ int64_t wsum[SCAN_BLOCK / 64];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nb; base += SCAN_BLOCK) {
    int64_t i = base + threadIdx.x;
    int64_t v = (i < nb) ? bsum[i] : 0;
    // inclusive wave scan
    int lane = threadIdx.x & 63;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      int64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t acc = 0;
      for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        int64_t t = wsum[w];
        wsum[w] = acc;
        acc += t;
      }
    }
    __syncthreads();
    int64_t excl = carry + wsum[threadIdx.x >> 6] + x - v;
    if (i < nb) bsum[i] = excl;
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ void scan_local(const int32_t* __restrict__ counts, int64_t n, const int64_t* __restrict__ bsum,
                           int64_t* __restrict__ offsets) {
  __shared__ int64_t wsum[SCAN_BLOCK / 64];
  __shared__ int64_t carry;
  int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK * SCAN_ITEMS;
  if (threadIdx.x == 0) carry = bsum[blockIdx.x];
  __syncthreads();
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)k * SCAN_BLOCK + threadIdx.x;
    int64_t v = (i < n) ? counts[i] : 0;
    int lane = threadIdx.x & 63;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      int64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t acc = 0;
      for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        int64_t t = wsum[w];
        wsum[w] = acc;
        acc += t;
      }
    }
    __syncthreads();
    int64_t excl = carry + wsum[threadIdx.x >> 6] + x - v;
    if (i < n) offsets[i] = excl;
    __syncthreads();
    if (threadIdx.x == SCAN_BLOCK - 1) carry = excl + v;
    __syncthreads();
  }
}

int qd_scan_counts(qd_ctx* ctx, const int32_t* d_counts, int64_t n, int64_t* d_offsets) {
  int64_t per = (int64_t)SCAN_BLOCK * SCAN_ITEMS;
  int64_t nb = (n + per - 1) / per;
  if (nb < 1) nb = 1;
  void* sp = nullptr;
  int rc = qd_ctx_scratch(ctx, sizeof(int64_t) * (nb + 1), &sp);
  if (rc) return rc;
  int64_t* bsum = (int64_t*)sp;
  scan_block_sums<<<(unsigned)nb, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, bsum);
  QD_CHECK_LAUNCH(ctx);
  scan_bsums<<<1, SCAN_BLOCK, 0, ctx->stream>>>(bsum, nb, d_offsets + n);
  QD_CHECK_LAUNCH(ctx);
  scan_local<<<(unsigned)nb, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, bsum, d_offsets);
  QD_CHECK_LAUNCH(ctx);
  return QD_OK;
}

// ------------------------------------------------------------ bandwidth
// Streaming kernels of known traffic: the ceiling that HBM-bound kernels are
// read against (non-temporal loads, 4 in flight per lane, grid-stride).
namespace {
typedef double bw_d2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) bw_read_kernel(const bw_d2* __restrict__ a, int64_t n, double* __restrict__ sink) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256 * 4;
  for (int64_t base = (int64_t)blockIdx.x * 256 * 4 + threadIdx.x; base < n; base += stride) {
    bw_d2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = base + (int64_t)u * 256;
      v[u] = i < n ? __builtin_nontemporal_load(a + i) : bw_d2{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) s += v[u].x + v[u].y;
  }
  if (s == 1.25e300) sink[blockIdx.x] = s;  // keeps the loads; never true for the zeroed input
}

__global__ void __launch_bounds__(256) bw_copy_kernel(const bw_d2* __restrict__ a, bw_d2* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256 * 4;
  for (int64_t base = (int64_t)blockIdx.x * 256 * 4 + threadIdx.x; base < n; base += stride) {
    bw_d2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n) v[u] = __builtin_nontemporal_load(a + i);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n) __builtin_nontemporal_store(v[u], b + i);
    }
  }
}
}  // namespace

extern "C" int qd_stream_bandwidth(qd_ctx* ctx, size_t bytes, int reps, double* read_gbs, double* copy_gbs) {
  if (!ctx || bytes < 4096 || reps < 1 || !read_gbs || !copy_gbs)
    return qd_set_error(ctx, QD_E_ARG, "qd_stream_bandwidth: bad arguments");
  hipSetDevice(ctx->device);
  const int64_t n = (int64_t)(bytes / 16);
  bw_d2 *a = nullptr, *b = nullptr;
  double* sink = nullptr;
  const unsigned grid = 8192;
  if (hipMalloc(&a, 16 * (size_t)n) != hipSuccess) return qd_set_error(ctx, QD_E_OOM, "bandwidth buffers");
  if (hipMalloc(&b, 16 * (size_t)n) != hipSuccess || hipMalloc(&sink, sizeof(double) * grid) != hipSuccess) {
    hipFree(a);
    hipFree(b);
    return qd_set_error(ctx, QD_E_OOM, "bandwidth buffers");
  }
  hipStream_t st = ctx->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = QD_OK;
  float ms_read = 0.f, ms_copy = 0.f;
  if (hipMemsetAsync(a, 0, 16 * (size_t)n, st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = QD_E_HIP;
  } else {
    bw_read_kernel<<<grid, 256, 0, st>>>(a, n, sink);
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) bw_read_kernel<<<grid, 256, 0, st>>>(a, n, sink);
    hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms_read, e0, e1) != hipSuccess) rc = QD_E_HIP;
    bw_copy_kernel<<<grid, 256, 0, st>>>(a, b, n);
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) bw_copy_kernel<<<grid, 256, 0, st>>>(a, b, n);
    hipEventRecord(e1, st);
    if (rc == QD_OK &&
        (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms_copy, e0, e1) != hipSuccess))
      rc = QD_E_HIP;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(st);
  hipFree(a);
  hipFree(b);
  hipFree(sink);
  if (rc) return qd_set_error(ctx, rc, "qd_stream_bandwidth: HIP call failed");
  const double moved = 16.0 * (double)n;
  *read_gbs = moved * reps / (ms_read * 1e-3) / 1e9;
  *copy_gbs = 2.0 * moved * reps / (ms_copy * 1e-3) / 1e9;
  return QD_OK;
}
