// basis.hip -- sector bases: combinatorial-number-system unrank/rank (K1/K2)
// and a predicate-scan fallback for constraint sets that are not disjoint
// contiguous ranges.
//
// Reference semantics (conservation.hpp:212-231, index.hpp:187-261): the
// sector is every x in [0, 2^nbits) satisfying all popcount constraints, in
// ascending x; rank(x) is the position in that list, a state outside the
// sector ranks to range().  For one constraint of k ones on bits [lo, lo+len)
// the position equals sum_j C(p_j, j) over the set bits p_1 < ... < p_k of
// the field (colex order); disjoint blocks combine mixed-radix with the
// highest block most significant, because ascending x compares high bits
// first.  Sectors that also fix the doubly occupied or empty sites (Nd / Nh,
// conservation.hpp:117-151,248-256) enumerate the sector of their Nset part
// and keep, in order, the states whose quantities match; they rank by binary
// search in that table.
#include <algorithm>
#include <vector>

#include "qd_internal.hpp"

static const BinomTable g_binom_host = qd_make_binom();
uint64_t qd_binom_host(int n, int k) {
  if (n < 0 || n > 64 || k < 0 || k > n) return 0;
  return g_binom_host.v[n][k];
}

// ------------------------------------------------------------ analysis
static bool contiguous(uint64_t m, int* lo, int* len) {
  if (m == 0) return false;
  int l = __builtin_ctzll(m);
  uint64_t s = m >> l;
  if ((s & (s + 1)) != 0) return false;  // not a run of ones
  *lo = l;
  *len = 64 - __builtin_clzll(s);
  return true;
}

// Fill sec from the constraints.  Returns 1 if the CNS form applies, 0 if
// the general (scan) form is needed, -1 if the sector is provably empty.
static int analyse_sector(int nbits, const qd_constraint* cons, int ncons, SectorDesc* sec) {
  std::memset(sec, 0, sizeof(*sec));
  sec->nbits = nbits;
  uint64_t all = (nbits >= 64) ? ~0ull : ((1ull << nbits) - 1ull);
  struct Blk {
    int lo, len, cnt;
  };
  std::vector<Blk> blks;
  uint64_t used = 0;
  bool cns_ok = true;
  for (int c = 0; c < ncons; ++c) {
    uint64_t m = cons[c].mask & all;  // bits >= nbits are always 0 in x
    if (m == 0) {
      if (cons[c].count != 0) return -1;
      continue;
    }
    int lo = 0, len = 0;
    if (!contiguous(m, &lo, &len) || (m & used)) {
      cns_ok = false;
      continue;
    }
    if (cons[c].count < 0 || cons[c].count > len) return -1;
    used |= m;
    blks.push_back({lo, len, (int)cons[c].count});
  }
  if (!cns_ok) return 0;
  // free gaps
  int pos = 0;
  std::sort(blks.begin(), blks.end(), [](const Blk& a, const Blk& b) { return a.lo < b.lo; });
  std::vector<Blk> full;
  for (auto& b : blks) {
    if (b.lo > pos) full.push_back({pos, b.lo - pos, -1});
    full.push_back(b);
    pos = b.lo + b.len;
  }
  if (pos < nbits) full.push_back({pos, nbits - pos, -1});
  if ((int)full.size() > QD_MAX_BLOCKS) return 0;
  sec->nblocks = (int)full.size();
  unsigned __int128 stride = 1;
  for (int b = 0; b < sec->nblocks; ++b) {
    sec->lo[b] = full[b].lo;
    sec->len[b] = full[b].len;
    sec->cnt[b] = full[b].cnt;
    unsigned __int128 sz;
    if (full[b].cnt >= 0)
      sz = qd_binom_host(full[b].len, full[b].cnt);
    else {
      if (full[b].len >= 63) return 0;
      sz = (unsigned __int128)1 << full[b].len;
    }
    sec->size[b] = (uint64_t)sz;
    sec->stride[b] = (uint64_t)stride;
    stride *= sz;
    if (stride > ((unsigned __int128)1 << 62)) return 0;
  }
  sec->total = (uint64_t)stride;
  sec->cns = 1;
  return 1;
}

// ------------------------------------------------------------ K1: unrank
__global__ void __launch_bounds__(256) unrank_kernel(SectorDesc sec, uint64_t* __restrict__ out) {
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 >= sec.total) return;
  const bool comp = sec.mdim > 1;  // composite: i0 = rank(q) * mdim + m
  uint64_t i = comp ? i0 / sec.mdim : i0;
  uint64_t x = 0;
  for (int b = 0; b < sec.nblocks; ++b) {
    uint64_t rb = (i / sec.stride[b]) % sec.size[b];
    uint64_t f = 0;
    if (sec.cnt[b] >= 0) {
      int p = sec.len[b];
      for (int j = sec.cnt[b]; j >= 1; --j) {
        // largest p' < p with C(p', j) <= rb  (p' >= j - 1 always qualifies)
        --p;
        while (qd_binom.v[p][j] > rb) --p;
        rb -= qd_binom.v[p][j];
        f |= (1ull << p);
      }
    } else {
      f = rb;
    }
    x |= f << sec.lo[b];
  }
  out[i0] = comp ? x * sec.mdim + i0 % sec.mdim : x;
}

// ------------------------------------------------------------ general path
// Quantities popcount(((x >> shift) ^ flip) & (x ^ flip) & sel) == cnt; a
// plain Nset has shift 0 and flip 0.
struct QList {
  int n;
  uint64_t sel[16];
  uint64_t flip[16];
  int shift[16];
  int64_t cnt[16];
};
__device__ inline bool q_ok(const QList& c, uint64_t x) {
  for (int k = 0; k < c.n; ++k) {
    const uint64_t f = c.flip[k];
    if ((int64_t)__popcll(((x >> c.shift[k]) ^ f) & (x ^ f) & c.sel[k]) != c.cnt[k]) return false;
  }
  return true;
}
// candidate i of the scan: the value i itself, or entry i of a table
template <bool TABLE>
__device__ inline uint64_t candidate(const uint64_t* __restrict__ src, uint64_t i) {
  return TABLE ? src[i] : i;
}
#define SCANCHUNK 4096
template <bool TABLE>
__global__ void __launch_bounds__(256) pred_count(QList c, const uint64_t* __restrict__ src, uint64_t range,
                                                  int32_t* __restrict__ counts) {
  uint64_t base = (uint64_t)blockIdx.x * SCANCHUNK;
  int s = 0;
  for (int k = threadIdx.x; k < SCANCHUNK; k += 256) {
    uint64_t i = base + k;
    if (i < range && q_ok(c, candidate<TABLE>(src, i))) ++s;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  __shared__ int red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
// ordered compaction of one chunk: each wave handles 64 consecutive
// candidates at a time, ballot gives in-order positions.
template <bool TABLE>
__global__ void __launch_bounds__(64) pred_fill(QList c, const uint64_t* __restrict__ src, uint64_t range,
                                                const int64_t* __restrict__ offs, uint64_t* __restrict__ out) {
  uint64_t base = (uint64_t)blockIdx.x * SCANCHUNK;
  int64_t pos = offs[blockIdx.x];
  for (int k = 0; k < SCANCHUNK; k += 64) {
    uint64_t i = base + k + threadIdx.x;
    uint64_t x = i < range ? candidate<TABLE>(src, i) : 0;
    bool ok = i < range && q_ok(c, x);
    uint64_t bal = __ballot(ok);
    if (ok) {
      int before = __popcll(bal & ((1ull << threadIdx.x) - 1ull));
      out[pos + before] = x;
    }
    pos += __popcll(bal);
  }
}

// In-order selection of the candidates that satisfy c: the values
// [0, range) (src == nullptr) or the entries of the ascending table src.
// Leaves b->d_states (ascending) and b->sec as a table-ranked sector.
static int select_states(qd_ctx* ctx, qd_basis* b, int nbits, const QList& c, const uint64_t* src,
                         uint64_t range) {
  int64_t nchunks = (int64_t)((range + SCANCHUNK - 1) / SCANCHUNK);
  if (nchunks < 1) nchunks = 1;
  int32_t* counts = nullptr;
  int64_t* offs = nullptr;
  if (hipMalloc(&counts, sizeof(int32_t) * nchunks) != hipSuccess ||
      hipMalloc(&offs, sizeof(int64_t) * (nchunks + 1)) != hipSuccess) {
    hipFree(counts);
    return qd_set_error(ctx, QD_E_OOM, "scan buffers");
  }
  if (src)
    pred_count<true><<<(unsigned)nchunks, 256, 0, ctx->stream>>>(c, src, range, counts);
  else
    pred_count<false><<<(unsigned)nchunks, 256, 0, ctx->stream>>>(c, src, range, counts);
  int rc = qd_scan_counts(ctx, counts, nchunks, offs);
  int64_t total = 0;
  if (rc == QD_OK) {
    hipMemcpyAsync(&total, offs + nchunks, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    if (hipMalloc(&b->d_states, sizeof(uint64_t) * (total ? total : 1)) != hipSuccess) rc = QD_E_OOM;
  }
  if (rc == QD_OK && total > 0) {
    if (src)
      pred_fill<true><<<(unsigned)nchunks, 64, 0, ctx->stream>>>(c, src, range, offs, b->d_states);
    else
      pred_fill<false><<<(unsigned)nchunks, 64, 0, ctx->stream>>>(c, src, range, offs, b->d_states);
  }
  hipError_t le = hipGetLastError();
  hipStreamSynchronize(ctx->stream);
  hipFree(counts);
  hipFree(offs);
  if (rc == QD_OK && le != hipSuccess) rc = qd_set_error(ctx, QD_E_HIP, "basis scan: %s", hipGetErrorString(le));
  if (rc != QD_OK) {
    hipFree(b->d_states);
    b->d_states = nullptr;
    return rc == QD_E_OOM ? qd_set_error(ctx, QD_E_OOM, "basis table") : rc;
  }
  b->sec.cns = 0;
  b->sec.nbits = nbits;
  b->sec.total = (uint64_t)total;
  return QD_OK;
}

// composite table of a table-ranked qubit sector: out[i] = q[i / mdim] * mdim + i % mdim
__global__ void __launch_bounds__(256) expand_modes_kernel(const uint64_t* __restrict__ qs, uint64_t nq,
                                                           uint64_t mdim, uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq * mdim) return;
  out[i] = qs[i / mdim] * mdim + i % mdim;
}

// ------------------------------------------------------------ K2: rank
__global__ void __launch_bounds__(256) rank_kernel(SectorDesc sec, const uint64_t* __restrict__ table,
                                                   const uint64_t* __restrict__ in, int64_t count,
                                                   int64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = (int64_t)qd_rank(sec, table, in[i]);
}

// ------------------------------------------------------------ ABI
extern "C" int qd_basis_create(qd_ctx* ctx, int nbits, const qd_constraint* cons, int ncons, qd_basis** out) {
  if (!ctx || !out || nbits < 0 || nbits > QD_MAX_BITS || ncons < 0 || (ncons > 0 && !cons))
    return qd_set_error(ctx, QD_E_ARG, "qd_basis_create: bad arguments");
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_basis* b = new qd_basis();
  b->ctx = ctx;
  int kind = analyse_sector(nbits, cons, ncons, &b->sec);
  if (kind == -1) {
    b->sec.nbits = nbits;
    b->sec.total = 0;
    b->sec.cns = 1;
    b->sec.nblocks = 0;
    *out = b;
    return QD_OK;
  }
  if (kind == 1) {
    uint64_t n = b->sec.total;
    hipError_t e = hipMalloc(&b->d_states, sizeof(uint64_t) * (n ? n : 1));
    if (e != hipSuccess) {
      delete b;
      return qd_set_error(ctx, QD_E_OOM, "basis table (%llu states): %s", (unsigned long long)n, hipGetErrorString(e));
    }
    if (n) {
      unrank_kernel<<<qd_grid((int64_t)n, 256), 256, 0, ctx->stream>>>(b->sec, b->d_states);
      hipError_t le = hipGetLastError();
      if (le != hipSuccess) {
        hipFree(b->d_states);
        delete b;
        return qd_set_error(ctx, QD_E_HIP, "unrank launch: %s", hipGetErrorString(le));
      }
    }
    *out = b;
    return QD_OK;
  }
  // general: predicate scan over [0, 2^nbits)
  if (nbits > 40 || ncons > 16) {
    delete b;
    return qd_set_error(ctx, QD_E_UNSUPPORTED,
                        "sector with overlapping/non-contiguous masks needs a scan of 2^%d states", nbits);
  }
  QList cl{};
  cl.n = ncons;
  for (int c = 0; c < ncons; ++c) {
    cl.sel[c] = cons[c].mask;
    cl.cnt[c] = cons[c].count;
  }
  int rc = select_states(ctx, b, nbits, cl, nullptr, 1ull << nbits);
  if (rc != QD_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return QD_OK;
}

extern "C" int qd_basis_create_q(qd_ctx* ctx, int nbits, const qd_quantity* q, int nq, qd_basis** out) {
  if (!ctx || !out || nbits < 0 || nbits > QD_MAX_BITS || nq < 0 || nq > 16 || (nq > 0 && !q))
    return qd_set_error(ctx, QD_E_ARG, "qd_basis_create_q: bad arguments");
  std::vector<qd_constraint> plain;
  QList gen{};
  for (int k = 0; k < nq; ++k) {
    if (q[k].shift < 0 || q[k].shift > 63) return qd_set_error(ctx, QD_E_ARG, "qd_basis_create_q: shift");
    if (q[k].shift == 0 && q[k].flip == 0) {
      plain.push_back(qd_constraint{q[k].sel, q[k].count});
    } else {
      gen.sel[gen.n] = q[k].sel;
      gen.flip[gen.n] = q[k].flip;
      gen.shift[gen.n] = q[k].shift;
      gen.cnt[gen.n] = q[k].count;
      ++gen.n;
    }
  }
  if (gen.n == 0) return qd_basis_create(ctx, nbits, plain.data(), (int)plain.size(), out);
  *out = nullptr;
  hipSetDevice(ctx->device);
  // the sector of the Nset part, enumerated when it has the CNS form
  qd_basis* base = nullptr;
  SectorDesc probe{};
  if (analyse_sector(nbits, plain.data(), (int)plain.size(), &probe) == 1) {
    int rc = qd_basis_create(ctx, nbits, plain.data(), (int)plain.size(), &base);
    if (rc != QD_OK) return rc;
  }
  qd_basis* b = new qd_basis();
  b->ctx = ctx;
  int rc;
  if (base) {
    if (base->sec.total == 0) {  // empty Nset sector: an empty table
      b->sec.nbits = nbits;
      b->sec.total = 0;
      b->sec.cns = 0;
      rc = QD_OK;
    } else {
      rc = select_states(ctx, b, nbits, gen, base->d_states, base->sec.total);
    }
    qd_basis_destroy(base);
  } else {
    if (nbits > 40) {
      delete b;
      return qd_set_error(ctx, QD_E_UNSUPPORTED, "constraint set needs a scan of 2^%d states", nbits);
    }
    QList all = gen;
    for (auto& c : plain) {
      if (all.n >= 16) break;
      all.sel[all.n] = c.mask;
      all.flip[all.n] = 0;
      all.shift[all.n] = 0;
      all.cnt[all.n] = c.count;
      ++all.n;
    }
    rc = select_states(ctx, b, nbits, all, nullptr, 1ull << nbits);
  }
  if (rc != QD_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return QD_OK;
}

extern "C" int qd_basis_create_modes(qd_ctx* ctx, int nbits, int mdim, const qd_quantity* q, int nq,
                                     qd_basis** out) {
  if (!ctx || !out || nbits < 0 || nbits > 48 || mdim < 2 || mdim > QD_MAX_MODE_DIM || nq < 0 || nq > 16 ||
      (nq > 0 && !q))
    return qd_set_error(ctx, QD_E_ARG, "qd_basis_create_modes: bad arguments");
  // the qubit sector first (its table, or its CNS form), then the mode digit
  qd_basis* qb = nullptr;
  int rc = qd_basis_create_q(ctx, nbits, q, nq, &qb);
  if (rc != QD_OK) return rc;
  *out = nullptr;
  hipSetDevice(ctx->device);
  qd_basis* b = new qd_basis();
  b->ctx = ctx;
  b->sec = qb->sec;
  b->sec.mdim = (uint64_t)mdim;
  const unsigned __int128 total = (unsigned __int128)qb->sec.total * (unsigned)mdim;
  if (total > ((unsigned __int128)1 << 62)) {
    qd_basis_destroy(qb);
    delete b;
    return qd_set_error(ctx, QD_E_UNSUPPORTED, "composite sector too large");
  }
  b->sec.total = (uint64_t)total;
  if (b->sec.total == 0) {
    qd_basis_destroy(qb);
    *out = b;
    return QD_OK;
  }
  if (hipMalloc(&b->d_states, sizeof(uint64_t) * b->sec.total) != hipSuccess) {
    qd_basis_destroy(qb);
    delete b;
    return qd_set_error(ctx, QD_E_OOM, "composite basis table");
  }
  if (b->sec.cns) {
    unrank_kernel<<<qd_grid((int64_t)b->sec.total, 256), 256, 0, ctx->stream>>>(b->sec, b->d_states);
  } else {
    expand_modes_kernel<<<qd_grid((int64_t)b->sec.total, 256), 256, 0, ctx->stream>>>(qb->d_states, qb->sec.total,
                                                                                    (uint64_t)mdim, b->d_states);
  }
  hipError_t le = hipGetLastError();
  hipStreamSynchronize(ctx->stream);
  qd_basis_destroy(qb);
  if (le != hipSuccess) {
    hipFree(b->d_states);
    delete b;
    return qd_set_error(ctx, QD_E_HIP, "composite basis: %s", hipGetErrorString(le));
  }
  *out = b;
  return QD_OK;
}

extern "C" int qd_basis_destroy(qd_basis* b) {
  if (!b) return QD_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  if (b->d_states) hipFree(b->d_states);
  delete b;
  return QD_OK;
}

extern "C" int64_t qd_basis_size(const qd_basis* b) { return b ? (int64_t)b->sec.total : -1; }

extern "C" int qd_basis_download(qd_basis* b, uint64_t* h_states) {
  if (!b || !h_states) return QD_E_ARG;
  if (b->sec.total == 0) return QD_OK;
  QD_HIP(b->ctx, hipMemcpyAsync(h_states, b->d_states, sizeof(uint64_t) * b->sec.total, hipMemcpyDeviceToHost,
                                b->ctx->stream));
  QD_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return QD_OK;
}

extern "C" const uint64_t* qd_basis_device_states(qd_basis* b) { return b ? b->d_states : nullptr; }

extern "C" int qd_basis_rank(qd_basis* b, const uint64_t* d_states, int64_t count, int64_t* d_ranks) {
  if (!b || count < 0 || (count > 0 && (!d_states || !d_ranks))) return QD_E_ARG;
  if (count == 0) return QD_OK;
  rank_kernel<<<qd_grid(count, 256), 256, 0, b->ctx->stream>>>(b->sec, b->d_states, d_states, count, d_ranks);
  QD_CHECK_LAUNCH(b->ctx);
  return QD_OK;
}
