// qd_internal.hpp -- shared declarations of the gfx950 implementation behind
// include/qudrip_hip.h.  Internal: nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/qudrip_hip.h"

// --------------------------------------------------------------- complex
// Complex128 as double2 (re, im).  The multiply is written out in the order
// std::complex<double> evaluates it without FMA:
//     (a+bi)(c+di) = (ac - bd) + (ad + bc) i
// The CSR fill kernels are compiled with -ffp-contract=off so that this
// stays two products and one add per component (bit-identical to the
// reference's host arithmetic); the bandwidth-bound kernels allow
// contraction.
typedef double2 cplx;

__host__ __device__ inline cplx cmake(double re, double im) {
  cplx r;
  r.x = re;
  r.y = im;
  return r;
}
__host__ __device__ inline cplx cmul(cplx a, cplx b) {
  return cmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return cmake(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx csub(cplx a, cplx b) { return cmake(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx cconj(cplx a) { return cmake(a.x, -a.y); }
__host__ __device__ inline cplx crmul(double d, cplx a) { return cmake(d * a.x, d * a.y); }
// conj(a) * b
__host__ __device__ inline cplx cdotmul(cplx a, cplx b) {
  return cmake(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// --------------------------------------------------------------- binomials
// C(n, k) for 0 <= k <= n <= 64 (C(64,32) < 2^64).
struct BinomTable {
  uint64_t v[65][65];
};
constexpr BinomTable qd_make_binom() {
  BinomTable t{};
  for (int n = 0; n <= 64; ++n) {
    t.v[n][0] = 1;
    for (int k = 1; k <= n; ++k) t.v[n][k] = t.v[n - 1][k - 1] + (k <= n - 1 ? t.v[n - 1][k] : 0);
  }
  return t;
}
// one copy per translation unit (no relocatable device code needed)
static __constant__ BinomTable qd_binom = qd_make_binom();
uint64_t qd_binom_host(int n, int k);

// --------------------------------------------------------------- sector
// A sector whose constraints are disjoint contiguous bit ranges is a mixed-
// radix product of per-block combinatorial number systems; ascending x
// order == block-major (highest block most significant).  Bits covered by
// no constraint form "free" blocks (any popcount).
#define QD_MAX_BLOCKS 8
struct SectorDesc {
  int nbits;
  int nblocks;
  int lo[QD_MAX_BLOCKS];
  int len[QD_MAX_BLOCKS];
  int cnt[QD_MAX_BLOCKS];  // -1: free block
  uint64_t size[QD_MAX_BLOCKS];
  uint64_t stride[QD_MAX_BLOCKS];
  uint64_t total;
  int cns;  // 1: CNS rank/unrank valid; 0: table + binary search
  uint64_t mdim;  // >= 2: composite states q * mdim + m (a boson mode); 0 or 1: none
};

// rank of x within the sector, or sec.total when x is outside it
// (index.hpp:234-240 semantics).  CNS form.
__device__ inline uint64_t qd_rank_cns(const SectorDesc& sec, uint64_t x) {
  if (sec.nbits < 64 && (x >> sec.nbits) != 0) return sec.total;
  uint64_t r = 0;
  for (int b = 0; b < sec.nblocks; ++b) {
    uint64_t f = (x >> sec.lo[b]);
    if (sec.len[b] < 64) f &= ((1ull << sec.len[b]) - 1ull);
    uint64_t rb;
    if (sec.cnt[b] >= 0) {
      if (__popcll(f) != sec.cnt[b]) return sec.total;
      rb = 0;
      int j = 1;
      while (f) {
        int p = __ffsll((unsigned long long)f) - 1;
        rb += qd_binom.v[p][j];
        f &= f - 1ull;
        ++j;
      }
    } else {
      rb = f;
    }
    r += rb * sec.stride[b];
  }
  return r;
}

// rank by binary search in the ascending state table (general sectors).
__device__ inline uint64_t qd_rank_table(const uint64_t* states, uint64_t n, uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = lo + ((hi - lo) >> 1);
    if (states[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && states[lo] == x) ? lo : n;
}

__device__ inline uint64_t qd_rank(const SectorDesc& sec, const uint64_t* states, uint64_t x) {
  if (sec.mdim > 1 && sec.cns) {  // composite CNS sector: rank(q) * mdim + m
    const uint64_t r = qd_rank_cns(sec, x / sec.mdim);
    return r >= sec.total ? sec.total : r * sec.mdim + x % sec.mdim;
  }
  return sec.cns ? qd_rank_cns(sec, x) : qd_rank_table(states, sec.total, x);
}

// --------------------------------------------------------------- term program
struct DevOp {
  int kind, a, b, string;
  cplx U[4];  // U_[i*2+j] = U(i,j)
};
struct DevTerm {
  cplx coeff;
  int op_begin, n_ops;
  int may_dup;  // 1 if two leaves of this term can reach the same target
  int n_leaves; // real leaves (bit: x2, hop: x1)
};

#define QD_LEAF_EPS 1e-10  // qudrip.hpp:18

// Leaves of one term from source state x, in the reference's order
// (operator.hpp:135-206): ops applied right-to-left, the choice of the
// first-applied bit op is the most significant digit of L.  bitOp children
// get (jw * U(b,o)) * val (operators/bit.hpp:57-63); hopBitOp gets one child
// (operators/hopbit.hpp:52-78).
// Returns false for a dead branch.
__device__ inline bool eval_leaf(const DevOp* __restrict__ ops, int n_ops, int nbitops, uint64_t x, int L,
                                 uint64_t* out_x, cplx* out_v) {
  cplx v = cmake(1.0, 0.0);
  int t = 0;
  for (int iop = n_ops - 1; iop >= 0; --iop) {
    const DevOp& op = ops[iop];
    if (op.kind == QD_OP_BIT) {
      int s = op.a;
      int o = (L >> (nbitops - 1 - t)) & 1;
      ++t;
      int b = (int)((x >> s) & 1ull);
      int prev_set = __popcll(x >> (s + 1));
      cplx jw = cmake((prev_set & 1) ? (double)op.string : 1.0, 0.0);
      v = cmul(cmul(jw, op.U[b * 2 + o]), v);
      x = (x & ~(1ull << s)) | ((uint64_t)o << s);
    } else {
      int a = op.a, b = op.b;
      if (a == b) {
        v = crmul((double)((x >> a) & 1ull), v);
      } else if (((x >> b) & 1ull) == 1ull && ((x >> a) & 1ull) == 0ull) {
        int jw1 = __popcll(x >> (b + 1));
        uint64_t x1 = x & ~(1ull << b);
        int jw2 = __popcll(x1 >> (a + 1));
        int sgn = ((jw1 & 1) ? op.string : 1) * ((jw2 & 1) ? op.string : 1);
        v = cmul(cmake((double)sgn, 0.0), v);
        x = x1 | (1ull << a);
      } else {
        return false;
      }
    }
  }
  *out_x = x;
  *out_v = v;
  return true;
}

// Leaves of one term over composite states x = q * mdim + m (a boson mode
// next to the qubits, Indices, index.hpp:112-184): bit and hop gates act on q
// as in eval_leaf, a mode gate (operators/mode.hpp:80-103) takes m to
// output j with value U(m, j) * val, where mats[op.a * mdim * mdim + m *
// mdim + j] = U(m, j).  Each op contributes a digit of L (bit 2, hop 1, mode
// mdim children), the first-applied op the most significant, as the tree of
// operator.hpp:135-206 orders its leaves.
__device__ inline bool eval_leaf_modes(const DevOp* __restrict__ ops, int n_ops, uint64_t x, int L, int mdim,
                                       const cplx* __restrict__ mats, uint64_t* out_x, cplx* out_v) {
  int digit[QD_MAX_OPS_PER_TERM];
  for (int iop = 0; iop < n_ops; ++iop) {  // least significant: the last-applied op
    const int r = ops[iop].kind == QD_OP_BIT ? 2 : ops[iop].kind == QD_OP_MODE ? mdim : 1;
    digit[iop] = L % r;
    L /= r;
  }
  uint64_t q = x / (uint64_t)mdim;
  int m = (int)(x % (uint64_t)mdim);
  cplx v = cmake(1.0, 0.0);
  for (int iop = n_ops - 1; iop >= 0; --iop) {
    const DevOp& op = ops[iop];
    if (op.kind == QD_OP_BIT) {
      const int s = op.a, o = digit[iop];
      const int b = (int)((q >> s) & 1ull);
      const int prev_set = __popcll(q >> (s + 1));
      const cplx jw = cmake((prev_set & 1) ? (double)op.string : 1.0, 0.0);
      v = cmul(cmul(jw, op.U[b * 2 + o]), v);
      q = (q & ~(1ull << s)) | ((uint64_t)o << s);
    } else if (op.kind == QD_OP_MODE) {
      const int j = digit[iop];
      v = cmul(mats[((size_t)op.a * mdim + m) * mdim + j], v);
      m = j;
    } else {
      const int a = op.a, b = op.b;
      if (a == b) {
        v = crmul((double)((q >> a) & 1ull), v);
      } else if (((q >> b) & 1ull) == 1ull && ((q >> a) & 1ull) == 0ull) {
        const int jw1 = __popcll(q >> (b + 1));
        const uint64_t q1 = q & ~(1ull << b);
        const int jw2 = __popcll(q1 >> (a + 1));
        const int sgn = ((jw1 & 1) ? op.string : 1) * ((jw2 & 1) ? op.string : 1);
        v = cmul(cmake((double)sgn, 0.0), v);
        q = q1 | (1ull << a);
      } else {
        return false;
      }
    }
  }
  *out_x = q * (uint64_t)mdim + (uint64_t)m;
  *out_v = v;
  return true;
}

// --------------------------------------------------------------- handles
// Caching pool for large per-solve device buffers (Krylov bases, request
// vectors).  Blocks freed into it are kept and handed to later requests of a
// size within 1/8 above theirs instead of going back to the driver: repeated
// allocation and release of multi-GB buffers fragments device memory, and
// kernels streaming through fragmented mappings lose bandwidth to address
// translation.  live + cached bytes stay within `cap` (cached blocks are
// released first when a new block would pass it).  Callers return a block
// only after the work that used it has completed on its stream.
struct qd_devpool {
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks;  // size -> block
  std::unordered_map<void*, size_t> live;    // block -> size
  size_t cached = 0, in_use = 0, cap = 0;
};

struct qd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  qd_devpool* pool = nullptr;  // not owned; nullptr: plain hipMalloc / hipFree
  // reusable device scratch for reductions
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // small pinned host buffer for results read back by the host
  double* pinned = nullptr;
  size_t pinned_count = 0;
  int num_cus = 256;
};

struct qd_basis {
  qd_ctx* ctx = nullptr;
  SectorDesc sec{};
  uint64_t* d_states = nullptr;  // ascending table (always materialised)
};

struct qd_termprog {
  qd_ctx* ctx = nullptr;
  int nbits = 0;
  std::vector<DevTerm> h_terms;
  std::vector<DevOp> h_ops;
  DevTerm* d_terms = nullptr;
  DevOp* d_ops = nullptr;
  // adjoint program (A^dagger = sum conj(c_k) P_k^dagger), for pull-form
  // products; built on demand.
  DevTerm* d_adj_terms = nullptr;
  DevOp* d_adj_ops = nullptr;
  int max_leaves = 1;
  // boson mode (qd_termprog_create_modes): mdim >= 2, mode-gate matrices
  // (nmats of the program, then their adjoints), mdim * mdim each
  int mdim = 1;
  int nmats = 0;
  std::vector<cplx> h_mats;
  cplx* d_mats = nullptr;
};

struct qd_csr {
  qd_ctx* ctx = nullptr;
  int64_t n = 0, nnz = 0;
  int32_t* d_outer = nullptr;
  int32_t* d_inner = nullptr;
  cplx* d_vals = nullptr;
  int hermitian = -1;  // -1 unknown, 0 no, 1 yes
  qd_basis* basis = nullptr;
  qd_termprog* prog = nullptr;
  // CSC of the adjoint program when the matrix is not Hermitian
  qd_csr* adj = nullptr;
};

enum { QD_LINOP_CSR = 0, QD_LINOP_MATFREE = 1 };
struct qd_kron;
struct qd_linop {
  qd_ctx* ctx = nullptr;
  int kind = QD_LINOP_CSR;
  int64_t n = 0;
  // CSR kind: pull over rows of the adjoint's CSC with conjugated values
  const int32_t* outer = nullptr;
  const int32_t* inner = nullptr;
  const cplx* vals = nullptr;
  int64_t nnz = 0;
  int lanes_per_row = 16;
  // matrix-free kind
  qd_basis* basis = nullptr;
  qd_termprog* prog = nullptr;
  qd_kron* kron = nullptr;  // two-block Kronecker form, when the program has it
  int use_kron = 1;
  // extra diagonal added by every product (qd_linop_with_diagonal); a view
  // linop borrows the arrays (and Kronecker form) of its base and owns none
  const cplx* diag0 = nullptr;
  bool view = false;
};
qd_kron* qd_kron_try_build(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog);
void qd_kron_destroy(qd_kron* k);
int qd_kron_apply(qd_ctx* ctx, qd_kron* k, const cplx* x, cplx* y, cplx shift, const cplx* ext_diag = nullptr);

struct qd_krylov {
  qd_ctx* ctx = nullptr;
  int64_t n = 0;
  int kry_dim = 0;
  int herm = 0;
  int act_dim = 0;
  cplx* d_V = nullptr;       // (kry_dim + 1) columns of n (padded)
  int64_t ld = 0;            // column stride in complex entries
  cplx* d_w = nullptr;       // scratch vector
  double* d_scal = nullptr;  // device scalars (projection coefficients, norms)
  std::vector<cplx> KR;      // kry_dim x kry_dim Hessenberg, column-major
  cplx n_last{0.0, 0.0};
  std::vector<double> evals;
  std::vector<cplx> evs_prev;  // convergeEV targets
  double err = 0.0;
  double beta_last = 0.0;
  double ritz_resid = 0.0;  // |(H - E0) psi| of the last ground-state Ritz pair
  double ritz_gap = 0.0;    // second minus lowest Ritz value
  int breakdown = 0;
  int64_t n_spmv = 0;
  int method = 1;  // 0 = CGS2, 1 = DCGS2, 2 = windowed + partial, 3 = windowed
  int full_steps = 0;  // steps of the partial scheme that covered the whole basis
  bool pending = false;  // d_w holds the next pending vector of a resumable expansion
};

// --------------------------------------------------------------- errors
int qd_set_error(qd_ctx* ctx, int code, const char* fmt, ...);
#define QD_HIP(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return qd_set_error((ctx), QD_E_HIP, "%s failed: %s (%s:%d)", #call,                  \
                          hipGetErrorString(e_), __FILE__, __LINE__);                       \
  } while (0)
#define QD_CHECK_LAUNCH(ctx) QD_HIP(ctx, hipGetLastError())

// large per-solve buffers: through ctx->pool when set, else hipMalloc/hipFree
int qd_dev_alloc(qd_ctx* ctx, size_t bytes, void** p);
void qd_dev_free(qd_ctx* ctx, void* p);
void qd_devpool_release_all(qd_devpool* pool);

// scratch growth (device), returns QD_OK or error
int qd_ctx_scratch(qd_ctx* ctx, size_t bytes, void** p);
int qd_ctx_pinned(qd_ctx* ctx, size_t count, double** p);

// --------------------------------------------------------------- internal API
// exclusive scan of int32 counts[n] into int64 offsets[n+1] (device)
int qd_scan_counts(qd_ctx* ctx, const int32_t* d_counts, int64_t n, int64_t* d_offsets);
// build the CSC of a program (prog or its adjoint) over a basis
int qd_csr_build_impl(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog, int adjoint, qd_csr** out);
int qd_termprog_ensure_adjoint(qd_termprog* prog);
// device vector helpers (krylov.hip)
int qd_dev_dot_many(qd_ctx* ctx, int64_t n, const cplx* V, int64_t ld, int ncols,
                    const double* col_scale, const cplx* w, cplx* d_out);
// FP64 host numerics (host_linalg.cpp)
int qd_tridiag_eig(int m, const double* d, const double* e, double* w, double* z, int want_z);
int qd_expm_pade13(int m, const cplx* A, cplx* E);
int qd_krylov_gs_step(qd_krylov* kr, qd_linop* op, double* d_psi, int add_dim, double* h_E0, bool want_vector);

// stored-matrix product fused with a dot (spmv.hip): y = (H + diag) x and,
// per block of 256 rows, the sum of conj(x_r) y_r into part[block] for
// qd_linop_dot_blocks(n) blocks; QD_E_UNSUPPORTED (nothing launched) for
// other operators.  write = false skips the store of y.
int64_t qd_linop_dot_blocks(int64_t n);
int qd_linop_dot_rows_per_block();  // rows summed into one partial
int qd_linop_apply_dot(qd_linop* op, const cplx* x, cplx* y, cplx* part, bool write);

// grid helpers
static inline unsigned qd_grid(int64_t work, int per_block) {
  int64_t g = (work + per_block - 1) / per_block;
  return (unsigned)(g < 1 ? 1 : g);
}
