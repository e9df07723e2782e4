// serve.hip -- continuous-batching solve server over the Krylov solvers.
//
// No reference counterpart: the reference runs one solve per process
// (programs/spin.cpp, programs/hubbard_diss.cpp).  The server keeps, per
// model, the sector basis, the stored matrix of the model's first operator A
// and the diagonal of its second operator D on the device; a request asks for
// the ground state or a real-time evolution of H = A + c D for its own c
// (Jz of an XXZ chain, U of a Hubbard lattice).  Its operator is a view of the
// cached matrix plus its own diagonal c D, so no request builds a matrix.
//
// Workers (one thread and one HIP stream each) take requests in arrival
// order as soon as a worker is free and the request's device memory fits in
// the HBM budget; a request that does not fit waits while later, smaller
// requests may pass it (at most QD_SERVE_MAX_BYPASS times), and a request that
// could never fit is rejected.  The Krylov state of every running request
// stays on the device until it finishes.  Each request runs the same kernels
// with the same reduction order whatever else is running, so its result does
// not depend on the mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mrhs_kron.hpp"

namespace {

constexpr int QD_SERVE_MAX_BYPASS = 8;

__global__ void scale_diag_kernel(int64_t n, double c, const cplx* __restrict__ src, cplx* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = cmake(c * src[i].x, c * src[i].y);
}

__global__ void real_part_kernel(int64_t n, const cplx* __restrict__ d, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = d[i].x;
}

// columns of a trajectory-minor block X[s * ld + t]
__global__ void col_gather_kernel(int64_t n, int ld, int t, const cplx* __restrict__ X, cplx* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) out[s] = X[s * ld + t];
}
__global__ void col_scatter_kernel(int64_t n, int ld, int t, const cplx* __restrict__ in, cplx* __restrict__ X) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) X[s * ld + t] = in[s];
}
__global__ void col_copy_kernel(int64_t n, int ld, int src, int dst, cplx* __restrict__ X) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) X[s * ld + dst] = X[s * ld + src];
}
__global__ void inv_sqrt_kernel(int m, const double* __restrict__ n2, double* __restrict__ f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) f[t] = n2[t] > 0.0 ? 1.0 / sqrt(n2[t]) : 0.0;
}

// diag[j] = value of column j's diagonal entry; flag = 1 if any entry is off the diagonal
__global__ void extract_diag_kernel(int64_t n, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                                    const cplx* __restrict__ vals, cplx* __restrict__ diag, int* __restrict__ flag) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cplx d = cmake(0.0, 0.0);
  for (int32_t e = outer[j]; e < outer[j + 1]; ++e) {
    if (inner[e] == j)
      d = cadd(d, vals[e]);
    else
      *flag = 1;
  }
  diag[j] = d;
}

// sum_s Re(d_s) |psi_s|^2 and sum_s |psi_s|^2: block partials, then one block in fixed order
constexpr int DX_THREADS = 256;
__global__ void __launch_bounds__(DX_THREADS)
    diag_expect_partial(int64_t n, const cplx* __restrict__ d, const cplx* __restrict__ psi, double2* __restrict__ part) {
  __shared__ double2 sh[DX_THREADS / 64];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * DX_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * DX_THREADS) {
    const double p = psi[i].x * psi[i].x + psi[i].y * psi[i].y;
    a += d[i].x * p;
    b += p;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = make_double2(a, b);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 s = sh[0];
    for (int w = 1; w < DX_THREADS / 64; ++w) {
      s.x += sh[w].x;
      s.y += sh[w].y;
    }
    part[blockIdx.x] = s;
  }
}
__global__ void diag_expect_final(int nb, const double2* __restrict__ part, double2* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double2 s = make_double2(0.0, 0.0);
  for (int b = 0; b < nb; ++b) {
    s.x += part[b].x;
    s.y += part[b].y;
  }
  *out = s;
}

unsigned blocks_for(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

struct Model {
  int nbits = 0;
  int64_t n = 0;
  int structured = 0;
  qd_basis* basis = nullptr;
  qd_termprog* progA = nullptr;
  qd_csr* csrA = nullptr;
  qd_linop* opA = nullptr;
  qd_linop* opK = nullptr;   // structured models: matrix-free A (two-block form when it has one)
  qd_kron* kronA = nullptr;  // opK's two-block form with both packed tables (borrowed; lockstep batches)
  cplx* d_diagD = nullptr;
  double* d_diagR = nullptr;  // its real part (the diagonal of a Hermitian D)
  size_t bytes = 0;
  std::vector<qd_term> tA, tD;
  std::vector<qd_op> oA, oD;
};

struct Job {
  qd_request req;
  const Model* model = nullptr;  // resolved under the server lock at submission
  double t_submit = 0.0;
  size_t footprint = 0;
  int bypassed = 0;
  int gs_method = 2;  // Krylov scheme of a ground state, fixed at submission
};

}  // namespace

struct qd_server {
  int device = 0;
  size_t budget = 0;
  size_t used = 0;  // models + admitted requests
  qd_ctx* setup = nullptr;
  std::vector<Model*> models;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> queue;
  std::vector<qd_result> results;
  // cancellation: running requests per tag, and the tags asked to stop
  // (ncancel lets a worker skip the lock while nothing is cancelled)
  std::unordered_map<int64_t, int> active;
  std::unordered_map<int64_t, int> cancel;
  std::atomic<int> ncancel{0};
  int64_t submitted = 0, finished = 0, rejected = 0;
  int running = 0, max_running = 0;
  size_t peak_used = 0;
  bool stopping = false;
  int batch_slots = 1;             // evolve requests of one model stepped together (1: no batching)
  int gs_method = 2;               // qd_krylov_set_method scheme of ground-state requests
  qd_devpool pool;                 // request buffers (cap: the HBM budget)
  std::vector<char> batch_active;  // per model: a batch owner admits its evolve requests
  std::vector<std::thread> threads;
  std::chrono::steady_clock::time_point t0;
  std::string err;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// A request with this tag was asked to stop (qd_server_cancel).
bool stop_requested(qd_server* S, int64_t tag) {
  if (S->ncancel.load(std::memory_order_acquire) == 0) return false;
  std::lock_guard<std::mutex> lk(S->mu);
  return S->cancel.count(tag) != 0;
}

// Bookkeeping of a request that starts / ends running.  Lock held.
void mark_running(qd_server* S, int64_t tag) { ++S->active[tag]; }
void mark_ended(qd_server* S, int64_t tag) {
  auto it = S->active.find(tag);
  if (it != S->active.end() && --it->second == 0) {
    S->active.erase(it);
    if (S->cancel.erase(tag)) S->ncancel.fetch_sub(1, std::memory_order_release);
  }
}

int cancelled(qd_ctx* ctx) { return qd_set_error(ctx, QD_E_CANCELLED, "request cancelled"); }

// Device contexts (one stream each) of servers live as long as the process
// and pass from a destroyed server to the next one: a process has a fixed
// number of hardware queues, and streams created after others were destroyed
// can land on a shared queue, which serialises workers meant to overlap.
std::mutex g_ctx_pool_mu;
std::vector<qd_ctx*> g_ctx_pool;

int ctx_acquire(int device, qd_ctx** out) {
  {
    std::lock_guard<std::mutex> lk(g_ctx_pool_mu);
    for (size_t i = 0; i < g_ctx_pool.size(); ++i) {
      if (g_ctx_pool[i]->device == device) {
        *out = g_ctx_pool[i];
        g_ctx_pool.erase(g_ctx_pool.begin() + (std::ptrdiff_t)i);
        return QD_OK;
      }
    }
  }
  return qd_ctx_create(device, out);
}

void ctx_release(qd_ctx* c) {
  if (!c) return;
  hipStreamSynchronize(c->stream);
  c->pool = nullptr;
  std::lock_guard<std::mutex> lk(g_ctx_pool_mu);
  g_ctx_pool.push_back(c);
}

size_t request_footprint(const Model& M, const qd_request& r) {
  const size_t vec = sizeof(cplx) * (size_t)M.n;
  const size_t slack = 8u << 20;
  if (r.kind == QD_REQ_GROUND_STATE) return (size_t)(r.k + 6) * vec + slack;  // basis k+1, scratch 2, psi, diag
  return (size_t)(r.k + 8) * vec + slack;  // basis, scratch, psi, next, diag
}

// <D> in the normalised state psi and |psi|^2
int diag_expect(qd_ctx* ctx, int64_t n, const cplx* d, const cplx* psi, double* dexp, double* nrm2) {
  const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + DX_THREADS - 1) / DX_THREADS));
  // partials from the server's pool: a hipMalloc / hipFree per request would
  // synchronise the whole device and stall the other workers' streams
  double2* part = nullptr;
  if (qd_dev_alloc(ctx, sizeof(double2) * (nb + 1), (void**)&part) != QD_OK)
    return qd_set_error(ctx, QD_E_OOM, "diag_expect partials");
  diag_expect_partial<<<nb, DX_THREADS, 0, ctx->stream>>>(n, d, psi, part);
  diag_expect_final<<<1, 64, 0, ctx->stream>>>(nb, part, part + nb);
  double2 h;
  hipError_t e1 = hipMemcpyAsync(&h, part + nb, sizeof(double2), hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  qd_dev_free(ctx, part);
  if (e1 != hipSuccess || e2 != hipSuccess) return qd_set_error(ctx, QD_E_HIP, "diag_expect");
  *dexp = h.y > 0.0 ? h.x / h.y : 0.0;
  *nrm2 = h.y;
  return QD_OK;
}

int run_ground_state(qd_server* S, qd_ctx* ctx, const Model& M, const qd_request& r, int method, qd_result* out) {
  const int64_t n = M.n;
  cplx *diag = nullptr, *psi = nullptr;
  qd_linop* op = nullptr;
  qd_krylov* kr = nullptr;
  int rc = QD_OK;
  auto done = [&](int code) {
    if (kr) qd_krylov_destroy(kr);
    if (op) qd_linop_destroy(op);
    hipStreamSynchronize(ctx->stream);
    qd_dev_free(ctx, diag);
    qd_dev_free(ctx, psi);
    return code;
  };
  if (stop_requested(S, r.tag)) return done(cancelled(ctx));
  if (qd_dev_alloc(ctx, sizeof(cplx) * n, (void**)&diag) != QD_OK ||
      qd_dev_alloc(ctx, sizeof(cplx) * n, (void**)&psi) != QD_OK)
    return done(qd_set_error(ctx, QD_E_OOM, "request vectors"));
  scale_diag_kernel<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(n, r.coeff, M.d_diagD, diag);
  if ((rc = qd_linop_with_diagonal(ctx, M.opA, (const double*)diag, &op))) return done(rc);
  if ((rc = qd_seeded_state(ctx, n, r.seed, (double*)psi))) return done(rc);
  if ((rc = qd_krylov_create(ctx, n, r.k, 1, &kr))) return done(rc);
  // 2 (default): local window + partial reorthogonalisation, whose tridiagonal
  // matches a fully reorthogonalised basis to O(eps |H|) at a fraction of the
  // passes; 1: full delayed CGS2, the drop-in path's scheme
  if ((rc = qd_krylov_set_method(kr, method))) return done(rc);
  double E = 0.0, Eprev = 0.0;
  bool have_prev = false;
  qd_krylov_info info{};
  if (r.tol > 0.0) {
    const int chunk = r.chunk > 0 ? r.chunk : 10;
    for (;;) {
      if ((rc = qd_krylov_gs_step(kr, op, (double*)psi, chunk, &E, false))) return done(rc);
      qd_krylov_get_info(kr, &info);
      if (stop_requested(S, r.tag)) return done(cancelled(ctx));  // between chunks
      // Ritz-residual bound of the eigenvalue error, r^2 / gap (Kato-Temple
      // with the second Ritz value for the gap), or the move of E0
      const double bound = info.ritz_gap > 0.0 ? info.ritz_resid * info.ritz_resid / info.ritz_gap : 1e300;
      if (info.act_dim >= 2 * chunk && bound <= r.tol * std::fabs(E)) break;
      if (have_prev && std::fabs(E - Eprev) <= r.tol * std::fabs(E)) break;
      if (info.act_dim >= r.k || info.breakdown) break;
      Eprev = E;
      have_prev = true;
    }
    if ((rc = qd_krylov_gs_step(kr, op, (double*)psi, 0, &E, true))) return done(rc);  // the Ritz vector, once
  } else {
    if ((rc = qd_krylov_ground_state(kr, op, (double*)psi, -1, &E))) return done(rc);
    qd_krylov_get_info(kr, &info);
    if (stop_requested(S, r.tag)) return done(cancelled(ctx));  // one call: stopped after it
  }
  double dexp = 0.0, nrm2 = 0.0;
  if ((rc = diag_expect(ctx, n, M.d_diagD, psi, &dexp, &nrm2))) return done(rc);
  out->energy = E;
  out->energy_start = E;
  out->observable = dexp;
  out->iterations = info.act_dim;
  out->krylov_mean = info.act_dim;
  return done(QD_OK);
}

int run_evolve(qd_server* S, qd_ctx* ctx, const Model& M, const qd_request& r, int method, qd_result* out) {
  const int64_t n = M.n;
  cplx *diag = nullptr, *psi = nullptr, *nxt = nullptr;
  qd_linop* op = nullptr;
  qd_krylov* kr = nullptr;
  int rc = QD_OK;
  auto done = [&](int code) {
    if (kr) qd_krylov_destroy(kr);
    if (op) qd_linop_destroy(op);
    hipStreamSynchronize(ctx->stream);
    qd_dev_free(ctx, diag);
    qd_dev_free(ctx, psi);
    qd_dev_free(ctx, nxt);
    return code;
  };
  if (stop_requested(S, r.tag)) return done(cancelled(ctx));
  if (qd_dev_alloc(ctx, sizeof(cplx) * n, (void**)&psi) != QD_OK ||
      qd_dev_alloc(ctx, sizeof(cplx) * n, (void**)&nxt) != QD_OK)
    return done(qd_set_error(ctx, QD_E_OOM, "request vectors"));
  // H = A + c D: a diagonal view of the model's A (for structured models its
  // matrix-free two-block form, else its stored matrix)
  if (qd_dev_alloc(ctx, sizeof(cplx) * n, (void**)&diag) != QD_OK)
    return done(qd_set_error(ctx, QD_E_OOM, "request diagonal"));
  scale_diag_kernel<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(n, r.coeff, M.d_diagD, diag);
  if ((rc = qd_linop_with_diagonal(ctx, M.structured ? M.opK : M.opA, (const double*)diag, &op))) return done(rc);
  if ((rc = qd_seeded_state(ctx, n, r.seed, (double*)psi))) return done(rc);
  // the basis scheme of the request (A + c D is Hermitian, so the partial
  // scheme applies to the real-time basis as to a ground state's)
  if ((rc = qd_krylov_create(ctx, n, r.k, method >= 2 ? 1 : 0, &kr))) return done(rc);
  if (method != 1 && (rc = qd_krylov_set_method(kr, method))) return done(rc);
  double h0[2] = {0.0, 0.0}, n0 = 0.0;
  if ((rc = qd_expect(ctx, op, (const double*)psi, h0))) return done(rc);
  if ((rc = qd_norm2(ctx, n, (const double*)psi, &n0))) return done(rc);
  int64_t msum = 0;
  int m_start = r.chunk > 0 ? r.chunk : 12;
  for (int s = 0; s < r.steps; ++s) {
    if (stop_requested(S, r.tag)) {  // between time steps
      out->iterations = s;
      return done(cancelled(ctx));
    }
    if (r.tol > 0.0) {
      int mu = 0, me = 0;
      rc = qd_krylov_evolve_tol(kr, op, (const double*)psi, (double*)nxt, r.dt, m_start, r.k, r.tol, &mu, &me);
      msum += mu;
      m_start = std::max(2, me);  // the next step starts where this one's estimate was met
    } else {
      rc = qd_krylov_evolve(kr, op, (const double*)psi, (double*)nxt, r.dt, r.k);
      msum += r.k - 1;
    }
    if (rc) return done(rc);
    std::swap(psi, nxt);
  }
  out->krylov_mean = r.steps > 0 ? (double)msum / r.steps : 0.0;
  double h1[2] = {0.0, 0.0}, n1 = 0.0, dexp = 0.0;
  if ((rc = qd_expect(ctx, op, (const double*)psi, h1))) return done(rc);
  if ((rc = diag_expect(ctx, n, M.d_diagD, psi, &dexp, &n1))) return done(rc);
  out->energy_start = h0[0] / n0;
  out->energy = h1[0] / n1;
  out->observable = dexp;
  out->iterations = r.steps;
  return done(QD_OK);
}

// Takes the first queued request that fits (arrival order; the head may be
// passed at most QD_SERVE_MAX_BYPASS times).  Requests that can never fit are
// answered with QD_E_OOM.  Called with the lock held.
bool pick(qd_server* S, Job* job) {
  size_t model_bytes = 0;
  for (Model* m : S->models) model_bytes += m->bytes;
  bool head_waiting = false;
  for (auto it = S->queue.begin(); it != S->queue.end();) {
    if (it->footprint + model_bytes > S->budget) {
      qd_result res{};
      res.tag = it->req.tag;
      res.status = QD_E_OOM;
      res.t_submit = it->t_submit;
      res.t_start = res.t_end = S->now();
      res.worker = -1;
      S->results.push_back(res);
      ++S->finished;
      ++S->rejected;
      it = S->queue.erase(it);
      S->cv.notify_all();
      continue;
    }
    if (it->req.kind == QD_REQ_EVOLVE && S->batch_active[it->req.model]) {
      ++it;  // its model's batch owner takes it at the next step
      continue;
    }
    if (it->footprint <= S->budget - S->used) {
      if (head_waiting) {
        Job& head = S->queue.front();
        if (head.bypassed >= QD_SERVE_MAX_BYPASS) return false;
        ++head.bypassed;
      }
      *job = *it;
      S->queue.erase(it);
      S->used += job->footprint;
      S->peak_used = std::max(S->peak_used, S->used);
      return true;
    }
    if (it == S->queue.begin()) head_waiting = true;
    ++it;
  }
  return false;
}

// ---------------------------------------------------------- evolve batches
// Evolve requests of one model with the same kry_dim_rt and dt advance
// together: one multi-RHS Arnoldi step (delayed CGS2 over all slots, the
// stored A read once per product for every slot, each slot's own diagonal
// coeff * D) per time step.  Requests that arrive while the batch runs join it
// at the next step; finished ones leave.  The Krylov dimension is shared: it
// grows by 2 (the step is redone) while the largest per-slot estimate of the
// step error exceeds the smallest tol, and shrinks by 1 when all are 100 times
// below it.  The batch reserves `footprint` per slot of the HBM budget.
// The product of the batch gathers one row of all slots per stored entry; for
// large sectors with few slots that moves more bytes than the matrix-free
// single-vector products of the same requests, so batching is off by default.
struct BatchSlot {
  Job job;
  int steps_done = 0;
  double e_start = 0.0;
  double t_start = 0.0;
  int64_t msum = 0;
};

void finish_result(qd_server* S, const qd_result& res) {
  std::lock_guard<std::mutex> lk(S->mu);
  mark_ended(S, res.tag);
  --S->running;
  S->results.push_back(res);
  ++S->finished;
}

int view_energy(qd_ctx* ctx, const Model& M, double coeff, const cplx* psi, cplx* diagc, double* e) {
  const int64_t n = M.n;
  scale_diag_kernel<<<blocks_for(n, 256), 256, 0, ctx->stream>>>(n, coeff, M.d_diagD, diagc);
  qd_linop* op = nullptr;
  int rc = qd_linop_with_diagonal(ctx, M.opA, (const double*)diagc, &op);
  if (rc) return rc;
  double h[2] = {0.0, 0.0}, n2 = 0.0;
  rc = qd_expect(ctx, op, (const double*)psi, h);
  if (!rc) rc = qd_norm2(ctx, n, (const double*)psi, &n2);
  qd_linop_destroy(op);
  *e = n2 > 0.0 ? h[0] / n2 : 0.0;
  return rc;
}

void run_evolve_batch(qd_server* S, qd_ctx* ctx, const Job& first, int wi) {
  const int model = first.req.model;
  const Model& M = *first.model;
  const int64_t n = M.n;
  const int Kmax = std::max(1, std::min(first.req.k - 1, MAXK));
  const double dt = first.req.dt;
  hipStream_t st = ctx->stream;
  const KronDesc* kd = qd_kron_desc(M.kronA);
  const int cmax = kd ? 64 : S->batch_slots;  // the Kronecker product takes up to 64 slots
  int C = 1;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    while (C < S->batch_slots && C < cmax && S->used + first.footprint <= S->budget) {
      S->used += first.footprint;
      ++C;
    }
    S->peak_used = std::max(S->peak_used, S->used);
    S->batch_active[model] = 1;
  }
  const int max_nch = (int)std::min<int64_t>(QD_TRAJ_MAX_CHUNKS, (n + 15) / 16);
  const int PMAX = 2 * Kmax + 2;
  std::vector<void*> allocs;
  bool ok = true;
  auto dalloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (qd_dev_alloc(ctx, bytes ? bytes : 16, &p) != QD_OK) {
      ok = false;
      return nullptr;
    }
    allocs.push_back(p);
    return p;
  };
  const size_t vecC = sizeof(cplx) * (size_t)n * C;
  cplx* X = (cplx*)dalloc(vecC);
  MrhsBufs mb{};
  mb.V = (cplx*)dalloc(vecC * Kmax);
  mb.U = (cplx*)dalloc(vecC);
  mb.W = (cplx*)dalloc(vecC);
  mb.part = (cplx*)dalloc(sizeof(cplx) * (size_t)max_nch * PMAX * C);
  mb.h = (cplx*)dalloc(sizeof(cplx) * C);
  mb.hess = (cplx*)dalloc(sizeof(cplx) * (size_t)Kmax * (Kmax + 1) * C);
  mb.red = (cplx*)dalloc(sizeof(cplx) * (size_t)PMAX * C);
  mb.coef2 = (cplx*)dalloc(sizeof(cplx) * (size_t)2 * Kmax * C);
  mb.sc = (cplx*)dalloc(sizeof(cplx) * 2 * C);
  mb.coef = (cplx*)dalloc(sizeof(cplx) * (size_t)Kmax * C);
  cplx* d_alpha = (cplx*)dalloc(sizeof(cplx) * C);
  cplx* d_beta = (cplx*)dalloc(sizeof(cplx) * C);
  double* d_scale = (double*)dalloc(sizeof(double) * C);
  double* d_est = (double*)dalloc(sizeof(double) * C);
  double* d_n2 = (double*)dalloc(sizeof(double) * C);
  int* d_ids = (int*)dalloc(sizeof(int) * C);
  GSec* d_gsec = (GSec*)dalloc(sizeof(GSec));
  cplx* tmp = (cplx*)dalloc(sizeof(cplx) * n);
  cplx* diagc = (cplx*)dalloc(sizeof(cplx) * n);
  std::vector<BatchSlot> slots;
  std::vector<int> ids(C);
  for (int i = 0; i < C; ++i) ids[i] = i;
  std::vector<cplx> h_alpha(C), h_beta(C, cmake(0.0, 0.0));
  std::vector<double> h_est(C);
  if (ok) {
    hipMemcpyAsync(d_ids, ids.data(), sizeof(int) * C, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_beta, h_beta.data(), sizeof(cplx) * C, hipMemcpyHostToDevice, st);
  }
  auto fail_slot = [&](BatchSlot& sl, int code) {
    qd_result res{};
    res.tag = sl.job.req.tag;
    res.status = code;
    res.t_submit = sl.job.t_submit;
    res.t_start = sl.t_start;
    res.t_end = S->now();
    res.worker = wi;
    finish_result(S, res);
  };
  std::vector<Job> pending{first};
  int Kcur = std::min(Kmax, first.req.chunk > 0 ? first.req.chunk : 12);
  int calm = 0;  // consecutive steps whose estimate had a wide margin
  int rc = ok ? QD_OK : QD_E_OOM;
  for (;;) {
    // admit: this model's queued evolve requests with the same k and dt
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (auto it = S->queue.begin(); it != S->queue.end() && (int)(slots.size() + pending.size()) < C;) {
        if (it->req.kind == QD_REQ_EVOLVE && it->req.model == model && it->req.k == first.req.k &&
            it->req.dt == dt) {
          pending.push_back(*it);
          mark_running(S, it->req.tag);
          it = S->queue.erase(it);
          ++S->running;
          S->max_running = std::max(S->max_running, S->running);
        } else {
          ++it;
        }
      }
      if (pending.empty() && slots.empty()) {
        S->batch_active[model] = 0;  // new arrivals start a new batch
        break;
      }
    }
    for (const Job& j : pending) {
      BatchSlot sl;
      sl.job = j;
      sl.t_start = S->now();
      if (rc == QD_OK) rc = qd_seeded_state(ctx, n, j.req.seed, (double*)tmp);
      if (rc == QD_OK) rc = view_energy(ctx, M, j.req.coeff, tmp, diagc, &sl.e_start);
      if (rc != QD_OK) {
        fail_slot(sl, rc);
        continue;
      }
      col_scatter_kernel<<<blocks_for(n, 256), 256, 0, st>>>(n, C, (int)slots.size(), tmp, X);
      slots.push_back(sl);
    }
    pending.clear();
    // requests with nothing (left) to do, or asked to stop
    for (int t = (int)slots.size() - 1; rc == QD_OK && t >= 0; --t) {
      BatchSlot& sl = slots[t];
      if (stop_requested(S, sl.job.req.tag)) {
        qd_result res{};
        res.tag = sl.job.req.tag;
        res.status = QD_E_CANCELLED;
        res.t_submit = sl.job.t_submit;
        res.t_start = sl.t_start;
        res.t_end = S->now();
        res.worker = wi;
        res.iterations = sl.steps_done;
        finish_result(S, res);
        const int last = (int)slots.size() - 1;
        if (t != last) {
          col_copy_kernel<<<blocks_for(n, 256), 256, 0, st>>>(n, C, last, t, X);
          slots[t] = slots[last];
        }
        slots.pop_back();
        continue;
      }
      if (sl.steps_done < sl.job.req.steps) continue;
      qd_result res{};
      res.tag = sl.job.req.tag;
      res.t_submit = sl.job.t_submit;
      res.t_start = sl.t_start;
      res.worker = wi;
      res.iterations = sl.steps_done;
      res.krylov_mean = sl.steps_done ? (double)sl.msum / sl.steps_done : 0.0;
      res.energy_start = sl.e_start;
      col_gather_kernel<<<blocks_for(n, 256), 256, 0, st>>>(n, C, t, X, tmp);
      double e = 0.0, dexp = 0.0, n2 = 0.0;
      rc = view_energy(ctx, M, sl.job.req.coeff, tmp, diagc, &e);
      if (rc == QD_OK) rc = diag_expect(ctx, n, M.d_diagD, tmp, &dexp, &n2);
      res.status = rc;
      res.energy = e;
      res.observable = dexp;
      res.t_end = S->now();
      finish_result(S, res);
      const int last = (int)slots.size() - 1;
      if (t != last) {
        col_copy_kernel<<<blocks_for(n, 256), 256, 0, st>>>(n, C, last, t, X);
        slots[t] = slots[last];
      }
      slots.pop_back();
    }
    if (rc != QD_OK) break;
    const int m = (int)slots.size();
    if (m == 0) continue;
    // one time step for every slot
    GSec g{};
    g.n = n;
    g.m = m;
    g.ldx = C;
    g.groups = (m + TT - 1) / TT;
    g.mp = TT;
    if (m <= TT / 2) {  // a few slots: 64 / mp rows per wave, mp >= m lanes per row
      g.mp = 1;
      while (g.mp < m) g.mp <<= 1;
    }
    g.rpc = rows_per_chunk(n, g.groups, 2048);
    g.nch = (int)((n + g.rpc - 1) / g.rpc);
    g.X = X;
    g.outer = M.csrA->d_outer;
    g.inner = M.csrA->d_inner;
    g.vals = M.csrA->d_vals;
    g.D = M.d_diagR;
    g.states = M.basis->d_states;
    const int64_t tiles = (int64_t)g.nch * g.groups;
    const int64_t blocks = (n + SPMM_RB - 1) / SPMM_RB * g.groups;
    const int64_t ngroups = g.groups;
    const int64_t VS = n * (int64_t)m;
    double tol = 1e300;
    for (int t = 0; t < m; ++t) {
      h_alpha[t] = cmake(slots[t].job.req.coeff, 0.0);
      if (slots[t].job.req.tol > 0.0) tol = std::min(tol, slots[t].job.req.tol);
    }
    hipMemcpyAsync(d_gsec, &g, sizeof(GSec), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_alpha, h_alpha.data(), sizeof(cplx) * m, hipMemcpyHostToDevice, st);
    // 1 / |x| per slot (the density kernels with no modes give |x|^2)
    g_dens_partial_kernel<<<(unsigned)tiles, TT, 0, st>>>(d_gsec, 1, 0, (double*)mb.part);
    g_dens_reduce_kernel<<<(unsigned)ngroups, RW * TT, 0, st>>>(d_gsec, 1, ngroups, 0, (const double*)mb.part, d_ids,
                                                                d_n2);
    inv_sqrt_kernel<<<g1(m, 256), 256, 0, st>>>(m, d_n2, d_scale);
    double maxest = 0.0;
    for (;;) {
      if (kd)
        mrhs_arnoldi_with(st, d_gsec, 1, tiles, ngroups, VS, m, Kcur, dt, d_scale, d_alpha, d_beta, mb,
                          KronProduct{*kd, d_gsec, m});
      else
        mrhs_arnoldi(st, d_gsec, 1, tiles, blocks, ngroups, VS, m, Kcur, dt, d_scale, d_alpha, d_beta, mb);
      g_step_estimate_kernel<<<g1(m, 256), 256, 0, st>>>(m, Kcur, mb.hess, mb.coef, d_est);
      hipMemcpyAsync(h_est.data(), d_est, sizeof(double) * m, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) {
        rc = QD_E_HIP;
        break;
      }
      maxest = 0.0;
      for (int t = 0; t < m; ++t) maxest = std::max(maxest, h_est[t]);
      if (tol < 1e300 && maxest > tol && Kcur < Kmax) {
        Kcur = std::min(Kmax, Kcur + 2);  // redo the step with a larger basis (X is untouched)
        continue;
      }
      break;
    }
    if (rc != QD_OK) break;
    mrhs_combine(st, d_gsec, 1, VS, m, Kcur, mb);
    for (auto& sl : slots) {
      ++sl.steps_done;
      sl.msum += Kcur;
    }
    // shrink only after a sustained margin: a step redone at a larger
    // dimension costs a whole Arnoldi step
    calm = (tol < 1e300 && maxest < 1e-3 * tol) ? calm + 1 : 0;
    if (calm >= 10 && Kcur > 2) {
      --Kcur;
      calm = 0;
    }
  }
  if (rc != QD_OK) {
    for (auto& sl : slots) fail_slot(sl, rc);
    for (auto& j : pending) {
      BatchSlot sl;
      sl.job = j;
      fail_slot(sl, rc);
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->batch_active[model] = 0;
  }
  hipStreamSynchronize(st);
  for (void* p : allocs) qd_dev_free(ctx, p);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->used -= (size_t)C * first.footprint;
  }
  S->cv.notify_all();
}

void worker_main(qd_server* S, int wi) {
  qd_ctx* ctx = nullptr;
  int crc = ctx_acquire(S->device, &ctx);
  if (ctx) ctx->pool = &S->pool;
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (crc == QD_OK && pick(S, &job)) break;
        if (S->stopping && S->queue.empty()) {
          lk.unlock();
          ctx_release(ctx);
          return;
        }
        if (crc != QD_OK && !S->queue.empty()) {
          // no device context on this worker: answer with the error
          job = S->queue.front();
          S->queue.pop_front();
          qd_result res{};
          res.tag = job.req.tag;
          res.status = crc;
          res.t_submit = job.t_submit;
          res.t_start = res.t_end = S->now();
          res.worker = wi;
          S->results.push_back(res);
          ++S->finished;
          S->cv.notify_all();
          continue;
        }
        S->cv.wait(lk);
      }
      ++S->running;
      S->max_running = std::max(S->max_running, S->running);
      mark_running(S, job.req.tag);
    }
    if (job.req.kind == QD_REQ_EVOLVE && S->batch_slots > 1) {
      run_evolve_batch(S, ctx, job, wi);
      continue;
    }
    qd_result res{};
    res.tag = job.req.tag;
    res.t_submit = job.t_submit;
    res.t_start = S->now();
    res.worker = wi;
    const Model& M = *job.model;
    res.status = job.req.kind == QD_REQ_GROUND_STATE ? run_ground_state(S, ctx, M, job.req, job.gs_method, &res)
                                                     : run_evolve(S, ctx, M, job.req, job.gs_method, &res);
    res.t_end = S->now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      mark_ended(S, job.req.tag);
      S->used -= job.footprint;
      --S->running;
      S->results.push_back(res);
      ++S->finished;
    }
    S->cv.notify_all();
  }
}

}  // namespace

extern "C" int qd_server_create(int device, size_t hbm_budget, int workers, qd_server** out) {
  if (!out || workers < 1 || workers > 64 || hbm_budget == 0) return QD_E_ARG;
  *out = nullptr;
  qd_server* S = new qd_server();
  S->device = device;
  S->budget = hbm_budget;
  S->pool.cap = hbm_budget;
  S->t0 = std::chrono::steady_clock::now();
  int rc = ctx_acquire(device, &S->setup);
  if (rc) {
    delete S;
    return rc;
  }
  for (int w = 0; w < workers; ++w) S->threads.emplace_back(worker_main, S, w);
  *out = S;
  return QD_OK;
}

extern "C" const char* qd_server_last_error(qd_server* s) { return s ? s->err.c_str() : "null server"; }

extern "C" int qd_server_add_model(qd_server* S, int nbits, const qd_constraint* cons, int ncons, const qd_term* tA,
                                   int ntA, const qd_op* oA, int noA, const qd_term* tD, int ntD, const qd_op* oD,
                                   int noD, int structured, int* model_id) {
  if (!S || !model_id || !tA || !oA || (ntD && (!tD || !oD))) return QD_E_ARG;
  qd_ctx* ctx = S->setup;
  Model* M = new Model();
  M->nbits = nbits;
  M->structured = structured;
  M->tA.assign(tA, tA + ntA);
  M->oA.assign(oA, oA + noA);
  if (ntD) {
    M->tD.assign(tD, tD + ntD);
    M->oD.assign(oD, oD + noD);
  }
  auto fail = [&](int code) {
    S->err = qd_last_error(ctx);
    if (M->opK) qd_linop_destroy(M->opK);
    if (M->opA) qd_linop_destroy(M->opA);
    if (M->csrA) qd_csr_destroy(M->csrA);
    if (M->progA) qd_termprog_destroy(M->progA);
    if (M->basis) qd_basis_destroy(M->basis);
    hipStreamSynchronize(ctx->stream);
    if (M->d_diagD) hipFree(M->d_diagD);
    if (M->d_diagR) hipFree(M->d_diagR);
    delete M;
    return code;
  };
  int rc = qd_basis_create(ctx, nbits, cons, ncons, &M->basis);
  if (rc) return fail(rc);
  M->n = qd_basis_size(M->basis);
  if ((rc = qd_termprog_create(ctx, nbits, tA, ntA, oA, noA, &M->progA))) return fail(rc);
  if ((rc = qd_csr_build(ctx, M->basis, M->progA, &M->csrA))) return fail(rc);
  if ((rc = qd_linop_from_csr(ctx, M->csrA, &M->opA))) return fail(rc);
  if (structured) {
    // A's two-block form, built once: real-time requests add c D as a
    // diagonal view of it; its launch variants are chosen here, before
    // several workers share it
    if ((rc = qd_linop_matfree(ctx, M->basis, M->progA, &M->opK))) return fail(rc);
    if (M->opK->kron) {
      cplx *sx = nullptr, *sy = nullptr;
      if (hipMalloc(&sx, sizeof(cplx) * M->n) != hipSuccess || hipMalloc(&sy, sizeof(cplx) * M->n) != hipSuccess) {
        hipFree(sx);
        return fail(qd_set_error(ctx, QD_E_OOM, "model scratch"));
      }
      rc = qd_seeded_state(ctx, M->n, 1, (double*)sx);
      if (rc == QD_OK) rc = qd_kron_prepare(ctx, M->opK->kron, sx, sy);
      hipStreamSynchronize(ctx->stream);
      hipFree(sx);
      hipFree(sy);
      if (rc) return fail(rc);
      const KronDesc* kd = qd_kron_desc(M->opK->kron);
      if (kd->ell && kd->ellB) M->kronA = M->opK->kron;  // the batched product needs both packed tables
    }
  }
  if (hipMalloc(&M->d_diagD, sizeof(cplx) * M->n) != hipSuccess) return fail(qd_set_error(ctx, QD_E_OOM, "model diagonal"));
  if (ntD) {
    qd_termprog* pD = nullptr;
    qd_csr* cD = nullptr;
    if ((rc = qd_termprog_create(ctx, nbits, tD, ntD, oD, noD, &pD))) return fail(rc);
    rc = qd_csr_build(ctx, M->basis, pD, &cD);
    if (rc) {
      qd_termprog_destroy(pD);
      return fail(rc);
    }
    int* flag = nullptr;
    int hflag = 0;
    if (hipMalloc(&flag, sizeof(int)) == hipSuccess) {
      hipMemsetAsync(flag, 0, sizeof(int), ctx->stream);
      extract_diag_kernel<<<blocks_for(M->n, 256), 256, 0, ctx->stream>>>(M->n, cD->d_outer, cD->d_inner, cD->d_vals,
                                                                          M->d_diagD, flag);
      hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
      hipStreamSynchronize(ctx->stream);
      hipFree(flag);
    } else {
      hflag = 2;
    }
    qd_csr_destroy(cD);
    qd_termprog_destroy(pD);
    if (hflag) return fail(qd_set_error(ctx, hflag == 1 ? QD_E_ARG : QD_E_OOM, "the second operator must be diagonal"));
  } else {
    hipMemsetAsync(M->d_diagD, 0, sizeof(cplx) * M->n, ctx->stream);
  }
  if (hipMalloc(&M->d_diagR, sizeof(double) * M->n) != hipSuccess)
    return fail(qd_set_error(ctx, QD_E_OOM, "model diagonal"));
  real_part_kernel<<<blocks_for(M->n, 256), 256, 0, ctx->stream>>>(M->n, M->d_diagD, M->d_diagR);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(qd_set_error(ctx, QD_E_HIP, "model set-up"));
  const qd_csr* c = M->csrA;
  M->bytes = sizeof(uint64_t) * (size_t)M->n + sizeof(int32_t) * (size_t)(M->n + 1) +
             (sizeof(int32_t) + sizeof(cplx)) * (size_t)c->nnz + (sizeof(cplx) + sizeof(double)) * (size_t)M->n;
  std::lock_guard<std::mutex> lk(S->mu);
  *model_id = (int)S->models.size();
  S->models.push_back(M);
  S->batch_active.push_back(0);
  S->used += M->bytes;
  S->peak_used = std::max(S->peak_used, S->used);
  return QD_OK;
}

extern "C" int qd_server_submit(qd_server* S, const qd_request* r) {
  if (!S || !r) return QD_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  if (r->model < 0 || r->model >= (int)S->models.size() || r->k < 2 ||
      (r->kind != QD_REQ_GROUND_STATE && r->kind != QD_REQ_EVOLVE) || (r->kind == QD_REQ_EVOLVE && r->steps < 0)) {
    S->err = "qd_server_submit: bad request";
    return QD_E_ARG;
  }
  Job j;
  j.req = *r;
  j.model = S->models[r->model];
  j.t_submit = S->now();
  j.footprint = request_footprint(*j.model, *r);
  j.gs_method = S->gs_method;
  S->queue.push_back(j);
  ++S->submitted;
  S->cv.notify_all();
  return QD_OK;
}

extern "C" int qd_server_cancel(qd_server* S, int64_t tag) {
  if (!S) return QD_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  bool found = false;
  for (auto it = S->queue.begin(); it != S->queue.end();) {
    if (it->req.tag != tag) {
      ++it;
      continue;
    }
    qd_result res{};
    res.tag = tag;
    res.status = QD_E_CANCELLED;
    res.t_submit = it->t_submit;
    res.t_start = res.t_end = S->now();
    res.worker = -1;
    S->results.push_back(res);
    ++S->finished;
    it = S->queue.erase(it);
    found = true;
  }
  if (S->active.count(tag)) {
    if (S->cancel.emplace(tag, 1).second) S->ncancel.fetch_add(1, std::memory_order_release);
    found = true;
  }
  if (found) S->cv.notify_all();
  if (!found) {
    S->err = "qd_server_cancel: no queued or running request with this tag";
    return QD_E_ARG;
  }
  return QD_OK;
}

extern "C" int qd_server_poll(qd_server* S, qd_result* out, int cap, int* count) {
  if (!S || !count || (cap > 0 && !out)) return QD_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  int k = std::min<int>(cap, (int)S->results.size());
  for (int i = 0; i < k; ++i) out[i] = S->results[i];
  S->results.erase(S->results.begin(), S->results.begin() + k);
  *count = k;
  return QD_OK;
}

extern "C" int qd_server_drain(qd_server* S) {
  if (!S) return QD_E_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv.wait(lk, [&] { return S->finished == S->submitted; });
  return QD_OK;
}

extern "C" int qd_server_set_batch(qd_server* S, int slots) {
  if (!S || slots < 1 || slots > 4096) return QD_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  S->batch_slots = slots;
  return QD_OK;
}

extern "C" int qd_server_set_gs_method(qd_server* S, int method) {
  if (!S || method < 0 || method > 2) return QD_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  S->gs_method = method;
  return QD_OK;
}

extern "C" double qd_server_clock(qd_server* S) { return S ? S->now() : 0.0; }

extern "C" int qd_server_get_stats(qd_server* S, qd_server_stats* st) {
  if (!S || !st) return QD_E_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  st->submitted = S->submitted;
  st->finished = S->finished;
  st->rejected = S->rejected;
  st->queued = (int64_t)S->queue.size();
  st->running = S->running;
  st->max_running = S->max_running;
  st->bytes_in_use = (int64_t)S->used;
  st->peak_bytes = (int64_t)S->peak_used;
  return QD_OK;
}

extern "C" int qd_server_destroy(qd_server* S) {
  if (!S) return QD_E_ARG;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv.notify_all();
  for (auto& t : S->threads) t.join();
  for (Model* M : S->models) {
    if (M->opK) qd_linop_destroy(M->opK);
    if (M->opA) qd_linop_destroy(M->opA);
    if (M->csrA) qd_csr_destroy(M->csrA);
    if (M->progA) qd_termprog_destroy(M->progA);
    if (M->basis) qd_basis_destroy(M->basis);
    if (M->d_diagD) hipFree(M->d_diagD);
    if (M->d_diagR) hipFree(M->d_diagR);
    delete M;
  }
  ctx_release(S->setup);
  qd_devpool_release_all(&S->pool);
  delete S;
  return QD_OK;
}
