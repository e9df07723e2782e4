// traj.hip -- K11: batched quantum-jump trajectories (programs/hubbard_diss.cpp:440-482,
// mc_manager.hpp:88-128), generalised to dephasing and loss channels with
// sector migration.
//
// Trajectories live in per-sector blocks, trajectory-minor: X[s * ld + t] is
// state s of the trajectory in slot t (slots 0..m-1 packed).  All blocks of
// one sector share its Hamiltonian, so the multi-RHS product reads each row
// of H once for every trajectory of the sector, and a wave's 64 lanes read 64
// consecutive trajectories (1 KB) per stored entry.  Each trajectory runs its
// own Arnoldi step (the reference's evolve with H_eff, hubbard_diss.cpp:467-469)
// with CGS2, per-trajectory Hessenberg and a per-trajectory exponential on the
// device.  Jumps are drawn on the device from a counter-based stream per
// trajectory, applied from the pre-step state into a staging area, and
// committed after the sector's evolution (moving lost trajectories to their
// new sector).  Every reduction has a fixed order, so a trajectory's result
// does not depend on which other trajectories share its batch.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qd_internal.hpp"

namespace {

constexpr int TT = 64;      // trajectories per wave
constexpr int MAXM = 32;    // modes (2 * nsite) supported
constexpr int MAXK = 32;    // Krylov dimension limit for the batched step

__device__ inline uint64_t tsm64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double traj_uniform(uint64_t seed, int64_t traj, int64_t step, int k) {
  uint64_t key = tsm64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)traj);
  return (double)(tsm64(key + 4ull * (uint64_t)step + (uint64_t)k) >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------- kernels
// D[s] = sum_c rate_c * n_{mode_c}(s)   (sum_c L_c^+ L_c on the sector)
__global__ void diag_rates_kernel(int64_t n, const uint64_t* __restrict__ states, const int* __restrict__ jmode,
                                  const double* __restrict__ jrate, int nj, double* __restrict__ D) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint64_t x = states[s];
  double d = 0.0;
  for (int c = 0; c < nj; ++c)
    if ((x >> jmode[c]) & 1ull) d += jrate[c];
  D[s] = d;
}

// column partials over row chunks.  Grid: (ceil(m/64), nchunk).
// part[(chunk * nq + q) * m + t]
// mode 0: q < j+1: conj(V_q[s,t]) W[s,t]          (Gram-Schmidt dots)
// mode 1: q = 0: |X[s,t]|^2                        (norms)
__global__ void __launch_bounds__(TT)
    col_dots_kernel(int64_t n, int m, int ldv, const cplx* __restrict__ V, int64_t vstride, int nq,
                    const cplx* __restrict__ W, int rows_per_chunk, cplx* __restrict__ part) {
  const int t = blockIdx.x * TT + threadIdx.x;
  const int chunk = blockIdx.y;
  if (t >= m) return;
  const int64_t r0 = (int64_t)chunk * rows_per_chunk;
  const int64_t r1 = min<int64_t>(n, r0 + rows_per_chunk);
  cplx acc[MAXK + 1];
  for (int q = 0; q < nq; ++q) acc[q] = cmake(0.0, 0.0);
  for (int64_t s = r0; s < r1; ++s) {
    const cplx w = W[s * ldv + t];
    for (int q = 0; q < nq; ++q) {
      const cplx v = V[q * vstride + s * ldv + t];
      acc[q] = cadd(acc[q], cdotmul(v, w));
    }
  }
  for (int q = 0; q < nq; ++q) part[((int64_t)chunk * nq + q) * m + t] = acc[q];
}

// out[q * m + t] = sum_chunk part[(chunk * nq + q) * m + t]  (fixed order)
__global__ void col_reduce_kernel(int m, int nq, int nchunk, const cplx* __restrict__ part, cplx* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * nq) return;
  const int q = idx / m, t = idx % m;
  cplx s = cmake(0.0, 0.0);
  for (int c = 0; c < nchunk; ++c) s = cadd(s, part[((int64_t)c * nq + q) * m + t]);
  out[q * m + t] = s;
}

// W[s,t] -= sum_q V_q[s,t] h[q*m+t]; Hess column j accumulates h
__global__ void __launch_bounds__(256)
    col_update_kernel(int64_t n, int m, int ldv, const cplx* __restrict__ V, int64_t vstride, int nq,
                      const cplx* __restrict__ h, cplx* __restrict__ W) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(idx % ldv);
  const int64_t s = idx / ldv;
  if (s >= n || t >= m) return;
  cplx w = W[s * ldv + t];
  for (int q = 0; q < nq; ++q) {
    const cplx v = V[q * vstride + s * ldv + t];
    const cplx c = h[q * m + t];
    w.x -= v.x * c.x - v.y * c.y;
    w.y -= v.x * c.y + v.y * c.x;
  }
  W[s * ldv + t] = w;
}

// Hess[(j * (K+1) + i) * m + t] (+)= h[i * m + t] for i <= j
__global__ void hess_acc_kernel(int m, int j, int K, const cplx* __restrict__ h, cplx* __restrict__ hess, int acc) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * (j + 1)) return;
  const int i = idx / m, t = idx % m;
  cplx* p = hess + ((int64_t)(j * (K + 1) + i)) * m + t;
  *p = acc ? cadd(*p, h[i * m + t]) : h[i * m + t];
}

// beta_t = sqrt(nrm[t]); Hess[j+1][j][t] = beta; scale[t] = 1/beta (0 on breakdown)
__global__ void beta_kernel(int m, int j, int K, const cplx* __restrict__ nrm, cplx* __restrict__ hess,
                            double* __restrict__ scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const double b = sqrt(nrm[t].x);
  hess[((int64_t)(j * (K + 1) + j + 1)) * m + t] = cmake(b, 0.0);
  scale[t] = (b >= 1e-10) ? 1.0 / b : 0.0;
}

// scale[t] = 1 / sqrt(nrm[t]) (0 for a zero column)
__global__ void inv_norm_kernel(int m, const cplx* __restrict__ nrm, double* __restrict__ scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const double b = sqrt(nrm[t].x);
  scale[t] = (b > 0.0) ? 1.0 / b : 0.0;
}

// dst[s,t] = src[s,t] * scale[t]
__global__ void col_scale_kernel(int64_t n, int m, int lds, const cplx* __restrict__ src, int ldd, cplx* __restrict__ dst,
                                 const double* __restrict__ scale) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(idx % m);
  const int64_t s = idx / m;
  if (s >= n) return;
  const cplx v = src[s * lds + t];
  const double f = scale[t];
  dst[s * ldd + t] = cmake(v.x * f, v.y * f);
}

// multi-RHS product with the per-trajectory non-Hermitian shift:
// W[s,t] = sum_e conj(val_e) V[col_e,t] - i/2 (D[s] - Gamma_t) V[s,t]
// One wave per (row s, 64-trajectory group): the row's entries are uniform
// across the wave (scalar loads), the gathers are 1 KB coalesced.
__global__ void __launch_bounds__(TT)
    spmm_tm_kernel(int64_t n, int m, int ldv, const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                   const cplx* __restrict__ vals, const double* __restrict__ D, const double* __restrict__ gamma,
                   const cplx* __restrict__ V, cplx* __restrict__ W) {
  const int64_t s = blockIdx.x;
  const int t = blockIdx.y * TT + threadIdx.x;
  if (t >= m) return;
  double sx = 0.0, sy = 0.0;
  const int32_t e0 = outer[s], e1 = outer[s + 1];
  for (int32_t e = e0; e < e1; ++e) {
    const cplx v = vals[e];
    const cplx x = V[(int64_t)inner[e] * ldv + t];
    sx += v.x * x.x + v.y * x.y;
    sy += v.x * x.y - v.y * x.x;
  }
  const cplx x = V[s * ldv + t];
  const double d = -0.5 * (D[s] - gamma[t]);  // (0, d) * x
  W[s * ldv + t] = cmake(sx - d * x.y, sy + d * x.x);
}

// y_t = expm(-i dt K_t) e_1 for the k x k Hessenberg of each trajectory:
// scaling and squaring with a truncated Taylor series applied to vectors
// (Moler & Van Loan, "Nineteen dubious ways to compute the exponential of a
// matrix", methods 1 and 3).  One thread per trajectory.
__global__ void small_expm_kernel(int m, int k, int K, const cplx* __restrict__ hess, double dt, cplx* __restrict__ coef) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  auto Hget = [&](int i, int j) -> cplx { return hess[((int64_t)(j * (K + 1) + i)) * m + t]; };
  // 1-norm of A = -i dt K (upper Hessenberg: rows 0 .. j+1 of column j)
  double nrm = 0.0;
  for (int j = 0; j < k; ++j) {
    double c = 0.0;
    for (int i = 0; i <= min(j + 1, k - 1); ++i) {
      cplx h = Hget(i, j);
      c += sqrt(h.x * h.x + h.y * h.y);
    }
    nrm = fmax(nrm, c * fabs(dt));
  }
  int sq = 0;
  while (nrm > 0.5 && sq < 16) {
    nrm *= 0.5;
    ++sq;
  }
  const double sc = ldexp(dt, -sq);
  cplx y[MAXK], term[MAXK], nt[MAXK];
  for (int i = 0; i < k; ++i) y[i] = cmake(i == 0 ? 1.0 : 0.0, 0.0);
  const int reps = 1 << sq;
  for (int r = 0; r < reps; ++r) {
    for (int i = 0; i < k; ++i) term[i] = y[i];
    for (int p = 1; p <= 24; ++p) {
      // nt = (-i sc K) term / p
      double tn = 0.0;
      for (int i = 0; i < k; ++i) {
        cplx a = cmake(0.0, 0.0);
        for (int j = (i > 0 ? i - 1 : 0); j < k; ++j) a = cadd(a, cmul(Hget(i, j), term[j]));
        // -i * sc * a / p
        nt[i] = cmake(sc * a.y / p, -sc * a.x / p);
      }
      for (int i = 0; i < k; ++i) {
        term[i] = nt[i];
        y[i] = cadd(y[i], nt[i]);
        tn += fabs(nt[i].x) + fabs(nt[i].y);
      }
      if (tn < 1e-18) break;
    }
  }
  for (int i = 0; i < k; ++i) coef[i * m + t] = y[i];
}

// X[s,t] = sum_j V_j[s,t] coef[j*m+t]
__global__ void combine_tm_kernel(int64_t n, int m, int ldv, const cplx* __restrict__ V, int64_t vstride, int k,
                                  const cplx* __restrict__ coef, int ldx, cplx* __restrict__ X) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(idx % m);
  const int64_t s = idx / m;
  if (s >= n) return;
  cplx a = cmake(0.0, 0.0);
  for (int j = 0; j < k; ++j) a = cadd(a, cmul(V[j * vstride + s * ldv + t], coef[j * m + t]));
  X[s * ldx + t] = a;
}

// densities: part[(chunk * (M+1) + q) * m + t] = sum_{s in chunk} |X[s,t]|^2 [bit q of state s]
// (q = M: plain norm^2)
__global__ void __launch_bounds__(TT)
    dens_partial_kernel(int64_t n, int m, int ld, const cplx* __restrict__ X, const uint64_t* __restrict__ states, int M,
                        int rows_per_chunk, double* __restrict__ part) {
  const int t = blockIdx.x * TT + threadIdx.x;
  const int chunk = blockIdx.y;
  if (t >= m) return;
  const int64_t r0 = (int64_t)chunk * rows_per_chunk;
  const int64_t r1 = min<int64_t>(n, r0 + rows_per_chunk);
  double acc[MAXM + 1];
  for (int q = 0; q <= M; ++q) acc[q] = 0.0;
  for (int64_t s = r0; s < r1; ++s) {
    const cplx v = X[s * ld + t];
    const double p = v.x * v.x + v.y * v.y;
    const uint64_t x = states[s];
    for (int q = 0; q < M; ++q)
      if ((x >> q) & 1ull) acc[q] += p;
    acc[M] += p;
  }
  for (int q = 0; q <= M; ++q) part[((int64_t)chunk * (M + 1) + q) * m + t] = acc[q];
}

// per slot: normalised densities into dens_all[traj * (M+1) + q] (q = M: norm^2)
__global__ void dens_reduce_kernel(int m, int M, int nchunk, const double* __restrict__ part,
                                   const int* __restrict__ slot2traj, double* __restrict__ dens_all) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  double v[MAXM + 1];
  for (int q = 0; q <= M; ++q) {
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[((int64_t)c * (M + 1) + q) * m + t];
    v[q] = s;
  }
  const int tr = slot2traj[t];
  for (int q = 0; q < M; ++q) dens_all[(int64_t)tr * (M + 1) + q] = v[q] / v[M];
  dens_all[(int64_t)tr * (M + 1) + M] = v[M];
}

// dens_sum[step * M + q] = sum over trajectories (index order) of dens_all; dens_sq likewise
__global__ void dens_accumulate_kernel(int ntraj, int M, const double* __restrict__ dens_all,
                                       double* __restrict__ dsum, double* __restrict__ dsq) {
  const int q = blockIdx.x;
  __shared__ double sh[2][4];
  double s = 0.0, s2 = 0.0;
  for (int tr = threadIdx.x; tr < ntraj; tr += 256) {
    double v = dens_all[(int64_t)tr * (M + 1) + q];
    s += v;
    s2 += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = s;
    sh[1][threadIdx.x >> 6] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    dsum[q] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    dsq[q] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// jump test per trajectory (hubbard_diss.cpp:450-463 generalised)
__global__ void decide_kernel(int ntraj, int M, const double* __restrict__ dens_all, const int* __restrict__ jmode,
                              const double* __restrict__ jrate, int nj, uint64_t seed, int64_t step, double dt,
                              int* __restrict__ action, double* __restrict__ gamma_traj) {
  const int tr = blockIdx.x * blockDim.x + threadIdx.x;
  if (tr >= ntraj) return;
  const double* d = dens_all + (int64_t)tr * (M + 1);
  double G = 0.0;
  for (int c = 0; c < nj; ++c) G += jrate[c] * d[jmode[c]];
  gamma_traj[tr] = G;
  const double r0 = traj_uniform(seed, tr, step, 0);
  const double r1 = traj_uniform(seed, tr, step, 1);
  int ch = -1;
  if (r0 < G * dt) {
    const double target = r1 * G;
    double cum = 0.0;
    for (int c = 0; c < nj; ++c) {
      const double w = jrate[c] * d[jmode[c]];
      cum += w;
      if (w > 0.0) ch = c;
      if (target < cum && w > 0.0) break;
    }
  }
  action[tr] = ch;
}

// gather per-slot Gamma from the per-trajectory array
__global__ void gather_gamma_kernel(int m, const int* __restrict__ slot2traj, const double* __restrict__ gtraj,
                                    double* __restrict__ g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) g[t] = gtraj[slot2traj[t]];
}

struct JumpRec {
  int slot;       // source slot
  int mode;       // mode m
  int loss;       // 0 dephasing, 1 loss
  int staging;    // staging vector index
  double scale;   // 1 / |L psi| (from the recorded densities and norm)
};

// staged[j][target] = L psi_slot (normalised); one block per jump
__global__ void jump_apply_kernel(int64_t n, int ld, const cplx* __restrict__ X, const uint64_t* __restrict__ states,
                                  const JumpRec* __restrict__ jr, SectorDesc tup, SectorDesc tdn, int nsite,
                                  int64_t stg_ld, cplx* __restrict__ staging) {
  const JumpRec J = jr[blockIdx.x];
  cplx* out = staging + (int64_t)J.staging * stg_ld;
  for (int64_t s = threadIdx.x; s < n; s += blockDim.x) {
    const uint64_t x = states[s];
    if (!((x >> J.mode) & 1ull)) continue;
    cplx v = X[s * ld + J.slot];
    v = cmake(v.x * J.scale, v.y * J.scale);
    if (!J.loss) {
      out[s] = v;
    } else {
      const uint64_t y = x & ~(1ull << J.mode);
      const double sg = (__popcll(x >> (J.mode + 1)) & 1) ? -1.0 : 1.0;
      const uint64_t r = qd_rank_cns(J.mode < nsite ? tup : tdn, y);
      out[r] = cmake(sg * v.x, sg * v.y);
    }
  }
}

// X[s, dst_slot] = staging[j][s] for each (staging index, slot) pair
__global__ void commit_kernel(int64_t n, int ld, cplx* __restrict__ X, const int2* __restrict__ pairs,
                              int64_t stg_ld, const cplx* __restrict__ staging) {
  const int2 p = pairs[blockIdx.x];
  for (int64_t s = threadIdx.x; s < n; s += blockDim.x) X[s * ld + p.y] = staging[(int64_t)p.x * stg_ld + s];
}

// column moves inside a block (swap-remove compaction): X[s, dst] = X[s, src]
__global__ void move_cols_kernel(int64_t n, int ld, cplx* __restrict__ X, const int2* __restrict__ moves, int nmoves) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int k = 0; k < nmoves; ++k) X[s * ld + moves[k].y] = X[s * ld + moves[k].x];
}

__global__ void zero_kernel(int64_t count, cplx* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = cmake(0.0, 0.0);
}

// block[s * ld + slot] = psi0[s] for slots 0..m-1
__global__ void broadcast_kernel(int64_t n, int m, int ld, const cplx* __restrict__ psi0, cplx* __restrict__ X) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(idx % m);
  const int64_t s = idx / m;
  if (s < n) X[s * ld + t] = psi0[s];
}

unsigned g1(int64_t work, int b) { return (unsigned)std::max<int64_t>(1, (work + b - 1) / b); }

// ---------------------------------------------------------------- host state
struct Sector {
  int a = -1, b = -1;
  int64_t n = 0;
  qd_basis* basis = nullptr;
  qd_csr* csr = nullptr;
  double* D = nullptr;
  cplx* X = nullptr;  // n x ld, trajectory-minor
  int ld = 0;
  int m = 0;                 // active slots
  std::vector<int> slot2traj;
  int* d_slot2traj = nullptr;
};

}  // namespace

extern "C" int qd_traj_run(qd_ctx* ctx, int nsite, const qd_term* terms, int nterms, const qd_op* ops, int nops,
                           int nup0, int ndn0, const double* d_psi0, const qd_jump* jumps, int njumps, int ntraj,
                           int nsteps, double dt, int k_rt, uint64_t seed, double* h_dens_mean, double* h_dens_var,
                           int* h_jump_log, qd_traj_stats* stats) {
  if (!ctx || nsite < 1 || 2 * nsite > MAXM || !terms || !ops || !d_psi0 || ntraj < 1 || nsteps < 0 || k_rt < 2 ||
      k_rt - 1 > MAXK || nup0 < 0 || nup0 > nsite || ndn0 < 0 || ndn0 > nsite || !h_dens_mean || (njumps && !jumps))
    return qd_set_error(ctx, QD_E_ARG, "qd_traj_run: bad arguments");
  const int M = 2 * nsite;
  const int K = k_rt - 1;  // Krylov vectors used (m = kry_dim_rt - 1, sparse_algorithm.hpp:557)
  hipStream_t st = ctx->stream;
  hipSetDevice(ctx->device);
  qd_termprog* prog = nullptr;
  int rc = qd_termprog_create(ctx, M, terms, nterms, ops, nops, &prog);
  if (rc) return rc;

  std::vector<int> jk(njumps), jm(njumps);
  std::vector<double> jrt(njumps);
  for (int c = 0; c < njumps; ++c) {
    jk[c] = jumps[c].kind;
    jm[c] = jumps[c].mode;
    jrt[c] = jumps[c].rate;
    if ((jk[c] != QD_JUMP_DEPHASE && jk[c] != QD_JUMP_LOSS) || jm[c] < 0 || jm[c] >= M || jrt[c] < 0) {
      qd_termprog_destroy(prog);
      return qd_set_error(ctx, QD_E_ARG, "jump channel %d invalid", c);
    }
  }
  std::vector<Sector> sec((nsite + 1) * (nsite + 1));
  auto sid = [&](int a, int b) { return a * (nsite + 1) + b; };
  // device scratch
  int *d_jmode = nullptr, *d_action = nullptr;
  double *d_jrate = nullptr, *d_dens_all = nullptr, *d_gtraj = nullptr, *d_dsum = nullptr, *d_dsq = nullptr;
  double *d_gslot = nullptr, *d_scale = nullptr, *d_dpart = nullptr;
  cplx *d_V = nullptr, *d_W = nullptr, *d_part = nullptr, *d_h = nullptr, *d_hess = nullptr, *d_coef = nullptr;
  cplx* d_staging = nullptr;
  JumpRec* d_jr = nullptr;
  int2 *d_pairs = nullptr, *d_moves = nullptr;
  std::vector<double> dsum((size_t)(nsteps + 1) * M, 0.0), dsq((size_t)(nsteps + 1) * M, 0.0);
  int64_t n_jumps = 0, n_loss = 0;
  std::vector<void*> allocs;
  auto dalloc = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) return false;
    allocs.push_back(*p);
    return true;
  };
  auto cleanup = [&]() {
    hipStreamSynchronize(st);
    for (auto& S : sec) {
      if (S.csr) qd_csr_destroy(S.csr);
      if (S.basis) qd_basis_destroy(S.basis);
    }
    for (void* p : allocs) hipFree(p);
    qd_termprog_destroy(prog);
  };
  auto ensure_sector = [&](int a, int b) -> int {
    Sector& S = sec[sid(a, b)];
    if (S.basis) return QD_OK;
    S.a = a;
    S.b = b;
    qd_constraint cons[2] = {{(1ull << nsite) - 1ull, a}, {((1ull << nsite) - 1ull) << nsite, b}};
    int r = qd_basis_create(ctx, M, cons, 2, &S.basis);
    if (r) return r;
    S.n = qd_basis_size(S.basis);
    r = qd_csr_build(ctx, S.basis, prog, &S.csr);
    if (r) return r;
    if (S.csr->hermitian != 1) return qd_set_error(ctx, QD_E_UNSUPPORTED, "trajectory Hamiltonian must be Hermitian");
    S.ld = ntraj;
    S.slot2traj.reserve(ntraj);
    if (!dalloc((void**)&S.D, sizeof(double) * S.n) || !dalloc((void**)&S.X, sizeof(cplx) * S.n * S.ld) ||
        !dalloc((void**)&S.d_slot2traj, sizeof(int) * ntraj))
      return qd_set_error(ctx, QD_E_OOM, "sector (%d,%d) buffers", a, b);
    diag_rates_kernel<<<g1(S.n, 256), 256, 0, st>>>(S.n, S.basis->d_states, d_jmode, d_jrate, njumps, S.D);
    return QD_OK;
  };

  // largest sector reachable by losses from (nup0, ndn0): C(n, a) is not monotone in a
  int64_t nmax = 1;
  for (int a = 0; a <= nup0; ++a)
    for (int b = 0; b <= ndn0; ++b)
      nmax = std::max<int64_t>(nmax, (int64_t)qd_binom_host(nsite, a) * (int64_t)qd_binom_host(nsite, b));
  const int max_chunks = 256;
  if (!dalloc((void**)&d_jmode, sizeof(int) * std::max(1, njumps)) ||
      !dalloc((void**)&d_jrate, sizeof(double) * std::max(1, njumps)) ||
      !dalloc((void**)&d_action, sizeof(int) * ntraj) || !dalloc((void**)&d_dens_all, sizeof(double) * ntraj * (M + 1)) ||
      !dalloc((void**)&d_gtraj, sizeof(double) * ntraj) || !dalloc((void**)&d_dsum, sizeof(double) * M) ||
      !dalloc((void**)&d_dsq, sizeof(double) * M) || !dalloc((void**)&d_gslot, sizeof(double) * ntraj) ||
      !dalloc((void**)&d_scale, sizeof(double) * ntraj) ||
      !dalloc((void**)&d_dpart, sizeof(double) * (size_t)max_chunks * (M + 1) * ntraj) ||
      !dalloc((void**)&d_V, sizeof(cplx) * (size_t)(K + 1) * nmax * ntraj) ||
      !dalloc((void**)&d_W, sizeof(cplx) * (size_t)nmax * ntraj) ||
      !dalloc((void**)&d_part, sizeof(cplx) * (size_t)max_chunks * (K + 1) * ntraj) ||
      !dalloc((void**)&d_h, sizeof(cplx) * (size_t)(K + 1) * ntraj) ||
      !dalloc((void**)&d_hess, sizeof(cplx) * (size_t)K * (K + 1) * ntraj) ||
      !dalloc((void**)&d_coef, sizeof(cplx) * (size_t)K * ntraj) ||
      !dalloc((void**)&d_jr, sizeof(JumpRec) * ntraj) || !dalloc((void**)&d_pairs, sizeof(int2) * ntraj) ||
      !dalloc((void**)&d_moves, sizeof(int2) * ntraj)) {
    cleanup();
    return qd_set_error(ctx, QD_E_OOM, "trajectory workspace (%d trajectories, dim %lld, k %d)", ntraj,
                        (long long)nmax, k_rt);
  }
  if (njumps) {
    hipMemcpyAsync(d_jmode, jm.data(), sizeof(int) * njumps, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_jrate, jrt.data(), sizeof(double) * njumps, hipMemcpyHostToDevice, st);
  }
  rc = ensure_sector(nup0, ndn0);
  if (rc) {
    cleanup();
    return rc;
  }
  {
    Sector& S = sec[sid(nup0, ndn0)];
    S.m = ntraj;
    S.slot2traj.resize(ntraj);
    for (int t = 0; t < ntraj; ++t) S.slot2traj[t] = t;
    hipMemcpyAsync(S.d_slot2traj, S.slot2traj.data(), sizeof(int) * ntraj, hipMemcpyHostToDevice, st);
    broadcast_kernel<<<g1(S.n * ntraj, 256), 256, 0, st>>>(S.n, ntraj, S.ld, (const cplx*)d_psi0, S.X);
  }
  std::vector<int> action(ntraj);
  // per-step jump lists; reserved for ntraj entries so that the async copies
  // from them stay valid until the next step's synchronisation
  std::vector<JumpRec> h_jr;
  std::vector<int2> h_pairs, h_moves;
  h_jr.reserve(ntraj);
  h_pairs.reserve(ntraj);
  h_moves.reserve(ntraj);
  std::vector<double> dens_host((size_t)ntraj * (M + 1));
  int stg_cap = 0;

  // row chunks of the column reductions depend on the sector size only, so a
  // trajectory's sums (and its result) do not depend on how many others share
  // its sector
  auto chunks_for = [&](int64_t n, int) {
    int64_t c = std::min<int64_t>((n + 63) / 64, (int64_t)max_chunks);
    return (int)std::max<int64_t>(1, c);
  };

  // densities of every active trajectory -> d_dens_all; returns the sums
  auto densities = [&](int step) -> int {
    for (auto& S : sec) {
      if (!S.basis || S.m == 0) continue;
      int nch = chunks_for(S.n, S.m);
      int rpc = (int)((S.n + nch - 1) / nch);
      nch = (int)((S.n + rpc - 1) / rpc);
      dim3 g((S.m + TT - 1) / TT, nch);
      dens_partial_kernel<<<g, TT, 0, st>>>(S.n, S.m, S.ld, S.X, S.basis->d_states, M, rpc, d_dpart);
      dens_reduce_kernel<<<g1(S.m, 256), 256, 0, st>>>(S.m, M, nch, d_dpart, S.d_slot2traj, d_dens_all);
    }
    dens_accumulate_kernel<<<M, 256, 0, st>>>(ntraj, M, d_dens_all, d_dsum, d_dsq);
    double hs[2 * MAXM];
    QD_HIP(ctx, hipMemcpyAsync(hs, d_dsum, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    QD_HIP(ctx, hipMemcpyAsync(hs + MAXM, d_dsq, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    QD_HIP(ctx, hipStreamSynchronize(st));
    for (int q = 0; q < M; ++q) {
      dsum[(size_t)step * M + q] = hs[q];
      dsq[(size_t)step * M + q] = hs[MAXM + q];
    }
    return QD_OK;
  };

  // one Arnoldi step for every active slot of a sector (X updated in place)
  auto evolve_sector = [&](Sector& S) -> int {
    const int m = S.m;
    if (m == 0) return QD_OK;
    const int ldv = m;
    const int64_t vstride = S.n * (int64_t)ldv;
    int nch = chunks_for(S.n, m);
    int rpc = (int)((S.n + nch - 1) / nch);
    nch = (int)((S.n + rpc - 1) / rpc);
    dim3 gd((m + TT - 1) / TT, nch);
    gather_gamma_kernel<<<g1(m, 256), 256, 0, st>>>(m, S.d_slot2traj, d_gtraj, d_gslot);
    // V_0 = X / |X|
    col_dots_kernel<<<gd, TT, 0, st>>>(S.n, m, S.ld, S.X, 0, 1, S.X, rpc, d_part);
    col_reduce_kernel<<<g1(m, 256), 256, 0, st>>>(m, 1, nch, d_part, d_h);
    inv_norm_kernel<<<g1(m, 256), 256, 0, st>>>(m, d_h, d_scale);
    col_scale_kernel<<<g1(S.n * m, 256), 256, 0, st>>>(S.n, m, S.ld, S.X, ldv, d_V, d_scale);
    for (int j = 0; j < K; ++j) {
      cplx* Vj = d_V + (int64_t)j * vstride;
      dim3 gs((unsigned)S.n, (m + TT - 1) / TT);
      spmm_tm_kernel<<<gs, TT, 0, st>>>(S.n, m, ldv, S.csr->d_outer, S.csr->d_inner, S.csr->d_vals, S.D, d_gslot,
                                        Vj, d_W);
      for (int pass = 0; pass < 2; ++pass) {
        col_dots_kernel<<<gd, TT, 0, st>>>(S.n, m, ldv, d_V, vstride, j + 1, d_W, rpc, d_part);
        col_reduce_kernel<<<g1((int64_t)m * (j + 1), 256), 256, 0, st>>>(m, j + 1, nch, d_part, d_h);
        hess_acc_kernel<<<g1((int64_t)m * (j + 1), 256), 256, 0, st>>>(m, j, K, d_h, d_hess, pass);
        col_update_kernel<<<g1(S.n * ldv, 256), 256, 0, st>>>(S.n, m, ldv, d_V, vstride, j + 1, d_h, d_W);
      }
      if (j + 1 < K) {
        col_dots_kernel<<<gd, TT, 0, st>>>(S.n, m, ldv, d_W, 0, 1, d_W, rpc, d_part);
        col_reduce_kernel<<<g1(m, 256), 256, 0, st>>>(m, 1, nch, d_part, d_h);
        beta_kernel<<<g1(m, 256), 256, 0, st>>>(m, j, K, d_h, d_hess, d_scale);
        col_scale_kernel<<<g1(S.n * m, 256), 256, 0, st>>>(S.n, m, ldv, d_W, ldv, d_V + (int64_t)(j + 1) * vstride,
                                                          d_scale);
      }
    }
    small_expm_kernel<<<g1(m, 64), 64, 0, st>>>(m, K, K, d_hess, dt, d_coef);
    combine_tm_kernel<<<g1(S.n * m, 256), 256, 0, st>>>(S.n, m, ldv, d_V, vstride, K, d_coef, S.ld, S.X);
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  };

  for (int step = 0; step <= nsteps; ++step) {
    rc = densities(step);
    if (rc) break;
    if (step == nsteps) break;
    decide_kernel<<<g1(ntraj, 256), 256, 0, st>>>(ntraj, M, d_dens_all, d_jmode, d_jrate, njumps, seed, step, dt,
                                                  d_action, d_gtraj);
    hipMemcpyAsync(action.data(), d_action, sizeof(int) * ntraj, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(dens_host.data(), d_dens_all, sizeof(double) * dens_host.size(), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
      rc = qd_set_error(ctx, QD_E_HIP, "trajectory step %d", step);
      break;
    }
    if (h_jump_log)
      for (int t = 0; t < ntraj; ++t) h_jump_log[(size_t)t * nsteps + step] = action[t];
    // jumps from the pre-step states into staging
    struct Pending {
      int traj, a, b, sidx, src_slot;
      int loss;
    };
    std::vector<Pending> pend;
    h_jr.clear();
    h_pairs.clear();
    h_moves.clear();
    for (auto& S : sec) {
      if (!S.basis || S.m == 0) continue;
      const size_t jr0 = h_jr.size();
      for (int slot = 0; slot < S.m; ++slot) {
        int tr = S.slot2traj[slot];
        int ch = action[tr];
        if (ch < 0) continue;
        int mode = jm[ch];
        int loss = jk[ch] == QD_JUMP_LOSS;
        double nd = dens_host[(size_t)tr * (M + 1) + mode] * dens_host[(size_t)tr * (M + 1) + M];
        JumpRec r{slot, mode, loss, (int)pend.size(), 1.0 / std::sqrt(nd)};
        int na = S.a, nb = S.b;
        if (loss) {
          if (mode < nsite) --na;
          else --nb;
        }
        pend.push_back({tr, na, nb, (int)pend.size(), slot, loss});
        h_jr.push_back(r);
      }
      const int njr = (int)(h_jr.size() - jr0);
      if (njr == 0) continue;
      // staging capacity grows on demand
      if (!d_staging || (int)pend.size() > stg_cap) {
        int ncap = std::max<int>((int)pend.size() * 2, 64);
        cplx* ns = nullptr;
        if (hipMalloc(&ns, sizeof(cplx) * (size_t)ncap * nmax) != hipSuccess) {
          rc = qd_set_error(ctx, QD_E_OOM, "jump staging");
          break;
        }
        if (d_staging) {
          hipMemcpyAsync(ns, d_staging, sizeof(cplx) * (size_t)stg_cap * nmax, hipMemcpyDeviceToDevice, st);
          hipStreamSynchronize(st);
          hipFree(d_staging);
        }
        d_staging = ns;
        stg_cap = ncap;
      }
      // target sector descriptors for losses out of S
      SectorDesc tup{}, tdn{};
      if (S.a > 0) {
        if ((rc = ensure_sector(S.a - 1, S.b))) break;
        tup = sec[sid(S.a - 1, S.b)].basis->sec;
      }
      if (S.b > 0) {
        if ((rc = ensure_sector(S.a, S.b - 1))) break;
        tdn = sec[sid(S.a, S.b - 1)].basis->sec;
      }
      const int first = (int)pend.size() - njr;
      zero_kernel<<<g1((int64_t)njr * nmax, 256), 256, 0, st>>>((int64_t)njr * nmax, d_staging + (int64_t)first * nmax);
      hipMemcpyAsync(d_jr + jr0, h_jr.data() + jr0, sizeof(JumpRec) * njr, hipMemcpyHostToDevice, st);
      jump_apply_kernel<<<(unsigned)njr, 256, 0, st>>>(S.n, S.ld, S.X, S.basis->d_states, d_jr + jr0, tup, tdn, nsite,
                                                       nmax, d_staging);
    }
    if (rc) break;
    // evolve every sector (jumpers too; their columns are replaced below)
    for (auto& S : sec) {
      if (!S.basis || S.m == 0) continue;
      if ((rc = evolve_sector(S))) break;
    }
    if (rc) break;
    // commit jumps: dephasing in place; losses move to the target sector
    if (!pend.empty()) {
      n_jumps += (int64_t)pend.size();
      // removals per source sector (losses), in descending slot order
      std::vector<std::vector<int>> removed(sec.size());
      std::vector<std::vector<int2>> commits(sec.size());
      for (auto& S : sec) {
        if (!S.basis || S.m == 0) continue;
        for (int slot = 0; slot < S.m; ++slot) {
          int tr = S.slot2traj[slot];
          if (action[tr] < 0) continue;
          // find the pending record of tr
          for (const Pending& p : pend)
            if (p.traj == tr) {
              if (!p.loss) {
                commits[sid(S.a, S.b)].push_back(make_int2(p.sidx, slot));
              } else {
                removed[sid(S.a, S.b)].push_back(slot);
              }
              break;
            }
        }
      }
      // losses: append to targets
      for (const Pending& p : pend) {
        if (!p.loss) continue;
        ++n_loss;
        Sector& T = sec[sid(p.a, p.b)];
        int slot = T.m++;
        T.slot2traj.resize(std::max<size_t>(T.slot2traj.size(), (size_t)T.m));
        T.slot2traj[slot] = p.traj;
        commits[sid(p.a, p.b)].push_back(make_int2(p.sidx, slot));
      }
      for (size_t i = 0; i < sec.size(); ++i) {
        Sector& S = sec[i];
        if (!S.basis || commits[i].empty()) continue;
        const size_t p0 = h_pairs.size();
        h_pairs.insert(h_pairs.end(), commits[i].begin(), commits[i].end());
        hipMemcpyAsync(d_pairs + p0, h_pairs.data() + p0, sizeof(int2) * commits[i].size(), hipMemcpyHostToDevice, st);
        commit_kernel<<<(unsigned)commits[i].size(), 256, 0, st>>>(S.n, S.ld, S.X, d_pairs + p0, nmax, d_staging);
      }
      // compaction of the source sectors (swap-remove, highest removed slot first)
      for (size_t i = 0; i < sec.size(); ++i) {
        Sector& S = sec[i];
        if (!S.basis || removed[i].empty()) continue;
        std::sort(removed[i].begin(), removed[i].end(), std::greater<int>());
        const size_t m0 = h_moves.size();
        for (int slot : removed[i]) {
          int last = S.m - 1;
          if (slot != last) {
            h_moves.push_back(make_int2(last, slot));
            S.slot2traj[slot] = S.slot2traj[last];
          }
          --S.m;
        }
        const int nmv = (int)(h_moves.size() - m0);
        if (nmv > 0) {
          hipMemcpyAsync(d_moves + m0, h_moves.data() + m0, sizeof(int2) * nmv, hipMemcpyHostToDevice, st);
          move_cols_kernel<<<g1(S.n, 256), 256, 0, st>>>(S.n, S.ld, S.X, d_moves + m0, nmv);
        }
      }
      for (auto& S : sec)
        if (S.basis && S.m > 0)
          hipMemcpyAsync(S.d_slot2traj, S.slot2traj.data(), sizeof(int) * S.m, hipMemcpyHostToDevice, st);
    }
    QD_CHECK_LAUNCH(ctx);
  }
  if (rc == QD_OK) {
    for (int step = 0; step <= nsteps; ++step)
      for (int q = 0; q < M; ++q) {
        double mean = dsum[(size_t)step * M + q] / ntraj;
        h_dens_mean[(size_t)step * M + q] = mean;
        if (h_dens_var) h_dens_var[(size_t)step * M + q] = dsq[(size_t)step * M + q] / ntraj - mean * mean;
      }
    if (stats) {
      stats->jumps = n_jumps;
      stats->losses = n_loss;
      int used = 0;
      for (auto& S : sec) used += S.basis ? 1 : 0;
      stats->sectors = used;
    }
  }
  if (d_staging) hipFree(d_staging);
  cleanup();
  return rc;
}
