// traj.hip -- batched quantum-jump trajectories (programs/hubbard_diss.cpp:440-482,
// mc_manager.hpp:88-128), generalised to dephasing and loss channels with
// sector migration.
//
// Trajectories live in per-sector blocks, trajectory-minor: X[s * ld + t] is
// state s of the trajectory in slot t (slots 0..m-1 packed).  All slots of one
// sector share its Hamiltonian, so the multi-RHS product reads each row of H
// once for every trajectory of the sector, and a wave's 64 lanes read 64
// consecutive trajectories (1 KB) per stored entry.
//
// Every phase of a time step runs as ONE grouped launch over all occupied
// sectors: a small table (GSec) gives each sector's rows, slots, pool offsets
// and tile ranges, and each workgroup finds its sector by binary search.  The
// sectors become uneven as trajectories lose particles (4,900 rows down to 1),
// so grouping keeps the small ones from running latency-bound and one at a
// time.  The reduction tile (rows per chunk) of each sector follows an
// occupancy target that a short timed search picks at the start of a run and
// again once the trajectories have spread over smaller sectors.
//
// Each trajectory runs its own Arnoldi step (the reference's evolve with
// H_eff, hubbard_diss.cpp:467-469; sparse_algorithm.hpp:548-562) with delayed
// CGS2 (two passes over the basis per Krylov vector), a per-trajectory
// Hessenberg and a per-trajectory exponential on the device.  Jumps are drawn on the device from
// a counter-based stream per trajectory, applied from the pre-step state into
// a staging area, and committed after the evolution (moving lost trajectories
// to their new sector).  All reductions have a fixed order for a given tile
// choice; a different choice moves sums by rounding only.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mrhs.hpp"

namespace {


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

// per global slot: the diagonal shift of H_eff,t = H - i/2 (D - Gamma_t) as
// alpha D + beta with alpha = -i/2, beta = i Gamma_t / 2, and 1/|psi_t| from the
// recorded densities and norm^2
__global__ void slot_params_kernel(int Mtot, int M, const int* __restrict__ slot2traj, const double* __restrict__ gtraj,
                                   const double* __restrict__ dens_all, cplx* __restrict__ alpha,
                                   cplx* __restrict__ beta, double* __restrict__ scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Mtot) return;
  const int tr = slot2traj[t];
  alpha[t] = cmake(0.0, -0.5);
  beta[t] = cmake(0.0, 0.5 * gtraj[tr]);
  const double n2 = dens_all[(int64_t)tr * (M + 1) + M];
  scale[t] = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
}

// dens_sum[q] = sum over trajectories (index order) of dens_all; dens_sq likewise
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
                              int* __restrict__ action, double* __restrict__ gamma_traj,
                              double* __restrict__ jscale) {
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
  // 1 / |L psi| of a jumper: |L psi|^2 = <n_m> |psi|^2 from the recorded densities
  if (ch >= 0) jscale[tr] = 1.0 / sqrt(d[jmode[ch]] * d[M]);
}

// every sector's block and basis, indexed by sector id (the per-step jump
// tables name sectors by this index, so each jump phase is one launch over
// all sectors)
struct SecTab {
  int64_t n;
  int ld;
  cplx* X;
  const uint64_t* states;
  SectorDesc sec;
};

struct JumpRec {
  int slot;      // source slot
  int mode;      // mode m
  int loss;      // 0 dephasing, 1 loss
  int staging;   // staging vector index
  double scale;  // 1 / |L psi| (from the recorded densities and norm)
  int src;       // source sector
  int tgt;       // target sector of a loss (the source for dephasing)
};

// staged[j][target] = L psi_slot (normalised); grid (jumps of every sector, row chunks)
__global__ void jump_apply_kernel(const SecTab* __restrict__ tab, const JumpRec* __restrict__ jr, int64_t stg_ld,
                                  cplx* __restrict__ staging) {
  const JumpRec J = jr[blockIdx.x];
  const SecTab& S = tab[J.src];
  const int64_t n = S.n;
  const int ld = S.ld;
  const cplx* __restrict__ X = S.X;
  const uint64_t* __restrict__ states = S.states;
  cplx* out = staging + (int64_t)J.staging * stg_ld;
  for (int64_t s = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.y * blockDim.x) {
    const uint64_t x = states[s];
    if (!((x >> J.mode) & 1ull)) continue;
    cplx v = X[s * ld + J.slot];
    v = cmake(v.x * J.scale, v.y * J.scale);
    if (!J.loss) {
      out[s] = v;
    } else {
      const uint64_t y = x & ~(1ull << J.mode);
      const double sg = (__popcll(x >> (J.mode + 1)) & 1) ? -1.0 : 1.0;
      const uint64_t r = qd_rank_cns(tab[J.tgt].sec, y);
      out[r] = cmake(sg * v.x, sg * v.y);
    }
  }
}

// X[s, slot] = staging[j][s] for each (staging index j, slot, sector) of the
// step; grid (commits of every sector, row chunks)
__global__ void commit_kernel(const SecTab* __restrict__ tab, const int4* __restrict__ pairs, int64_t stg_ld,
                              const cplx* __restrict__ staging) {
  const int4 p = pairs[blockIdx.x];
  const SecTab& S = tab[p.z];
  for (int64_t s = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; s < S.n; s += (int64_t)gridDim.y * blockDim.x)
    S.X[s * S.ld + p.y] = staging[(int64_t)p.x * stg_ld + s];
}

// column moves of the swap-remove compaction, in list order, X[s, dst] =
// X[s, src]: list j = (sector, first move, move count); grid (row blocks, lists)
__global__ void move_cols_kernel(const SecTab* __restrict__ tab, const int4* __restrict__ lists,
                                 const int2* __restrict__ moves) {
  const int4 L = lists[blockIdx.y];
  const SecTab& S = tab[L.x];
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.n) return;
  cplx* row = S.X + s * S.ld;
  for (int k = 0; k < L.z; ++k) row[moves[L.y + k].y] = row[moves[L.y + k].x];
}

// row chunks of 256 for the per-jump kernels: a jump's column spread over up
// to 64 workgroups instead of one walking all rows
inline unsigned col_chunks(int64_t n) { return (unsigned)std::min<int64_t>(64, std::max<int64_t>(1, (n + 255) / 256)); }

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

// ---------------------------------------------------------------- host state
struct Sector {
  int a = -1, b = -1;
  int64_t n = 0;
  qd_basis* basis = nullptr;
  qd_csr* csr = nullptr;
  double* D = nullptr;
  cplx* X = nullptr;  // n x ld, trajectory-minor
  int ld = 0;
  int m = 0;  // active slots
  std::vector<int> slot2traj;
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
  // Lanczos form (mrhs.hpp) when sum_c L_c^+ L_c is constant on every sector:
  // D = sum_m r_m n_m with r_m the total rate of mode m, constant on a sector
  // (fixed up and down counts) when r_m is one value over the up modes and one
  // over the down modes.  QD_TRAJ_KRYLOV=arnoldi keeps the Arnoldi form.
  bool use_lz = true;
  {
    std::vector<double> rm(M, 0.0);
    for (int c = 0; c < njumps; ++c) rm[jm[c]] += jrt[c];
    for (int half = 0; half < 2; ++half) {
      const auto mm = std::minmax_element(rm.begin() + half * nsite, rm.begin() + (half + 1) * nsite);
      if (*mm.second - *mm.first > 1e-13 * std::fabs(*mm.second)) use_lz = false;
    }
    const char* e = std::getenv("QD_TRAJ_KRYLOV");
    if (e && std::string(e) == "arnoldi") use_lz = false;
  }
  std::vector<Sector> sec((nsite + 1) * (nsite + 1));
  auto sid = [&](int a, int b) { return a * (nsite + 1) + b; };
  // largest sector reachable by losses from (nup0, ndn0): C(n, a) is not monotone in a
  int64_t nmax = 1;
  for (int a = 0; a <= nup0; ++a)
    for (int b = 0; b <= ndn0; ++b)
      nmax = std::max<int64_t>(nmax, (int64_t)qd_binom_host(nsite, a) * (int64_t)qd_binom_host(nsite, b));
  const int NSMAX = (int)sec.size();
  // chunk counts are at most min(QD_TRAJ_MAX_CHUNKS, ceil(n / 16)) per sector
  const int max_nch = (int)std::min<int64_t>(QD_TRAJ_MAX_CHUNKS, (nmax + 15) / 16);
  const int PMAX = std::max(2 * K + 2, M + 1);
  // a step's commit table: at most one commit per trajectory, one move list
  // per sector, one move per trajectory
  const size_t pm_bytes = sizeof(int4) * (size_t)ntraj + sizeof(int4) * NSMAX + sizeof(int2) * (size_t)ntraj;

  // device buffers
  int *d_jmode = nullptr, *d_action = nullptr, *d_slot2traj = nullptr;
  double *d_jrate = nullptr, *d_dens_all = nullptr, *d_gtraj = nullptr, *d_dsum = nullptr, *d_dsq = nullptr;
  double *d_scale = nullptr, *d_jscale = nullptr;
  cplx *d_alpha = nullptr, *d_beta = nullptr, *d_red = nullptr, *d_coef2 = nullptr, *d_sc = nullptr;
  cplx *d_V = nullptr, *d_W = nullptr, *d_W2 = nullptr, *d_part = nullptr, *d_h = nullptr, *d_hess = nullptr;
  cplx *d_coef = nullptr, *d_staging = nullptr;
  LzState lz{};
  cplx* d_apart = nullptr;  // the product's per-item alpha partials (Lanczos form)
  GSec* d_gsec = nullptr;
  JumpRec* d_jr = nullptr;
  char* d_pm = nullptr;  // a step's commits, move lists and column moves
  SecTab* d_sectab = nullptr;
  std::vector<double> dsum((size_t)(nsteps + 1) * M, 0.0), dsq((size_t)(nsteps + 1) * M, 0.0);
  int64_t n_jumps = 0, n_loss = 0;
  std::vector<void*> allocs;
  std::vector<SecTab> h_sectab(NSMAX);  // uploaded before a jump phase after sectors were added
  bool tab_dirty = false;
  auto dalloc = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) return false;
    allocs.push_back(*p);
    return true;
  };
  // Pinned host regions for the per-step tables and readbacks.  Copies from
  // pageable memory are staged by the runtime and hold the host until the
  // stream reaches them, which left the GPU idle around every small launch of
  // the jump phase.  An upload region is rewritten only after its previous
  // upload has completed (its event).
  struct PinRegion {
    char* p = nullptr;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  char* pin_base = nullptr;
  hipEvent_t ev_rb = nullptr, ev_flag = nullptr;  // a step's readback of the jump decisions; of the flag
  PinRegion pin_gsec, pin_s2t, pin_jr, pin_pm, pin_tab;
  int* h_action = nullptr;
  double* h_jscale = nullptr;
  int* h_flag = nullptr;
  auto upload = [&](PinRegion& r, void* dst, const void* src, size_t bytes) {
    if (r.pending) hipEventSynchronize(r.ev);
    if (bytes) std::memcpy(r.p, src, bytes);
    hipMemcpyAsync(dst, r.p, bytes, hipMemcpyHostToDevice, st);
    hipEventRecord(r.ev, st);
    r.pending = true;
  };
  auto cleanup = [&]() {
    hipStreamSynchronize(st);
    for (PinRegion* r : {&pin_gsec, &pin_s2t, &pin_jr, &pin_pm, &pin_tab})
      if (r->ev) hipEventDestroy(r->ev);
    if (ev_rb) hipEventDestroy(ev_rb);
    if (ev_flag) hipEventDestroy(ev_flag);
    if (pin_base) hipHostFree(pin_base);
    for (auto& S : sec) {
      if (S.csr) qd_csr_destroy(S.csr);
      if (S.basis) qd_basis_destroy(S.basis);
    }
    if (d_staging) hipFree(d_staging);
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
    if (!dalloc((void**)&S.D, sizeof(double) * S.n) || !dalloc((void**)&S.X, sizeof(cplx) * S.n * S.ld))
      return qd_set_error(ctx, QD_E_OOM, "sector (%d,%d) buffers", a, b);
    diag_rates_kernel<<<g1(S.n, 256), 256, 0, st>>>(S.n, S.basis->d_states, d_jmode, d_jrate, njumps, S.D);
    h_sectab[sid(a, b)] = SecTab{S.n, S.ld, S.X, S.basis->d_states, S.basis->sec};
    tab_dirty = true;
    return QD_OK;
  };

  if (!dalloc((void**)&d_jmode, sizeof(int) * std::max(1, njumps)) ||
      !dalloc((void**)&d_jrate, sizeof(double) * std::max(1, njumps)) ||
      !dalloc((void**)&d_action, sizeof(int) * ntraj) || !dalloc((void**)&d_slot2traj, sizeof(int) * ntraj) ||
      !dalloc((void**)&d_dens_all, sizeof(double) * ntraj * (M + 1)) ||
      !dalloc((void**)&d_gtraj, sizeof(double) * ntraj) || !dalloc((void**)&d_jscale, sizeof(double) * ntraj) ||
      !dalloc((void**)&d_dsum, sizeof(double) * (size_t)(nsteps + 1) * M) ||
      !dalloc((void**)&d_dsq, sizeof(double) * (size_t)(nsteps + 1) * M) ||
      !dalloc((void**)&d_scale, sizeof(double) * (ntraj + 1)) ||
      !dalloc((void**)&d_alpha, sizeof(cplx) * ntraj) || !dalloc((void**)&d_beta, sizeof(cplx) * ntraj) ||
      !dalloc((void**)&d_red, sizeof(cplx) * (size_t)(2 * K + 2) * ntraj) ||
      !dalloc((void**)&d_coef2, sizeof(cplx) * (size_t)(2 * K) * ntraj) ||
      !dalloc((void**)&d_sc, sizeof(cplx) * 2 * (size_t)ntraj) ||
      !dalloc((void**)&d_V, sizeof(cplx) * (size_t)K * nmax * ntraj) ||
      !dalloc((void**)&d_W, sizeof(cplx) * (size_t)nmax * ntraj) ||
      !dalloc((void**)&d_W2, sizeof(cplx) * (size_t)nmax * ntraj) ||
      !dalloc((void**)&d_part, sizeof(cplx) * (size_t)max_nch * PMAX * ntraj) ||
      !dalloc((void**)&d_h, sizeof(cplx) * (size_t)K * ntraj) ||
      !dalloc((void**)&d_hess, sizeof(cplx) * (size_t)K * (K + 1) * ntraj) ||
      !dalloc((void**)&d_coef, sizeof(cplx) * (size_t)K * ntraj) ||
      !dalloc((void**)&d_gsec, sizeof(GSec) * NSMAX) || !dalloc((void**)&d_jr, sizeof(JumpRec) * ntraj) ||
      !dalloc((void**)&d_pm, pm_bytes) || !dalloc((void**)&d_sectab, sizeof(SecTab) * NSMAX) ||
      !dalloc((void**)&lz.beta, sizeof(double) * (size_t)K * ntraj) ||
      !dalloc((void**)&lz.inv, sizeof(double) * (size_t)K * ntraj) ||
      !dalloc((void**)&lz.ip, sizeof(cplx) * (size_t)K * ntraj) || !dalloc((void**)&lz.ca, sizeof(cplx) * ntraj) ||
      !dalloc((void**)&lz.cb, sizeof(cplx) * ntraj) ||
      !dalloc((void**)&lz.om, sizeof(double) * 3 * (size_t)(K + 1) * ntraj) ||
      !dalloc((void**)&lz.tn, sizeof(double) * ntraj) || !dalloc((void**)&lz.flag, sizeof(cplx)) ||
      // product items over all sectors: sum of ceil(n / rows) ceil(m / 64) <= (nmax / 4 + 1) (ntraj / 64 + sectors)
      (use_lz && !dalloc((void**)&d_apart, sizeof(cplx) * TT * (size_t)(nmax / SPMM_RB + 1) *
                                               (size_t)((ntraj + TT - 1) / TT + NSMAX)))) {
    cleanup();
    return qd_set_error(ctx, QD_E_OOM, "trajectory workspace (%d trajectories, dim %lld, k %d)", ntraj,
                        (long long)nmax, k_rt);
  }
  {
    auto up256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_gsec = up256(sizeof(GSec) * NSMAX), b_s2t = up256(sizeof(int) * ntraj),
                 b_jr = up256(sizeof(JumpRec) * ntraj), b_pm = up256(pm_bytes), b_tab = up256(sizeof(SecTab) * NSMAX),
                 b_act = up256(sizeof(int) * ntraj), b_js = up256(sizeof(double) * ntraj), b_flag = 256;
    if (hipHostMalloc((void**)&pin_base, b_gsec + b_s2t + b_jr + b_pm + b_tab + b_act + b_js + b_flag) != hipSuccess) {
      pin_base = nullptr;
      cleanup();
      return qd_set_error(ctx, QD_E_OOM, "trajectory host tables (%d trajectories)", ntraj);
    }
    char* q = pin_base;
    pin_gsec.p = q;
    q += b_gsec;
    pin_s2t.p = q;
    q += b_s2t;
    pin_jr.p = q;
    q += b_jr;
    pin_pm.p = q;
    q += b_pm;
    pin_tab.p = q;
    q += b_tab;
    h_action = reinterpret_cast<int*>(q);
    q += b_act;
    h_jscale = reinterpret_cast<double*>(q);
    q += b_js;
    h_flag = reinterpret_cast<int*>(q);
    for (hipEvent_t* e : {&pin_gsec.ev, &pin_s2t.ev, &pin_jr.ev, &pin_pm.ev, &pin_tab.ev, &ev_rb, &ev_flag})
      if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
        *e = nullptr;
        cleanup();
        return qd_set_error(ctx, QD_E_HIP, "trajectory events");
      }
  }
  // per-slot operands of the product, defined before the first tuning run reads them
  zero_kernel<<<g1((ntraj + 1) / 2, 256), 256, 0, st>>>((ntraj + 1) / 2, (cplx*)d_scale);  // ntraj + 1 doubles
  zero_kernel<<<g1(ntraj, 256), 256, 0, st>>>(ntraj, d_alpha);
  zero_kernel<<<g1(ntraj, 256), 256, 0, st>>>(ntraj, d_beta);
  // the Lanczos form's trigger for recomputing a step fully reorthogonalised:
  // a quarter of sqrt(eps / K), as the single-vector partial scheme
  // (QD_TRAJ_OMEGA overrides it)
  lz.trigger = 0.25 * std::sqrt(2.220446049250313e-16 / K);
  if (const char* e = std::getenv("QD_TRAJ_OMEGA")) lz.trigger = std::atof(e);
  int64_t lz_fallback = 0;
  if (use_lz) {
    // the Lanczos form writes only the three diagonals of each Hessenberg
    zero_kernel<<<g1((int64_t)K * (K + 1) * ntraj, 256), 256, 0, st>>>((int64_t)K * (K + 1) * ntraj, d_hess);
    zero_kernel<<<g1(ntraj, 256), 256, 0, st>>>(ntraj, lz.ca);
    zero_kernel<<<g1(ntraj, 256), 256, 0, st>>>(ntraj, (cplx*)lz.cb);  // allocated as ntraj cplx
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
    broadcast_kernel<<<g1(S.n * ntraj, 256), 256, 0, st>>>(S.n, ntraj, S.ld, (const cplx*)d_psi0, S.X);
  }
  // per-step host tables (uploaded through the pinned regions)
  std::vector<JumpRec> h_jr;
  std::vector<int> pend_of(ntraj, -1);  // a jumping trajectory's index in the step's pending list
  std::vector<char> h_pm;
  std::vector<GSec> h_gsec;
  std::vector<int> h_slot2traj;
  h_jr.reserve(ntraj);
  h_pm.reserve(pm_bytes);
  h_gsec.reserve(NSMAX);
  h_slot2traj.reserve(ntraj);
  int stg_cap = 0;
  int ns = 0;
  int64_t VS = 0, tiles = 0, blocks = 0, ngroups = 0;
  int occupancy_target = 2048;  // waves per sector for the reduction tiles (tuned below)
  int spmm_var = 0;              // product kernel (SpmmVarProduct; tuned below)
  const char* var_env = std::getenv("QD_TRAJ_SPMM");
  const int var_forced = var_env ? std::atoi(var_env) : -1;
  if (var_env && (var_forced < 0 || var_forced >= SPMM_VARIANTS)) {
    cleanup();
    return qd_set_error(ctx, QD_E_ARG, "QD_TRAJ_SPMM must be 0..%d", SPMM_VARIANTS - 1);
  }
  if (var_forced >= 0) spmm_var = var_forced;

  // grouped-launch tables from the current sector membership
  auto build_groups = [&]() {
    h_gsec.clear();
    h_slot2traj.clear();
    int64_t vo = 0, co = 0;
    tiles = blocks = ngroups = 0;
    for (auto& S : sec) {
      if (!S.basis || S.m == 0) continue;
      GSec g{};
      g.n = S.n;
      g.m = S.m;
      g.ldx = S.ld;
      g.groups = (S.m + TT - 1) / TT;
      g.mp = TT;  // one slot per lane (sector results do not depend on the mix of sizes)
      g.rpc = rows_per_chunk(S.n, g.groups, occupancy_target);
      g.nch = (int)((S.n + g.rpc - 1) / g.rpc);
      g.so = (int)h_slot2traj.size();
      g.vo = vo;
      g.co = co;
      g.tile0 = tiles;
      g.blk0 = blocks;
      g.grp0 = ngroups;
      g.X = S.X;
      g.outer = S.csr->d_outer;
      g.inner = S.csr->d_inner;
      g.vals = S.csr->d_vals;
      g.D = S.D;
      g.states = S.basis->d_states;
      h_gsec.push_back(g);
      h_slot2traj.insert(h_slot2traj.end(), S.slot2traj.begin(), S.slot2traj.begin() + S.m);
      vo += S.n * S.m;
      co += (int64_t)g.nch * S.m;
      tiles += (int64_t)g.nch * g.groups;
      blocks += (S.n + spmm_rows(spmm_var) - 1) / spmm_rows(spmm_var) * g.groups;
      ngroups += g.groups;
    }
    ns = (int)h_gsec.size();
    VS = vo;
    upload(pin_gsec, d_gsec, h_gsec.data(), sizeof(GSec) * ns);
    upload(pin_s2t, d_slot2traj, h_slot2traj.data(), sizeof(int) * h_slot2traj.size());
  };

  // densities of every trajectory -> d_dens_all; sums over trajectories ->
  // row `step` of d_dsum, d_dsq (read back once, after the last step)
  auto densities = [&](int step) -> int {
    g_dens_partial_kernel<<<(unsigned)tiles, TT, 0, st>>>(d_gsec, ns, M, (double*)d_part);
    g_dens_reduce_kernel<<<(unsigned)(ngroups * (M + 1)), RW * TT, 0, st>>>(d_gsec, ns, ngroups, M, (double*)d_part,
                                                                            d_slot2traj, d_dens_all);
    dens_accumulate_kernel<<<M, 256, 0, st>>>(ntraj, M, d_dens_all, d_dsum + (size_t)step * M,
                                              d_dsq + (size_t)step * M);
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  };

  // one Arnoldi step (delayed CGS2, K vectors, or the Lanczos form) for every
  // slot of every occupied sector, X updated in place.  evolve_begin enqueues
  // the expansion, which only reads X; evolve_end enqueues the combination
  // that overwrites X.  The jump phase (which reads the pre-step X) is
  // enqueued in between, so the host's jump bookkeeping overlaps the
  // expansion on the GPU.  In the Lanczos form the combination is enqueued
  // before the host reads the orthogonality flag and writes nothing when the
  // flag is set; the host then enqueues the fully reorthogonalised step and
  // its combination.
  const MrhsBufs mb{d_V, d_W2, d_W, d_part, d_h, d_hess, d_red, d_coef2, d_sc, d_coef};
  auto evolve_begin = [&]() -> int {
    const int Mtot = ntraj;
    slot_params_kernel<<<g1(Mtot, 256), 256, 0, st>>>(Mtot, M, d_slot2traj, d_gtraj, d_dens_all, d_alpha, d_beta,
                                                       d_scale);
    const SpmmVarProduct prod{d_gsec, ns, blocks, spmm_var, use_lz ? d_apart : nullptr};
    if (use_lz) {
      zero_kernel<<<1, 1, 0, st>>>(1, reinterpret_cast<cplx*>(lz.flag));
      mrhs_lanczos_with(st, d_gsec, ns, tiles, ngroups, VS, Mtot, K, dt, d_scale, d_alpha, d_beta, mb, lz, prod,
                        d_apart, spmm_rows(spmm_var));
      QD_HIP(ctx, hipMemcpyAsync(h_flag, lz.flag, sizeof(int), hipMemcpyDeviceToHost, st));
      QD_HIP(ctx, hipEventRecord(ev_flag, st));
    } else {
      mrhs_arnoldi_with(st, d_gsec, ns, tiles, ngroups, VS, Mtot, K, dt, d_scale, d_alpha, d_beta, mb, prod);
    }
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  };
  auto evolve_end = [&]() -> int {
    const int Mtot = ntraj;
    if (!use_lz) {
      mrhs_combine(st, d_gsec, ns, VS, Mtot, K, mb);
      QD_CHECK_LAUNCH(ctx);
      return QD_OK;
    }
    mrhs_combine(st, d_gsec, ns, VS, Mtot, K, mb, lz.flag);
    QD_HIP(ctx, hipEventSynchronize(ev_flag));
    if (*h_flag) {
      // a slot's basis is no longer semi-orthogonal: the combination above
      // wrote nothing; this step again, fully reorthogonalised, from the same X
      ++lz_fallback;
      const SpmmVarProduct prod_plain{d_gsec, ns, blocks, spmm_var, nullptr};
      mrhs_arnoldi_with(st, d_gsec, ns, tiles, ngroups, VS, Mtot, K, dt, d_scale, d_alpha, d_beta, mb, prod_plain);
      mrhs_combine(st, d_gsec, ns, VS, Mtot, K, mb);
      // the Arnoldi step filled the Hessenbergs above the diagonals
      zero_kernel<<<g1((int64_t)K * (K + 1) * ntraj, 256), 256, 0, st>>>((int64_t)K * (K + 1) * ntraj, d_hess);
    }
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  };

  // Tile autotuner for the uneven sectors: the occupancy target of the
  // reduction tiles (rows per chunk of every sector) is timed on the widest
  // reduction pass over the current sector membership (a read-only pass,
  // results discarded; in the Lanczos form the update pass, whose output
  // slot is rewritten before it is read).  It runs at the start and again once the loss jumps
  // have spread the trajectories over smaller sectors.
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEventCreate(&ev0);
  hipEventCreate(&ev1);
  const bool tune_log = std::getenv("QD_TUNE_LOG") != nullptr;
  auto tune_tiles = [&]() {
    const int cand[7] = {512, 1024, 2048, 4096, 8192, 16384, 32768};
    float best = 1e30f;
    int best_c = occupancy_target;
    for (int c : cand) {
      occupancy_target = c;
      build_groups();
      const int nq = std::max(1, K - 1);
      auto pass = [&]() {
        if (!use_lz)
          launch_red<2>(nq, (unsigned)tiles, st, d_gsec, ns, d_V, VS, d_W, d_W2, d_part);
        else if (K >= 2)
          g_lz_update_kernel<<<(unsigned)tiles, TT, 0, st>>>(d_gsec, ns, d_V, VS, K >= 3 ? 1 : 0, d_W, lz.ca, lz.cb,
                                                             d_part);
        else
          g_pass_kernel<1, 0><<<(unsigned)tiles, TT, 0, st>>>(d_gsec, ns, d_V, VS, d_W, nullptr, ntraj, d_part);
      };
      pass();
      hipEventRecord(ev0, st);
      for (int r = 0; r < 2; ++r) pass();
      hipEventRecord(ev1, st);
      hipEventSynchronize(ev1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev0, ev1);
      if (tune_log)
        std::fprintf(stderr, "[qd] trajectory tiles: target %d -> %lld tiles over %d sectors, reduction pass %.1f us\n", c,
                     (long long)tiles, ns, 1000.0 * ms / 2);
      if (ms < best) {
        best = ms;
        best_c = c;
      }
    }
    occupancy_target = best_c;
  };
  tune_tiles();
  // Product kernel, timed the same way on the product from X into the pools
  // (X is only read; the pools are rewritten by the next step).
  auto tune_spmm = [&]() {
    if (var_forced >= 0) return;
    float best = 1e30f;
    int best_var = spmm_var;
    for (int var = 0; var < SPMM_VARIANTS; ++var) {
      spmm_var = var;
      build_groups();
      const SpmmVarProduct prod{d_gsec, ns, blocks, var, use_lz ? d_apart : nullptr};
      prod(st, 1, nullptr, d_scale, d_alpha, d_beta, d_V, d_W);
      hipEventRecord(ev0, st);
      for (int r = 0; r < 2; ++r) prod(st, 1, nullptr, d_scale, d_alpha, d_beta, d_V, d_W);
      hipEventRecord(ev1, st);
      hipEventSynchronize(ev1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev0, ev1);
      if (tune_log)
        std::fprintf(stderr, "[qd] trajectory product: kernel %d (%d-row items), %lld items over %d sectors, %.1f us\n",
                     var, spmm_rows(var), (long long)blocks, ns, 1000.0 * ms / 2);
      if (ms < best) {
        best = ms;
        best_var = var;
      }
    }
    spmm_var = best_var;
    build_groups();
  };
  tune_spmm();

  for (int step = 0; step <= nsteps; ++step) {
    if (step > 0 && step == nsteps / 4) {
      tune_tiles();
      tune_spmm();
    }
    build_groups();
    rc = densities(step);
    if (rc) break;
    if (step == nsteps) break;
    decide_kernel<<<g1(ntraj, 256), 256, 0, st>>>(ntraj, M, d_dens_all, d_jmode, d_jrate, njumps, seed, step, dt,
                                                  d_action, d_gtraj, d_jscale);
    hipMemcpyAsync(h_action, d_action, sizeof(int) * ntraj, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(h_jscale, d_jscale, sizeof(double) * ntraj, hipMemcpyDeviceToHost, st);
    hipEventRecord(ev_rb, st);
    // the expansion of every occupied sector (jumpers too; their columns are
    // replaced below) goes to the GPU before the host waits for the jump
    // decisions
    if ((rc = evolve_begin())) break;
    if (hipEventSynchronize(ev_rb) != hipSuccess) {
      rc = qd_set_error(ctx, QD_E_HIP, "trajectory step %d", step);
      break;
    }
    const int* action = h_action;
    if (h_jump_log)
      for (int t = 0; t < ntraj; ++t) h_jump_log[(size_t)t * nsteps + step] = action[t];
    // jumps from the pre-step states into staging: the records of every
    // sector (target sectors built on first use), then one upload, one
    // clearing of the staging vectors and one launch
    struct Pending {
      int traj, a, b, sidx, src_slot;
      int loss;
    };
    std::vector<Pending> pend;
    h_jr.clear();
    for (auto& S : sec) {
      if (!S.basis || S.m == 0) continue;
      const int src = sid(S.a, S.b);
      const size_t jr0 = h_jr.size();
      for (int slot = 0; slot < S.m; ++slot) {
        int tr = S.slot2traj[slot];
        int ch = action[tr];
        if (ch < 0) continue;
        int mode = jm[ch];
        int loss = jk[ch] == QD_JUMP_LOSS;
        int na = S.a, nb = S.b;
        if (loss) {
          if (mode < nsite)
            --na;
          else
            --nb;
        }
        pend_of[tr] = (int)pend.size();
        h_jr.push_back(JumpRec{slot, mode, loss, (int)pend.size(), h_jscale[tr], src, loss ? sid(na, nb) : src});
        pend.push_back({tr, na, nb, (int)pend.size(), slot, loss});
      }
      // the sectors a loss out of S reaches, built on the first jump out of S
      if (h_jr.size() > jr0) {
        if (S.a > 0 && (rc = ensure_sector(S.a - 1, S.b))) break;
        if (S.b > 0 && (rc = ensure_sector(S.a, S.b - 1))) break;
      }
    }
    if (rc) break;
    if (!pend.empty()) {
      // staging capacity grows on demand (nothing in it outlives a step)
      if ((int)pend.size() > stg_cap) {
        if (d_staging) hipFree(d_staging);
        d_staging = nullptr;
        stg_cap = 0;
        const int ncap = std::min(ntraj, std::max<int>((int)pend.size() * 2, 64));
        if (hipMalloc(&d_staging, sizeof(cplx) * (size_t)ncap * nmax) != hipSuccess) {
          d_staging = nullptr;
          rc = qd_set_error(ctx, QD_E_OOM, "jump staging");
          break;
        }
        stg_cap = ncap;
      }
      if (tab_dirty) {
        upload(pin_tab, d_sectab, h_sectab.data(), sizeof(SecTab) * h_sectab.size());
        tab_dirty = false;
      }
      zero_kernel<<<g1((int64_t)pend.size() * nmax, 256), 256, 0, st>>>((int64_t)pend.size() * nmax, d_staging);
      upload(pin_jr, d_jr, h_jr.data(), sizeof(JumpRec) * h_jr.size());
      jump_apply_kernel<<<dim3((unsigned)pend.size(), col_chunks(nmax)), 256, 0, st>>>(d_sectab, d_jr, nmax,
                                                                                         d_staging);
    }
    if ((rc = evolve_end())) break;
    // commit jumps: dephasing in place; losses move to the target sector.
    // The host tables first, then one upload (commits, move lists, moves) and
    // two launches: all commits before all moves
    if (!pend.empty()) {
      n_jumps += (int64_t)pend.size();
      std::vector<std::vector<int>> removed(sec.size());
      std::vector<int4> h_commit;
      for (auto& S : sec) {
        if (!S.basis || S.m == 0) continue;
        const int si = sid(S.a, S.b);
        for (int slot = 0; slot < S.m; ++slot) {
          const int tr = S.slot2traj[slot];
          if (action[tr] < 0) continue;
          const Pending& p = pend[pend_of[tr]];
          if (!p.loss)
            h_commit.push_back(make_int4(p.sidx, slot, si, 0));
          else
            removed[si].push_back(slot);
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
        h_commit.push_back(make_int4(p.sidx, slot, sid(p.a, p.b), 0));
      }
      // compaction of the source sectors (swap-remove, highest removed slot first)
      std::vector<int4> h_lists;
      std::vector<int2> h_mv;
      for (size_t i = 0; i < sec.size(); ++i) {
        Sector& S = sec[i];
        if (!S.basis || removed[i].empty()) continue;
        std::sort(removed[i].begin(), removed[i].end(), std::greater<int>());
        const int m0 = (int)h_mv.size();
        for (int slot : removed[i]) {
          int last = S.m - 1;
          if (slot != last) {
            h_mv.push_back(make_int2(last, slot));
            S.slot2traj[slot] = S.slot2traj[last];
          }
          --S.m;
        }
        const int nmv = (int)h_mv.size() - m0;
        if (nmv > 0) h_lists.push_back(make_int4((int)i, m0, nmv, 0));
      }
      for (const Pending& p : pend) pend_of[p.traj] = -1;
      // one buffer: commits (int4), move lists (int4), moves (int2)
      const size_t b_c = sizeof(int4) * h_commit.size(), b_l = sizeof(int4) * h_lists.size(),
                   b_m = sizeof(int2) * h_mv.size();
      h_pm.resize(b_c + b_l + b_m);
      if (b_c) std::memcpy(h_pm.data(), h_commit.data(), b_c);
      if (b_l) std::memcpy(h_pm.data() + b_c, h_lists.data(), b_l);
      if (b_m) std::memcpy(h_pm.data() + b_c + b_l, h_mv.data(), b_m);
      upload(pin_pm, d_pm, h_pm.data(), h_pm.size());
      if (!h_commit.empty())
        commit_kernel<<<dim3((unsigned)h_commit.size(), col_chunks(nmax)), 256, 0, st>>>(
            d_sectab, reinterpret_cast<const int4*>(d_pm), nmax, d_staging);
      if (!h_lists.empty())
        move_cols_kernel<<<dim3((unsigned)((nmax + 255) / 256), (unsigned)h_lists.size()), 256, 0, st>>>(
            d_sectab, reinterpret_cast<const int4*>(d_pm + b_c), reinterpret_cast<const int2*>(d_pm + b_c + b_l));
    }
    QD_CHECK_LAUNCH(ctx);
  }
  if (rc == QD_OK) {
    // the per-step sums over trajectories, kept on the device during the run
    hipError_t e = hipMemcpyAsync(dsum.data(), d_dsum, sizeof(double) * dsum.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dsq.data(), d_dsq, sizeof(double) * dsq.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = qd_set_error(ctx, QD_E_HIP, "trajectory density sums: %s", hipGetErrorString(e));
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
      stats->lanczos = use_lz ? 1 : 0;
      stats->lanczos_fallback_steps = (int32_t)lz_fallback;
    }
  }
  hipEventDestroy(ev0);
  hipEventDestroy(ev1);
  cleanup();
  return rc;
}
