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
#include <vector>

#include "qd_internal.hpp"

namespace {

constexpr int TT = 64;    // trajectories per wave
constexpr int MAXM = 32;  // modes (2 * nsite) supported
constexpr int MAXK = 32;  // Krylov dimension limit for the batched step

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

// One occupied sector in a grouped launch.
struct GSec {
  int64_t n;      // rows (sector dimension)
  int m;          // active slots
  int ldx;        // leading dimension of X
  int rpc, nch;   // rows per reduction chunk, chunks
  int groups;     // ceil(m / 64)
  int so;         // first global slot
  int64_t vo;     // offset of the sector's n x m block in the V / W pools
  int64_t co;     // sum of nch * m over the earlier sectors (partials unit)
  int64_t tile0;  // first reduction tile (tiles = nch * groups)
  int64_t blk0;   // first product block (blocks = ceil(n / SPMM_RB) * groups)
  int64_t grp0;   // first 64-slot group (sum of groups over the earlier sectors)
  cplx* X;
  const int32_t* outer;
  const int32_t* inner;
  const cplx* vals;
  const double* D;
  const uint64_t* states;
};

// index of the last sector whose key <= v (keys ascending, first key 0)
template <typename F>
__device__ inline int find_sec(int ns, int64_t v, F key) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key(mid) <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
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

// Reduction tile -> (sector, chunk, slot of this lane); one wave per tile.
struct Tile {
  int sec, chunk, t;
};
__device__ inline Tile tile_of(const GSec* __restrict__ gs, int ns) {
  Tile r;
  const int64_t b = blockIdx.x;
  r.sec = find_sec(ns, b, [&](int i) { return gs[i].tile0; });
  const GSec& S = gs[r.sec];
  const int64_t local = b - S.tile0;
  r.chunk = (int)(local / S.groups);
  r.t = (int)(local % S.groups) * TT + threadIdx.x;
  return r;
}

// One pass over a row chunk of NQ basis vectors and W for every sector
// (trajectory-minor blocks in the V / W pools, lds = m).
//   MODE 0: part_q = sum_s conj(V_q[s,t]) W[s,t]                  (first Gram-Schmidt pass)
//   MODE 1: W <- W - sum_q V_q h_q, then part_q as in MODE 0      (update + second pass)
//   MODE 2: W <- W - sum_q V_q h_q, then part_0 = sum_s |W[s,t]|^2 (update + norm)
// part[P co + (chunk P + q) m + t], P = NQ (modes 0, 1) or 1 (mode 2); h[q Mtot + slot].
// NQ is a template parameter so the per-lane accumulators stay in registers.
template <int NQ, int MODE>
__global__ void __launch_bounds__(TT)
    g_pass_kernel(const GSec* __restrict__ gs, int ns, const cplx* __restrict__ V, int64_t VS, cplx* __restrict__ W,
                  const cplx* __restrict__ h, int Mtot, cplx* __restrict__ part) {
  constexpr int NP = MODE == 2 ? 1 : NQ;
  const Tile tl = tile_of(gs, ns);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  if (t >= m) return;
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = min<int64_t>(S.n, r0 + S.rpc);
  const cplx* __restrict__ Vb = V + S.vo;
  cplx* __restrict__ Wb = W + S.vo;
  cplx hc[NQ];
  if (MODE != 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) hc[q] = h[(int64_t)q * Mtot + S.so + t];
  }
  cplx acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = cmake(0.0, 0.0);
#pragma unroll 2
  for (int64_t s = r0; s < r1; ++s) {
    const int64_t o = s * m + t;
    cplx v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = Vb[q * VS + o];
    cplx w = Wb[o];
    if (MODE != 0) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        w.x -= v[q].x * hc[q].x - v[q].y * hc[q].y;
        w.y -= v[q].x * hc[q].y + v[q].y * hc[q].x;
      }
      Wb[o] = w;
    }
    if (MODE == 2) {
      acc[0].x += w.x * w.x + w.y * w.y;
    } else {
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc[q] = cadd(acc[q], cdotmul(v[q], w));
    }
  }
  cplx* pb = part + NP * S.co;
#pragma unroll
  for (int q = 0; q < NP; ++q) pb[((int64_t)tl.chunk * NP + q) * m + t] = acc[q];
}

// Sums the chunk partials of every slot (fixed order) and files them:
//   RMODE 0: h = sum, Hess[0..nq-1][j] = sum;  RMODE 1: h = sum, Hess[..][j] += sum;
//   RMODE 2: beta = sqrt(sum), Hess[j+1][j] = beta, scale = 1/beta (0 below 1e-10).
// One 256-thread block per (q, 64-slot group): wave w sums chunks w, w+4, ...,
// then the four wave sums are added in wave order.  Hess[(j (K+1) + i) Mtot + slot].
__global__ void __launch_bounds__(256)
    g_reduce_kernel(const GSec* __restrict__ gs, int ns, int64_t ngroups, int Mtot, int nq,
                    const cplx* __restrict__ part, cplx* __restrict__ h, cplx* __restrict__ hess, int j, int K,
                    int rmode, double* __restrict__ scale) {
  __shared__ cplx sh[4][TT];
  const int q = (int)(blockIdx.x / ngroups);
  const int64_t gi = blockIdx.x % ngroups;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  cplx s = cmake(0.0, 0.0);
  if (t < S.m) {
    const cplx* pb = part + (int64_t)nq * S.co;
    for (int c = wv; c < S.nch; c += 4) s = cadd(s, pb[((int64_t)c * nq + q) * S.m + t]);
  }
  sh[wv][lane] = s;
  __syncthreads();
  if (wv != 0 || t >= S.m) return;
  s = cadd(cadd(sh[0][lane], sh[1][lane]), cadd(sh[2][lane], sh[3][lane]));
  const int g = S.so + t;
  if (rmode == 2) {
    const double b = sqrt(s.x);
    hess[((int64_t)(j * (K + 1) + j + 1)) * Mtot + g] = cmake(b, 0.0);
    scale[g] = (b >= 1e-10) ? 1.0 / b : 0.0;
    return;
  }
  h[(int64_t)q * Mtot + g] = s;
  cplx* p = hess + ((int64_t)(j * (K + 1) + q)) * Mtot + g;
  *p = rmode ? cadd(*p, s) : s;
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

// Multi-RHS product with the normalisation of the source folded in, all sectors:
//   x = f_t src[s,t] (f = NULL: 1),  Vout[s,t] = x (Vout may be NULL),
//   Wout[s,t] = (H x)[s] + (alpha_t D[s] + beta_t) x[s]
// H from the ColMajor arrays of a Hermitian operator (pull form: row s of H is
// column s conjugated); alpha, beta per slot.
// src is the sector's X (from_x) or the W pool.  One wave per (SPMM_RB rows,
// 64 slots): the wave loads the rows' pointers and up to 64 entries of a row
// with one vector load each and broadcasts them with v_readlane, so all of a
// row's gathers (64 consecutive trajectories each) are in flight together.
//
// Block order: workgroups are dealt to the 8 XCDs round-robin, so block b
// runs on XCD b % 8; it is given work item (b % 8) * (nb / 8) + b / 8, which
// hands every XCD a contiguous range of (sector, group, rows) items.  Within a
// sector the items run group-major (all rows of one 64-slot group, then the
// next), so the rows an XCD gathers from belong to the one or two groups it is
// working on and stay in its L2.
constexpr int SPMM_RB = 4;
__device__ inline double readlane_d(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}
__global__ void __launch_bounds__(TT)
    g_spmm_kernel(const GSec* __restrict__ gs, int ns, int64_t nblocks, int from_x, const cplx* __restrict__ Wsrc,
                  const double* __restrict__ f, const cplx* __restrict__ alpha, const cplx* __restrict__ beta,
                  cplx* __restrict__ Vout, cplx* __restrict__ Wout) {
  const int64_t per_xcd = (int64_t)gridDim.x / 8;
  const int64_t b = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (b >= nblocks) return;
  const int si = find_sec(ns, b, [&](int i) { return gs[i].blk0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x;
  const int64_t nrb = (S.n + SPMM_RB - 1) / SPMM_RB;
  const int64_t local = b - S.blk0;
  const int64_t s0 = (local % nrb) * SPMM_RB;
  const int nr = (int)min<int64_t>(SPMM_RB, S.n - s0);
  const int t = (int)(local / nrb) * TT + lane;
  const int m = S.m;
  const bool active = t < m;
  const cplx* __restrict__ src = from_x ? S.X : Wsrc + S.vo;
  const int64_t lds = from_x ? S.ldx : m;
  const int g = S.so + (active ? t : 0);
  const double ft = active ? (f ? f[g] : 1.0) : 0.0;
  const cplx al = active ? alpha[g] : cmake(0.0, 0.0);
  const cplx be = active ? beta[g] : cmake(0.0, 0.0);
  const int optr = (lane <= nr) ? S.outer[s0 + lane] : 0;
  for (int r = 0; r < nr; ++r) {
    const int64_t s = s0 + r;
    const int e0 = __builtin_amdgcn_readlane(optr, r);
    const int e1 = __builtin_amdgcn_readlane(optr, r + 1);
    double sx = 0.0, sy = 0.0;
    for (int base = e0; base < e1; base += TT) {
      const int cnt = min(TT, e1 - base);
      const int ci = (lane < cnt) ? S.inner[base + lane] : 0;
      const cplx cv = (lane < cnt) ? S.vals[base + lane] : cmake(0.0, 0.0);
      // batches of 8 entries: all 8 gathers are issued before the sums; the
      // padding entries of the last batch repeat a valid column with weight 0
      for (int k0 = 0; k0 < cnt; k0 += 8) {
        cplx x[8];
        double vx[8], vy[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = min(k0 + i, cnt - 1);
          const int col = __builtin_amdgcn_readlane(ci, k);
          const bool real = k0 + i < cnt;
          vx[i] = real ? readlane_d(cv.x, k) : 0.0;
          vy[i] = real ? readlane_d(cv.y, k) : 0.0;
          x[i] = active ? src[(int64_t)col * lds + t] : cmake(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          sx += vx[i] * x[i].x + vy[i] * x[i].y;
          sy += vx[i] * x[i].y - vy[i] * x[i].x;
        }
      }
    }
    if (active) {
      const cplx x0 = src[s * lds + t];
      const cplx x = cmake(x0.x * ft, x0.y * ft);
      const cplx dg = cmake(al.x * S.D[s] + be.x, al.y * S.D[s] + be.y);
      const cplx dx = cmul(dg, x);
      if (Vout) Vout[S.vo + s * m + t] = x;
      Wout[S.vo + s * m + t] = cmake(sx * ft + dx.x, sy * ft + dx.y);
    }
  }
}

// ---------------------------------------------------------- delayed CGS2
// Batched form of the delayed classical Gram-Schmidt with one
// reorthogonalisation (the single-vector form and its derivation are in
// krylov.hip): step j holds, per slot, the pending vector u (A q_{j-1}
// projected once) and w = A u.  One pass over Q_{j-1} gives c = Q^H u,
// d = Q^H w, |u|^2 and u^H w (g_red_kernel); the small algebra of each slot
// (g_dc_small_kernel) finishes column j-1 of its Hessenberg and the first pass
// of column j; a second pass writes q_j and the next pending vector
// (g_dc_update_kernel).  Two passes over the basis per vector instead of
// CGS2's three fused ones.

// part[P co + (chunk P + p) m + t], P = NR NQ + 2:
//   p < NQ: conj(V_p) u;  NR == 2: NQ + p: conj(V_p) w;  P-2: |u|^2;  P-1: conj(u) w
template <int NQ, int NR>
__global__ void __launch_bounds__(TT)
    g_red_kernel(const GSec* __restrict__ gs, int ns, const cplx* __restrict__ V, int64_t VS,
                 const cplx* __restrict__ U, const cplx* __restrict__ W, cplx* __restrict__ part) {
  constexpr int P = NR * NQ + 2;
  const Tile tl = tile_of(gs, ns);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  if (t >= m) return;
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = min<int64_t>(S.n, r0 + S.rpc);
  const cplx* __restrict__ Vb = V + S.vo;
  const cplx* __restrict__ Ub = U + S.vo;
  const cplx* __restrict__ Wb = W + S.vo;
  cplx ac[NQ], ad[NR == 2 ? NQ : 1];
  cplx uu = cmake(0.0, 0.0), uw = cmake(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < NQ; ++q) ac[q] = cmake(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < (NR == 2 ? NQ : 1); ++q) ad[q] = cmake(0.0, 0.0);
#pragma unroll 2
  for (int64_t s = r0; s < r1; ++s) {
    const int64_t o = s * m + t;
    const cplx u = Ub[o];
    const cplx w = (NR == 2) ? Wb[o] : cmake(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const cplx v = Vb[q * VS + o];
      ac[q] = cadd(ac[q], cdotmul(v, u));
      if (NR == 2) ad[q] = cadd(ad[q], cdotmul(v, w));
    }
    uu = cadd(uu, cdotmul(u, u));
    if (NR == 2) uw = cadd(uw, cdotmul(u, w));
  }
  cplx* pb = part + P * S.co + (int64_t)tl.chunk * P * m + t;
#pragma unroll
  for (int q = 0; q < NQ; ++q) pb[(int64_t)q * m] = ac[q];
  if (NR == 2) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) pb[(int64_t)(NQ + q) * m] = ad[q];
  }
  pb[(int64_t)(P - 2) * m] = uu;
  pb[(int64_t)(P - 1) * m] = uw;
}

// red[p Mtot + slot] = sum over chunks (fixed order) of the partials, P per slot.
// One 256-thread block per (p, 64-slot group); wave w sums chunks w, w+4, ...
__global__ void __launch_bounds__(256)
    g_reduce_p_kernel(const GSec* __restrict__ gs, int ns, int64_t ngroups, int Mtot, int P,
                      const cplx* __restrict__ part, cplx* __restrict__ red) {
  __shared__ cplx sh[4][TT];
  const int p = (int)(blockIdx.x / ngroups);
  const int64_t gi = blockIdx.x % ngroups;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  cplx s = cmake(0.0, 0.0);
  if (t < S.m) {
    const cplx* pb = part + (int64_t)P * S.co;
    for (int c = wv; c < S.nch; c += 4) s = cadd(s, pb[((int64_t)c * P + p) * S.m + t]);
  }
  sh[wv][lane] = s;
  __syncthreads();
  if (wv != 0 || t >= S.m) return;
  red[(int64_t)p * Mtot + S.so + t] = cadd(cadd(sh[0][lane], sh[1][lane]), cadd(sh[2][lane], sh[3][lane]));
}

// Small algebra of step j for every slot (one thread each; Hess[(col (K+1) + row) Mtot + slot]).
// last: finalisation (red holds c and |u|^2 only).  coef[q Mtot + slot]: a (= c) at
// q < j, b at K + q; sc[slot] = {1/alpha, f_j/alpha}.  Below eps the column ends
// the Krylov space: every coefficient is zero, so the later vectors vanish.
__global__ void g_dc_small_kernel(int Mtot, int j, int K, int last, const cplx* __restrict__ red,
                                  cplx* __restrict__ hess, cplx* __restrict__ coef, cplx* __restrict__ sc, double eps) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const int NR = last ? 1 : 2;
  const int P = NR * j + 2;
  auto R = [&](int p) { return red[(int64_t)p * Mtot + g]; };
  auto H = [&](int row, int col) -> cplx& { return hess[((int64_t)col * (K + 1) + row) * Mtot + g]; };
  double cc = 0.0;
  cplx chd = cmake(0.0, 0.0);
  for (int i = 0; i < j; ++i) {
    const cplx c = R(i);
    cc += c.x * c.x + c.y * c.y;
    if (!last) chd = cadd(chd, cdotmul(c, R(j + i)));
  }
  const double a2 = R(P - 2).x - cc;
  const double alpha = a2 > 0.0 ? sqrt(a2) : 0.0;
  // finish column j-1
  for (int i = 0; i < j; ++i) H(i, j - 1) = cadd(H(i, j - 1), R(i));
  if (j <= K) H(j, j - 1) = cmake(alpha, 0.0);
  if (last) return;
  if (!(alpha >= eps)) {
    for (int i = 0; i < j; ++i) {
      coef[(int64_t)i * Mtot + g] = cmake(0.0, 0.0);
      coef[(int64_t)(K + i) * Mtot + g] = cmake(0.0, 0.0);
    }
    for (int i = 0; i <= j; ++i) H(i, j) = cmake(0.0, 0.0);
    sc[2 * g] = cmake(0.0, 0.0);
    sc[2 * g + 1] = cmake(0.0, 0.0);
    return;
  }
  const double inv = 1.0 / alpha;
  const cplx fj = crmul(inv * inv, csub(R(P - 1), chd));
  // e = H[0:j+1, 0:j] c;  column j first pass = f - e / alpha
  for (int i = 0; i <= j; ++i) {
    cplx e = cmake(0.0, 0.0);
    for (int k = (i > 0 ? i - 1 : 0); k < j; ++k) e = cadd(e, cmul(H(i, k), R(k)));
    const cplx f = (i < j) ? crmul(inv, R(j + i)) : fj;
    H(i, j) = csub(f, crmul(inv, e));
    if (i < j) {
      coef[(int64_t)i * Mtot + g] = R(i);
      coef[(int64_t)(K + i) * Mtot + g] = csub(f, crmul(inv, cmul(R(i), fj)));
    }
  }
  sc[2 * g] = cmake(inv, 0.0);
  sc[2 * g + 1] = crmul(inv, fj);
}

// q_j = (u - sum_q V_q a_q) / alpha -> V_j;  u <- w / alpha - sum_q V_q b_q - u f_j / alpha
template <int NQ>
__global__ void __launch_bounds__(TT)
    g_dc_update_kernel(const GSec* __restrict__ gs, int ns, cplx* __restrict__ V, int64_t VS, int j,
                       cplx* __restrict__ U, const cplx* __restrict__ W, const cplx* __restrict__ coef,
                       const cplx* __restrict__ sc, int Mtot, int K) {
  const Tile tl = tile_of(gs, ns);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  if (t >= m) return;
  const int g = S.so + t;
  cplx a[NQ], b[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    a[q] = coef[(int64_t)q * Mtot + g];
    b[q] = coef[(int64_t)(K + q) * Mtot + g];
  }
  const double inv = sc[2 * g].x;
  const cplx f = sc[2 * g + 1];
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = min<int64_t>(S.n, r0 + S.rpc);
  cplx* __restrict__ Vb = V + S.vo;
  cplx* __restrict__ Ub = U + S.vo;
  const cplx* __restrict__ Wb = W + S.vo;
#pragma unroll 2
  for (int64_t s = r0; s < r1; ++s) {
    const int64_t o = s * m + t;
    cplx s1 = cmake(0.0, 0.0), s2 = cmake(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const cplx v = Vb[q * VS + o];
      s1 = cadd(s1, cmul(v, a[q]));
      s2 = cadd(s2, cmul(v, b[q]));
    }
    const cplx u = Ub[o];
    const cplx w = Wb[o];
    Vb[(int64_t)j * VS + o] = crmul(inv, csub(u, s1));
    const cplx uf = cmul(u, f);
    Ub[o] = cmake(w.x * inv - s2.x - uf.x, w.y * inv - s2.y - uf.y);
  }
}

// y_t = expm(-i dt K_t) e_1 for the k x k Hessenberg of each trajectory:
// scaling and squaring with a truncated Taylor series applied to vectors
// (Moler & Van Loan, "Nineteen dubious ways to compute the exponential of a
// matrix", methods 1 and 3): A = -i dt K / 2^s with |A|_1 <= 1/2, then
// y <- T_16(A) y, 2^s times (remainder below 1e-19 relative).
// One wave per trajectory; lane i owns row i, the matrix and the current
// Taylor term live in LDS.
__global__ void __launch_bounds__(64)
    small_expm_kernel(int m, int k, int K, const cplx* __restrict__ hess, double dt, cplx* __restrict__ coef) {
  __shared__ cplx Hs[MAXK][MAXK + 1];
  __shared__ cplx tv[MAXK];
  __shared__ double cn[MAXK];
  const int t = blockIdx.x;
  const int i = threadIdx.x;
  for (int e = i; e < k * k; e += 64) {
    const int r = e % k, c = e / k;
    Hs[r][c] = (r <= c + 1) ? hess[((int64_t)(c * (K + 1) + r)) * m + t] : cmake(0.0, 0.0);
  }
  __syncthreads();
  if (i < k) {
    double c = 0.0;
    for (int r = 0; r <= min(i + 1, k - 1); ++r) c += sqrt(Hs[r][i].x * Hs[r][i].x + Hs[r][i].y * Hs[r][i].y);
    cn[i] = c * fabs(dt);
  }
  __syncthreads();
  double nrm = 0.0;
  for (int c = 0; c < k; ++c) nrm = fmax(nrm, cn[c]);
  int sq = 0;
  while (nrm > 0.5 && sq < 16) {
    nrm *= 0.5;
    ++sq;
  }
  const double sc = ldexp(dt, -sq);
  cplx y = cmake(i == 0 ? 1.0 : 0.0, 0.0);
  const int reps = 1 << sq;
  for (int rep = 0; rep < reps; ++rep) {
    cplx term = y;
    for (int p = 1; p <= 16; ++p) {
      __syncthreads();
      if (i < k) tv[i] = term;
      __syncthreads();
      cplx a = cmake(0.0, 0.0);
      if (i < k)
        for (int j = (i > 0 ? i - 1 : 0); j < k; ++j) a = cadd(a, cmul(Hs[i][j], tv[j]));
      term = cmake(sc * a.y / p, -sc * a.x / p);  // -i sc a / p
      y = cadd(y, term);
    }
  }
  if (i < k) coef[i * m + t] = y;
}

// X[s,t] = sum_j V_j[s,t] coef[j Mtot + slot], all sectors (one thread per element of the pools)
__global__ void g_combine_kernel(const GSec* __restrict__ gs, int ns, int64_t total, const cplx* __restrict__ V,
                                 int64_t VS, int k, const cplx* __restrict__ coef, int Mtot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int si = find_sec(ns, e, [&](int i) { return gs[i].vo; });
  const GSec& S = gs[si];
  const int64_t local = e - S.vo;
  const int64_t s = local / S.m;
  const int t = (int)(local % S.m);
  const int g = S.so + t;
  cplx a = cmake(0.0, 0.0);
  for (int j = 0; j < k; ++j) a = cadd(a, cmul(V[j * VS + e], coef[(int64_t)j * Mtot + g]));
  S.X[s * S.ldx + t] = a;
}

// densities: part[(M+1) co + (chunk (M+1) + q) m + t] = sum_{s in chunk} |X[s,t]|^2 [bit q of state s]
// (q = M: plain norm^2), all sectors
__global__ void __launch_bounds__(TT)
    g_dens_partial_kernel(const GSec* __restrict__ gs, int ns, int M, double* __restrict__ part) {
  const Tile tl = tile_of(gs, ns);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  if (t >= m) return;
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = min<int64_t>(S.n, r0 + S.rpc);
  double acc[MAXM + 1];
#pragma unroll
  for (int q = 0; q <= MAXM; ++q) acc[q] = 0.0;
  for (int64_t s = r0; s < r1; ++s) {
    const cplx v = S.X[s * S.ldx + t];
    const double p = v.x * v.x + v.y * v.y;
    const uint64_t x = S.states[s];
#pragma unroll
    for (int q = 0; q < MAXM; ++q)
      if ((x >> q) & 1ull) acc[q] += p;
    acc[MAXM] += p;
  }
  double* pb = part + (int64_t)(M + 1) * S.co;
#pragma unroll
  for (int q = 0; q < MAXM; ++q)
    if (q < M) pb[((int64_t)tl.chunk * (M + 1) + q) * m + t] = acc[q];
  pb[((int64_t)tl.chunk * (M + 1) + M) * m + t] = acc[MAXM];
}

// dens_all[traj (M+1) + q] = density q / norm^2 (q < M), norm^2 (q = M); one thread per (q, slot)
__global__ void g_dens_reduce_kernel(const GSec* __restrict__ gs, int ns, int Mtot, int M,
                                     const double* __restrict__ part, const int* __restrict__ slot2traj,
                                     double* __restrict__ dens_all) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)Mtot * (M + 1)) return;
  const int q = (int)(idx / Mtot), g = (int)(idx % Mtot);
  const int si = find_sec(ns, g, [&](int i) { return (int64_t)gs[i].so; });
  const GSec& S = gs[si];
  const int t = g - S.so;
  const double* pb = part + (int64_t)(M + 1) * S.co;
  double v = 0.0, nn = 0.0;
  for (int c = 0; c < S.nch; ++c) {
    v += pb[((int64_t)c * (M + 1) + q) * S.m + t];
    if (q < M) nn += pb[((int64_t)c * (M + 1) + M) * S.m + t];
  }
  dens_all[(int64_t)slot2traj[g] * (M + 1) + q] = (q < M) ? v / nn : v;
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

struct JumpRec {
  int slot;      // source slot
  int mode;      // mode m
  int loss;      // 0 dephasing, 1 loss
  int staging;   // staging vector index
  double scale;  // 1 / |L psi| (from the recorded densities and norm)
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

template <int NR>
void launch_red(int nq, unsigned tiles, hipStream_t st, const GSec* gs, int ns, const cplx* V, int64_t VS,
                const cplx* U, const cplx* W, cplx* part) {
  switch (nq) {
#define QD_RED_CASE(N)                                                            \
  case N:                                                                         \
    g_red_kernel<N, NR><<<tiles, TT, 0, st>>>(gs, ns, V, VS, U, W, part);         \
    break;
    QD_RED_CASE(1) QD_RED_CASE(2) QD_RED_CASE(3) QD_RED_CASE(4) QD_RED_CASE(5) QD_RED_CASE(6)
    QD_RED_CASE(7) QD_RED_CASE(8) QD_RED_CASE(9) QD_RED_CASE(10) QD_RED_CASE(11) QD_RED_CASE(12)
    QD_RED_CASE(13) QD_RED_CASE(14) QD_RED_CASE(15) QD_RED_CASE(16) QD_RED_CASE(17) QD_RED_CASE(18)
    QD_RED_CASE(19) QD_RED_CASE(20) QD_RED_CASE(21) QD_RED_CASE(22) QD_RED_CASE(23) QD_RED_CASE(24)
    QD_RED_CASE(25) QD_RED_CASE(26) QD_RED_CASE(27) QD_RED_CASE(28) QD_RED_CASE(29) QD_RED_CASE(30)
    QD_RED_CASE(31) QD_RED_CASE(32)
#undef QD_RED_CASE
    default:
      break;
  }
}

void launch_dc_update(int nq, unsigned tiles, hipStream_t st, const GSec* gs, int ns, cplx* V, int64_t VS, int j,
                      cplx* U, const cplx* W, const cplx* coef, const cplx* sc, int Mtot, int K) {
  switch (nq) {
#define QD_UPD_CASE(N)                                                                         \
  case N:                                                                                      \
    g_dc_update_kernel<N><<<tiles, TT, 0, st>>>(gs, ns, V, VS, j, U, W, coef, sc, Mtot, K);    \
    break;
    QD_UPD_CASE(1) QD_UPD_CASE(2) QD_UPD_CASE(3) QD_UPD_CASE(4) QD_UPD_CASE(5) QD_UPD_CASE(6)
    QD_UPD_CASE(7) QD_UPD_CASE(8) QD_UPD_CASE(9) QD_UPD_CASE(10) QD_UPD_CASE(11) QD_UPD_CASE(12)
    QD_UPD_CASE(13) QD_UPD_CASE(14) QD_UPD_CASE(15) QD_UPD_CASE(16) QD_UPD_CASE(17) QD_UPD_CASE(18)
    QD_UPD_CASE(19) QD_UPD_CASE(20) QD_UPD_CASE(21) QD_UPD_CASE(22) QD_UPD_CASE(23) QD_UPD_CASE(24)
    QD_UPD_CASE(25) QD_UPD_CASE(26) QD_UPD_CASE(27) QD_UPD_CASE(28) QD_UPD_CASE(29) QD_UPD_CASE(30)
    QD_UPD_CASE(31)
#undef QD_UPD_CASE
    default:
      break;
  }
}

unsigned g1(int64_t work, int b) { return (unsigned)std::max<int64_t>(1, (work + b - 1) / b); }

// Rows per reduction chunk of a sector: about `target` waves for the sector
// alone (target / groups chunks), 16 to 1024 rows, at most QD_TRAJ_MAX_CHUNKS
// chunks.
constexpr int QD_TRAJ_MAX_CHUNKS = 256;
int rows_per_chunk(int64_t n, int groups, int target) {
  const int64_t want = std::max<int64_t>(1, target / std::max(1, groups));
  int64_t rpc = std::max<int64_t>(16, (n + want - 1) / want);
  rpc = std::min<int64_t>(rpc, 1024);
  rpc = std::max<int64_t>(rpc, (n + QD_TRAJ_MAX_CHUNKS - 1) / QD_TRAJ_MAX_CHUNKS);
  return (int)rpc;
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

  // device buffers
  int *d_jmode = nullptr, *d_action = nullptr, *d_slot2traj = nullptr;
  double *d_jrate = nullptr, *d_dens_all = nullptr, *d_gtraj = nullptr, *d_dsum = nullptr, *d_dsq = nullptr;
  double *d_scale = nullptr;
  cplx *d_alpha = nullptr, *d_beta = nullptr, *d_red = nullptr, *d_coef2 = nullptr, *d_sc = nullptr;
  cplx *d_V = nullptr, *d_W = nullptr, *d_W2 = nullptr, *d_part = nullptr, *d_h = nullptr, *d_hess = nullptr;
  cplx *d_coef = nullptr, *d_staging = nullptr;
  GSec* d_gsec = nullptr;
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
    return QD_OK;
  };

  if (!dalloc((void**)&d_jmode, sizeof(int) * std::max(1, njumps)) ||
      !dalloc((void**)&d_jrate, sizeof(double) * std::max(1, njumps)) ||
      !dalloc((void**)&d_action, sizeof(int) * ntraj) || !dalloc((void**)&d_slot2traj, sizeof(int) * ntraj) ||
      !dalloc((void**)&d_dens_all, sizeof(double) * ntraj * (M + 1)) ||
      !dalloc((void**)&d_gtraj, sizeof(double) * ntraj) || !dalloc((void**)&d_dsum, sizeof(double) * M) ||
      !dalloc((void**)&d_dsq, sizeof(double) * M) || !dalloc((void**)&d_scale, sizeof(double) * ntraj) ||
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
      !dalloc((void**)&d_pairs, sizeof(int2) * ntraj) || !dalloc((void**)&d_moves, sizeof(int2) * ntraj)) {
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
    broadcast_kernel<<<g1(S.n * ntraj, 256), 256, 0, st>>>(S.n, ntraj, S.ld, (const cplx*)d_psi0, S.X);
  }
  std::vector<int> action(ntraj);
  std::vector<double> dens_host((size_t)ntraj * (M + 1));
  // per-step host tables, reserved so that the async copies from them stay
  // valid until the next synchronisation
  std::vector<JumpRec> h_jr;
  std::vector<int2> h_pairs, h_moves;
  std::vector<GSec> h_gsec;
  std::vector<int> h_slot2traj;
  h_jr.reserve(ntraj);
  h_pairs.reserve(ntraj);
  h_moves.reserve(ntraj);
  h_gsec.reserve(NSMAX);
  h_slot2traj.reserve(ntraj);
  int stg_cap = 0;
  int ns = 0;
  int64_t VS = 0, tiles = 0, blocks = 0, ngroups = 0;
  int occupancy_target = 2048;  // waves per sector for the reduction tiles (tuned below)

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
      blocks += (S.n + SPMM_RB - 1) / SPMM_RB * g.groups;
      ngroups += g.groups;
    }
    ns = (int)h_gsec.size();
    VS = vo;
    hipMemcpyAsync(d_gsec, h_gsec.data(), sizeof(GSec) * ns, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_slot2traj, h_slot2traj.data(), sizeof(int) * ntraj, hipMemcpyHostToDevice, st);
  };

  // densities of every trajectory -> d_dens_all; sums over trajectories -> host (step >= 0)
  auto densities = [&](int step) -> int {
    g_dens_partial_kernel<<<(unsigned)tiles, TT, 0, st>>>(d_gsec, ns, M, (double*)d_part);
    g_dens_reduce_kernel<<<g1((int64_t)ntraj * (M + 1), 256), 256, 0, st>>>(d_gsec, ns, ntraj, M, (double*)d_part,
                                                                           d_slot2traj, d_dens_all);
    dens_accumulate_kernel<<<M, 256, 0, st>>>(ntraj, M, d_dens_all, d_dsum, d_dsq);
    double hs[2 * MAXM];
    QD_HIP(ctx, hipMemcpyAsync(hs, d_dsum, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    QD_HIP(ctx, hipMemcpyAsync(hs + MAXM, d_dsq, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    QD_HIP(ctx, hipStreamSynchronize(st));
    if (step >= 0)
      for (int q = 0; q < M; ++q) {
        dsum[(size_t)step * M + q] = hs[q];
        dsq[(size_t)step * M + q] = hs[MAXM + q];
      }
    return QD_OK;
  };

  // one Arnoldi step (delayed CGS2, K vectors) for every slot of every occupied sector, X updated in place
  auto evolve_all = [&]() -> int {
    const int Mtot = ntraj;
    slot_params_kernel<<<g1(Mtot, 256), 256, 0, st>>>(Mtot, M, d_slot2traj, d_gtraj, d_dens_all, d_alpha, d_beta,
                                                       d_scale);
    const unsigned sgrid = (unsigned)((blocks + 7) / 8 * 8);
    const unsigned tl = (unsigned)tiles;
    cplx* U = d_W2;
    cplx* W = d_W;
    // V_0 = X / |X|, U = H_eff V_0, then its first projection: H[0][0] = V_0^H U, U -= V_0 H[0][0]
    g_spmm_kernel<<<sgrid, TT, 0, st>>>(d_gsec, ns, blocks, 1, nullptr, d_scale, d_alpha, d_beta, d_V, U);
    g_pass_kernel<1, 0><<<tl, TT, 0, st>>>(d_gsec, ns, d_V, VS, U, nullptr, Mtot, d_part);
    g_reduce_kernel<<<(unsigned)ngroups, 256, 0, st>>>(d_gsec, ns, ngroups, Mtot, 1, d_part, d_h, d_hess, 0, K, 0,
                                                        nullptr);
    g_pass_kernel<1, 2><<<tl, TT, 0, st>>>(d_gsec, ns, d_V, VS, U, d_h, Mtot, d_part);
    for (int j = 1; j < K; ++j) {
      const int P = 2 * j + 2;
      g_spmm_kernel<<<sgrid, TT, 0, st>>>(d_gsec, ns, blocks, 0, U, nullptr, d_alpha, d_beta, nullptr, W);
      launch_red<2>(j, tl, st, d_gsec, ns, d_V, VS, U, W, d_part);
      g_reduce_p_kernel<<<(unsigned)(ngroups * P), 256, 0, st>>>(d_gsec, ns, ngroups, Mtot, P, d_part, d_red);
      g_dc_small_kernel<<<g1(Mtot, 256), 256, 0, st>>>(Mtot, j, K, 0, d_red, d_hess, d_coef2, d_sc, 1e-10);
      launch_dc_update(j, tl, st, d_gsec, ns, d_V, VS, j, U, W, d_coef2, d_sc, Mtot, K);
    }
    // the pending vector's second projection completes column K-1
    launch_red<1>(K, tl, st, d_gsec, ns, d_V, VS, U, nullptr, d_part);
    g_reduce_p_kernel<<<(unsigned)(ngroups * (K + 2)), 256, 0, st>>>(d_gsec, ns, ngroups, Mtot, K + 2, d_part, d_red);
    g_dc_small_kernel<<<g1(Mtot, 256), 256, 0, st>>>(Mtot, K, K, 1, d_red, d_hess, d_coef2, d_sc, 1e-10);
    small_expm_kernel<<<(unsigned)Mtot, 64, 0, st>>>(Mtot, K, K, d_hess, dt, d_coef);
    g_combine_kernel<<<g1(VS, 256), 256, 0, st>>>(d_gsec, ns, VS, d_V, VS, K, d_coef, Mtot);
    QD_CHECK_LAUNCH(ctx);
    return QD_OK;
  };

  // Tile autotuner for the uneven sectors: the occupancy target of the
  // reduction tiles (rows per chunk of every sector) is timed on the widest
  // reduction pass over the current sector membership (a read-only pass,
  // results discarded).  It runs at the start and again once the loss jumps
  // have spread the trajectories over smaller sectors.
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEventCreate(&ev0);
  hipEventCreate(&ev1);
  auto tune_tiles = [&]() {
    const int cand[5] = {512, 1024, 2048, 4096, 8192};
    float best = 1e30f;
    int best_c = occupancy_target;
    for (int c : cand) {
      occupancy_target = c;
      build_groups();
      const int nq = std::max(1, K - 1);
      launch_red<2>(nq, (unsigned)tiles, st, d_gsec, ns, d_V, VS, d_W, d_W2, d_part);
      hipEventRecord(ev0, st);
      for (int r = 0; r < 2; ++r) launch_red<2>(nq, (unsigned)tiles, st, d_gsec, ns, d_V, VS, d_W, d_W2, d_part);
      hipEventRecord(ev1, st);
      hipEventSynchronize(ev1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev0, ev1);
      if (ms < best) {
        best = ms;
        best_c = c;
      }
    }
    occupancy_target = best_c;
  };
  tune_tiles();

  for (int step = 0; step <= nsteps; ++step) {
    if (step > 0 && step == nsteps / 4) tune_tiles();
    build_groups();
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
          if (mode < nsite)
            --na;
          else
            --nb;
        }
        pend.push_back({tr, na, nb, (int)pend.size(), slot, loss});
        h_jr.push_back(r);
      }
      const int njr = (int)(h_jr.size() - jr0);
      if (njr == 0) continue;
      // staging capacity grows on demand
      if (!d_staging || (int)pend.size() > stg_cap) {
        int ncap = std::max<int>((int)pend.size() * 2, 64);
        cplx* nsg = nullptr;
        if (hipMalloc(&nsg, sizeof(cplx) * (size_t)ncap * nmax) != hipSuccess) {
          rc = qd_set_error(ctx, QD_E_OOM, "jump staging");
          break;
        }
        if (d_staging) {
          hipMemcpyAsync(nsg, d_staging, sizeof(cplx) * (size_t)stg_cap * nmax, hipMemcpyDeviceToDevice, st);
          hipStreamSynchronize(st);
          hipFree(d_staging);
        }
        d_staging = nsg;
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
    // evolve every occupied sector (jumpers too; their columns are replaced below)
    if ((rc = evolve_all())) break;
    // commit jumps: dephasing in place; losses move to the target sector
    if (!pend.empty()) {
      n_jumps += (int64_t)pend.size();
      std::vector<std::vector<int>> removed(sec.size());
      std::vector<std::vector<int2>> commits(sec.size());
      for (auto& S : sec) {
        if (!S.basis || S.m == 0) continue;
        for (int slot = 0; slot < S.m; ++slot) {
          int tr = S.slot2traj[slot];
          if (action[tr] < 0) continue;
          for (const Pending& p : pend)
            if (p.traj == tr) {
              if (!p.loss)
                commits[sid(S.a, S.b)].push_back(make_int2(p.sidx, slot));
              else
                removed[sid(S.a, S.b)].push_back(slot);
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
  hipEventDestroy(ev0);
  hipEventDestroy(ev1);
  cleanup();
  return rc;
}
