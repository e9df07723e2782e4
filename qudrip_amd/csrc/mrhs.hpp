// mrhs.hpp -- grouped multi-RHS Krylov machinery shared by the batched
// trajectory engine (traj.hip) and the solve server (serve.hip).
//
// A batch is a set of "sectors" (blocks of slots that share one operator and
// one basis); a slot is one right-hand side (a trajectory, a request).  Slots
// are stored trajectory-minor, X[s * ld + t], and every kernel runs as one
// grouped launch over all sectors: a GSec table gives each sector's rows,
// slots, pool offsets and tile ranges, and each workgroup finds its sector by
// binary search.  One Arnoldi step (delayed CGS2, per-slot Hessenberg and
// exponential) advances every slot at once.
#pragma once
#include <algorithm>
#include <cmath>

#include "qd_internal.hpp"

namespace {

constexpr int TT = 64;    // trajectories per wave
constexpr int MAXM = 32;  // modes (2 * nsite) supported
constexpr int MAXK = 32;  // Krylov dimension limit for the batched step

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
  int mp;         // lanes per row in a reduction tile: 64 (one slot per lane), or for
                  // a single-group sector a power of two >= m (64 / mp rows at a time)
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


// Reduction tile -> (sector, chunk, slot of this lane); one wave per tile.
// With mp < 64 the wave covers rstep = 64 / mp rows at a time: lane = sub mp + t.
struct Tile {
  int sec, chunk, t, sub, rstep;
};
__device__ inline Tile tile_of(const GSec* __restrict__ gs, int ns, int lane) {
  Tile r;
  const int64_t b = blockIdx.x;
  r.sec = find_sec(ns, b, [&](int i) { return gs[i].tile0; });
  const GSec& S = gs[r.sec];
  const int64_t local = b - S.tile0;
  r.chunk = (int)(local / S.groups);
  if (S.mp < TT) {
    r.t = lane % S.mp;
    r.sub = lane / S.mp;
    r.rstep = TT / S.mp;
  } else {
    r.t = (int)(local % S.groups) * TT + lane;
    r.sub = 0;
    r.rstep = 1;
  }
  return r;
}
__device__ inline Tile tile_of(const GSec* __restrict__ gs, int ns) { return tile_of(gs, ns, (int)threadIdx.x); }

// sum over the rstep lanes of a tile that hold the same slot (lanes t, t + mp, ...)
__device__ inline double sub_sum(double v, int mp) {
  for (int o = mp; o < TT; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline cplx sub_sum(cplx v, int mp) { return cmake(sub_sum(v.x, mp), sub_sum(v.y, mp)); }

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
  const bool act = t < m;
  if (tl.rstep == 1 && !act) return;  // no cross-lane sums in one-slot-per-lane tiles
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = act ? min<int64_t>(S.n, r0 + S.rpc) : r0;
  const cplx* __restrict__ Vb = V + S.vo;
  cplx* __restrict__ Wb = W + S.vo;
  cplx hc[NQ];
  if (MODE != 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) hc[q] = act ? h[(int64_t)q * Mtot + S.so + t] : cmake(0.0, 0.0);
  }
  cplx acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = cmake(0.0, 0.0);
#pragma unroll 2
  for (int64_t s = r0 + tl.sub; s < r1; s += tl.rstep) {
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
  if (tl.rstep > 1) {
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = sub_sum(acc[q], S.mp);
  }
  if (!act || tl.sub != 0) return;
  cplx* pb = part + NP * S.co;
#pragma unroll
  for (int q = 0; q < NP; ++q) pb[((int64_t)tl.chunk * NP + q) * m + t] = acc[q];
}

// Chunk partials are summed by RW waves per (value, 64-slot group): wave w
// adds chunks w, w + RW, ... and the wave sums are added in wave order.
constexpr int RW = 16;

// Sums the chunk partials of every slot (fixed order) and files them:
//   RMODE 0: h = sum, Hess[0..nq-1][j] = sum;  RMODE 1: h = sum, Hess[..][j] += sum;
//   RMODE 2: beta = sqrt(sum), Hess[j+1][j] = beta, scale = 1/beta (0 below 1e-10).
// One block per (q, 64-slot group).  Hess[(j (K+1) + i) Mtot + slot].
__global__ void __launch_bounds__(RW * TT)
    g_reduce_kernel(const GSec* __restrict__ gs, int ns, int64_t ngroups, int Mtot, int nq,
                    const cplx* __restrict__ part, cplx* __restrict__ h, cplx* __restrict__ hess, int j, int K,
                    int rmode, double* __restrict__ scale) {
  __shared__ cplx sh[RW][TT];
  const int q = (int)(blockIdx.x / ngroups);
  const int64_t gi = blockIdx.x % ngroups;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  cplx s = cmake(0.0, 0.0);
  if (t < S.m) {
    const cplx* pb = part + (int64_t)nq * S.co;
#pragma unroll 4
    for (int c = wv; c < S.nch; c += RW) s = cadd(s, pb[((int64_t)c * nq + q) * S.m + t]);
  }
  sh[wv][lane] = s;
  __syncthreads();
  if (wv != 0 || t >= S.m) return;
  s = sh[0][lane];
  for (int w = 1; w < RW; ++w) s = cadd(s, sh[w][lane]);
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

// Multi-RHS product with the normalisation of the source folded in, all sectors:
//   x = f_t src[s,t] (f = NULL: 1),  Vout[s,t] = x (Vout may be NULL),
//   Wout[s,t] = (H x)[s] + (alpha_t D[s] + beta_t) x[s],
//   apart[item 64 + lane] = sum over the item's rows s of conj(x[s,t]) Wout[s,t]
//   (apart may be NULL; item = the block's work item, slot t = its lane)
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
                  cplx* __restrict__ Vout, cplx* __restrict__ Wout, cplx* __restrict__ apart) {
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
  cplx ap = cmake(0.0, 0.0);  // sum over the item's rows of conj(x) W (apart)
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
      // both outputs stream out non-temporally: the L2 is left to the gathered rows of src
      if (Vout) {
        __builtin_nontemporal_store(x.x, &Vout[S.vo + s * m + t].x);
        __builtin_nontemporal_store(x.y, &Vout[S.vo + s * m + t].y);
      }
      const cplx wv = cmake(sx * ft + dx.x, sy * ft + dx.y);
      __builtin_nontemporal_store(wv.x, &Wout[S.vo + s * m + t].x);
      __builtin_nontemporal_store(wv.y, &Wout[S.vo + s * m + t].y);
      ap = cadd(ap, cdotmul(x, wv));
    }
  }
  if (apart) apart[b * TT + lane] = ap;
}

// p[0] through a global-space pointer (a global load counts in vmcnt only;
// a flat load also counts in lgkmcnt)
typedef double qd_d2 __attribute__((ext_vector_type(2)));
__device__ inline cplx load_global(const cplx* p) {
  const qd_d2 v = *(const __attribute__((address_space(1))) qd_d2*)p;
  return cmake(v.x, v.y);
}

// p[0] = v, one 16-byte non-temporal global store
__device__ inline void store_nt(cplx* p, cplx v) {
  qd_d2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, (__attribute__((address_space(1))) qd_d2*)p);
}

// g_spmm_kernel with the entries of a wave's RB rows taken as one stream:
// they are contiguous in the CSR arrays, so the batches of 8 gathers run
// across row boundaries and only the last batch of the wave is padded (rows
// of about 9 entries otherwise pad 7 of every 16 gathers).  A row's sum is
// flushed to its outputs when the stream passes its last entry (a uniform
// compare).  The kernel is bound by the texture path's instruction rate, so
// the per-row loads and stores are few and wide: the rows' diagonal values
// come in one load, the outputs go out as 16-byte stores, and the own-row
// values of the diagonal terms are loaded up front (PRE) or at the flush.
// Items are RB rows of one 64-slot group (blk0 / nblocks count ceil(n / RB)
// * groups per sector).  Same per-row entry order and sums as g_spmm_kernel
// (pads added none), and the same apart.
template <int RB, bool PRE>
__global__ void __launch_bounds__(TT)
    g_spmm_flat_kernel(const GSec* __restrict__ gs, int ns, int64_t nblocks, int from_x,
                       const cplx* __restrict__ Wsrc, const double* __restrict__ f, const cplx* __restrict__ alpha,
                       const cplx* __restrict__ beta, cplx* __restrict__ Vout, cplx* __restrict__ Wout,
                       cplx* __restrict__ apart) {
  static_assert(RB < TT, "row pointers of an item fit in one wave");
  const int64_t per_xcd = (int64_t)gridDim.x / 8;
  const int64_t b = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (b >= nblocks) return;
  const int si = find_sec(ns, b, [&](int i) { return gs[i].blk0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x;
  const int64_t nrb = (S.n + RB - 1) / RB;
  const int64_t local = b - S.blk0;
  const int64_t s0 = (local % nrb) * RB;
  const int nr = (int)min<int64_t>(RB, S.n - s0);
  const int t = (int)(local / nrb) * TT + lane;
  const int m = S.m;
  const bool active = t < m;
  const cplx* __restrict__ src = from_x ? S.X : Wsrc + S.vo;
  const int64_t lds = from_x ? S.ldx : m;
  const int g = S.so + (active ? t : 0);
  const int tl = active ? t : 0;  // inactive lanes read slot 0 (in range) and drop it
  const double ft = active ? (f ? f[g] : 1.0) : 0.0;
  const cplx al = active ? alpha[g] : cmake(0.0, 0.0);
  const cplx be = active ? beta[g] : cmake(0.0, 0.0);
  const int optr = (lane <= nr) ? S.outer[s0 + lane] : 0;
  const double dv = (lane < nr) ? S.D[s0 + lane] : 0.0;  // lane r: D of row r
  cplx xo[PRE ? RB : 1];
  if (PRE) {
#pragma unroll
    for (int r = 0; r < RB; ++r) xo[r] = load_global(src + min<int64_t>(s0 + r, S.n - 1) * lds + tl);
  }
  const int E0 = __builtin_amdgcn_readlane(optr, 0);
  const int E1 = __builtin_amdgcn_readlane(optr, nr);
  int r = 0;
  int rend = __builtin_amdgcn_readlane(optr, 1);
  double sx = 0.0, sy = 0.0;
  cplx ap = cmake(0.0, 0.0);  // sum over the item's rows of conj(x) W (apart)
  // outputs of row rr (uniform, rr < nr)
  auto flush = [&](int rr) {
    const int64_t s = s0 + rr;
    const double d = readlane_d(dv, rr);
    if (active) {
      cplx x0;
      if (PRE) {
        x0 = xo[0];
#pragma unroll
        for (int q = 1; q < (PRE ? RB : 1); ++q)
          if (rr == q) x0 = xo[q];
      } else {
        x0 = load_global(src + s * lds + t);
      }
      const cplx x = cmake(x0.x * ft, x0.y * ft);
      const cplx dg = cmake(al.x * d + be.x, al.y * d + be.y);
      const cplx dx = cmul(dg, x);
      if (Vout) store_nt(Vout + S.vo + s * m + t, x);
      const cplx wv = cmake(sx * ft + dx.x, sy * ft + dx.y);
      store_nt(Wout + S.vo + s * m + t, wv);
      ap = cadd(ap, cdotmul(x, wv));
    }
    sx = 0.0;
    sy = 0.0;
  };
  for (int base = E0; base < E1; base += TT) {
    const int cnt = min(TT, E1 - base);
    const int ci = (lane < cnt) ? S.inner[base + lane] : 0;
    const cplx cv = (lane < cnt) ? S.vals[base + lane] : cmake(0.0, 0.0);
    for (int k0 = 0; k0 < cnt; k0 += 8) {
      cplx x[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int col = __builtin_amdgcn_readlane(ci, min(k0 + i, cnt - 1));
        x[i] = load_global(src + (int64_t)col * lds + tl);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = k0 + i;
        if (k < cnt) {
          while (base + k >= rend) {  // rows that end here (empty rows included)
            flush(r);
            ++r;
            rend = __builtin_amdgcn_readlane(optr, r + 1);
          }
          const double vx = readlane_d(cv.x, k);
          const double vy = readlane_d(cv.y, k);
          sx += vx * x[i].x + vy * x[i].y;
          sy += vx * x[i].y - vy * x[i].x;
        }
      }
    }
  }
  for (; r < nr; ++r) flush(r);
  if (apart) apart[b * TT + lane] = ap;
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
  const bool act = t < m;
  if (tl.rstep == 1 && !act) return;  // no cross-lane sums in one-slot-per-lane tiles
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = act ? min<int64_t>(S.n, r0 + S.rpc) : r0;
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
  for (int64_t s = r0 + tl.sub; s < r1; s += tl.rstep) {
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
  if (tl.rstep > 1) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) ac[q] = sub_sum(ac[q], S.mp);
    if (NR == 2) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) ad[q] = sub_sum(ad[q], S.mp);
    }
    uu = sub_sum(uu, S.mp);
    uw = sub_sum(uw, S.mp);
  }
  if (!act || tl.sub != 0) return;
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

// g_red_kernel with the basis split over QW waves of one block: all QW waves
// walk the same rows and slots of a tile, wave w takes the vectors q = w,
// w + QW, ... (its accumulators only).  Rows go in rounds of QW: wave w
// loads u and w of row w of the round into LDS (double-buffered, one
// barrier per round), and every wave reads them from there, so each line of
// u and w is requested from L2 once per block.  The last wave also sums
// |u|^2 and u^H w.  Per-lane registers hold ceil(NQ / QW) pairs of
// accumulators instead of NQ, so more waves stay resident to hide the
// loads.  Same partials, same layout, rows added in the same order.
// Every wave of a block runs the same number of rounds (the loop bounds do
// not depend on the lane), so the barriers are uniform.
constexpr int QW = 4;
typedef double mr_dbl2 __attribute__((ext_vector_type(2)));
__device__ inline cplx nt_load(const cplx* p) {
  const mr_dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const mr_dbl2*>(p));
  return cmake(v.x, v.y);
}
template <int NQ, int NR>
__global__ void __launch_bounds__(QW * TT)
    g_red_split_kernel(const GSec* __restrict__ gs, int ns, const cplx* __restrict__ V, int64_t VS,
                       const cplx* __restrict__ U, const cplx* __restrict__ W, cplx* __restrict__ part) {
  constexpr int P = NR * NQ + 2;
  constexpr int NQW = (NQ + QW - 1) / QW;
  __shared__ cplx su[2][QW][TT], sw[2][QW][TT];
  const int wv = threadIdx.x / TT, lane = threadIdx.x % TT;
  const Tile tl = tile_of(gs, ns, lane);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  const bool act = t < m;
  const bool sums = wv == QW - 1;
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = min<int64_t>(S.n, r0 + S.rpc);  // uniform over the block
  const cplx* __restrict__ Vb = V + S.vo;
  const cplx* __restrict__ Ub = U + S.vo;
  const cplx* __restrict__ Wb = W + S.vo;
  cplx ac[NQW], ad[NR == 2 ? NQW : 1];
  cplx uu = cmake(0.0, 0.0), uw = cmake(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < NQW; ++i) ac[i] = cmake(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < (NR == 2 ? NQW : 1); ++i) ad[i] = cmake(0.0, 0.0);
  const int64_t round = (int64_t)QW * tl.rstep;
  int buf = 0;
  for (int64_t sb = r0; sb < r1; sb += round, buf ^= 1) {
    {
      const int64_t s = sb + tl.sub + (int64_t)wv * tl.rstep;
      cplx u = cmake(0.0, 0.0), w = cmake(0.0, 0.0);
      if (act && s < r1) {
        u = Ub[s * m + t];
        if (NR == 2) w = Wb[s * m + t];
      }
      su[buf][wv][lane] = u;
      if (NR == 2) sw[buf][wv][lane] = w;
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < QW; ++rr) {
      const int64_t s = sb + tl.sub + (int64_t)rr * tl.rstep;
      if (!act || s >= r1) continue;
      const int64_t o = s * m + t;
      const cplx u = su[buf][rr][lane];
      const cplx w = (NR == 2) ? sw[buf][rr][lane] : cmake(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < NQW; ++i) {
        const int q = wv + QW * i;
        if (q < NQ) {
          const cplx v = nt_load(Vb + q * VS + o);
          ac[i] = cadd(ac[i], cdotmul(v, u));
          if (NR == 2) ad[i] = cadd(ad[i], cdotmul(v, w));
        }
      }
      if (sums) {
        uu = cadd(uu, cdotmul(u, u));
        if (NR == 2) uw = cadd(uw, cdotmul(u, w));
      }
    }
  }
  if (tl.rstep > 1) {
#pragma unroll
    for (int i = 0; i < NQW; ++i) {
      ac[i] = sub_sum(ac[i], S.mp);
      if (NR == 2) ad[i] = sub_sum(ad[i], S.mp);
    }
    uu = sub_sum(uu, S.mp);
    uw = sub_sum(uw, S.mp);
  }
  if (!act || tl.sub != 0) return;
  cplx* pb = part + P * S.co + (int64_t)tl.chunk * P * m + t;
#pragma unroll
  for (int i = 0; i < NQW; ++i) {
    const int q = wv + QW * i;
    if (q < NQ) {
      pb[(int64_t)q * m] = ac[i];
      if (NR == 2) pb[(int64_t)(NQ + q) * m] = ad[i];
    }
  }
  if (sums) {
    pb[(int64_t)(P - 2) * m] = uu;
    pb[(int64_t)(P - 1) * m] = uw;
  }
}

// red[p Mtot + slot] = sum over chunks (fixed order) of the partials, P per slot.
// One block per (p, 64-slot group).
__global__ void __launch_bounds__(RW * TT)
    g_reduce_p_kernel(const GSec* __restrict__ gs, int ns, int64_t ngroups, int Mtot, int P,
                      const cplx* __restrict__ part, cplx* __restrict__ red) {
  __shared__ cplx sh[RW][TT];
  const int p = (int)(blockIdx.x / ngroups);
  const int64_t gi = blockIdx.x % ngroups;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  cplx s = cmake(0.0, 0.0);
  if (t < S.m) {
    const cplx* pb = part + (int64_t)P * S.co;
#pragma unroll 4
    for (int c = wv; c < S.nch; c += RW) s = cadd(s, pb[((int64_t)c * P + p) * S.m + t]);
  }
  sh[wv][lane] = s;
  __syncthreads();
  if (wv != 0 || t >= S.m) return;
  s = sh[0][lane];
  for (int w = 1; w < RW; ++w) s = cadd(s, sh[w][lane]);
  red[(int64_t)p * Mtot + S.so + t] = s;
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
  for (int64_t s = r0 + tl.sub; s < r1; s += tl.rstep) {
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

// g_dc_update_kernel with the basis split over QW waves: wave w sums
// v_q a_q and v_q b_q over its vectors q = w, w + QW, ... for QW rows at a
// time and parks the two partial sums in LDS; after a barrier, wave w
// finishes row w of the round (partials of waves 0..QW-1 added in that
// order, then the stores of q_j and the next pending vector).  The
// coefficients of a lane's vectors stay in registers (ceil(NQ / QW) pairs).
// Every wave of a block runs the same number of rounds (the loop bounds do
// not depend on the lane), so the barriers are uniform.
template <int NQ>
__global__ void __launch_bounds__(QW * TT)
    g_dc_update_split_kernel(const GSec* __restrict__ gs, int ns, cplx* __restrict__ V, int64_t VS, int j,
                             cplx* __restrict__ U, const cplx* __restrict__ W, const cplx* __restrict__ coef,
                             const cplx* __restrict__ sc, int Mtot, int K) {
  constexpr int NQW = (NQ + QW - 1) / QW;
  __shared__ cplx sh[QW][2][QW][TT];  // [wave][s1, s2][row of the round][lane]
  const int wv = threadIdx.x / TT, lane = threadIdx.x % TT;
  const Tile tl = tile_of(gs, ns, lane);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  const bool act = t < m;
  const int g = S.so + (act ? t : 0);
  cplx a[NQW], b[NQW];
#pragma unroll
  for (int i = 0; i < NQW; ++i) {
    const int q = wv + QW * i;
    a[i] = (act && q < NQ) ? coef[(int64_t)q * Mtot + g] : cmake(0.0, 0.0);
    b[i] = (act && q < NQ) ? coef[(int64_t)(K + q) * Mtot + g] : cmake(0.0, 0.0);
  }
  const double inv = act ? sc[2 * g].x : 0.0;
  const cplx f = act ? sc[2 * g + 1] : cmake(0.0, 0.0);
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = min<int64_t>(S.n, r0 + S.rpc);  // uniform over the block
  cplx* __restrict__ Vb = V + S.vo;
  cplx* __restrict__ Ub = U + S.vo;
  const cplx* __restrict__ Wb = W + S.vo;
  const int64_t round = (int64_t)QW * tl.rstep;
  for (int64_t sb = r0; sb < r1; sb += round) {
    // this wave's row of the round: u and w requested before the basis loads
    const int64_t sf = sb + tl.sub + (int64_t)wv * tl.rstep;  // this wave's row of the round
    const bool fin = act && sf < r1;
#pragma unroll
    for (int rr = 0; rr < QW; ++rr) {
      const int64_t s = sb + tl.sub + (int64_t)rr * tl.rstep;
      cplx s1 = cmake(0.0, 0.0), s2 = cmake(0.0, 0.0);
      if (act && s < r1) {
        const int64_t o = s * m + t;
#pragma unroll
        for (int i = 0; i < NQW; ++i) {
          const int q = wv + QW * i;
          if (q < NQ) {
            const cplx v = nt_load(Vb + q * VS + o);
            s1 = cadd(s1, cmul(v, a[i]));
            s2 = cadd(s2, cmul(v, b[i]));
          }
        }
      }
      sh[wv][0][rr][lane] = s1;
      sh[wv][1][rr][lane] = s2;
    }
    __syncthreads();
    if (fin) {
      const int rr = wv;
      cplx s1 = sh[0][0][rr][lane], s2 = sh[0][1][rr][lane];
#pragma unroll
      for (int w2 = 1; w2 < QW; ++w2) {
        s1 = cadd(s1, sh[w2][0][rr][lane]);
        s2 = cadd(s2, sh[w2][1][rr][lane]);
      }
      {
        const int64_t o = sf * m + t;
        const cplx u_row = Ub[o];
        const cplx w_row = Wb[o];
        const cplx qj = crmul(inv, csub(u_row, s1));
        __builtin_nontemporal_store(qj.x, &Vb[(int64_t)j * VS + o].x);
        __builtin_nontemporal_store(qj.y, &Vb[(int64_t)j * VS + o].y);
        const cplx uf = cmul(u_row, f);
        Ub[o] = cmake(w_row.x * inv - s2.x - uf.x, w_row.y * inv - s2.y - uf.y);
      }
    }
    __syncthreads();
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

// X[s,t] = sum_j V_j[s,t] coef[j Mtot + slot], all sectors (one thread per
// element of the pools); nothing is written when skip is given and nonzero
// (a combination enqueued before the host has read a flag that voids it)
__global__ void g_combine_kernel(const GSec* __restrict__ gs, int ns, int64_t total, const cplx* __restrict__ V,
                                 int64_t VS, int k, const cplx* __restrict__ coef, int Mtot,
                                 const int* __restrict__ skip) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total || (skip && *skip)) return;
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
  const bool act = t < m;
  if (tl.rstep == 1 && !act) return;
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = act ? min<int64_t>(S.n, r0 + S.rpc) : r0;
  double acc[MAXM + 1];
#pragma unroll
  for (int q = 0; q <= MAXM; ++q) acc[q] = 0.0;
  for (int64_t s = r0 + tl.sub; s < r1; s += tl.rstep) {
    const cplx v = S.X[s * S.ldx + t];
    const double p = v.x * v.x + v.y * v.y;
    const uint64_t x = S.states[s];
#pragma unroll
    for (int q = 0; q < MAXM; ++q)
      if ((x >> q) & 1ull) acc[q] += p;
    acc[MAXM] += p;
  }
  if (tl.rstep > 1) {
#pragma unroll
    for (int q = 0; q < MAXM; ++q)
      if (q < M) acc[q] = sub_sum(acc[q], S.mp);
    acc[MAXM] = sub_sum(acc[MAXM], S.mp);
  }
  if (!act || tl.sub != 0) return;
  double* pb = part + (int64_t)(M + 1) * S.co;
#pragma unroll
  for (int q = 0; q < MAXM; ++q)
    if (q < M) pb[((int64_t)tl.chunk * (M + 1) + q) * m + t] = acc[q];
  pb[((int64_t)tl.chunk * (M + 1) + M) * m + t] = acc[MAXM];
}

// dens_all[traj (M+1) + q] = density q / norm^2 (q < M), norm^2 (q = M).
// One block per (q, 64-slot group), chunks summed as in g_reduce_kernel.
__global__ void __launch_bounds__(RW * TT)
    g_dens_reduce_kernel(const GSec* __restrict__ gs, int ns, int64_t ngroups, int M, const double* __restrict__ part,
                         const int* __restrict__ slot2traj, double* __restrict__ dens_all) {
  __shared__ double sh[2][RW][TT];
  const int q = (int)(blockIdx.x / ngroups);
  const int64_t gi = blockIdx.x % ngroups;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  double v = 0.0, nn = 0.0;
  if (t < S.m) {
    const double* pb = part + (int64_t)(M + 1) * S.co;
#pragma unroll 4
    for (int c = wv; c < S.nch; c += RW) {
      v += pb[((int64_t)c * (M + 1) + q) * S.m + t];
      if (q < M) nn += pb[((int64_t)c * (M + 1) + M) * S.m + t];
    }
  }
  sh[0][wv][lane] = v;
  sh[1][wv][lane] = nn;
  __syncthreads();
  if (wv != 0 || t >= S.m) return;
  v = sh[0][0][lane];
  nn = sh[1][0][lane];
  for (int w = 1; w < RW; ++w) {
    v += sh[0][w][lane];
    nn += sh[1][w][lane];
  }
  dens_all[(int64_t)slot2traj[S.so + t] * (M + 1) + q] = (q < M) ? v / nn : v;
}

// basis sizes from which the reduction and update passes split the basis
// over QW waves (below it the one-wave kernels hold few registers anyway)
constexpr int QD_SPLIT_MIN = 4;
template <int NR>
void launch_red(int nq, unsigned tiles, hipStream_t st, const GSec* gs, int ns, const cplx* V, int64_t VS,
                const cplx* U, const cplx* W, cplx* part) {
  switch (nq) {
#define QD_RED_CASE(N)                                                                  \
  case N:                                                                               \
    if (N >= QD_SPLIT_MIN)                                                              \
      g_red_split_kernel<N, NR><<<tiles, QW * TT, 0, st>>>(gs, ns, V, VS, U, W, part);  \
    else                                                                                \
      g_red_kernel<N, NR><<<tiles, TT, 0, st>>>(gs, ns, V, VS, U, W, part);             \
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
#define QD_UPD_CASE(N)                                                                                   \
  case N:                                                                                                \
    if (N >= QD_SPLIT_MIN)                                                                               \
      g_dc_update_split_kernel<N><<<tiles, QW * TT, 0, st>>>(gs, ns, V, VS, j, U, W, coef, sc, Mtot, K);  \
    else                                                                                                 \
      g_dc_update_kernel<N><<<tiles, TT, 0, st>>>(gs, ns, V, VS, j, U, W, coef, sc, Mtot, K);            \
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
constexpr int QD_TRAJ_MAX_CHUNKS = 2048;
int rows_per_chunk(int64_t n, int groups, int target) {
  const int64_t want = std::max<int64_t>(1, target / std::max(1, groups));
  int64_t rpc = std::max<int64_t>(16, (n + want - 1) / want);
  rpc = std::min<int64_t>(rpc, 1024);
  rpc = std::max<int64_t>(rpc, (n + QD_TRAJ_MAX_CHUNKS - 1) / QD_TRAJ_MAX_CHUNKS);
  return (int)rpc;
}


// per-slot a-posteriori error estimate of the step (Saad 1992, section 5):
// est = h_{K+1,K} |y_K| with y = expm(-i dt K_K) e_1 (coef) and h_{K+1,K} the
// norm the finalisation left in Hess[K][K-1]
__global__ void g_step_estimate_kernel(int Mtot, int K, const cplx* __restrict__ hess, const cplx* __restrict__ coef,
                                       double* __restrict__ est) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const cplx h = hess[((int64_t)(K - 1) * (K + 1) + K) * Mtot + g];
  const cplx y = coef[(int64_t)(K - 1) * Mtot + g];
  est[g] = sqrt(h.x * h.x + h.y * h.y) * sqrt(y.x * y.x + y.y * y.y);
}

// Work buffers of one batched Arnoldi step, sized for the largest batch.
struct MrhsBufs {
  cplx *V, *U, *W;      // K vector pools (VS each), pending vector, product
  cplx* part;           // chunk partials
  cplx *h, *hess;       // first projection; Hessenbergs [(col (K+1) + row) Mtot + slot]
  cplx *red, *coef2;    // reduced partials; update coefficients
  cplx *sc, *coef;      // per-slot scalars; y = expm(...) e_1 [q Mtot + slot]
};

// The multi-RHS product of mrhs_arnoldi: (st, from_x, Wsrc, f, alpha, beta,
// Vout, Wout) with g_spmm_kernel's meaning.  This one is g_spmm_kernel on the
// sectors' stored matrices.
struct SpmmProduct {
  const GSec* gs;
  int ns;
  int64_t blocks;
  void operator()(hipStream_t st, int from_x, const cplx* Wsrc, const double* f, const cplx* alpha,
                  const cplx* beta, cplx* Vout, cplx* Wout) const {
    const unsigned sgrid = (unsigned)((blocks + 7) / 8 * 8);
    g_spmm_kernel<<<sgrid, TT, 0, st>>>(gs, ns, blocks, from_x, Wsrc, f, alpha, beta, Vout, Wout, nullptr);
  }
};

// The same product by one of three kernels, chosen by timing: 0
// g_spmm_kernel, 1 g_spmm_flat_kernel<4, true>, 2 g_spmm_flat_kernel<8,
// false>; blocks counts ceil(n / rows) * groups items per sector, as the
// sectors' blk0 do.  apart (or NULL): the kernels' per-item sums of conj(x)
// W, blocks x 64 values.  (Products over 32- and 16-slot row segments and
// 16-row streams were slower on the L = 8 trajectory sectors.)
constexpr int SPMM_VARIANTS = 3;
inline int spmm_rows(int var) { return var == 2 ? 8 : SPMM_RB; }
struct SpmmVarProduct {
  const GSec* gs;
  int ns;
  int64_t blocks;
  int var;
  cplx* apart;
  void operator()(hipStream_t st, int from_x, const cplx* Wsrc, const double* f, const cplx* alpha,
                  const cplx* beta, cplx* Vout, cplx* Wout) const {
    const unsigned sgrid = (unsigned)((blocks + 7) / 8 * 8);
    switch (var) {
      case 1:
        g_spmm_flat_kernel<4, true>
            <<<sgrid, TT, 0, st>>>(gs, ns, blocks, from_x, Wsrc, f, alpha, beta, Vout, Wout, apart);
        break;
      case 2:
        g_spmm_flat_kernel<8, false>
            <<<sgrid, TT, 0, st>>>(gs, ns, blocks, from_x, Wsrc, f, alpha, beta, Vout, Wout, apart);
        break;
      default:
        g_spmm_kernel<<<sgrid, TT, 0, st>>>(gs, ns, blocks, from_x, Wsrc, f, alpha, beta, Vout, Wout, apart);
    }
  }
};

// K Arnoldi vectors (delayed CGS2) for every slot: V_0 = f X (f = per-slot
// 1/|x|), H_eff = H + alpha D + beta per slot, then y = expm(-i dt K) e_1.
// X is not touched; mrhs_combine writes V y into it.
template <class Prod>
inline void mrhs_arnoldi_with(hipStream_t st, const GSec* d_gsec, int ns, int64_t tiles, int64_t ngroups, int64_t VS,
                              int Mtot, int K, double dt, const double* f, const cplx* alpha, const cplx* beta,
                              const MrhsBufs& b, const Prod& prod) {
  const unsigned tl = (unsigned)tiles;
  cplx* U = b.U;
  cplx* W = b.W;
  // V_0 = f X, U = H_eff V_0, then its first projection: H[0][0] = V_0^H U, U -= V_0 H[0][0]
  prod(st, 1, nullptr, f, alpha, beta, b.V, U);
  g_pass_kernel<1, 0><<<tl, TT, 0, st>>>(d_gsec, ns, b.V, VS, U, nullptr, Mtot, b.part);
  g_reduce_kernel<<<(unsigned)ngroups, RW * TT, 0, st>>>(d_gsec, ns, ngroups, Mtot, 1, b.part, b.h, b.hess, 0, K,
                                                          0, nullptr);
  g_pass_kernel<1, 2><<<tl, TT, 0, st>>>(d_gsec, ns, b.V, VS, U, b.h, Mtot, b.part);
  for (int j = 1; j < K; ++j) {
    const int P = 2 * j + 2;
    prod(st, 0, U, nullptr, alpha, beta, nullptr, W);
    launch_red<2>(j, tl, st, d_gsec, ns, b.V, VS, U, W, b.part);
    g_reduce_p_kernel<<<(unsigned)(ngroups * P), RW * TT, 0, st>>>(d_gsec, ns, ngroups, Mtot, P, b.part, b.red);
    g_dc_small_kernel<<<g1(Mtot, 256), 256, 0, st>>>(Mtot, j, K, 0, b.red, b.hess, b.coef2, b.sc, 1e-10);
    launch_dc_update(j, tl, st, d_gsec, ns, b.V, VS, j, U, W, b.coef2, b.sc, Mtot, K);
  }
  // the pending vector's second projection completes column K-1 (and h_{K+1,K})
  launch_red<1>(K, tl, st, d_gsec, ns, b.V, VS, U, nullptr, b.part);
  g_reduce_p_kernel<<<(unsigned)(ngroups * (K + 2)), RW * TT, 0, st>>>(d_gsec, ns, ngroups, Mtot, K + 2, b.part,
                                                                         b.red);
  g_dc_small_kernel<<<g1(Mtot, 256), 256, 0, st>>>(Mtot, K, K, 1, b.red, b.hess, b.coef2, b.sc, 1e-10);
  small_expm_kernel<<<(unsigned)Mtot, 64, 0, st>>>(Mtot, K, K, b.hess, dt, b.coef);
}

inline void mrhs_arnoldi(hipStream_t st, const GSec* d_gsec, int ns, int64_t tiles, int64_t blocks, int64_t ngroups,
                         int64_t VS, int Mtot, int K, double dt, const double* f, const cplx* alpha, const cplx* beta,
                         const MrhsBufs& b) {
  mrhs_arnoldi_with(st, d_gsec, ns, tiles, ngroups, VS, Mtot, K, dt, f, alpha, beta, b, SpmmProduct{d_gsec, ns, blocks});
}

// ---------------------------------------------------------- Lanczos form
// When H is Hermitian and alpha_t D + beta_t is one number c_t on every
// sector (each mode carries the same total jump rate r, so sum_c L_c^+ L_c =
// r N is constant on a sector), H_eff = H + c_t there.  Its Arnoldi process
// then reduces to the three-term recurrence: in exact arithmetic the
// Hessenberg is tridiagonal, h_jj = alpha_j (Paige's form: v_j^H (H_eff v_j
// - beta_j v_{j-1}), which includes c_t), h_{j+1,j} = h_{j,j+1} = beta_{j+1},
// and the Krylov vectors are the same.  It is the recurrence with which the
// reference propagates Hermitian operators (Hermitian::LanczosSolver::evolve,
// sparse_algorithm.hpp:261-285), here on H with c_t added to the diagonal.
// Per Krylov vector: the product (which can also write the alpha sums), one
// update pass; no pass over the basis.
// The vectors stay unnormalised in the pool (u_0 = v_0, u_j = beta_j v_j),
// so nothing is rewritten: the product takes 1/beta_j as its source factor,
// the update and the final combination fold it into their coefficients.
// Finite-precision Lanczos vectors lose orthogonality as Ritz values
// converge, but the projected exponential stays accurate (Druskin,
// Greenbaum & Knizhnerman, SIAM J. Sci. Comput. 19 (1998) 38-54).
// Semi-orthogonality is monitored per slot with Simon's estimate of the loss
// of orthogonality (Simon, Math. Comp. 42 (1984) 115-142), the recurrence
// the single-vector Lanczos with partial reorthogonalisation uses:
//   beta_{j+1} w_{j+1,k} = beta_{k+1} w_{j,k+1} + (alpha_k - alpha_j) w_{j,k}
//                          + beta_k w_{j,k-1} - beta_j w_{j-1,k} +- eps1 |T|,
// w_{j+1,j} = eps1 (the recurrence's local orthogonality), w_{j+1,j+1} = 1,
// eps1 = sqrt(n) eps / 2.  A slot whose max_k |w_{j+1,k}| passes `trigger`
// sets *flag; the caller then recomputes the step fully reorthogonalised.
struct LzState {
  double* beta;  // [j Mtot + slot] beta_j = |u_j|
  double* inv;   // 1 / beta_j (inv_0 = 1; 0 once the Krylov space has ended)
  cplx* ip;      // <v_j, v_{j-1}>
  cplx* ca;      // update coefficients of the current step: alpha_j / beta_j
  double* cb;    //                                          beta_j / beta_{j-1}
  double* om;    // three rotating rows of the estimate: [(row (K + 1) + k) Mtot + slot]
  double* tn;    // [slot] running estimate of |T|
  int* flag;     // set when a slot's estimate passes trigger
  double trigger;
};

// u_{j+1} = W - ca_t u_j - cb_t u_{j-1} (no u_{j-1} term for j = 0) into pool
// slot j + 1, with part[2 co + (chunk 2 + p) m + t]: p = 0: |u_{j+1}|^2,
// p = 1: conj(u_j) u_{j+1}.  Tiles as in g_pass_kernel.
__global__ void __launch_bounds__(TT)
    g_lz_update_kernel(const GSec* __restrict__ gs, int ns, cplx* __restrict__ V, int64_t VS, int j,
                       const cplx* __restrict__ W, const cplx* __restrict__ ca, const double* __restrict__ cb,
                       cplx* __restrict__ part) {
  const Tile tl = tile_of(gs, ns);
  const GSec& S = gs[tl.sec];
  const int t = tl.t;
  const int m = S.m;
  const bool act = t < m;
  if (tl.rstep == 1 && !act) return;
  const int64_t r0 = (int64_t)tl.chunk * S.rpc;
  const int64_t r1 = act ? min<int64_t>(S.n, r0 + S.rpc) : r0;
  const cplx* __restrict__ Uj = V + (int64_t)j * VS + S.vo;
  const cplx* __restrict__ Um = V + (int64_t)(j > 0 ? j - 1 : 0) * VS + S.vo;
  cplx* __restrict__ Un = V + (int64_t)(j + 1) * VS + S.vo;
  const cplx* __restrict__ Wb = W + S.vo;
  const cplx a = act ? ca[S.so + t] : cmake(0.0, 0.0);
  const double b = (act && j > 0) ? cb[S.so + t] : 0.0;
  double n2 = 0.0;
  cplx c = cmake(0.0, 0.0);
#pragma unroll 2
  for (int64_t s = r0 + tl.sub; s < r1; s += tl.rstep) {
    const int64_t o = s * m + t;
    const cplx w = Wb[o];
    const cplx u = Uj[o];
    const cplx um = j > 0 ? Um[o] : cmake(0.0, 0.0);
    const cplx un = cmake(w.x - (a.x * u.x - a.y * u.y) - b * um.x, w.y - (a.x * u.y + a.y * u.x) - b * um.y);
    Un[o] = un;
    n2 += un.x * un.x + un.y * un.y;
    c = cadd(c, cdotmul(u, un));
  }
  if (tl.rstep > 1) {
    n2 = sub_sum(n2, S.mp);
    c = sub_sum(c, S.mp);
  }
  if (!act || tl.sub != 0) return;
  cplx* pb = part + 2 * S.co;
  pb[((int64_t)tl.chunk * 2 + 0) * m + t] = cmake(n2, 0.0);
  pb[((int64_t)tl.chunk * 2 + 1) * m + t] = c;
}

// sum over the chunks of value q (of NP per chunk) for slot t of sector S, by
// the RW waves of the block (same order as g_reduce_kernel); valid in wave 0
__device__ inline cplx lz_chunk_sum(const GSec& S, const cplx* __restrict__ part, int NP, int q, int t, int wv,
                                    int lane, cplx (*sh)[TT]) {
  cplx s = cmake(0.0, 0.0);
  if (t < S.m) {
    const cplx* pb = part + (int64_t)NP * S.co;
#pragma unroll 4
    for (int c = wv; c < S.nch; c += RW) s = cadd(s, pb[((int64_t)c * NP + q) * S.m + t]);
  }
  __syncthreads();
  sh[wv][lane] = s;
  __syncthreads();
  s = sh[0][lane];
  for (int w = 1; w < RW; ++w) s = cadd(s, sh[w][lane]);
  return s;
}

// Lanczos step j for slot g from s = <v_j, W> (= inv_j <u_j, W>): alpha_j =
// s - beta_j <v_j, v_{j-1}> -> Hess[j][j]; ca = alpha_j inv_j, cb = beta_j
// inv_{j-1} (0 for j = 0 and once the space has ended)
__device__ inline void lz_alpha_finish(cplx s, int g, int Mtot, int j, int K, cplx* __restrict__ hess,
                                       const LzState& lz) {
  double iv = 1.0, cbv = 0.0;
  cplx al = s;
  if (j == 0) {
    lz.inv[g] = 1.0;
    lz.om[g] = 1.0;  // w_{0,0} (row 0)
    lz.tn[g] = 0.0;
  } else {
    iv = lz.inv[(int64_t)j * Mtot + g];
    const double bj = lz.beta[(int64_t)j * Mtot + g];
    const cplx ip = lz.ip[(int64_t)j * Mtot + g];
    al = cmake(s.x - bj * ip.x, s.y - bj * ip.y);
    cbv = iv > 0.0 ? bj * lz.inv[(int64_t)(j - 1) * Mtot + g] : 0.0;
  }
  hess[((int64_t)(j * (K + 1) + j)) * Mtot + g] = al;
  lz.ca[g] = crmul(iv, al);
  lz.cb[g] = cbv;
}

// Lanczos step j after the pass over (u_j, W): s = inv_j <u_j, W>, then
// lz_alpha_finish.  One block per 64-slot group.
__global__ void __launch_bounds__(RW * TT)
    g_lz_alpha_kernel(const GSec* __restrict__ gs, int ns, int Mtot, const cplx* __restrict__ part, int j, int K,
                      cplx* __restrict__ hess, LzState lz) {
  __shared__ cplx sh[RW][TT];
  const int64_t gi = blockIdx.x;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  const cplx s = lz_chunk_sum(S, part, 1, 0, t, wv, lane, sh);
  if (wv != 0 || t >= S.m) return;
  const int g = S.so + t;
  const double iv = j == 0 ? 1.0 : lz.inv[(int64_t)j * Mtot + g];
  lz_alpha_finish(crmul(iv, s), g, Mtot, j, K, hess, lz);
}

// The same from the product's per-item sums of conj(x) W (x = v_j, the
// product's normalised source): s = the sum over the items of the slot's
// 64-slot group, item order, RW waves as in lz_chunk_sum.  Items of `rows`
// rows, group-major from the sector's blk0.  One block per 64-slot group.
__global__ void __launch_bounds__(RW * TT)
    g_lz_alpha_sp_kernel(const GSec* __restrict__ gs, int ns, int Mtot, const cplx* __restrict__ apart, int rows,
                         int j, int K, cplx* __restrict__ hess, LzState lz) {
  __shared__ cplx sh[RW][TT];
  const int64_t gi = blockIdx.x;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int gr = (int)(gi - S.grp0);
  const int t = gr * TT + lane;
  const int64_t nrb = (S.n + rows - 1) / rows;
  const cplx* __restrict__ pb = apart + (S.blk0 + (int64_t)gr * nrb) * TT + lane;
  // 8 loads in flight per wave, added in item order (rb, rb + RW, ...)
  cplx s = cmake(0.0, 0.0);
  for (int64_t rb = wv; rb < nrb; rb += 8 * RW) {
    cplx v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = rb + RW * i < nrb ? pb[(rb + RW * i) * TT] : cmake(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < 8; ++i) s = cadd(s, v[i]);
  }
  sh[wv][lane] = s;
  __syncthreads();
  if (wv != 0 || t >= S.m) return;
  s = sh[0][lane];
  for (int w = 1; w < RW; ++w) s = cadd(s, sh[w][lane]);
  lz_alpha_finish(s, S.so + t, Mtot, j, K, hess, lz);
}

// after the update of step j: beta_{j+1} = |u_{j+1}|, inv_{j+1} (0 below
// 1e-10 or after an earlier end, as the Arnoldi form), <v_{j+1}, v_j>;
// Hess[j+1][j] = beta_{j+1}, Hess[j][j+1] = beta_{j+1} (0 if the space ended).
__global__ void __launch_bounds__(RW * TT)
    g_lz_beta_kernel(const GSec* __restrict__ gs, int ns, int Mtot, const cplx* __restrict__ part, int j, int K,
                     cplx* __restrict__ hess, LzState lz) {
  __shared__ cplx sh[RW][TT];
  const int64_t gi = blockIdx.x;
  const int si = find_sec(ns, gi, [&](int i) { return gs[i].grp0; });
  const GSec& S = gs[si];
  const int lane = threadIdx.x & (TT - 1), wv = threadIdx.x / TT;
  const int t = (int)(gi - S.grp0) * TT + lane;
  const cplx n2 = lz_chunk_sum(S, part, 2, 0, t, wv, lane, sh);
  const cplx c = lz_chunk_sum(S, part, 2, 1, t, wv, lane, sh);
  if (wv != 0 || t >= S.m) return;
  const int g = S.so + t;
  const double b = sqrt(n2.x);
  const double ivj = lz.inv[(int64_t)j * Mtot + g];
  const double iv = (b >= 1e-10 && ivj > 0.0) ? 1.0 / b : 0.0;
  lz.beta[(int64_t)(j + 1) * Mtot + g] = b;
  lz.inv[(int64_t)(j + 1) * Mtot + g] = iv;
  lz.ip[(int64_t)(j + 1) * Mtot + g] = cmake(c.x * iv * ivj, -c.y * iv * ivj);
  hess[((int64_t)(j * (K + 1) + j + 1)) * Mtot + g] = cmake(b, 0.0);
  hess[((int64_t)((j + 1) * (K + 1) + j)) * Mtot + g] = cmake(iv > 0.0 ? b : 0.0, 0.0);
  if (iv == 0.0) return;  // the space has ended: no further vector to estimate
  // Simon's estimate of row j + 1 from rows j and j - 1 (alpha_k - alpha_j:
  // the real parts; the slot's shift c_t is imaginary and cancels)
  const int64_t ld = (int64_t)(K + 1) * Mtot;
  double* rn = lz.om + (int64_t)((j + 1) % 3) * ld + g;  // row j + 1
  const double* r1 = lz.om + (int64_t)(j % 3) * ld + g;  // row j
  const double* r2 = lz.om + (int64_t)((j + 2) % 3) * ld + g;  // row j - 1
  const double aj = hess[((int64_t)(j * (K + 1) + j)) * Mtot + g].x;
  const double bj = j > 0 ? lz.beta[(int64_t)j * Mtot + g] : 0.0;
  const double T = fmax(lz.tn[g], fabs(aj) + bj + b);
  lz.tn[g] = T;
  const double eps1 = 0.5 * sqrt((double)S.n) * 2.220446049250313e-16;
  double mx = 0.0;
  for (int k = 0; k < j; ++k) {
    const double ak = hess[((int64_t)(k * (K + 1) + k)) * Mtot + g].x;
    const double bk1 = lz.beta[(int64_t)(k + 1) * Mtot + g];
    double tt = bk1 * r1[(int64_t)(k + 1) * Mtot] + (ak - aj) * r1[(int64_t)k * Mtot] - bj * r2[(int64_t)k * Mtot];
    if (k > 0) tt += lz.beta[(int64_t)k * Mtot + g] * r1[(int64_t)(k - 1) * Mtot];
    const double v = (tt + copysign(eps1 * T, tt)) / b;
    rn[(int64_t)k * Mtot] = v;
    mx = fmax(mx, fabs(v));
  }
  rn[(int64_t)j * Mtot] = eps1;
  rn[(int64_t)(j + 1) * Mtot] = 1.0;
  if (mx > lz.trigger) atomicOr(lz.flag, 1);
}

// coef[j Mtot + slot] *= inv_j: coefficients of the unnormalised pool vectors
__global__ void g_lz_scale_kernel(int Mtot, int K, const double* __restrict__ inv, cplx* __restrict__ coef) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)K * Mtot) return;
  coef[e] = crmul(inv[e], coef[e]);
}

// K Lanczos vectors for every slot (the Lanczos form above): V_0 = f X, then
// y = expm(-i dt T_t) e_1 scaled for the unnormalised vectors, so that
// mrhs_combine writes V y into X as after mrhs_arnoldi.  Hess must hold zeros
// off the three diagonals (only those are written here).
// apart (or NULL): the product writes per-item sums of conj(x) W there
// (items of apart_rows rows), which give alpha_j without a pass over (u_j,
// W); otherwise that pass runs.
template <class Prod>
inline void mrhs_lanczos_with(hipStream_t st, const GSec* d_gsec, int ns, int64_t tiles, int64_t ngroups,
                              int64_t VS, int Mtot, int K, double dt, const double* f, const cplx* alpha,
                              const cplx* beta, const MrhsBufs& b, const LzState& lz, const Prod& prod,
                              const cplx* apart = nullptr, int apart_rows = 0) {
  const unsigned tl = (unsigned)tiles;
  cplx* W = b.W;
  for (int j = 0; j < K; ++j) {
    if (j == 0)
      prod(st, 1, nullptr, f, alpha, beta, b.V, W);
    else
      prod(st, 0, b.V + (int64_t)j * VS, lz.inv + (int64_t)j * Mtot, alpha, beta, nullptr, W);
    if (apart) {
      g_lz_alpha_sp_kernel<<<(unsigned)ngroups, RW * TT, 0, st>>>(d_gsec, ns, Mtot, apart, apart_rows, j, K, b.hess,
                                                                  lz);
    } else {
      g_pass_kernel<1, 0><<<tl, TT, 0, st>>>(d_gsec, ns, b.V + (int64_t)j * VS, VS, W, nullptr, Mtot, b.part);
      g_lz_alpha_kernel<<<(unsigned)ngroups, RW * TT, 0, st>>>(d_gsec, ns, Mtot, b.part, j, K, b.hess, lz);
    }
    if (j + 1 < K) {
      g_lz_update_kernel<<<tl, TT, 0, st>>>(d_gsec, ns, b.V, VS, j, W, lz.ca, lz.cb, b.part);
      g_lz_beta_kernel<<<(unsigned)ngroups, RW * TT, 0, st>>>(d_gsec, ns, Mtot, b.part, j, K, b.hess, lz);
    }
  }
  small_expm_kernel<<<(unsigned)Mtot, 64, 0, st>>>(Mtot, K, K, b.hess, dt, b.coef);
  g_lz_scale_kernel<<<g1((int64_t)K * Mtot, 256), 256, 0, st>>>(Mtot, K, lz.inv, b.coef);
}

inline void mrhs_combine(hipStream_t st, const GSec* d_gsec, int ns, int64_t VS, int Mtot, int K, const MrhsBufs& b,
                         const int* skip = nullptr) {
  g_combine_kernel<<<g1(VS, 256), 256, 0, st>>>(d_gsec, ns, VS, b.V, VS, K, b.coef, Mtot, skip);
}

}  // namespace
