/*
 * traj.c -- CPU restatement of the quantum-jump trajectory loop.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.c).
 *
 * Reference: programs/hubbard_diss.cpp:440-482 (loop), :514-517
 * (QJ_operator: H_eff = H - i/2 (sum_c L_c^+ L_c - Gamma I)), :519-529
 * (jump weights), :565-577 (bsearch), mc_manager.hpp:88-128 (averaging).
 * Generalised to the channels of the batched workload: dephasing
 * L = sqrt(g) n_m and single-particle loss L = sqrt(g) c_m (fermionic, Jordan-
 * Wigner sign from the occupied modes above m).  A loss moves the trajectory
 * from sector (N_up, N_dn) to (N_up - 1, N_dn) or (N_up, N_dn - 1).
 *
 * Per step (dt), for a normalised psi:
 *   w_c = g_c <n_{m_c}>, Gamma = sum_c w_c           (L^+L = g n_m for both kinds)
 *   jump iff r0 < Gamma dt; channel = first c with r1 Gamma < sum_{c'<=c} w_c'
 *   jump:   psi <- L_c psi / |L_c psi|
 *   else:   psi <- Arnoldi step (k_rt, expandBasis<0> + expm) with
 *           H_eff = H - i/2 (D - Gamma),  D = sum_c g_c n_{m_c} (diagonal),
 *           then psi <- psi / |psi|          (hubbard_diss.cpp:467-470)
 * Densities <n_m> of the normalised state are recorded at t_0 .. t_nsteps.
 *
 * Random numbers (shared definition with the GPU engine; the reference seeds
 * mt19937 from random_device and is not reproducible):
 *   key = splitmix64(seed * 0x9E3779B97F4A7C15 + traj)
 *   r_k(step) = (splitmix64(key + 4 step + k) >> 11) * 2^-53,  k = 0, 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/qudrip_hip.h"

typedef struct {
  double re, im;
} tc;

int64_t orc_basis_scan(int nbits, const qd_constraint* cons, int ncons, uint64_t* out, int64_t cap);
int64_t orc_rank(const uint64_t* basis, int64_t n, uint64_t x);
int64_t orc_build_csc(const uint64_t* basis, int64_t n, const qd_term* terms, int nterms, const qd_op* ops,
                      int32_t* colptr, int32_t* rowidx, double* vals, int64_t cap);
int orc_expm(int m, const double* A_in, double* E_out);
int64_t orc_csr_add(int64_t n, const int32_t* ao, const int32_t* ai, const double* av, const int32_t* bo,
                    const int32_t* bi, const double* bv, int sign, int32_t* co, int32_t* ci, double* cv);
void orc_csr_scale(int64_t nnz, double c_re, double c_im, int real_scalar, const double* av, double* cv);
void orc_csr_adjoint(int64_t n, const int32_t* ao, const int32_t* ai, const double* av, int32_t* co, int32_t* ci,
                     double* cv);
int64_t orc_csr_multiply(int64_t n, const int32_t* ao, const int32_t* ai, const double* av, const int32_t* bo,
                         const int32_t* bi, const double* bv, int32_t* co, int32_t* ci, double* cv);
void orc_spmv_csc(int64_t n, const int32_t* colptr, const int32_t* rowidx, const double* vals, const double* x,
                  double* y);

static inline uint64_t sm64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double orc_traj_uniform(uint64_t seed, int64_t traj, int64_t step, int k) {
  uint64_t key = sm64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)traj);
  return (double)(sm64(key + 4ull * (uint64_t)step + (uint64_t)k) >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
  int built;
  int64_t n;
  uint64_t* basis;
  int32_t *cp, *ri;
  double* va;
  double* D; /* sum_c g_c n_m per state */
  /* faithful mode: qj_tot = sum_c L_c^+ L_c, built once per sector like the
   * reference's qj_tot before its time loop */
  int qj_built;
  int64_t qj_nnz;
  int32_t *qo, *qi;
  double* qv;
} sector_t;

static int build_sector(sector_t* S, int nsite, int a, int b, const qd_term* terms, int nterms, const qd_op* ops,
                        const int* jmode, const double* jrate, int nj) {
  qd_constraint cons[2];
  cons[0].mask = (1ull << nsite) - 1ull;
  cons[0].count = a;
  cons[1].mask = ((1ull << nsite) - 1ull) << nsite;
  cons[1].count = b;
  int64_t n = orc_basis_scan(2 * nsite, cons, 2, NULL, 0);
  S->n = n;
  S->basis = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
  orc_basis_scan(2 * nsite, cons, 2, S->basis, n);
  S->cp = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  int64_t nnz = orc_build_csc(S->basis, n, terms, nterms, ops, S->cp, NULL, NULL, 0);
  if (nnz < 0) return -1;
  S->ri = (int32_t*)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  S->va = (double*)malloc(sizeof(double) * 2 * (nnz ? nnz : 1));
  orc_build_csc(S->basis, n, terms, nterms, ops, S->cp, S->ri, S->va, nnz);
  S->D = (double*)calloc(n ? n : 1, sizeof(double));
  for (int64_t s = 0; s < n; ++s)
    for (int c = 0; c < nj; ++c)
      if ((S->basis[s] >> jmode[c]) & 1ull) S->D[s] += jrate[c];
  S->built = 1;
  return 0;
}

/* y = H_eff x with H_eff = H - i/2 (D - Gamma); H from the ColMajor arrays
 * of a Hermitian operator (scatter form, like Eigen's ColMajor product). */
static void heff_apply(const sector_t* S, double Gamma, const tc* x, tc* y) {
  int64_t n = S->n;
  memset(y, 0, sizeof(tc) * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    tc xj = x[j];
    for (int32_t e = S->cp[j]; e < S->cp[j + 1]; ++e) {
      double vr = S->va[2 * e], vi = S->va[2 * e + 1];
      y[S->ri[e]].re += vr * xj.re - vi * xj.im;
      y[S->ri[e]].im += vr * xj.im + vi * xj.re;
    }
  }
  for (int64_t s = 0; s < n; ++s) {
    double d = -0.5 * (S->D[s] - Gamma); /* -i/2 (D - Gamma) x = (0, d) * x */
    y[s].re += -d * x[s].im;
    y[s].im += d * x[s].re;
  }
}

static double vnorm2(int64_t n, const tc* x) {
  double s = 0;
  for (int64_t i = 0; i < n; ++i) s += x[i].re * x[i].re + x[i].im * x[i].im;
  return s;
}

/* A stored matrix in the reference's ColMajor arrays. */
typedef struct {
  int64_t n, nnz;
  int32_t *o, *i;
  double* v;
} csc_t;
static void csc_free(csc_t* a) {
  free(a->o);
  free(a->i);
  free(a->v);
  memset(a, 0, sizeof(*a));
}
static void csc_alloc(csc_t* a, int64_t n, int64_t nnz) {
  a->n = n;
  a->nnz = nnz;
  a->o = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
  a->i = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  a->v = (double*)malloc(sizeof(double) * 2 * (size_t)(nnz ? nnz : 1));
}

/* y = H_eff x for the Arnoldi step: the implicit form (stored H plus the
 * diagonal -i/2 (D - Gamma)) or an assembled stored matrix. */
typedef struct {
  const sector_t* S;
  double Gamma;
  const csc_t* A; /* non-NULL: the assembled H_eff */
} heff_t;
static void heff_op(const heff_t* op, const tc* x, tc* y) {
  if (op->A)
    orc_spmv_csc(op->A->n, op->A->o, op->A->i, op->A->v, (const double*)x, (double*)y);
  else
    heff_apply(op->S, op->Gamma, x, y);
}

/* Arnoldi step of ArnoldiSolver<T,0>::evolve (sparse_algorithm.hpp:548-562):
 * m = k_rt - 1 MGS steps, psi_out = V_m expm(-i h K_m) e_1. */
static int arnoldi_heff(const heff_t* op, int k_rt, double h, const tc* psi, tc* out) {
  int64_t n = op->S->n;
  int m = k_rt - 1;
  tc** V = (tc**)malloc(sizeof(tc*) * (m + 1));
  for (int j = 0; j <= m; ++j) V[j] = (tc*)malloc(sizeof(tc) * (size_t)n);
  tc* K = (tc*)calloc((size_t)(m + 1) * (m + 1), sizeof(tc));
  double nr = sqrt(vnorm2(n, psi));
  for (int64_t r = 0; r < n; ++r) {
    V[0][r].re = psi[r].re / nr;
    V[0][r].im = psi[r].im / nr;
  }
  int idx;
  for (idx = 1; idx <= m; ++idx) {
    heff_op(op, V[idx - 1], V[idx]);
    for (int i = 0; i < idx; ++i) {
      tc hh = {0, 0};
      for (int64_t r = 0; r < n; ++r) {
        hh.re += V[i][r].re * V[idx][r].re + V[i][r].im * V[idx][r].im;
        hh.im += V[i][r].re * V[idx][r].im - V[i][r].im * V[idx][r].re;
      }
      for (int64_t r = 0; r < n; ++r) {
        V[idx][r].re -= hh.re * V[i][r].re - hh.im * V[i][r].im;
        V[idx][r].im -= hh.re * V[i][r].im + hh.im * V[i][r].re;
      }
      K[i + (idx - 1) * (m + 1)] = hh;
    }
    double nn = sqrt(vnorm2(n, V[idx]));
    for (int64_t r = 0; r < n; ++r) {
      V[idx][r].re /= nn;
      V[idx][r].im /= nn;
    }
    K[idx + (idx - 1) * (m + 1)].re = nn;
    if (nn < 1e-10) {
      ++idx;
      break;
    }
  }
  int act = idx - 1;
  tc* A = (tc*)calloc((size_t)(act ? act * act : 1), sizeof(tc));
  tc* E = (tc*)malloc(sizeof(tc) * act * act);
  for (int j = 0; j < act; ++j)
    for (int i = 0; i < act; ++i) {
      tc kv = K[i + j * (m + 1)];
      A[i + j * act].re = h * kv.im;
      A[i + j * act].im = -h * kv.re;
    }
  int rc = orc_expm(act, (const double*)A, (double*)E);
  if (rc == 0)
    for (int64_t r = 0; r < n; ++r) {
      tc s = {0, 0};
      for (int j = 0; j < act; ++j) {
        s.re += V[j][r].re * E[j].re - V[j][r].im * E[j].im;
        s.im += V[j][r].re * E[j].im + V[j][r].im * E[j].re;
      }
      out[r] = s;
    }
  for (int j = 0; j <= m; ++j) free(V[j]);
  free(V);
  free(K);
  free(A);
  free(E);
  return rc;
}

/* `program >> sector` (operator_abel.hpp:25-34): the ColMajor arrays of one
 * channel's L_c over the sector's basis. */
static void build_prog(const sector_t* S, const qd_term* t, int nt, const qd_op* ops, csc_t* out) {
  int32_t* cp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(S->n + 1));
  int64_t nnz = orc_build_csc(S->basis, S->n, t, nt, ops, cp, NULL, NULL, 0);
  free(cp);
  csc_alloc(out, S->n, nnz < 0 ? 0 : nnz);
  orc_build_csc(S->basis, S->n, t, nt, ops, out->o, out->i, out->v, out->nnz);
}

/* SpMatrix(L.adjoint()) * L (hubbard_diss.cpp:527). */
static void lhl(const csc_t* L, csc_t* out) {
  csc_t La;
  csc_alloc(&La, L->n, L->nnz);
  orc_csr_adjoint(L->n, L->o, L->i, L->v, La.o, La.i, La.v);
  int64_t nnz = orc_csr_multiply(L->n, La.o, La.i, La.v, L->o, L->i, L->v, NULL, NULL, NULL);
  csc_alloc(out, L->n, nnz);
  orc_csr_multiply(L->n, La.o, La.i, La.v, L->o, L->i, L->v, out->o, out->i, out->v);
  csc_free(&La);
}

/* C = A + sign B, allocated. */
static void add_into(const csc_t* A, const csc_t* B, int sign, csc_t* C) {
  int64_t nnz = orc_csr_add(A->n, A->o, A->i, A->v, B->o, B->i, B->v, sign, NULL, NULL, NULL);
  csc_alloc(C, A->n, nnz);
  orc_csr_add(A->n, A->o, A->i, A->v, B->o, B->i, B->v, sign, C->o, C->i, C->v);
}

/* C = c A, allocated (c real when c_im == 0 and real_scalar). */
static void scale_into(const csc_t* A, double c_re, double c_im, int real_scalar, csc_t* C) {
  csc_alloc(C, A->n, A->nnz);
  memcpy(C->o, A->o, sizeof(int32_t) * (size_t)(A->n + 1));
  memcpy(C->i, A->i, sizeof(int32_t) * (size_t)(A->nnz ? A->nnz : 1));
  orc_csr_scale(A->nnz, c_re, c_im, real_scalar, A->v, C->v);
}

/* <psi|M|psi> (real part) by the stored product, like State::eval. */
static double expect_re(const csc_t* M, const tc* psi, tc* tmp) {
  orc_spmv_csc(M->n, M->o, M->i, M->v, (const double*)psi, (double*)tmp);
  double s = 0.0;
  for (int64_t r = 0; r < M->n; ++r) s += psi[r].re * tmp[r].re + psi[r].im * tmp[r].im;
  return s;
}

/* qj_tot = sum_c L_c^+ L_c over the sector (once per sector). */
static void build_qj(sector_t* S, int nj, const qd_term* lterms, const int* loff, const qd_op* lops) {
  csc_t acc;
  csc_alloc(&acc, S->n, 0);
  memset(acc.o, 0, sizeof(int32_t) * (size_t)(S->n + 1));
  for (int c = 0; c < nj; ++c) {
    csc_t L, M, next;
    build_prog(S, lterms + loff[c], loff[c + 1] - loff[c], lops, &L);
    lhl(&L, &M);
    add_into(&acc, &M, 1, &next);
    csc_free(&L);
    csc_free(&M);
    csc_free(&acc);
    acc = next;
  }
  S->qj_nnz = acc.nnz;
  S->qo = acc.o;
  S->qi = acc.i;
  S->qv = acc.v;
  S->qj_built = 1;
}

/*
 * Runs trajectories traj_offset .. traj_offset + ntraj - 1 from psi0 (sector
 * (nup0, ndn0), ascending basis order).  dens_sum[(nsteps+1) * 2 nsite]
 * receives the sum over trajectories of <n_m>(t_i); jump_log (optional,
 * ntraj * nsteps) the channel taken at each step or -1.  Returns 0 or -1.
 */
static int traj_impl(int nsite, const qd_term* terms, int nterms, const qd_op* ops, int nup0, int ndn0,
                     const double* psi0, const int* jkind, const int* jmode, const double* jrate, int nj, int ntraj,
                     int traj_offset, int nsteps, double dt, int k_rt, uint64_t seed, double* dens_sum,
                     int* jump_log, const qd_term* lterms, const int* loff, const qd_op* lops) {
  int NS = (nsite + 1) * (nsite + 1);
  sector_t* sec = (sector_t*)calloc(NS, sizeof(sector_t));
  int M = 2 * nsite;
  double* dens = (double*)malloc(sizeof(double) * M);
  double* w = (double*)malloc(sizeof(double) * (nj ? nj : 1));
  int rc = 0;
  for (int tr = 0; tr < ntraj && rc == 0; ++tr) {
    int a = nup0, b = ndn0;
    sector_t* S = &sec[a * (nsite + 1) + b];
    if (!S->built && build_sector(S, nsite, a, b, terms, nterms, ops, jmode, jrate, nj)) {
      rc = -1;
      break;
    }
    tc* psi = (tc*)malloc(sizeof(tc) * (size_t)S->n);
    memcpy(psi, psi0, sizeof(tc) * (size_t)S->n);
    double nr = sqrt(vnorm2(S->n, psi));
    for (int64_t s = 0; s < S->n; ++s) {
      psi[s].re /= nr;
      psi[s].im /= nr;
    }
    for (int step = 0; step <= nsteps; ++step) {
      /* densities of the normalised state at t_step */
      double n2 = vnorm2(S->n, psi);
      for (int m = 0; m < M; ++m) dens[m] = 0.0;
      for (int64_t s = 0; s < S->n; ++s) {
        double p = psi[s].re * psi[s].re + psi[s].im * psi[s].im;
        uint64_t x = S->basis[s];
        for (int m = 0; m < M; ++m)
          if ((x >> m) & 1ull) dens[m] += p;
      }
      for (int m = 0; m < M; ++m) {
        dens[m] /= n2;
        dens_sum[(size_t)step * M + m] += dens[m];
      }
      if (step == nsteps) break;
      /* weights and the jump test */
      double G = 0.0;
      if (lterms) {
        /* lindblad_weight (hubbard_diss.cpp:519-529): every step, each L_c
         * is built by `>>`, multiplied by its adjoint and evaluated */
        tc* tmp = (tc*)malloc(sizeof(tc) * (size_t)(S->n ? S->n : 1));
        for (int c = 0; c < nj; ++c) {
          csc_t L, M;
          build_prog(S, lterms + loff[c], loff[c + 1] - loff[c], lops, &L);
          lhl(&L, &M);
          w[c] = expect_re(&M, psi, tmp) / n2;
          G += w[c];
          csc_free(&L);
          csc_free(&M);
        }
        free(tmp);
      } else {
        for (int c = 0; c < nj; ++c) {
          w[c] = jrate[c] * dens[jmode[c]];
          G += w[c];
        }
      }
      int64_t gtraj = traj_offset + tr;
      double r0 = orc_traj_uniform(seed, gtraj, step, 0);
      double r1 = orc_traj_uniform(seed, gtraj, step, 1);
      int ch = -1;
      if (r0 < G * dt) {
        double target = r1 * G, cum = 0.0;
        for (int c = 0; c < nj; ++c) {
          cum += w[c];
          if (w[c] > 0.0) ch = c;
          if (target < cum && w[c] > 0.0) break;
        }
      }
      if (jump_log) jump_log[(size_t)tr * nsteps + step] = ch;
      if (ch >= 0) {
        int m = jmode[ch];
        if (jkind[ch] == 0 && lterms) { /* jump_op = L >> sector; psi <- L L psi (:462-463) */
          csc_t L;
          build_prog(S, lterms + loff[ch], loff[ch + 1] - loff[ch], lops, &L);
          tc* t1 = (tc*)malloc(sizeof(tc) * (size_t)(S->n ? S->n : 1));
          orc_spmv_csc(S->n, L.o, L.i, L.v, (const double*)psi, (double*)t1);
          orc_spmv_csc(S->n, L.o, L.i, L.v, (const double*)t1, (double*)psi);
          free(t1);
          csc_free(&L);
        } else if (jkind[ch] == 0) { /* dephasing: n_m psi */
          for (int64_t s = 0; s < S->n; ++s)
            if (!((S->basis[s] >> m) & 1ull)) psi[s].re = psi[s].im = 0.0;
        } else { /* loss: c_m psi into the neighbouring sector */
          int na = a, nb = b;
          if (m < nsite) na = a - 1;
          else nb = b - 1;
          sector_t* T = &sec[na * (nsite + 1) + nb];
          if (!T->built && build_sector(T, nsite, na, nb, terms, nterms, ops, jmode, jrate, nj)) {
            rc = -1;
            break;
          }
          tc* np = (tc*)calloc((size_t)(T->n ? T->n : 1), sizeof(tc));
          for (int64_t s = 0; s < S->n; ++s) {
            uint64_t x = S->basis[s];
            if (!((x >> m) & 1ull)) continue;
            int sg = (__builtin_popcountll(x >> (m + 1)) & 1) ? -1 : 1;
            int64_t r = orc_rank(T->basis, T->n, x & ~(1ull << m));
            np[r].re = sg * psi[s].re;
            np[r].im = sg * psi[s].im;
          }
          free(psi);
          psi = np;
          S = T;
          a = na;
          b = nb;
        }
        double nn = sqrt(vnorm2(S->n, psi));
        for (int64_t s = 0; s < S->n; ++s) {
          psi[s].re /= nn;
          psi[s].im /= nn;
        }
      } else {
        tc* out = (tc*)malloc(sizeof(tc) * (size_t)S->n);
        heff_t op = {S, G, NULL};
        csc_t Heff = {0, 0, NULL, NULL, NULL};
        if (lterms) {
          /* Htot assembled (a linear combination, :443), then QJ_operator
           * (:514-517): H -= 0.5 i (qj_tot - Gamma Id) */
          if (!S->qj_built) build_qj(S, nj, lterms, loff, lops);
          csc_t H = {S->n, S->cp[S->n], S->cp, S->ri, S->va};
          csc_t Q = {S->n, S->qj_nnz, S->qo, S->qi, S->qv};
          csc_t Id, GId, A, B, Htot;
          csc_alloc(&Id, S->n, S->n);
          for (int64_t r = 0; r <= S->n; ++r) Id.o[r] = (int32_t)r;
          for (int64_t r = 0; r < S->n; ++r) {
            Id.i[r] = (int32_t)r;
            Id.v[2 * r] = 1.0;
            Id.v[2 * r + 1] = 0.0;
          }
          scale_into(&Id, G, 0.0, 1, &GId);
          add_into(&Q, &GId, -1, &A);
          scale_into(&A, 0.0, 0.5, 0, &B);
          scale_into(&H, 1.0, 0.0, 1, &Htot);
          add_into(&Htot, &B, -1, &Heff);
          csc_free(&Id);
          csc_free(&GId);
          csc_free(&A);
          csc_free(&B);
          csc_free(&Htot);
          op.A = &Heff;
        }
        int erc = arnoldi_heff(&op, k_rt, dt, psi, out);
        if (lterms) csc_free(&Heff);
        if (erc) {
          free(out);
          rc = -1;
          break;
        }
        double nn = sqrt(vnorm2(S->n, out));
        for (int64_t s = 0; s < S->n; ++s) {
          psi[s].re = out[s].re / nn;
          psi[s].im = out[s].im / nn;
        }
        free(out);
      }
    }
    free(psi);
  }
  for (int i = 0; i < NS; ++i)
    if (sec[i].built) {
      free(sec[i].basis);
      free(sec[i].cp);
      free(sec[i].ri);
      free(sec[i].va);
      free(sec[i].D);
      if (sec[i].qj_built) {
        free(sec[i].qo);
        free(sec[i].qi);
        free(sec[i].qv);
      }
    }
  free(sec);
  free(dens);
  free(w);
  return rc;
}

int orc_traj_run(int nsite, const qd_term* terms, int nterms, const qd_op* ops, int nup0, int ndn0,
                 const double* psi0, const int* jkind, const int* jmode, const double* jrate, int nj, int ntraj,
                 int traj_offset, int nsteps, double dt, int k_rt, uint64_t seed, double* dens_sum, int* jump_log) {
  return traj_impl(nsite, terms, nterms, ops, nup0, ndn0, psi0, jkind, jmode, jrate, nj, ntraj, traj_offset, nsteps,
                   dt, k_rt, seed, dens_sum, jump_log, NULL, NULL, NULL);
}

/* The same loop as the reference runs it: every step the jump operators
 * L_c (programs lterms[loff[c] .. loff[c+1]), ops lops, e.g. sqrt(g) n_m) are
 * built by `>>`, multiplied by their adjoints and evaluated for the weights
 * (hubbard_diss.cpp:519-529); a dephasing jump applies the built L_c twice
 * (:462-463); a non-jump step assembles H_eff = Htot - 0.5 i (qj_tot -
 * Gamma Id) as a stored matrix (:443, :514-517) and runs the Arnoldi step on
 * it.  qj_tot is built once per sector.  Loss jumps move between sectors by
 * rank lookups as above.  Same random streams, so the jump logs agree with
 * orc_traj_run's up to rounding of the weights. */
int orc_traj_run_rebuild(int nsite, const qd_term* terms, int nterms, const qd_op* ops, int nup0, int ndn0,
                         const double* psi0, const int* jkind, const int* jmode, const double* jrate, int nj,
                         int ntraj, int traj_offset, int nsteps, double dt, int k_rt, uint64_t seed, double* dens_sum,
                         int* jump_log, const qd_term* lterms, const int* loff, const qd_op* lops) {
  if (!lterms || !loff || !lops) return -1;
  return traj_impl(nsite, terms, nterms, ops, nup0, ndn0, psi0, jkind, jmode, jrate, nj, ntraj, traj_offset, nsteps,
                   dt, k_rt, seed, dens_sum, jump_log, lterms, loff, lops);
}
