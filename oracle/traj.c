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

/* Arnoldi step of ArnoldiSolver<T,0>::evolve (sparse_algorithm.hpp:548-562):
 * m = k_rt - 1 MGS steps, psi_out = V_m expm(-i h K_m) e_1. */
static int arnoldi_heff(const sector_t* S, double Gamma, int k_rt, double h, const tc* psi, tc* out) {
  int64_t n = S->n;
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
    heff_apply(S, Gamma, V[idx - 1], V[idx]);
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
  tc* A = (tc*)malloc(sizeof(tc) * act * act);
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

/*
 * Runs trajectories traj_offset .. traj_offset + ntraj - 1 from psi0 (sector
 * (nup0, ndn0), ascending basis order).  dens_sum[(nsteps+1) * 2 nsite]
 * receives the sum over trajectories of <n_m>(t_i); jump_log (optional,
 * ntraj * nsteps) the channel taken at each step or -1.  Returns 0 or -1.
 */
int orc_traj_run(int nsite, const qd_term* terms, int nterms, const qd_op* ops, int nup0, int ndn0,
                 const double* psi0, const int* jkind, const int* jmode, const double* jrate, int nj, int ntraj,
                 int traj_offset, int nsteps, double dt, int k_rt, uint64_t seed, double* dens_sum, int* jump_log) {
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
      for (int c = 0; c < nj; ++c) {
        w[c] = jrate[c] * dens[jmode[c]];
        G += w[c];
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
        if (jkind[ch] == 0) { /* dephasing: n_m psi */
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
        if (arnoldi_heff(S, G, k_rt, dt, psi, out)) {
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
    }
  free(sec);
  free(dens);
  free(w);
  return rc;
}
