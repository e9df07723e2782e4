/*
 * oracle.c -- CPU restatement of the reference algorithms on the Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (qudrip_amd/, include/)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg do, as the checker or the timed CPU baseline.
 *
 * Complex arithmetic is written out component-wise so the operation order is
 * the one std::complex<double> performs on x86-64 without FMA (the reference's
 * build: g++ -O3, no -march, CMakeLists.txt:6,26):
 *     (a+bi)(c+di) = (ac - bd) + (ad + bc) i,   d*(a+bi) = (da) + (db) i.
 * Build with -ffp-contract=off so no multiply-add is fused.
 *
 * Pinning (how each part is checked independently of this file; the tests
 * named here live in tests/test_oracle.py):
 *   basis scan / rank ............ closed-form binomial dimensions, CNS rank
 *                                  identity, brute-force popcount predicate.
 *   tree expansion / CSC build ... independent dense construction from the
 *                                  physics definition (numpy), closed-form nnz,
 *                                  free-fermion / Bethe-ansatz energies.
 *   spmv / map_acc ............... dense numpy matrix-vector products.
 *   Lanczos ground state ......... closed-form E0 (XY free fermions, XXX Bethe
 *                                  ansatz) and dense numpy eigvalsh.
 *   tridiagonal eigensolver ...... numpy.linalg.eigh on random tridiagonals.
 *   expm ......................... closed forms (diagonal, rotation, nilpotent)
 *                                  and scipy.linalg.expm fixtures.
 *   Arnoldi step ................. dense scipy expm propagation, norm and
 *                                  energy conservation.
 *   trajectories ................. exact Lindblad master equation (scipy) for
 *                                  small systems, statistically.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/qudrip_hip.h"

#define ORC_EPS 1e-10 /* qudrip.hpp:18 */

typedef struct {
  double re, im;
} oc;

static inline oc c_make(double re, double im) {
  oc r;
  r.re = re;
  r.im = im;
  return r;
}
static inline oc c_mul(oc a, oc b) {
  return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline oc c_add(oc a, oc b) { return c_make(a.re + b.re, a.im + b.im); }
static inline oc c_sub(oc a, oc b) { return c_make(a.re - b.re, a.im - b.im); }
static inline oc c_rmul(double d, oc a) { return c_make(d * a.re, d * a.im); }
static inline oc c_conj(oc a) { return c_make(a.re, -a.im); }
static inline double c_abs(oc a) { return hypot(a.re, a.im); }

static inline int popc(uint64_t x) { return __builtin_popcountll(x); }

/* ======================================================================
 * Basis.  conservation.hpp:212-231 (LooseConstrain/Constrain): scan every x
 * in [0, 2^nbits) in ascending order, keep x iff every constraint holds
 * (Evaluate<0>, :201-210: |popcount(x & mask) - value| <= 0, Nset::eval
 * :102-104).  Returns the sector size; writes min(size, cap) states.
 * ====================================================================== */
int64_t orc_basis_scan(int nbits, const qd_constraint* cons, int ncons, uint64_t* out,
                       int64_t cap) {
  uint64_t range = (nbits >= 64) ? 0 : (1ull << nbits);
  int64_t count = 0;
  for (uint64_t x = 0; x < range; ++x) {
    int ok = 1;
    for (int c = 0; c < ncons && ok; ++c)
      if ((int64_t)popc(x & cons[c].mask) != cons[c].count) ok = 0;
    if (ok) {
      if (count < cap && out) out[count] = x;
      ++count;
    }
  }
  return count;
}

/* Constrain over Nset and Nd quantities (conservation.hpp:103,129-132,
 * 212-231): Nd::eval is popcount(((x >> Nsite) ^ mask) & (x ^ mask)) with
 * mask = bits 0..Nsite-1 for getNh, 0 for getNd; an Nset is the same form
 * with shift 0, flip 0 and sel = its mask.  q[k] = {sel, flip, shift, count}
 * as qd_quantity. */
typedef struct {
  uint64_t sel, flip;
  int32_t shift, reserved;
  int64_t count;
} orc_quantity;
int64_t orc_basis_scan_q(int nbits, const orc_quantity* q, int nq, uint64_t* out, int64_t cap) {
  uint64_t range = (nbits >= 64) ? 0 : (1ull << nbits);
  int64_t count = 0;
  for (uint64_t x = 0; x < range; ++x) {
    int ok = 1;
    for (int c = 0; c < nq && ok; ++c) {
      const uint64_t f = q[c].flip;
      if ((int64_t)popc(((x >> q[c].shift) ^ f) & (x ^ f) & q[c].sel) != q[c].count) ok = 0;
    }
    if (ok) {
      if (count < cap && out) out[count] = x;
      ++count;
    }
  }
  return count;
}

/* index.hpp:234-240: inv_mapping lookup; a miss returns range() == n.  The
 * mapping is ascending, so a binary search finds the same entry. */
int64_t orc_rank(const uint64_t* basis, int64_t n, uint64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (basis[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < n && basis[lo] == x) return lo;
  return n;
}

uint64_t orc_binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  uint64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
  return r;
}

/* ======================================================================
 * Operator tree.  operator.hpp:135-206 (Operator::full_branch): the ops of a
 * term are applied right-to-left (:151), level by level; every node of a
 * level is branched in order, so the leaves come out in lexicographic order
 * of the child choices with the first-applied op most significant.  A node
 * holds (state + 1) or 0 for a dead branch (:143, :162).
 *   bitOp::branch   operators/bit.hpp:48-84   two children o = 0, 1
 *   hopBitOp::branch operators/hopbit.hpp:43-84 one child
 * The reference sizes the leaf scan by the declared n_output (2 per op,
 * operator.hpp:49-55,193); slots past the real leaves of a hop term are
 * never written and stay 0, so they are dropped -- only real leaves are
 * produced here.
 * Returns the number of leaves (<= cap) written to leaf_idx / leaf_val.
 * ====================================================================== */
int orc_expand_term(const qd_op* ops, int n_ops, uint64_t state, const double* init,
                    uint64_t* leaf_idx, double* leaf_val, int cap) {
  enum { MAXN = 1 << QD_MAX_OPS_PER_TERM };
  uint64_t idx_a[MAXN], idx_b[MAXN];
  oc val_a[MAXN], val_b[MAXN];
  int n = 1;
  idx_a[0] = state + 1;
  val_a[0] = c_make(init[0], init[1]);
  uint64_t* ci = idx_a;
  uint64_t* ni = idx_b;
  oc* cv = val_a;
  oc* nv = val_b;
  for (int iop = n_ops - 1; iop >= 0; --iop) {
    const qd_op* op = &ops[iop];
    int m = 0;
    for (int j = 0; j < n; ++j) {
      if (op->kind == QD_OP_BIT) {
        int s = op->site_a;
        if (ci[j] != 0) {
          uint64_t x = ci[j] - 1;
          int b = (int)((x >> s) & 1ull);
          int prev_set = popc(x >> (s + 1));
          oc jw = c_make((prev_set % 2) ? (double)op->string : 1.0, 0.0);
          oc u0 = c_make(op->U[(b * 2 + 0) * 2], op->U[(b * 2 + 0) * 2 + 1]);
          oc u1 = c_make(op->U[(b * 2 + 1) * 2], op->U[(b * 2 + 1) * 2 + 1]);
          nv[m] = c_mul(c_mul(jw, u0), cv[j]);
          ni[m] = (x & ~(1ull << s)) + 1;
          nv[m + 1] = c_mul(c_mul(jw, u1), cv[j]);
          ni[m + 1] = (x | (1ull << s)) + 1;
        } else {
          ni[m] = 0;
          nv[m] = c_make(0.0, 0.0);
          ni[m + 1] = 0;
          nv[m + 1] = c_make(0.0, 0.0);
        }
        m += 2;
      } else {
        int a = op->site_a; /* site2: created  */
        int b = op->site_b; /* site1: annihilated */
        if (ci[j] != 0) {
          uint64_t x = ci[j] - 1;
          if (a == b) {
            ni[m] = x + 1;
            nv[m] = c_rmul((double)((x >> a) & 1ull), cv[j]);
          } else if (((x >> b) & 1ull) == 1 && ((x >> a) & 1ull) == 0) {
            int jw1 = popc(x >> (b + 1));
            uint64_t x1 = x & ~(1ull << b);
            int jw2 = popc(x1 >> (a + 1));
            uint64_t x2 = x1 | (1ull << a);
            int sgn = ((jw1 % 2) ? op->string : 1) * ((jw2 % 2) ? op->string : 1);
            nv[m] = c_mul(c_make((double)sgn, 0.0), cv[j]);
            ni[m] = x2 + 1;
          } else {
            ni[m] = 0;
            nv[m] = c_make(0.0, 0.0);
          }
        } else {
          ni[m] = 0;
          nv[m] = c_make(0.0, 0.0);
        }
        m += 1;
      }
    }
    n = m;
    uint64_t* ti = ci;
    ci = ni;
    ni = ti;
    oc* tv = cv;
    cv = nv;
    nv = tv;
  }
  int w = n < cap ? n : cap;
  for (int j = 0; j < w; ++j) {
    leaf_idx[j] = ci[j];
    leaf_val[2 * j] = cv[j].re;
    leaf_val[2 * j + 1] = cv[j].im;
  }
  return n;
}

/* ======================================================================
 * Eager build, `ham >> index`.
 *  - operator.hpp:250-262: for each source column i, expand every term;
 *    keep leaf j iff idx != 0 and |v| > eps (std::abs == hypot), target
 *    row = rank of the leaf state (a miss is dropped, :196-199).
 *  - Eigen setFromTriplets (operator.hpp:264): duplicates of one term and
 *    one column are summed in insertion (leaf) order.
 *  - operator_abel.hpp:28-31: result += c_k * M_k in term order; Eigen's
 *    sparse sum treats a missing side as zero, so every stored entry is
 *    acc = (+0,+0); acc = acc + c_k * v_k over the terms that hold it.
 *    Entries that cancel to zero stay stored.
 *  - Storage ColMajor (matrix.hpp:11): colptr over source columns, rows
 *    ascending.
 * If rowidx/vals are NULL or cap < nnz only colptr is written.  Returns
 * nnz, or -1 on allocation failure / a term wider than the tree limit.
 * ====================================================================== */
typedef struct {
  int64_t row;
  oc acc;
} orc_entry;

int64_t orc_build_csc(const uint64_t* basis, int64_t n, const qd_term* terms, int nterms,
                      const qd_op* ops, int32_t* colptr, int32_t* rowidx, double* vals,
                      int64_t cap) {
  enum { MAXL = 1 << QD_MAX_OPS_PER_TERM };
  uint64_t lidx[MAXL];
  double lval[2 * MAXL];
  int64_t trow[MAXL];
  oc tval[MAXL];
  int64_t ecap = 256;
  orc_entry* ent = (orc_entry*)malloc(sizeof(orc_entry) * ecap);
  if (!ent) return -1;
  int64_t nnz = 0;
  const double one[2] = {1.0, 0.0};
  for (int64_t i = 0; i < n; ++i) {
    int64_t ne = 0;
    colptr[i] = (int32_t)nnz;
    for (int k = 0; k < nterms; ++k) {
      const qd_term* t = &terms[k];
      if (t->n_ops > QD_MAX_OPS_PER_TERM) {
        free(ent);
        return -1;
      }
      int nl = orc_expand_term(ops + t->op_begin, t->n_ops, basis[i], one, lidx, lval, MAXL);
      /* per-term triplets, duplicates merged in leaf order */
      int nt = 0;
      for (int j = 0; j < nl; ++j) {
        if (lidx[j] == 0) continue;
        oc v = c_make(lval[2 * j], lval[2 * j + 1]);
        if (!(c_abs(v) > ORC_EPS)) continue;
        int64_t r = orc_rank(basis, n, lidx[j] - 1);
        if (r >= n) continue;
        int found = -1;
        for (int q = 0; q < nt; ++q)
          if (trow[q] == r) {
            found = q;
            break;
          }
        if (found >= 0)
          tval[found] = c_add(tval[found], v);
        else {
          trow[nt] = r;
          tval[nt] = v;
          ++nt;
        }
      }
      /* result += c_k * M_k */
      oc ck = c_make(t->coeff[0], t->coeff[1]);
      for (int q = 0; q < nt; ++q) {
        oc prod = c_mul(ck, tval[q]);
        int64_t e;
        for (e = 0; e < ne; ++e)
          if (ent[e].row == trow[q]) break;
        if (e == ne) {
          if (ne == ecap) {
            ecap *= 2;
            orc_entry* ne2 = (orc_entry*)realloc(ent, sizeof(orc_entry) * ecap);
            if (!ne2) {
              free(ent);
              return -1;
            }
            ent = ne2;
          }
          ent[ne].row = trow[q];
          ent[ne].acc = c_make(0.0, 0.0);
          ++ne;
        }
        ent[e].acc = c_add(ent[e].acc, prod);
      }
    }
    /* rows ascending (insertion sort: entries per column are few) */
    for (int64_t a = 1; a < ne; ++a) {
      orc_entry key = ent[a];
      int64_t b = a - 1;
      while (b >= 0 && ent[b].row > key.row) {
        ent[b + 1] = ent[b];
        --b;
      }
      ent[b + 1] = key;
    }
    if (rowidx && vals && nnz + ne <= cap) {
      for (int64_t e = 0; e < ne; ++e) {
        rowidx[nnz + e] = (int32_t)ent[e].row;
        vals[2 * (nnz + e)] = ent[e].acc.re;
        vals[2 * (nnz + e) + 1] = ent[e].acc.im;
      }
    }
    nnz += ne;
  }
  colptr[n] = (int32_t)nnz;
  free(ent);
  return nnz;
}

/* ======================================================================
 * Composite index QbitsIndex * SingleModeIndex(mdim) (Indices,
 * index.hpp:112-184): a node holds the composite value q * mdim + m (+1).
 * Tree as orc_expand_term, level by level; bit and hop gates act on q, a
 * mode gate (operators/mode.hpp:80-103) gives every node mdim children,
 * child j with value U_[m * mdim + j] * val (U_ as elOpBase stores it,
 * operator.hpp:24-29), dead nodes mdim dead children.  mats holds the
 * program's matrices, mdim * mdim interleaved complex each; a mode gate's
 * site_a picks one.  Returns the leaf count; *out_idx / *out_val are
 * malloc'ed (caller frees).
 * ====================================================================== */
static int expand_term_modes(const qd_op* ops, int n_ops, int mdim, const double* mats, uint64_t state,
                             uint64_t** out_idx, oc** out_val) {
  int64_t total = 1;
  for (int iop = 0; iop < n_ops; ++iop)
    total *= ops[iop].kind == QD_OP_BIT ? 2 : ops[iop].kind == QD_OP_MODE ? mdim : 1;
  uint64_t* ci = (uint64_t*)malloc(sizeof(uint64_t) * total);
  uint64_t* ni = (uint64_t*)malloc(sizeof(uint64_t) * total);
  oc* cv = (oc*)malloc(sizeof(oc) * total);
  oc* nv = (oc*)malloc(sizeof(oc) * total);
  int n = 1;
  ci[0] = state + 1;
  cv[0] = c_make(1.0, 0.0);
  for (int iop = n_ops - 1; iop >= 0; --iop) {
    const qd_op* op = &ops[iop];
    int m = 0;
    for (int j = 0; j < n; ++j) {
      const int alive = ci[j] != 0;
      const uint64_t x = alive ? ci[j] - 1 : 0;
      const uint64_t q = x / (uint64_t)mdim;
      const uint64_t md = x % (uint64_t)mdim;
      if (op->kind == QD_OP_MODE) {
        const double* U = mats + 2 * (size_t)op->site_a * mdim * mdim;
        for (int o = 0; o < mdim; ++o) {
          if (alive) {
            oc u = c_make(U[2 * (md * mdim + o)], U[2 * (md * mdim + o) + 1]);
            nv[m + o] = c_mul(u, cv[j]);
            ni[m + o] = q * mdim + o + 1;
          } else {
            nv[m + o] = c_make(0.0, 0.0);
            ni[m + o] = 0;
          }
        }
        m += mdim;
      } else if (op->kind == QD_OP_BIT) {
        int s = op->site_a;
        if (alive) {
          int b = (int)((q >> s) & 1ull);
          int prev_set = popc(q >> (s + 1));
          oc jw = c_make((prev_set % 2) ? (double)op->string : 1.0, 0.0);
          oc u0 = c_make(op->U[(b * 2 + 0) * 2], op->U[(b * 2 + 0) * 2 + 1]);
          oc u1 = c_make(op->U[(b * 2 + 1) * 2], op->U[(b * 2 + 1) * 2 + 1]);
          nv[m] = c_mul(c_mul(jw, u0), cv[j]);
          ni[m] = (q & ~(1ull << s)) * mdim + md + 1;
          nv[m + 1] = c_mul(c_mul(jw, u1), cv[j]);
          ni[m + 1] = (q | (1ull << s)) * mdim + md + 1;
        } else {
          ni[m] = ni[m + 1] = 0;
          nv[m] = nv[m + 1] = c_make(0.0, 0.0);
        }
        m += 2;
      } else {
        int a = op->site_a, b = op->site_b;
        ni[m] = 0;
        nv[m] = c_make(0.0, 0.0);
        if (alive) {
          if (a == b) {
            ni[m] = x + 1;
            nv[m] = c_rmul((double)((q >> a) & 1ull), cv[j]);
          } else if (((q >> b) & 1ull) == 1 && ((q >> a) & 1ull) == 0) {
            int jw1 = popc(q >> (b + 1));
            uint64_t q1 = q & ~(1ull << b);
            int jw2 = popc(q1 >> (a + 1));
            int sgn = ((jw1 % 2) ? op->string : 1) * ((jw2 % 2) ? op->string : 1);
            nv[m] = c_mul(c_make((double)sgn, 0.0), cv[j]);
            ni[m] = (q1 | (1ull << a)) * mdim + md + 1;
          }
        }
        m += 1;
      }
    }
    n = m;
    uint64_t* ti = ci;
    ci = ni;
    ni = ti;
    oc* tv = cv;
    cv = nv;
    nv = tv;
  }
  free(ni);
  free(nv);
  *out_idx = ci;
  *out_val = cv;
  return n;
}

/* `ham >> index` over a composite index: as orc_build_csc (same triplet
 * merge and term sum), leaves from expand_term_modes; basis holds composite
 * values ascending.  Returns nnz or -1. */
int64_t orc_build_csc_modes(const uint64_t* basis, int64_t n, int mdim, const qd_term* terms, int nterms,
                            const qd_op* ops, const double* mats, int32_t* colptr, int32_t* rowidx, double* vals,
                            int64_t cap) {
  int64_t ecap = 256;
  orc_entry* ent = (orc_entry*)malloc(sizeof(orc_entry) * ecap);
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t ne = 0;
    colptr[i] = (int32_t)nnz;
    for (int k = 0; k < nterms; ++k) {
      const qd_term* t = &terms[k];
      uint64_t* lidx;
      oc* lv;
      int nl = expand_term_modes(ops + t->op_begin, t->n_ops, mdim, mats, basis[i], &lidx, &lv);
      int64_t* trow = (int64_t*)malloc(sizeof(int64_t) * (nl ? nl : 1));
      oc* tval = (oc*)malloc(sizeof(oc) * (nl ? nl : 1));
      int nt = 0;
      for (int j = 0; j < nl; ++j) {
        if (lidx[j] == 0) continue;
        if (!(c_abs(lv[j]) > ORC_EPS)) continue;
        int64_t r = orc_rank(basis, n, lidx[j] - 1);
        if (r >= n) continue;
        int found = -1;
        for (int q = 0; q < nt; ++q)
          if (trow[q] == r) {
            found = q;
            break;
          }
        if (found >= 0)
          tval[found] = c_add(tval[found], lv[j]);
        else {
          trow[nt] = r;
          tval[nt] = lv[j];
          ++nt;
        }
      }
      oc ck = c_make(t->coeff[0], t->coeff[1]);
      for (int q = 0; q < nt; ++q) {
        oc prod = c_mul(ck, tval[q]);
        int64_t e;
        for (e = 0; e < ne; ++e)
          if (ent[e].row == trow[q]) break;
        if (e == ne) {
          if (ne == ecap) {
            ecap *= 2;
            ent = (orc_entry*)realloc(ent, sizeof(orc_entry) * ecap);
          }
          ent[ne].row = trow[q];
          ent[ne].acc = c_make(0.0, 0.0);
          ++ne;
        }
        ent[e].acc = c_add(ent[e].acc, prod);
      }
      free(trow);
      free(tval);
      free(lidx);
      free(lv);
    }
    for (int64_t a = 1; a < ne; ++a) {
      orc_entry key = ent[a];
      int64_t b = a - 1;
      while (b >= 0 && ent[b].row > key.row) {
        ent[b + 1] = ent[b];
        --b;
      }
      ent[b + 1] = key;
    }
    if (rowidx && vals && nnz + ne <= cap) {
      for (int64_t e = 0; e < ne; ++e) {
        rowidx[nnz + e] = (int32_t)ent[e].row;
        vals[2 * (nnz + e)] = ent[e].acc.re;
        vals[2 * (nnz + e) + 1] = ent[e].acc.im;
      }
    }
    nnz += ne;
  }
  colptr[n] = (int32_t)nnz;
  free(ent);
  return nnz;
}

/* ======================================================================
 * Lazy apply, OperatorSum::map_acc (operator_abel.hpp:36-42) over
 * Operator::map_acc (operator.hpp:215-232): for each term k and source i,
 * expand with init = (c_k * coeff) * from_i and scatter to[target] += leaf,
 * dropping a leaf with |leaf| <= eps AFTER the scaling by from_i (:228).
 * ====================================================================== */
void orc_map_acc(const uint64_t* basis, int64_t n, const qd_term* terms, int nterms,
                 const qd_op* ops, const double* from, double* to, const double* coeff) {
  enum { MAXL = 1 << QD_MAX_OPS_PER_TERM };
  uint64_t lidx[MAXL];
  double lval[2 * MAXL];
  oc cf = c_make(coeff[0], coeff[1]);
  for (int k = 0; k < nterms; ++k) {
    const qd_term* t = &terms[k];
    oc ck = c_mul(c_make(t->coeff[0], t->coeff[1]), cf);
    for (int64_t i = 0; i < n; ++i) {
      oc init = c_mul(ck, c_make(from[2 * i], from[2 * i + 1]));
      double in[2] = {init.re, init.im};
      int nl = orc_expand_term(ops + t->op_begin, t->n_ops, basis[i], in, lidx, lval, MAXL);
      for (int j = 0; j < nl; ++j) {
        if (lidx[j] == 0) continue;
        int64_t r = orc_rank(basis, n, lidx[j] - 1);
        if (r >= n) continue;
        oc v = c_make(lval[2 * j], lval[2 * j + 1]);
        if (!(c_abs(v) > ORC_EPS)) continue;
        to[2 * r] += v.re;
        to[2 * r + 1] += v.im;
      }
    }
  }
}

/* ======================================================================
 * Dense-vector helpers (Eigen expressions used by sparse_algorithm.hpp).
 * ====================================================================== */

/* Eigen ColMajor sparse * dense vector: y = 0; for each column j, scatter
 * y[row] += val * x_j (sparse_algorithm.hpp:378,404; state.hpp:41). */
void orc_spmv_csc(int64_t n, const int32_t* colptr, const int32_t* rowidx, const double* vals,
                  const double* x, double* y) {
  memset(y, 0, sizeof(double) * 2 * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    oc xj = c_make(x[2 * j], x[2 * j + 1]);
    for (int32_t e = colptr[j]; e < colptr[j + 1]; ++e) {
      oc p = c_mul(c_make(vals[2 * e], vals[2 * e + 1]), xj);
      y[2 * rowidx[e]] += p.re;
      y[2 * rowidx[e] + 1] += p.im;
    }
  }
}

static oc v_dot(int64_t n, const oc* a, const oc* b) { /* a^H b */
  oc s = c_make(0.0, 0.0);
  for (int64_t i = 0; i < n; ++i) s = c_add(s, c_mul(c_conj(a[i]), b[i]));
  return s;
}
static double v_norm2(int64_t n, const oc* a) { return v_dot(n, a, a).re; }

/* Hermitian::norm (sparse_algorithm.hpp:7-16): Re(psi^H psi). */
double orc_norm2(int64_t n, const double* psi) { return v_norm2(n, (const oc*)psi); }

/* State::eval (state.hpp:33-44): psi^H (M psi), unnormalised. */
void orc_eval(int64_t n, const int32_t* colptr, const int32_t* rowidx, const double* vals,
              const double* psi, double* out) {
  double* y = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  orc_spmv_csc(n, colptr, rowidx, vals, psi, y);
  oc r = v_dot(n, (const oc*)psi, (const oc*)y);
  out[0] = r.re;
  out[1] = r.im;
  free(y);
}

/* ======================================================================
 * Seeded start vector replacing shuffle_state (utils.hpp:129-140), whose
 * random_device seeding is not reproducible: entry i gets
 *   re = U(2i), im = U(2i+1),  U(c) = ((splitmix64((seed<<32) ^ c) >> 11)
 *   * 2^-53) * 2 - 1, i.e. uniform on [-1, 1) like the reference's
 * uniform_real_distribution(-1, 1); then psi /= sqrt(Re psi^H psi).
 * ====================================================================== */
static inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double orc_uniform(uint64_t seed, uint64_t c) {
  uint64_t u = splitmix64((seed << 32) ^ c);
  return ((double)(u >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}
void orc_seeded_state(int64_t n, uint64_t seed, double* psi, int normalise) {
  for (int64_t i = 0; i < n; ++i) {
    psi[2 * i] = orc_uniform(seed, 2 * (uint64_t)i);
    psi[2 * i + 1] = orc_uniform(seed, 2 * (uint64_t)i + 1);
  }
  if (normalise) {
    double s = sqrt(v_norm2(n, (const oc*)psi));
    for (int64_t i = 0; i < 2 * n; ++i) psi[i] /= s;
  }
}

/* ======================================================================
 * Symmetric tridiagonal eigenproblem (the oracle's stand-in for Eigen's
 * SelfAdjointEigenSolver on the Krylov matrix, sparse_algorithm.hpp:571,
 * which reads the lower triangle: diag Re(alpha), sub-diagonal beta).
 * Method: Sturm-sequence bisection for every eigenvalue (ascending) and
 * inverse iteration for requested eigenvectors -- deliberately a different
 * algorithm from the product's implicit QL.
 * ====================================================================== */
static int sturm_count(int m, const double* d, const double* e, double x) {
  /* number of eigenvalues < x */
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < m; ++i) {
    if (q == 0.0) q = 1e-300;
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

int orc_tridiag_eigvals(int m, const double* d, const double* e, double* w) {
  if (m <= 0) return 0;
  double lo = d[0], hi = d[0];
  for (int i = 0; i < m; ++i) {
    double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < m - 1 ? fabs(e[i]) : 0.0);
    if (d[i] - r < lo) lo = d[i] - r;
    if (d[i] + r > hi) hi = d[i] + r;
  }
  if (!isfinite(lo) || !isfinite(hi)) return -1;
  double span = hi - lo;
  lo -= 1e-9 * span + 1e-300;
  hi += 1e-9 * span + 1e-300;
  for (int k = 0; k < m; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (a + b);
      if (mid == a || mid == b) break;
      if (sturm_count(m, d, e, mid) > k)
        b = mid;
      else
        a = mid;
    }
    w[k] = 0.5 * (a + b);
  }
  return 0;
}

/* eigenvector of the tridiagonal for eigenvalue lam by inverse iteration
 * (Thomas algorithm on T - (lam + delta) I), unit 2-norm, first nonzero
 * component positive. */
int orc_tridiag_eigvec(int m, const double* d, const double* e, double lam, double* z) {
  double scale = 0.0;
  for (int i = 0; i < m; ++i) scale = fmax(scale, fabs(d[i]) + (i < m - 1 ? fabs(e[i]) : 0));
  double shift = lam + 1e-13 * (scale > 0 ? scale : 1.0);
  double* c = (double*)malloc(sizeof(double) * m);
  double* dd = (double*)malloc(sizeof(double) * m);
  double* y = (double*)malloc(sizeof(double) * m);
  for (int i = 0; i < m; ++i) z[i] = 1.0 / sqrt((double)m) * (1.0 + 0.01 * i);
  for (int it = 0; it < 6; ++it) {
    /* solve (T - shift) y = z */
    for (int i = 0; i < m; ++i) {
      dd[i] = d[i] - shift;
      y[i] = z[i];
    }
    for (int i = 0; i < m; ++i) {
      if (fabs(dd[i]) < 1e-300) dd[i] = 1e-300;
      if (i < m - 1) {
        c[i] = e[i] / dd[i];
        dd[i + 1] -= c[i] * e[i];
        y[i + 1] -= c[i] * y[i];
      }
    }
    y[m - 1] /= dd[m - 1];
    for (int i = m - 2; i >= 0; --i) y[i] = (y[i] - e[i] * y[i + 1]) / dd[i];
    double s = 0;
    for (int i = 0; i < m; ++i) s += y[i] * y[i];
    s = sqrt(s);
    if (!(s > 0) || !isfinite(s)) {
      free(c);
      free(dd);
      free(y);
      return -1;
    }
    for (int i = 0; i < m; ++i) z[i] = y[i] / s;
  }
  int f = 0;
  while (f < m && fabs(z[f]) < 1e-300) ++f;
  if (f < m && z[f] < 0)
    for (int i = 0; i < m; ++i) z[i] = -z[i];
  free(c);
  free(dd);
  free(y);
  return 0;
}

/* ======================================================================
 * Lanczos ground state exactly as the reference runs it:
 * Arnoldi::getLanczosSolver(idx, k).getGroundState(H, psi, t)
 *   initKrylov (sparse_algorithm.hpp:514-519): v0 = psi / sqrt(norm)
 *   expandBasis<1> (:395-425): 3-term recurrence on a 3-slot ring, no
 *     reorthogonalisation; w -= <v_{i-2},w> v_{i-2}; w -= <v_{i-1},w> v_{i-1};
 *     beta = |w|; stop if beta < eps (act_dim = idx - 1, :421,424)
 *   eigen-solve tridiag(Re alpha, beta) (:571, lower triangle)
 *   getState<herm 1> (:458-473): rerun the recurrence from v0 and
 *     accumulate psi = sum_i y_i v_i (not renormalised).
 * Outputs E0 = lowest Ritz value, psi_out, all Ritz values (evals, >= k),
 * alpha/beta (optional, >= k).  Returns act_dim.
 * ====================================================================== */
static void lanczos_step(int64_t n, const int32_t* cp, const int32_t* ri, const double* va,
                         oc* V[3], int idx, oc* h_alpha, double* h_beta) {
  int i0 = idx % 3, im1 = (idx - 1) % 3, im2 = (idx - 2) % 3;
  orc_spmv_csc(n, cp, ri, va, (const double*)V[im1], (double*)V[i0]);
  if (idx >= 2) {
    oc nh = v_dot(n, V[im2], V[i0]);
    for (int64_t r = 0; r < n; ++r) V[i0][r] = c_sub(V[i0][r], c_mul(nh, V[im2][r]));
  }
  oc nh = v_dot(n, V[im1], V[i0]);
  for (int64_t r = 0; r < n; ++r) V[i0][r] = c_sub(V[i0][r], c_mul(nh, V[im1][r]));
  double nn = sqrt(v_norm2(n, V[i0]));
  for (int64_t r = 0; r < n; ++r) {
    V[i0][r].re /= nn;
    V[i0][r].im /= nn;
  }
  *h_alpha = nh;
  *h_beta = nn;
}

int orc_lanczos_gs(int64_t n, const int32_t* colptr, const int32_t* rowidx, const double* vals,
                   int k, const double* psi_in, double* psi_out, double* E0, double* evals,
                   double* alpha_out, double* beta_out) {
  oc* V[3];
  for (int s = 0; s < 3; ++s) V[s] = (oc*)malloc(sizeof(oc) * (size_t)n);
  oc* V0 = (oc*)malloc(sizeof(oc) * (size_t)n);
  double nrm = sqrt(v_norm2(n, (const oc*)psi_in));
  for (int64_t r = 0; r < n; ++r) {
    V[0][r].re = psi_in[2 * r] / nrm;
    V[0][r].im = psi_in[2 * r + 1] / nrm;
  }
  memcpy(V0, V[0], sizeof(oc) * (size_t)n);
  double* al = (double*)malloc(sizeof(double) * (k + 1));
  double* be = (double*)malloc(sizeof(double) * (k + 1));
  int idx;
  for (idx = 1; idx <= k; ++idx) {
    oc a;
    double b;
    lanczos_step(n, colptr, rowidx, vals, V, idx, &a, &b);
    al[idx - 1] = a.re;
    be[idx - 1] = b;
    if (b < ORC_EPS) break;
  }
  int act = idx - 1;
  if (alpha_out)
    for (int i = 0; i < act; ++i) alpha_out[i] = al[i];
  if (beta_out)
    for (int i = 0; i < act; ++i) beta_out[i] = be[i];
  int rc = act;
  if (act <= 0) {
    rc = 0;
    goto done;
  }
  {
    double* w = (double*)malloc(sizeof(double) * act);
    double* z = (double*)malloc(sizeof(double) * act);
    if (orc_tridiag_eigvals(act, al, be, w) != 0 || orc_tridiag_eigvec(act, al, be, w[0], z) != 0) {
      free(w);
      free(z);
      rc = -1;
      goto done;
    }
    if (E0) *E0 = w[0];
    if (evals)
      for (int i = 0; i < act; ++i) evals[i] = w[i];
    /* second pass (getState<herm 1>) */
    memcpy(V[0], V0, sizeof(oc) * (size_t)n);
    oc* psi = (oc*)psi_out;
    for (int64_t r = 0; r < n; ++r) psi[r] = c_rmul(z[0], V[0][r]);
    for (int i = 1; i < act; ++i) {
      oc a;
      double b;
      lanczos_step(n, colptr, rowidx, vals, V, i, &a, &b);
      for (int64_t r = 0; r < n; ++r) psi[r] = c_add(psi[r], c_rmul(z[i], V[i % 3][r]));
    }
    free(w);
    free(z);
  }
done:
  for (int s = 0; s < 3; ++s) free(V[s]);
  free(V0);
  free(al);
  free(be);
  return rc;
}

/* ======================================================================
 * Matrix exponential for the oracle: scaling and squaring around a Taylor
 * series summed to convergence (the product uses Pade-13; a different
 * method on purpose).  A, E: m x m complex, column-major, interleaved.
 * ====================================================================== */
static void cm_mul(int m, const oc* A, const oc* B, oc* C) {
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      oc s = c_make(0, 0);
      for (int l = 0; l < m; ++l) s = c_add(s, c_mul(A[i + l * m], B[l + j * m]));
      C[i + j * m] = s;
    }
}
int orc_expm(int m, const double* A_in, double* E_out) {
  const oc* A = (const oc*)A_in;
  oc* E = (oc*)E_out;
  double nrm = 0;
  for (int j = 0; j < m; ++j) {
    double s = 0;
    for (int i = 0; i < m; ++i) s += c_abs(A[i + j * m]);
    if (s > nrm) nrm = s;
  }
  if (!isfinite(nrm)) return -1;
  int sq = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    ++sq;
  }
  double sc = ldexp(1.0, -sq);
  oc* X = (oc*)malloc(sizeof(oc) * m * m);
  oc* T = (oc*)malloc(sizeof(oc) * m * m);
  oc* T2 = (oc*)malloc(sizeof(oc) * m * m);
  for (int i = 0; i < m * m; ++i) X[i] = c_rmul(sc, A[i]);
  for (int i = 0; i < m * m; ++i) {
    E[i] = c_make(0, 0);
    T[i] = c_make(0, 0);
  }
  for (int i = 0; i < m; ++i) {
    E[i + i * m] = c_make(1, 0);
    T[i + i * m] = c_make(1, 0);
  }
  for (int p = 1; p <= 40; ++p) {
    cm_mul(m, T, X, T2);
    double tn = 0;
    for (int i = 0; i < m * m; ++i) {
      T[i] = c_rmul(1.0 / p, T2[i]);
      E[i] = c_add(E[i], T[i]);
      tn += c_abs(T[i]);
    }
    if (tn < 1e-18) break;
  }
  for (int s = 0; s < sq; ++s) {
    cm_mul(m, E, E, T2);
    memcpy(E, T2, sizeof(oc) * m * m);
  }
  free(X);
  free(T);
  free(T2);
  return 0;
}

/* ======================================================================
 * Arnoldi real-time step as the reference runs it:
 * Arnoldi::getSolver(idx, k).evolve(H, psi, t, h, k_rt)
 *   (sparse_algorithm.hpp:548-562, expandBasis<0> :372-393)
 *   m = k_rt - 1 MGS-Arnoldi steps; on |w| < eps the loop stops with the
 *   current column kept (act_dim = idx); K = Hessenberg(act_dim);
 *   psi_out = V[:, :act] expm(-i h K)[:, 0] (not renormalised; V[:,0] is
 *   psi / |psi|, so the norm of psi_in is NOT carried, :517).
 * Returns act_dim, or -1.
 * ====================================================================== */
int orc_arnoldi_evolve(int64_t n, const int32_t* colptr, const int32_t* rowidx,
                       const double* vals, int k_rt, double h, const double* psi_in,
                       double* psi_out) {
  int m = k_rt - 1;
  if (m < 1) return -1;
  oc** V = (oc**)malloc(sizeof(oc*) * (m + 1));
  for (int j = 0; j <= m; ++j) V[j] = (oc*)malloc(sizeof(oc) * (size_t)n);
  oc* K = (oc*)calloc((size_t)(m + 1) * (m + 1), sizeof(oc));
  double nrm = sqrt(v_norm2(n, (const oc*)psi_in));
  for (int64_t r = 0; r < n; ++r) {
    V[0][r].re = psi_in[2 * r] / nrm;
    V[0][r].im = psi_in[2 * r + 1] / nrm;
  }
  int idx;
  for (idx = 1; idx <= m; ++idx) {
    orc_spmv_csc(n, colptr, rowidx, vals, (const double*)V[idx - 1], (double*)V[idx]);
    for (int i = 0; i < idx; ++i) {
      oc nh = v_dot(n, V[i], V[idx]);
      for (int64_t r = 0; r < n; ++r) V[idx][r] = c_sub(V[idx][r], c_mul(nh, V[i][r]));
      K[i + (idx - 1) * (m + 1)] = nh;
    }
    double nn = sqrt(v_norm2(n, V[idx]));
    for (int64_t r = 0; r < n; ++r) {
      V[idx][r].re /= nn;
      V[idx][r].im /= nn;
    }
    K[idx + (idx - 1) * (m + 1)] = c_make(nn, 0);
    if (nn < ORC_EPS) {
      ++idx;
      break;
    }
  }
  int act = idx - 1;
  oc* Km = (oc*)malloc(sizeof(oc) * act * act);
  oc* E = (oc*)malloc(sizeof(oc) * act * act);
  for (int j = 0; j < act; ++j)
    for (int i = 0; i < act; ++i) {
      oc kv = K[i + j * (m + 1)];
      /* -i * h * K */
      Km[i + j * act] = c_make(h * kv.im, -h * kv.re);
    }
  int rc = orc_expm(act, (const double*)Km, (double*)E);
  if (rc == 0) {
    oc* out = (oc*)psi_out;
    for (int64_t r = 0; r < n; ++r) {
      oc s = c_make(0, 0);
      for (int j = 0; j < act; ++j) s = c_add(s, c_mul(V[j][r], E[j]));
      out[r] = s;
    }
  }
  for (int j = 0; j <= m; ++j) free(V[j]);
  free(V);
  free(K);
  free(Km);
  free(E);
  return rc == 0 ? act : -1;
}

/* ======================================================================
 * Hermitian::LanczosSolver (sparse_algorithm.hpp:236-313), as the reference
 * runs it, two vectors v, w taking turns:
 *   pass 1, getKrylovReprAd (:67-121): v = psi / sqrt(norm(psi));
 *     a_0 = v^H H v; w = H v - a_0 v; n = |w|^2; w /= sqrt(n); then while
 *     n > eps (1e-10, qudrip.hpp:18; n is the SQUARED norm): with x the
 *     newer vector and y the other, a = x^H H x (complex), b = sqrt(n),
 *     y = H x - a x - sqrt(n) y, n = |y|^2, y /= sqrt(n); the loop breaks
 *     once the count of a exceeds kry_dim (idx > kry_dim, :104-106).
 *     err = sqrt(n) (:119).
 *   Krylov matrix: a on the diagonal, b beside it; the eigen-solver reads
 *     the real diagonal.
 *   mode 0, getGroundState (:285-300): coefficients = eigenvector of the
 *     lowest eigenvalue (inverse iteration, orc_tridiag_eigvec);
 *   mode 1, evolve (:257-283): coefficients = first column of
 *     exp(-i h T), from orc_expm on the dense tridiagonal (the reference
 *     goes through the eigen-decomposition; same matrix function).
 *   pass 2, getKrylovGS (:166-233): rerun the recurrence from psi with
 *     alpha = Re(x^H H x) and accumulate out = sum_i c_i q_i (the vector the
 *     reference computes after the last term is not used and not formed).
 * Outputs: psi_out; alpha (complex, dim); beta[i] = b_i (dim - 1 of them)
 * then beta[dim - 1] = err; evals (dim, ascending); *err.  Returns dim (the
 * count of a), or -1.
 * ====================================================================== */
int orc_hlanczos(int64_t n, const int32_t* colptr, const int32_t* rowidx, const double* vals,
                 int kry_dim, int mode, double h, const double* psi_in, double* psi_out,
                 double* alpha_out, double* beta_out, double* evals, double* err_out) {
  if (kry_dim < 1 || n < 1) return -1;
  oc* v = (oc*)calloc((size_t)n, sizeof(oc));
  oc* w = (oc*)calloc((size_t)n, sizeof(oc));
  oc* t = (oc*)malloc(sizeof(oc) * (size_t)n);
  oc* a = (oc*)malloc(sizeof(oc) * (size_t)(kry_dim + 2));
  double* b = (double*)malloc(sizeof(double) * (size_t)(kry_dim + 2));
  int rc = -1;
  double nrm = sqrt(v_norm2(n, (const oc*)psi_in));
  /* ---- pass 1 */
  for (int64_t r = 0; r < n; ++r) {
    v[r].re = psi_in[2 * r] / nrm;
    v[r].im = psi_in[2 * r + 1] / nrm;
  }
  int na = 0, nb = 0;
  orc_spmv_csc(n, colptr, rowidx, vals, (const double*)v, (double*)t);
  a[na++] = v_dot(n, v, t);
  for (int64_t r = 0; r < n; ++r) w[r] = c_sub(t[r], c_mul(a[0], v[r]));
  double nn = v_norm2(n, w);
  for (int64_t r = 0; r < n; ++r) {
    w[r].re /= sqrt(nn);
    w[r].im /= sqrt(nn);
  }
  int use_v = 0, idx = 1;
  while (nn > ORC_EPS) {
    oc* x = use_v ? v : w; /* the newer vector */
    oc* y = use_v ? w : v;
    orc_spmv_csc(n, colptr, rowidx, vals, (const double*)x, (double*)t);
    oc ai = v_dot(n, x, t);
    a[na++] = ai;
    b[nb++] = sqrt(nn);
    for (int64_t r = 0; r < n; ++r) y[r] = c_sub(c_sub(t[r], c_mul(ai, x[r])), c_rmul(sqrt(nn), y[r]));
    nn = v_norm2(n, y);
    for (int64_t r = 0; r < n; ++r) {
      y[r].re /= sqrt(nn);
      y[r].im /= sqrt(nn);
    }
    use_v = !use_v;
    ++idx;
    if (idx > kry_dim) break;
  }
  const int dim = na;
  const double err = sqrt(nn);
  double* d = (double*)malloc(sizeof(double) * dim);
  double* e = (double*)malloc(sizeof(double) * (dim > 1 ? dim : 1));
  double* ev = (double*)malloc(sizeof(double) * dim);
  oc* c = (oc*)calloc((size_t)dim, sizeof(oc));
  for (int i = 0; i < dim; ++i) d[i] = a[i].re;
  for (int i = 0; i + 1 < dim; ++i) e[i] = b[i];
  if (orc_tridiag_eigvals(dim, d, e, ev) != 0) goto done;
  if (mode == 0) {
    double* z = (double*)malloc(sizeof(double) * dim);
    int r2 = orc_tridiag_eigvec(dim, d, e, ev[0], z);
    for (int i = 0; i < dim; ++i) c[i] = c_make(z[i], 0.0);
    free(z);
    if (r2 != 0) goto done;
  } else {
    oc* T = (oc*)calloc((size_t)dim * dim, sizeof(oc));
    oc* E = (oc*)malloc(sizeof(oc) * (size_t)dim * dim);
    for (int i = 0; i < dim; ++i) {
      T[i + i * dim] = c_make(0.0, -h * d[i]); /* -i h T */
      if (i + 1 < dim) {
        T[(i + 1) + i * dim] = c_make(0.0, -h * e[i]);
        T[i + (i + 1) * dim] = c_make(0.0, -h * e[i]);
      }
    }
    int r2 = orc_expm(dim, (const double*)T, (double*)E);
    for (int i = 0; i < dim; ++i) c[i] = E[i]; /* column 0 */
    free(T);
    free(E);
    if (r2 != 0) goto done;
  }
  /* ---- pass 2 */
  {
    oc* out = (oc*)psi_out;
    for (int64_t r = 0; r < n; ++r) {
      v[r].re = psi_in[2 * r] / nrm;
      v[r].im = psi_in[2 * r + 1] / nrm;
    }
    for (int64_t r = 0; r < n; ++r) out[r] = c_mul(c[0], v[r]);
    if (dim > 1) {
      orc_spmv_csc(n, colptr, rowidx, vals, (const double*)v, (double*)t);
      double al = v_dot(n, v, t).re;
      for (int64_t r = 0; r < n; ++r) w[r] = c_sub(t[r], c_rmul(al, v[r]));
      nn = v_norm2(n, w);
      for (int64_t r = 0; r < n; ++r) {
        w[r].re /= sqrt(nn);
        w[r].im /= sqrt(nn);
      }
    }
    for (int i = 1; i < dim; ++i) {
      oc* x = (i % 2) ? w : v; /* odd: w is the newer vector */
      oc* y = (i % 2) ? v : w;
      for (int64_t r = 0; r < n; ++r) out[r] = c_add(out[r], c_mul(c[i], x[r]));
      if (i + 1 == dim) break;
      orc_spmv_csc(n, colptr, rowidx, vals, (const double*)x, (double*)t);
      double al = v_dot(n, x, t).re;
      for (int64_t r = 0; r < n; ++r) y[r] = c_sub(c_sub(t[r], c_rmul(al, x[r])), c_rmul(sqrt(nn), y[r]));
      nn = v_norm2(n, y);
      for (int64_t r = 0; r < n; ++r) {
        y[r].re /= sqrt(nn);
        y[r].im /= sqrt(nn);
      }
    }
  }
  if (alpha_out)
    for (int i = 0; i < dim; ++i) {
      alpha_out[2 * i] = a[i].re;
      alpha_out[2 * i + 1] = a[i].im;
    }
  if (beta_out) {
    for (int i = 0; i + 1 < dim; ++i) beta_out[i] = b[i];
    beta_out[dim - 1] = err;
  }
  if (evals)
    for (int i = 0; i < dim; ++i) evals[i] = ev[i];
  if (err_out) *err_out = err;
  rc = dim;
done:
  free(v);
  free(w);
  free(t);
  free(a);
  free(b);
  free(d);
  free(e);
  free(ev);
  free(c);
  return rc;
}
