/* spalg.c -- oracle: arithmetic between stored matrices (Eigen ColMajor
 * SparseMatrix<complex<double>>, the reference's SpMatrix, matrix.hpp:11),
 * as the reference's callers write it (programs/hubbard_diss.cpp:106-159,
 * 443-446, 514-529).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the device SpMatrix algebra; the
 * product never links it.
 *
 * Restated rules (Eigen 3.3/3.4, the reference's dependency, not vendored):
 *  - A + B / A - B (SparseCwiseBinaryOp.h, the union iterator): for every
 *    column, rows of both operands in ascending order; both present ->
 *    op(a, b); one side only -> op(a, 0) or op(0, b); nothing is pruned.
 *    `A += B` and `A -= B` assign A + B and A - B.
 *  - c * A (scalar_product_op): c * a per entry; a double c multiplies both
 *    parts (std::complex operator*(double, complex)).
 *  - SpMatrix(A.adjoint()) (assignment of a transposed expression): counts
 *    per row, then the columns of A in order append (j, conj a) to the new
 *    column of their row, so rows ascend inside each new column.
 *  - A * B (ConservativeSparseSparseProduct.h): for column j, each entry
 *    (k, y) of B's column j in order, each entry (i, x) of A's column k in
 *    order: the first visit of row i sets v_i = x * y, later visits add
 *    x * y; the column's rows are then sorted ascending.
 * Complex products are (ac - bd, ad + bc), sums component-wise, no FMA.
 * Pins (tests/test_oracle.py): dense numpy results of the same expressions
 * with dyadic values (every operation exact, so the order cannot matter);
 * hand-derived signed-zero and explicit-zero cases; the transpose and the
 * product against numpy on random patterns.  The summation order of the
 * product for non-dyadic values is Eigen's documented loop, restated, and
 * cannot be checked against Eigen here (parity unpinned for those bits).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double re, im;
} oc;

static oc oc_mul(oc a, oc b) {
  oc r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* A + B (sign = 1) or A - B (sign = -1); out arrays may be NULL (count
 * only).  Returns nnz. */
int64_t orc_csr_add(int64_t n, const int32_t* ao, const int32_t* ai, const double* av, const int32_t* bo,
                    const int32_t* bi, const double* bv, int sign, int32_t* co, int32_t* ci, double* cv) {
  int64_t w = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (co) co[j] = (int32_t)w;
    int32_t p = ao[j], pe = ao[j + 1], q = bo[j], qe = bo[j + 1];
    while (p < pe || q < qe) {
      int32_t r;
      oc a = {0.0, 0.0}, b = {0.0, 0.0};
      if (p < pe && q < qe && ai[p] == bi[q]) {
        r = ai[p];
        a.re = av[2 * p]; a.im = av[2 * p + 1]; ++p;
        b.re = bv[2 * q]; b.im = bv[2 * q + 1]; ++q;
      } else if (p < pe && (q >= qe || ai[p] < bi[q])) {
        r = ai[p];
        a.re = av[2 * p]; a.im = av[2 * p + 1]; ++p;
      } else {
        r = bi[q];
        b.re = bv[2 * q]; b.im = bv[2 * q + 1]; ++q;
      }
      if (ci) {
        ci[w] = r;
        cv[2 * w] = sign > 0 ? a.re + b.re : a.re - b.re;
        cv[2 * w + 1] = sign > 0 ? a.im + b.im : a.im - b.im;
      }
      ++w;
    }
  }
  if (co) co[n] = (int32_t)w;
  return w;
}

/* c * A over the values (pattern unchanged). */
void orc_csr_scale(int64_t nnz, double c_re, double c_im, int real_scalar, const double* av, double* cv) {
  const oc c = {c_re, c_im};
  for (int64_t e = 0; e < nnz; ++e) {
    const oc a = {av[2 * e], av[2 * e + 1]};
    oc r;
    if (real_scalar) {
      r.re = c_re * a.re;
      r.im = c_re * a.im;
    } else {
      r = oc_mul(c, a);
    }
    cv[2 * e] = r.re;
    cv[2 * e + 1] = r.im;
  }
}

/* SpMatrix(A.adjoint()); out arrays of n+1 and nnz. */
void orc_csr_adjoint(int64_t n, const int32_t* ao, const int32_t* ai, const double* av, int32_t* co, int32_t* ci,
                     double* cv) {
  const int64_t nnz = ao[n];
  int64_t* pos = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t e = 0; e < nnz; ++e) pos[ai[e] + 1]++;
  for (int64_t i = 0; i < n; ++i) pos[i + 1] += pos[i];
  for (int64_t i = 0; i <= n; ++i) co[i] = (int32_t)pos[i];
  for (int64_t j = 0; j < n; ++j)
    for (int32_t e = ao[j]; e < ao[j + 1]; ++e) {
      const int64_t w = pos[ai[e]]++;
      ci[w] = (int32_t)j;
      cv[2 * w] = av[2 * e];
      cv[2 * w + 1] = -av[2 * e + 1];
    }
  free(pos);
}

static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* A * B; out arrays may be NULL (count only).  Returns nnz. */
int64_t orc_csr_multiply(int64_t n, const int32_t* ao, const int32_t* ai, const double* av, const int32_t* bo,
                         const int32_t* bi, const double* bv, int32_t* co, int32_t* ci, double* cv) {
  char* mask = (char*)calloc((size_t)n + 1, 1);
  oc* val = (oc*)malloc(sizeof(oc) * ((size_t)n + 1));
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  int64_t w = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (co) co[j] = (int32_t)w;
    int64_t cnt = 0;
    for (int32_t q = bo[j]; q < bo[j + 1]; ++q) {
      const oc y = {bv[2 * q], bv[2 * q + 1]};
      const int32_t k = bi[q];
      for (int32_t p = ao[k]; p < ao[k + 1]; ++p) {
        const int32_t i = ai[p];
        const oc x = {av[2 * p], av[2 * p + 1]};
        const oc xy = oc_mul(x, y);
        if (!mask[i]) {
          mask[i] = 1;
          val[i] = xy;
          idx[cnt++] = i;
        } else {
          val[i].re += xy.re;
          val[i].im += xy.im;
        }
      }
    }
    qsort(idx, (size_t)cnt, sizeof(int32_t), cmp_i32);
    for (int64_t t = 0; t < cnt; ++t) {
      const int32_t i = idx[t];
      if (ci) {
        ci[w] = i;
        cv[2 * w] = val[i].re;
        cv[2 * w + 1] = val[i].im;
      }
      mask[i] = 0;
      ++w;
    }
  }
  if (co) co[n] = (int32_t)w;
  free(mask);
  free(val);
  free(idx);
  return w;
}
