// host_linalg.cpp -- H1/H2: the small dense FP64 problems of the Krylov
// methods, solved on the host (the Krylov dimension is <= a few hundred).
//
// H1  symmetric tridiagonal eigenproblem (replaces Eigen's
//     SelfAdjointEigenSolver on the Krylov matrix, sparse_algorithm.hpp:571):
//     implicit QL iteration with Wilkinson-type shifts on (d, e), eigenvector
//     rotations accumulated into Z, then eigenpairs sorted ascending.
// H2  matrix exponential (replaces Eigen's MatrixExp, matrix.hpp:20, used at
//     sparse_algorithm.hpp:559): scaling and squaring with the [13/13] Pade
//     approximant, N. J. Higham, "The scaling and squaring method for the
//     matrix exponential revisited", SIAM J. Matrix Anal. Appl. 26 (2005).
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

#include "qd_internal.hpp"

// d[m] diagonal, e[m-1] sub-diagonal.  w[m] eigenvalues ascending; when
// want_z, z[m*m] column-major eigenvectors (column i <-> w[i]), each with its
// largest-magnitude component positive.  Returns 0, or 1 if an eigenvalue
// did not converge in 60 sweeps (NaN input).
int qd_tridiag_eig(int m, const double* d_in, const double* e_in, double* w, double* z, int want_z) {
  if (m <= 0) return 0;
  std::vector<double> d(d_in, d_in + m), e(m, 0.0);
  for (int i = 0; i + 1 < m; ++i) e[i] = e_in[i];
  for (int i = 0; i < m; ++i)
    if (!std::isfinite(d[i]) || !std::isfinite(e[i])) return 1;
  std::vector<double> Z;
  if (want_z) {
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[(size_t)i * m + i] = 1.0;
  }
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    for (;;) {
      // find the first negligible off-diagonal at or after l
      int mm = l;
      for (; mm < m - 1; ++mm) {
        double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
        if (std::fabs(e[mm]) <= 1e-16 * dd || std::fabs(e[mm]) < 1e-300) break;
      }
      if (mm == l) break;
      if (++iter > 60) return 1;
      // shift from the leading 2x2 block
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      bool deflated = false;
      for (i = mm - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[mm] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (want_z) {
          double* zi = &Z[(size_t)i * m];
          double* zi1 = &Z[(size_t)(i + 1) * m];
          for (int k = 0; k < m; ++k) {
            double t = zi1[k];
            zi1[k] = s * zi[k] + c * t;
            zi[k] = c * zi[k] - s * t;
          }
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[mm] = 0.0;
    }
  }
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int i = 0; i < m; ++i) w[i] = d[ord[i]];
  if (want_z) {
    for (int i = 0; i < m; ++i) {
      const double* src = &Z[(size_t)ord[i] * m];
      int big = 0;
      for (int k = 1; k < m; ++k)
        if (std::fabs(src[k]) > std::fabs(src[big])) big = k;
      double sg = src[big] < 0 ? -1.0 : 1.0;
      for (int k = 0; k < m; ++k) z[(size_t)i * m + k] = sg * src[k];
    }
  }
  return 0;
}

// ------------------------------------------------------------------ expm
namespace {
using C = std::complex<double>;

void mat_mul(int m, const C* A, const C* B, C* Out) {
  for (int j = 0; j < m; ++j) {
    C* oc = Out + (size_t)j * m;
    for (int i = 0; i < m; ++i) oc[i] = 0.0;
    for (int l = 0; l < m; ++l) {
      C b = B[(size_t)j * m + l];
      if (b == C(0.0, 0.0)) continue;
      const C* al = A + (size_t)l * m;
      for (int i = 0; i < m; ++i) oc[i] += al[i] * b;
    }
  }
}

// solve P X = Q (m x m, column-major) by LU with partial pivoting; Q <- X
int lu_solve(int m, std::vector<C>& P, std::vector<C>& Q) {
  std::vector<int> piv(m);
  for (int k = 0; k < m; ++k) {
    int p = k;
    double best = std::abs(P[(size_t)k * m + k]);
    for (int i = k + 1; i < m; ++i) {
      double v = std::abs(P[(size_t)k * m + i]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (!(best > 0.0) || !std::isfinite(best)) return 1;
    piv[k] = p;
    if (p != k) {
      for (int j = 0; j < m; ++j) std::swap(P[(size_t)j * m + k], P[(size_t)j * m + p]);
      for (int j = 0; j < m; ++j) std::swap(Q[(size_t)j * m + k], Q[(size_t)j * m + p]);
    }
    C inv = 1.0 / P[(size_t)k * m + k];
    for (int i = k + 1; i < m; ++i) {
      C f = P[(size_t)k * m + i] * inv;
      P[(size_t)k * m + i] = f;
      if (f == C(0.0, 0.0)) continue;
      for (int j = k + 1; j < m; ++j) P[(size_t)j * m + i] -= f * P[(size_t)j * m + k];
      for (int j = 0; j < m; ++j) Q[(size_t)j * m + i] -= f * Q[(size_t)j * m + k];
    }
  }
  for (int j = 0; j < m; ++j) {
    C* q = &Q[(size_t)j * m];
    for (int i = m - 1; i >= 0; --i) {
      C s = q[i];
      for (int l = i + 1; l < m; ++l) s -= P[(size_t)l * m + i] * q[l];
      q[i] = s / P[(size_t)i * m + i];
    }
  }
  return 0;
}
}  // namespace

int qd_expm_pade13(int m, const cplx* A_in, cplx* E_out) {
  if (m <= 0) return 0;
  static const double b[14] = {64764752532480000.0,
                               32382376266240000.0,
                               7771770303897600.0,
                               1187353796428800.0,
                               129060195264000.0,
                               10559470521600.0,
                               670442572800.0,
                               33522128640.0,
                               1323241920.0,
                               40840800.0,
                               960960.0,
                               16380.0,
                               182.0,
                               1.0};
  const double theta13 = 5.371920351148152;
  size_t mm = (size_t)m * m;
  std::vector<C> A(mm);
  for (size_t i = 0; i < mm; ++i) A[i] = C(A_in[i].x, A_in[i].y);
  double norm1 = 0.0;
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += std::abs(A[(size_t)j * m + i]);
    norm1 = std::max(norm1, s);
  }
  if (!std::isfinite(norm1)) return 1;
  int sq = 0;
  if (norm1 > theta13) sq = (int)std::ceil(std::log2(norm1 / theta13));
  double sc = std::ldexp(1.0, -sq);
  for (auto& v : A) v *= sc;
  std::vector<C> A2(mm), A4(mm), A6(mm), T(mm), U(mm), V(mm);
  mat_mul(m, A.data(), A.data(), A2.data());
  mat_mul(m, A2.data(), A2.data(), A4.data());
  mat_mul(m, A2.data(), A4.data(), A6.data());
  // U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I]
  for (size_t i = 0; i < mm; ++i) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
  mat_mul(m, A6.data(), T.data(), U.data());
  for (size_t i = 0; i < mm; ++i) U[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
  for (int i = 0; i < m; ++i) U[(size_t)i * m + i] += b[1];
  mat_mul(m, A.data(), U.data(), T.data());
  U.swap(T);
  // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
  for (size_t i = 0; i < mm; ++i) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
  mat_mul(m, A6.data(), T.data(), V.data());
  for (size_t i = 0; i < mm; ++i) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
  for (int i = 0; i < m; ++i) V[(size_t)i * m + i] += b[0];
  // (V - U) X = (V + U)
  std::vector<C> P(mm), Q(mm);
  for (size_t i = 0; i < mm; ++i) {
    P[i] = V[i] - U[i];
    Q[i] = V[i] + U[i];
  }
  if (lu_solve(m, P, Q)) return 1;
  for (int s = 0; s < sq; ++s) {
    mat_mul(m, Q.data(), Q.data(), T.data());
    Q.swap(T);
  }
  for (size_t i = 0; i < mm; ++i) {
    if (!std::isfinite(Q[i].real()) || !std::isfinite(Q[i].imag())) return 1;
    E_out[i] = cmake(Q[i].real(), Q[i].imag());
  }
  return 0;
}

// ------------------------------------------------------------------ ABI
// Exposed for tests of the host numerics.
extern "C" int qd_host_tridiag_eig(int m, const double* d, const double* e, double* w, double* z) {
  return qd_tridiag_eig(m, d, e, w, z, z != nullptr) ? QD_E_EIGEN : QD_OK;
}
extern "C" int qd_host_expm(int m, const double* A, double* E) {
  return qd_expm_pade13(m, (const cplx*)A, (cplx*)E) ? QD_E_EIGEN : QD_OK;
}
