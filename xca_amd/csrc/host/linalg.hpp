// Small dense linear algebra on row-major double arrays, used for host-side setup
// (DLR operator construction, Dyson system matrices). Nothing here is on the hot path.
#pragma once
#include <cmath>
#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace xca::la {

// C[m x n] = A[m x k] * B[k x n] (row-major, no aliasing)
inline void matmul(const double *A, const double *B, double *C, int m, int k, int n) {
  for (int i = 0; i < m; ++i) {
    double *c = C + (size_t)i * n;
    for (int j = 0; j < n; ++j) c[j] = 0.0;
    for (int l = 0; l < k; ++l) {
      double a = A[(size_t)i * k + l];
      const double *b = B + (size_t)l * n;
      for (int j = 0; j < n; ++j) c[j] += a * b[j];
    }
  }
}

// In-place LU factorisation with partial pivoting (LAPACK getrf convention: A = P L U,
// unit lower L). piv[i] = row swapped with row i at step i. Throws on an exactly singular pivot.
inline void getrf(double *A, int n, std::vector<int> &piv) {
  piv.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = std::fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(A[(size_t)i * n + k]);
      if (v > amax) { amax = v; p = i; }
    }
    piv[k] = p;
    if (amax == 0.0) throw std::runtime_error("getrf: singular matrix");
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
    double inv = 1.0 / A[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double l = A[(size_t)i * n + k] * inv;
      A[(size_t)i * n + k] = l;
      if (l != 0.0)
        for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
    }
  }
}

// Solve (P L U) X = B in place for B[n x nrhs] row-major (each column is one right-hand side).
inline void getrs(const double *LU, int n, const std::vector<int> &piv, double *B, int nrhs) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k)
      for (int j = 0; j < nrhs; ++j) std::swap(B[(size_t)k * nrhs + j], B[(size_t)piv[k] * nrhs + j]);
  for (int i = 1; i < n; ++i)
    for (int k = 0; k < i; ++k) {
      double l = LU[(size_t)i * n + k];
      if (l != 0.0)
        for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] -= l * B[(size_t)k * nrhs + j];
    }
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) {
      double u = LU[(size_t)i * n + k];
      if (u != 0.0)
        for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] -= u * B[(size_t)k * nrhs + j];
    }
    double inv = 1.0 / LU[(size_t)i * n + i];
    for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] *= inv;
  }
}

// Pivoted, reorthogonalised Gram-Schmidt on the rows of A[m x n] (row-major, destroyed).
// Returns the selected row indices in pivot order. Stops after maxrank rows, or when the largest
// remaining residual row norm falls below tol times the largest initial row norm (tol <= 0
// disables the tolerance test).
inline std::vector<int> pivoted_gram_schmidt_rows(std::vector<double> A, int m, int n, double tol, int maxrank) {
  std::vector<int> perm(m);
  for (int i = 0; i < m; ++i) perm[i] = i;
  std::vector<double> norms(m);
  auto rownorm2 = [&](int i) {
    long double s = 0;
    const double *a = &A[(size_t)i * n];
    for (int j = 0; j < n; ++j) s += (long double)a[j] * a[j];
    return (double)s;
  };
  double nmax2 = 0;
  for (int i = 0; i < m; ++i) nmax2 = std::max(nmax2, norms[i] = rownorm2(i));
  const double abs_tol = tol * std::sqrt(nmax2);
  std::vector<int> sel;
  int kmax = std::min(m, maxrank > 0 ? maxrank : m);
  for (int k = 0; k < kmax; ++k) {
    int p = k;
    for (int i = k + 1; i < m; ++i)
      if (norms[i] > norms[p]) p = i;
    if (tol > 0 && std::sqrt(norms[p]) < abs_tol) break;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
      std::swap(norms[k], norms[p]);
      std::swap(perm[k], perm[p]);
    }
    double *q = &A[(size_t)k * n];
    // reorthogonalise the pivot row against the accepted rows (twice is enough)
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < k; ++i) {
        const double *qi = &A[(size_t)i * n];
        long double d = 0;
        for (int j = 0; j < n; ++j) d += (long double)qi[j] * q[j];
        for (int j = 0; j < n; ++j) q[j] -= (double)d * qi[j];
      }
    double nq = std::sqrt(rownorm2(k));
    if (nq == 0.0) break;
    for (int j = 0; j < n; ++j) q[j] /= nq;
    sel.push_back(perm[k]);
    for (int i = k + 1; i < m; ++i) {
      double *a = &A[(size_t)i * n];
      long double d = 0;
      for (int j = 0; j < n; ++j) d += (long double)q[j] * a[j];
      for (int j = 0; j < n; ++j) a[j] -= (double)d * q[j];
      norms[i] = rownorm2(i);
    }
  }
  return sel;
}

} // namespace xca::la
