// smalleig.hpp -- lowest eigenpair of the small symmetric matrices of the
// eigensolvers (the projected Davidson matrix, the Lanczos tridiagonal).
//
// Only the lowest eigenvector is needed, so instead of accumulating every
// eigenvector (O(n^3)): Householder reduction to tridiagonal form keeping
// the reflectors, the eigenvalues of the tridiagonal matrix by the QL
// iteration with implicit shifts (values only, O(n^2)), the lowest
// eigenvector of the tridiagonal matrix by inverse iteration, and the
// reflectors applied back (O(n^2)).  The methods are the textbook ones
// (Householder reduction and the implicit QL iteration: Martin, Reinsch and
// Wilkinson, Numer. Math. 11 (1968); inverse iteration: Wilkinson, The
// Algebraic Eigenvalue Problem, 1965).  Host code, header-only.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "qsx_internal.hpp"

namespace qsx {

// eigenvalues of the symmetric tridiagonal matrix (d diagonal, e[0..n-2]
// subdiagonal) by the QL iteration with implicit shifts; d and e are
// overwritten, d ends up holding the eigenvalues (unsorted)
inline void tql_values(int n, std::vector<double>& d, std::vector<double>& e) {
  if (n <= 1) return;
  e.resize(n);
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0;; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (iter >= 64) fail(QSX_ERR_ARG, "tridiagonal eigenproblem did not converge");
      // shift from the leading 2x2 block, chased up from the split at m
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      bool split = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::sqrt(f * f + g * g);
        e[i + 1] = r;
        if (r == 0.0) {  // the matrix splits here
          d[i + 1] -= p;
          e[m] = 0.0;
          split = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
      }
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

// lowest eigenvalue of the symmetric tridiagonal matrix (d, e[0..n-2])
inline double tri_min_eigenvalue(int n, const double* d, const double* e) {
  std::vector<double> w(d, d + n), f(e, e + std::max(0, n - 1));
  tql_values(n, w, f);
  return *std::min_element(w.begin(), w.end());
}

// lowest eigenvalue and its eigenvector y (unit norm) of the symmetric
// tridiagonal matrix (d, e[0..n-2])
inline void tri_lowest(int n, const double* d, const double* e, double& lam, std::vector<double>& y) {
  y.assign(n, 1.0);
  lam = tri_min_eigenvalue(n, d, e);
  if (n == 1) return;
  // inverse iteration with the shift just below the lowest eigenvalue:
  // T - sigma I is positive definite, so its LDL^T factors need no pivoting
  double scale = 0.0;
  for (int i = 0; i < n; ++i)
    scale = std::max(scale, std::fabs(d[i]) + (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0));
  const double sigma = lam - 1e-10 * std::max(scale, DBL_MIN);
  std::vector<double> D(n), L(n, 0.0);
  D[0] = d[0] - sigma;
  for (int i = 1; i < n; ++i) {
    L[i - 1] = e[i - 1] / D[i - 1];
    D[i] = d[i] - sigma - L[i - 1] * e[i - 1];
  }
  for (int it = 0; it < 3; ++it) {
    for (int i = 1; i < n; ++i) y[i] -= L[i - 1] * y[i - 1];   // L z = y
    for (int i = 0; i < n; ++i) y[i] /= D[i];                   // D t = z
    for (int i = n - 2; i >= 0; --i) y[i] -= L[i] * y[i + 1];  // L^T x = t
    double nrm = 0.0;
    for (double x : y) nrm += x * x;
    nrm = std::sqrt(nrm);
    for (double& x : y) x /= nrm;
  }
}

// lowest eigenvalue and eigenvector of the symmetric n x n matrix a
// (row-major)
inline void sym_lowest(int n, std::vector<double> a, double& lam, std::vector<double>& y) {
  std::vector<std::vector<double>> U;  // reflectors u_k (entries k+1..n-1)
  std::vector<double> H;               // u_k^T u_k / 2
  std::vector<double> u(n), p(n);
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += a[(size_t)i * n + k] * a[(size_t)i * n + k];
    alpha = std::sqrt(alpha);
    const double x0 = a[(size_t)(k + 1) * n + k];
    const double beta = x0 >= 0.0 ? -alpha : alpha;
    const double h = beta * beta - x0 * beta;
    if (h <= 0.0) continue;  // column already reduced
    std::fill(u.begin(), u.end(), 0.0);
    for (int i = k + 2; i < n; ++i) u[i] = a[(size_t)i * n + k];
    u[k + 1] = x0 - beta;
    // trailing block (rows / cols k..n-1): A <- H A H, H = I - u u^T / h:
    // p = A u / h, p <- p - (u^T p / 2h) u, A <- A - u p^T - p u^T
    double K = 0.0;
    for (int i = k; i < n; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < n; ++j) t += a[(size_t)i * n + j] * u[j];
      p[i] = t / h;
      K += u[i] * p[i];
    }
    K /= 2.0 * h;
    for (int i = k; i < n; ++i) p[i] -= K * u[i];
    for (int i = k; i < n; ++i)
      for (int j = k; j < n; ++j) a[(size_t)i * n + j] -= u[i] * p[j] + p[i] * u[j];
    U.push_back(u);
    H.push_back(h);
  }
  std::vector<double> d(n), e(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = a[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = a[(size_t)(i + 1) * n + i];
  tri_lowest(n, d.data(), e.data(), lam, y);
  // y <- H_0 H_1 ... y (the reflectors in reverse order)
  for (int r = (int)U.size() - 1; r >= 0; --r) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += U[r][i] * y[i];
    t /= H[r];
    for (int i = 0; i < n; ++i) y[i] -= t * U[r][i];
  }
}

}  // namespace qsx
