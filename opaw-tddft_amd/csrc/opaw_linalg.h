// Host dense linear algebra of the small subspace problems: eigenpairs of a
// Hermitian matrix by Householder reduction to a real tridiagonal matrix and
// the implicit QL iteration with Wilkinson shifts (the textbook route of
// Golub and Van Loan, sections 8.3-8.4), O(n^3).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

namespace opaw {

// A (n x n, row-major, Hermitian) -> ascending eigenvalues w and
// eigenvectors as the columns of V (row-major: V[i * n + k] is component i of
// eigenvector k)
inline void herm_eig_tridiag(int n, std::vector<std::complex<double>> A, std::vector<double>& w,
                             std::vector<std::complex<double>>& V) {
  using C = std::complex<double>;
  w.assign(n, 0.0);
  V.assign((size_t)n * n, C(0.0));
  if (n == 0) return;
  // Q accumulates the Householder reflections (row-major), A <- Q^H A Q
  std::vector<C> Q((size_t)n * n, C(0.0));
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  std::vector<C> v(n), p(n), q(n);
  for (int k = 0; k + 2 < n; ++k) {
    double xn = 0.0;
    for (int i = k + 1; i < n; ++i) xn += std::norm(A[(size_t)i * n + k]);
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const C x0 = A[(size_t)(k + 1) * n + k];
    const C ph = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : C(1.0);
    const C alpha = -ph * xn;
    std::fill(v.begin(), v.end(), C(0.0));
    for (int i = k + 1; i < n; ++i) v[i] = A[(size_t)i * n + k];
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
    if (vn == 0.0) continue;
    vn = std::sqrt(vn);
    for (int i = k + 1; i < n; ++i) v[i] /= vn;
    // H = I - 2 v v^H;  A <- H A H:  p = A v, K = v^H p, q = p - K v,
    // A <- A - 2 v q^H - 2 q v^H
    for (int i = 0; i < n; ++i) {
      C s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * v[j];
      p[i] = s;
    }
    C K = 0.0;
    for (int i = k + 1; i < n; ++i) K += std::conj(v[i]) * p[i];
    for (int i = 0; i < n; ++i) q[i] = p[i] - K * v[i];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        A[(size_t)i * n + j] -= 2.0 * (v[i] * std::conj(q[j]) + q[i] * std::conj(v[j]));
    // Q <- Q H
    for (int i = 0; i < n; ++i) {
      C s = 0.0;
      for (int j = k + 1; j < n; ++j) s += Q[(size_t)i * n + j] * v[j];
      for (int j = k + 1; j < n; ++j) Q[(size_t)i * n + j] -= 2.0 * s * std::conj(v[j]);
    }
  }
  // real tridiagonal: d (diagonal), e (sub-diagonal magnitudes) after the
  // phase change D with D_{i+1} = D_i * phase(A_{i+1,i})
  std::vector<double> d(n), e(n, 0.0);
  std::vector<C> D(n, C(1.0));
  for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i].real();
  for (int i = 0; i + 1 < n; ++i) {
    const C sub = A[(size_t)(i + 1) * n + i];
    const double m = std::abs(sub);
    e[i] = m;
    D[i + 1] = D[i] * (m > 0.0 ? sub / m : C(1.0));
  }
  // implicit QL on (d, e) with the eigenvectors of the tridiagonal in Z
  std::vector<double> Z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m == l || ++iter > 60) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      for (; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          f = Z[(size_t)k * n + i + 1];
          Z[(size_t)k * n + i + 1] = s * Z[(size_t)k * n + i] + c * f;
          Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * f;
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  // eigenvectors: (Q D) Z
  for (int k = 0; k < n; ++k) {
    const int o = ord[k];
    w[k] = d[o];
    for (int i = 0; i < n; ++i) {
      C s = 0.0;
      for (int j = 0; j < n; ++j) s += Q[(size_t)i * n + j] * D[j] * Z[(size_t)j * n + o];
      V[(size_t)i * n + k] = s;
    }
  }
}

}  // namespace opaw
