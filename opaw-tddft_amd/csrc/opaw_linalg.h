// Host dense linear algebra of the small subspace problems: eigenpairs of a
// Hermitian matrix by Householder reduction to a real tridiagonal matrix and
// the implicit QL iteration with Wilkinson shifts (the textbook route of
// Golub and Van Loan, sections 8.3-8.4), O(n^3), for orthog_phi and diagh.
//
// Complex products are written out in real arithmetic (a std::complex
// product compiles to a library call that handles infinities), the
// reflections touch only the trailing block they change, and the QL
// rotations run on rows of the transposed eigenvector matrix.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <numeric>
#include <vector>

namespace opaw {

namespace la {
using C = std::complex<double>;
// a b
inline C mul(C a, C b) { return C(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()); }
// conj(a) b
inline C cmul(C a, C b) { return C(a.real() * b.real() + a.imag() * b.imag(), a.real() * b.imag() - a.imag() * b.real()); }
// a conj(b)
inline C mulc(C a, C b) { return C(a.real() * b.real() + a.imag() * b.imag(), a.imag() * b.real() - a.real() * b.imag()); }
}  // namespace la

// Householder reduction of a Hermitian A (n x n, row-major, overwritten) to
// a real symmetric tridiagonal matrix: Q^H A Q = D T D^H with
// Q = H_0 H_1 ... H_{n-3}, H_k = I - 2 v_k v_k^H (v_k is row k of Vh, on
// indices k+1.., and used[k] = 0 when step k is the identity), D diagonal
// phases, T = tridiag(e, d, e) with e[i] >= 0 coupling i and i+1.
struct Tridiag {
  std::vector<std::complex<double>> Vh, D;
  std::vector<char> used;
  std::vector<double> d, e;
};

inline void tridiagonalize(int n, std::vector<std::complex<double>>& A, Tridiag& t) {
  using C = std::complex<double>;
  using la::cmul;
  using la::mul;
  using la::mulc;
  t.Vh.assign((size_t)n * n, C(0.0));
  t.used.assign(n, 0);
  std::vector<C>& Vh = t.Vh;
  std::vector<C> p(n), q(n);
  for (int k = 0; k + 2 < n; ++k) {
    C* v = &Vh[(size_t)k * n];
    const int k1 = k + 1;
    double xn = 0.0;
    for (int i = k1; i < n; ++i) xn += std::norm(A[(size_t)i * n + k]);
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const C x0 = A[(size_t)k1 * n + k];
    const C ph = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : C(1.0);
    const C alpha = -ph * xn;
    for (int i = k1; i < n; ++i) v[i] = A[(size_t)i * n + k];
    v[k1] -= alpha;
    double vn = 0.0;
    for (int i = k1; i < n; ++i) vn += std::norm(v[i]);
    if (vn == 0.0) continue;
    vn = 1.0 / std::sqrt(vn);
    for (int i = k1; i < n; ++i) v[i] *= vn;
    // H = I - 2 v v^H with v on [k1, n): column k becomes alpha e_k1, and the
    // trailing block B = A[k1:, k1:] becomes H B H:  p = B v, K = v^H p,
    // q = p - K v, B <- B - 2 v q^H - 2 q v^H
    for (int i = k1; i < n; ++i) {
      const C* row = &A[(size_t)i * n];
      double sr = 0.0, si = 0.0;
      for (int j = k1; j < n; ++j) {
        const C a = row[j], b = v[j];
        sr += a.real() * b.real() - a.imag() * b.imag();
        si += a.real() * b.imag() + a.imag() * b.real();
      }
      p[i] = C(sr, si);
    }
    C K = 0.0;
    for (int i = k1; i < n; ++i) K += cmul(v[i], p[i]);
    for (int i = k1; i < n; ++i) q[i] = p[i] - mul(K, v[i]);
    for (int i = k1; i < n; ++i) {
      C* row = &A[(size_t)i * n];
      const C vi2 = 2.0 * v[i], qi2 = 2.0 * q[i];
      for (int j = k1; j < n; ++j) row[j] -= mulc(vi2, q[j]) + mulc(qi2, v[j]);
    }
    A[(size_t)k1 * n + k] = alpha;
    A[(size_t)k * n + k1] = std::conj(alpha);
    for (int i = k1 + 1; i < n; ++i) {
      A[(size_t)i * n + k] = 0.0;
      A[(size_t)k * n + i] = 0.0;
    }
    t.used[k] = 1;
  }
  // real tridiagonal: d (diagonal), e (sub-diagonal magnitudes) after the
  // phase change D with D_{i+1} = D_i * phase(A_{i+1,i})
  t.d.assign(n, 0.0);
  t.e.assign(n, 0.0);
  t.D.assign(n, C(1.0));
  for (int i = 0; i < n; ++i) t.d[i] = A[(size_t)i * n + i].real();
  for (int i = 0; i + 1 < n; ++i) {
    const C sub = A[(size_t)(i + 1) * n + i];
    const double m = std::abs(sub);
    t.e[i] = m;
    t.D[i + 1] = mul(t.D[i], m > 0.0 ? sub / m : C(1.0));
  }
}

// Implicit QL with Wilkinson shifts on the symmetric tridiagonal (d, e)
// (e[i] couples i and i+1): d becomes the eigenvalues (unsorted).  With Zt,
// the rotations are applied to its rows (Zt[i * n + r] is component r of
// vector i, so a rotation of vectors i and i+1 runs over two contiguous
// rows); without, only the eigenvalues are formed, in O(n^2).
inline void tridiag_ql(int n, std::vector<double>& d, std::vector<double>& e, double* Zt) {
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
        if (Zt) {
          double* zi = &Zt[(size_t)i * n];
          double* zi1 = &Zt[(size_t)(i + 1) * n];
          for (int k = 0; k < n; ++k) {
            const double fz = zi1[k];
            zi1[k] = s * zi[k] + c * fz;
            zi[k] = c * zi[k] - s * fz;
          }
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

// A (n x n, row-major, Hermitian) -> ascending eigenvalues w and
// eigenvectors as the columns of V (row-major: V[i * n + k] is component i of
// eigenvector k)
inline void herm_eig_tridiag(int n, std::vector<std::complex<double>> A, std::vector<double>& w,
                             std::vector<std::complex<double>>& V) {
  using C = std::complex<double>;
  using la::mul;
  w.assign(n, 0.0);
  V.assign((size_t)n * n, C(0.0));
  if (n == 0) return;
  Tridiag td;
  tridiagonalize(n, A, td);
  const std::vector<C>& Vh = td.Vh;
  // Q = H_0 (H_1 (... (H_{n-3} I))): H_k Q = Q - 2 v (v^H Q) on rows and
  // columns k+1.. (the rest of Q is still the identity there)
  std::vector<C> Q((size_t)n * n, C(0.0));
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  std::vector<C> t(n);
  for (int k = n - 3; k >= 0; --k) {
    if (!td.used[k]) continue;
    const int k1 = k + 1;
    const C* v = &Vh[(size_t)k * n];
    for (int j = k1; j < n; ++j) t[j] = 0.0;
    for (int i = k1; i < n; ++i) {
      const C vi = std::conj(v[i]);
      const C* row = &Q[(size_t)i * n];
      for (int j = k1; j < n; ++j) t[j] += mul(vi, row[j]);
    }
    for (int i = k1; i < n; ++i) {
      const C vi2 = 2.0 * v[i];
      C* row = &Q[(size_t)i * n];
      for (int j = k1; j < n; ++j) row[j] -= mul(vi2, t[j]);
    }
  }
  std::vector<double>& d = td.d;
  std::vector<double>& e = td.e;
  const std::vector<C>& D = td.D;
  // implicit QL on (d, e); Zt holds the eigenvectors of the tridiagonal as
  // rows (Zt[i * n + r] is component r of vector i)
  std::vector<double> Zt((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Zt[(size_t)i * n + i] = 1.0;
  tridiag_ql(n, d, e, Zt.data());
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  // eigenvectors: (Q D) Z
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Q[(size_t)i * n + j] = mul(Q[(size_t)i * n + j], D[j]);
  for (int k = 0; k < n; ++k) w[k] = d[ord[k]];
  for (int i = 0; i < n; ++i) {
    const C* qi = &Q[(size_t)i * n];
    for (int k = 0; k < n; ++k) {
      const double* z = &Zt[(size_t)ord[k] * n];
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < n; ++j) {
        sr += qi[j].real() * z[j];
        si += qi[j].imag() * z[j];
      }
      V[(size_t)i * n + k] = C(sr, si);
    }
  }
}

}  // namespace opaw
