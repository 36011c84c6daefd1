// Host dense linear algebra of the small subspace problems: eigenpairs of a
// Hermitian matrix by Householder reduction to a real tridiagonal matrix and
// the implicit QL iteration with Wilkinson shifts (the textbook route of
// Golub and Van Loan, sections 8.3-8.4), O(n^3); the lowest few by QL
// eigenvalues and inverse iteration (sections 8.4-8.5); and the lowest
// pairs of a Hermitian pencil through Cholesky or, for a near-singular
// overlap, a rank-revealing eigendecomposition.
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

// (T - lam I) = P L U for the tridiagonal T = tridiag(e, d, e), Gaussian
// elimination with row interchanges; U has the diagonals u0, u1, u2, and
// diagonal entries smaller than pert in magnitude are raised to pert.
struct TriLU {
  std::vector<double> u0, u1, u2, l;
  std::vector<char> swp;
};

inline void tri_lu(int n, const double* d, const double* e, double lam, double pert, TriLU& f) {
  f.u0.assign(n, 0.0);
  f.u1.assign(n, 0.0);
  f.u2.assign(n, 0.0);
  f.l.assign(n, 0.0);
  f.swp.assign(n, 0);
  auto floor_to = [pert](double x) { return std::fabs(x) >= pert ? x : (x >= 0.0 ? pert : -pert); };
  double dg = d[0] - lam, sp = n > 1 ? e[0] : 0.0;   // the pending row: (dg, sp, 0)
  for (int i = 0; i + 1 < n; ++i) {
    const double c = e[i];                           // row i+1: (c, a1, b1)
    const double a1 = d[i + 1] - lam, b1 = i + 2 < n ? e[i + 1] : 0.0;
    if (std::fabs(c) > std::fabs(dg)) {
      const double mlt = dg / c;
      f.swp[i] = 1;
      f.l[i] = mlt;
      f.u0[i] = floor_to(c);
      f.u1[i] = a1;
      f.u2[i] = b1;
      dg = sp - mlt * a1;
      sp = -mlt * b1;
    } else {
      dg = floor_to(dg);
      const double mlt = c / dg;
      f.l[i] = mlt;
      f.u0[i] = dg;
      f.u1[i] = sp;
      dg = a1 - mlt * sp;
      sp = b1;
    }
  }
  f.u0[n - 1] = floor_to(dg);
}

// x <- (T - lam I)^-1 x with the factors of tri_lu
inline void tri_solve(int n, const TriLU& f, double* x) {
  for (int i = 0; i + 1 < n; ++i) {
    if (f.swp[i]) std::swap(x[i], x[i + 1]);
    x[i + 1] -= f.l[i] * x[i];
  }
  x[n - 1] /= f.u0[n - 1];
  if (n >= 2) x[n - 2] = (x[n - 2] - f.u1[n - 2] * x[n - 1]) / f.u0[n - 2];
  for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - f.u1[i] * x[i + 1] - f.u2[i] * x[i + 2]) / f.u0[i];
}

// The nout lowest eigenpairs of a Hermitian A (n x n, row-major): ascending
// w[nout] and the eigenvectors as the columns of V (row-major n x nout).
// Householder reduction as herm_eig_tridiag; the eigenvalues of the
// tridiagonal by QL without vectors; each wanted eigenvector of the
// tridiagonal by inverse iteration (three solves with T - lam I from a fixed
// pseudo-random start); vectors whose eigenvalues lie within 1e-3 ||T|| of
// each other form a cluster and are orthogonalised against the cluster's
// earlier vectors after every solve, and equal eigenvalues are separated by
// 10 eps |lam| (the scheme of LAPACK's dstein); finally the phases D and the
// reflections are applied to the nout vectors only.  O(n^3) for the
// reduction, O(n^2 nout) for the rest.
inline void herm_eig_lowest(int n, std::vector<std::complex<double>> A, int nout, std::vector<double>& w,
                            std::vector<std::complex<double>>& V) {
  using C = std::complex<double>;
  using la::cmul;
  using la::mul;
  nout = std::max(0, std::min(nout, n));
  w.assign(nout, 0.0);
  V.assign((size_t)n * nout, C(0.0));
  if (n == 0 || nout == 0) return;
  Tridiag td;
  tridiagonalize(n, A, td);
  std::vector<double> lam = td.d, ee = td.e;
  tridiag_ql(n, lam, ee, nullptr);
  std::sort(lam.begin(), lam.end());
  const double* d = td.d.data();
  const double* e = td.e.data();
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    tnorm = std::max(tnorm, std::fabs(d[i]) + (i > 0 ? e[i - 1] : 0.0) + (i + 1 < n ? e[i] : 0.0));
  const double eps = 2.220446049250313e-16;
  const double pert = tnorm > 0.0 ? eps * tnorm : 1e-300;
  const double ortol = 1e-3 * tnorm;
  std::vector<double> Z((size_t)nout * n);   // rows: eigenvectors of T
  TriLU f;
  int c0 = 0;                                // first vector of the current cluster
  double prev = 0.0;
  for (int j = 0; j < nout; ++j) {
    double x = lam[j];
    if (j == 0 || lam[j] - lam[j - 1] > ortol) {
      c0 = j;
    } else {
      const double pt = 10.0 * eps * std::fabs(x);
      if (x - prev < pt) x = prev + pt;
    }
    prev = x;
    tri_lu(n, d, e, x, pert, f);
    double* z = &Z[(size_t)j * n];
    uint64_t st = 0x9E3779B97F4A7C15ull * (uint64_t)(j + 1);
    for (int i = 0; i < n; ++i) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      z[i] = (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    for (int it = 0; it < 3; ++it) {
      tri_solve(n, f, z);
      for (int pass = 0; pass < 2; ++pass)
        for (int k = c0; k < j; ++k) {
          const double* zk = &Z[(size_t)k * n];
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += zk[i] * z[i];
          for (int i = 0; i < n; ++i) z[i] -= dot * zk[i];
        }
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) nrm += z[i] * z[i];
      nrm = 1.0 / std::sqrt(nrm);
      for (int i = 0; i < n; ++i) z[i] *= nrm;
    }
    w[j] = lam[j];
  }
  // y = Q D z: phases, then H_0 (H_1 (... (H_{n-3} y)))
  std::vector<C> y(n);
  for (int j = 0; j < nout; ++j) {
    const double* z = &Z[(size_t)j * n];
    for (int i = 0; i < n; ++i) y[i] = td.D[i] * z[i];
    for (int k = n - 3; k >= 0; --k) {
      if (!td.used[k]) continue;
      const C* v = &td.Vh[(size_t)k * n];
      C s = 0.0;
      for (int i = k + 1; i < n; ++i) s += cmul(v[i], y[i]);
      s *= 2.0;
      for (int i = k + 1; i < n; ++i) y[i] -= mul(s, v[i]);
    }
    for (int i = 0; i < n; ++i) V[(size_t)i * nout + j] = y[i];
  }
}

// Lowest nout eigenpairs of the Hermitian pencil (H, S), S positive
// semi-definite (row-major m x m): H c = w S c, c^H S c = 1.  The columns
// are first scaled to unit S-norm (D S D, D H D), so small but independent
// directions count as much as large ones.
// - Fast route: D S D = L L^H (Cholesky, L lower) and the standard problem
//   L^-1 (D H D) L^-H.  Taken when the factorisation exists and
//   1 / ||L^-1||_F^2 = 1 / trace((D S D)^-1), a lower bound of the smallest
//   eigenvalue of D S D, exceeds floor * m (m, the trace, bounds the
//   largest), i.e. when the route below would keep every direction.
// - Otherwise: eigenvectors of D S D with eigenvalues <= floor * the largest
//   are dropped and the problem is solved on the rest.
// Returns the number of directions kept (m on the fast route; below nout
// means no solution and the outputs are untouched).  Ccol is column-major
// m x nout, w ascending; *fast tells which route ran.
inline int pencil_lowest(int m, const std::vector<std::complex<double>>& S0,
                         const std::vector<std::complex<double>>& H0, int nout, double floor,
                         std::vector<std::complex<double>>& Ccol, std::vector<double>& w, bool* fast = nullptr) {
  using C = std::complex<double>;
  using la::cmul;
  using la::mul;
  using la::mulc;
  if (fast) *fast = false;
  std::vector<double> dsc(m);
  for (int i = 0; i < m; ++i) {
    const double sii = S0[(size_t)i * m + i].real();
    dsc[i] = sii > 0.0 ? 1.0 / std::sqrt(sii) : 0.0;
  }
  std::vector<C> S((size_t)m * m), H((size_t)m * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      S[(size_t)i * m + j] = S0[(size_t)i * m + j] * (dsc[i] * dsc[j]);
      H[(size_t)i * m + j] = H0[(size_t)i * m + j] * (dsc[i] * dsc[j]);
    }
  // Cholesky from the lower triangle: L[i][j] = (S[i][j] - sum_k L[i][k]
  // conj(L[j][k])) / L[j][j]
  std::vector<C> L((size_t)m * m, C(0.0));
  bool ok = true;
  for (int j = 0; j < m && ok; ++j) {
    const C* lj = &L[(size_t)j * m];
    double d = S[(size_t)j * m + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(lj[k]);
    if (!(d > 0.0)) {
      ok = false;
      break;
    }
    const double ljj = std::sqrt(d), inv = 1.0 / ljj;
    L[(size_t)j * m + j] = ljj;
    for (int i = j + 1; i < m; ++i) {
      const C* li = &L[(size_t)i * m];
      C s = S[(size_t)i * m + j];
      for (int k = 0; k < j; ++k) s -= mulc(li[k], lj[k]);
      L[(size_t)i * m + j] = s * inv;
    }
  }
  if (ok) {
    // M = L^-1 by rows: M[i] = (e_i - sum_{k<i} L[i][k] M[k]) / L[i][i]
    std::vector<C> M((size_t)m * m, C(0.0));
    double fro = 0.0;
    for (int i = 0; i < m; ++i) {
      C* mi = &M[(size_t)i * m];
      const C* li = &L[(size_t)i * m];
      for (int k = 0; k < i; ++k) {
        const C a = li[k];
        const C* mk = &M[(size_t)k * m];
        for (int c = 0; c <= k; ++c) mi[c] -= mul(a, mk[c]);
      }
      mi[i] += 1.0;
      const double inv = 1.0 / li[i].real();
      for (int c = 0; c <= i; ++c) {
        mi[c] *= inv;
        fro += std::norm(mi[c]);
      }
    }
    if (1.0 / fro > floor * m) {
      // T = M H (rows of M are zero beyond the diagonal), then the upper
      // triangle of Hr = T M^H, mirrored
      std::vector<C> T((size_t)m * m, C(0.0)), Hr((size_t)m * m);
      for (int i = 0; i < m; ++i) {
        C* ti = &T[(size_t)i * m];
        const C* mi = &M[(size_t)i * m];
        for (int k = 0; k <= i; ++k) {
          const C a = mi[k];
          const C* hk = &H[(size_t)k * m];
          for (int j = 0; j < m; ++j) ti[j] += mul(a, hk[j]);
        }
      }
      for (int i = 0; i < m; ++i) {
        const C* ti = &T[(size_t)i * m];
        for (int j = i; j < m; ++j) {
          const C* mj = &M[(size_t)j * m];
          double sr = 0.0, si = 0.0;
          for (int k = 0; k <= j; ++k) {
            const C p = mulc(ti[k], mj[k]);
            sr += p.real();
            si += p.imag();
          }
          Hr[(size_t)i * m + j] = C(sr, si);
        }
      }
      for (int i = 0; i < m; ++i) {
        Hr[(size_t)i * m + i] = C(Hr[(size_t)i * m + i].real(), 0.0);
        for (int j = i + 1; j < m; ++j) Hr[(size_t)j * m + i] = std::conj(Hr[(size_t)i * m + j]);
      }
      std::vector<double> wh;
      std::vector<C> U;
      herm_eig_lowest(m, std::move(Hr), nout, wh, U);
      // C = D M^H U (U: m x nout): C[i][j] = dsc[i] sum_{k>=i} conj(M[k][i]) U[k][j]
      Ccol.assign((size_t)m * nout, C(0.0));
      w = wh;
      for (int j = 0; j < nout; ++j)
        for (int i = 0; i < m; ++i) {
          double sr = 0.0, si = 0.0;
          for (int k = i; k < m; ++k) {
            const C p = cmul(M[(size_t)k * m + i], U[(size_t)k * nout + j]);
            sr += p.real();
            si += p.imag();
          }
          Ccol[i + (size_t)m * j] = C(sr, si) * dsc[i];
        }
      if (fast) *fast = true;
      return m;
    }
  }
  // rank-revealing route through the eigenvectors of D S D
  std::vector<double> ws;
  std::vector<C> Vs;
  herm_eig_tridiag(m, S, ws, Vs);
  const double top = ws.back();
  std::vector<int> keep;
  for (int k = 0; k < m; ++k)
    if (ws[k] > floor * top) keep.push_back(k);
  const int r = (int)keep.size();
  if (r < nout) return r;
  std::vector<C> X((size_t)m * r);   // row-major m x r
  for (int i = 0; i < m; ++i)
    for (int k = 0; k < r; ++k) X[(size_t)i * r + k] = Vs[(size_t)i * m + keep[k]] / std::sqrt(ws[keep[k]]);
  std::vector<C> HX((size_t)m * r), Hr((size_t)r * r);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const C h = H[(size_t)i * m + j];
      const C* xj = &X[(size_t)j * r];
      C* out = &HX[(size_t)i * r];
      for (int k = 0; k < r; ++k) out[k] += mul(h, xj[k]);
    }
  for (int i = 0; i < m; ++i)
    for (int a = 0; a < r; ++a) {
      const C xa = X[(size_t)i * r + a];
      const C* hx = &HX[(size_t)i * r];
      C* out = &Hr[(size_t)a * r];
      for (int b = 0; b < r; ++b) out[b] += cmul(xa, hx[b]);
    }
  for (int a = 0; a < r; ++a)
    for (int b = a + 1; b < r; ++b) {
      const C h = 0.5 * (Hr[(size_t)a * r + b] + std::conj(Hr[(size_t)b * r + a]));
      Hr[(size_t)a * r + b] = h;
      Hr[(size_t)b * r + a] = std::conj(h);
    }
  std::vector<double> wh;
  std::vector<C> U;
  herm_eig_lowest(r, std::move(Hr), nout, wh, U);
  Ccol.assign((size_t)m * nout, C(0.0));
  w = wh;
  for (int j = 0; j < nout; ++j)
    for (int i = 0; i < m; ++i) {
      C s = 0.0;
      for (int k = 0; k < r; ++k) s += mul(X[(size_t)i * r + k], U[(size_t)k * nout + j]);
      Ccol[i + (size_t)m * j] = s * dsc[i];
    }
  return r;
}

}  // namespace opaw
