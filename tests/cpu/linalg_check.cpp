// Reads a case from stdin and writes the result to stdout (binary, little
// endian):
// - n >= 0, then a row-major complex n x n matrix (doubles re, im):
//   opaw::herm_eig_tridiag -> w[n], then V row-major as re, im pairs;
// - n = -1, then m, nout, floor, S and H (row-major complex m x m):
//   opaw::pencil_lowest -> r (int), fast (int), w[nout], C column-major;
// - n = -2, then n, nout and A: opaw::herm_eig_lowest -> w[nout], then V
//   row-major n x nout.
#include <cstdio>
#include <vector>

#include "opaw_linalg.h"

using Cv = std::vector<std::complex<double>>;

static bool read_mat(Cv& A) {
  return A.empty() || std::fread(A.data(), sizeof(double) * 2, A.size(), stdin) == A.size();
}

int main() {
  int n = 0;
  if (std::fread(&n, sizeof(int), 1, stdin) != 1) return 1;
  if (n == -1) {
    int m = 0, nout = 0;
    double floor = 0.0;
    if (std::fread(&m, sizeof(int), 1, stdin) != 1 || std::fread(&nout, sizeof(int), 1, stdin) != 1 ||
        std::fread(&floor, sizeof(double), 1, stdin) != 1 || m < 1 || nout < 1 || nout > m)
      return 1;
    Cv S((size_t)m * m), H((size_t)m * m), Ccol;
    if (!read_mat(S) || !read_mat(H)) return 1;
    std::vector<double> w;
    bool fast = false;
    const int r = opaw::pencil_lowest(m, S, H, nout, floor, Ccol, w, &fast);
    const int f = fast ? 1 : 0;
    std::fwrite(&r, sizeof(int), 1, stdout);
    std::fwrite(&f, sizeof(int), 1, stdout);
    if (r >= nout) {
      std::fwrite(w.data(), sizeof(double), w.size(), stdout);
      std::fwrite(Ccol.data(), sizeof(double) * 2, Ccol.size(), stdout);
    }
    return 0;
  }
  if (n == -2) {
    int nout = 0;
    if (std::fread(&n, sizeof(int), 1, stdin) != 1 || std::fread(&nout, sizeof(int), 1, stdin) != 1 || n < 1 ||
        nout < 1 || nout > n)
      return 1;
    Cv A((size_t)n * n), V;
    if (!read_mat(A)) return 1;
    std::vector<double> w;
    opaw::herm_eig_lowest(n, A, nout, w, V);
    std::fwrite(w.data(), sizeof(double), w.size(), stdout);
    std::fwrite(V.data(), sizeof(double) * 2, V.size(), stdout);
    return 0;
  }
  if (n < 0) return 1;
  Cv A((size_t)n * n);
  if (!read_mat(A)) return 1;
  std::vector<double> w;
  Cv V;
  opaw::herm_eig_tridiag(n, A, w, V);
  std::fwrite(w.data(), sizeof(double), w.size(), stdout);
  std::fwrite(V.data(), sizeof(double) * 2, V.size(), stdout);
  return 0;
}
