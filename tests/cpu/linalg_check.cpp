// Reads a case from stdin and writes the result to stdout (binary, little
// endian):
// n >= 0, then a row-major complex n x n matrix (doubles re, im):
// opaw::herm_eig_tridiag -> w[n], then V row-major as re, im pairs.
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
