// Reads n and a row-major complex n x n matrix (doubles re, im) from stdin,
// writes the eigenvalues and eigenvectors of opaw::herm_eig_tridiag to stdout
// (binary doubles: w[n], then V row-major as re, im pairs).
#include <cstdio>
#include <vector>

#include "opaw_linalg.h"

int main() {
  int n = 0;
  if (std::fread(&n, sizeof(int), 1, stdin) != 1 || n < 0) return 1;
  std::vector<std::complex<double>> A((size_t)n * n);
  if (n && std::fread(A.data(), sizeof(double) * 2, A.size(), stdin) != A.size()) return 1;
  std::vector<double> w;
  std::vector<std::complex<double>> V;
  opaw::herm_eig_tridiag(n, A, w, V);
  std::fwrite(w.data(), sizeof(double), w.size(), stdout);
  std::fwrite(V.data(), sizeof(double) * 2, V.size(), stdout);
  return 0;
}
