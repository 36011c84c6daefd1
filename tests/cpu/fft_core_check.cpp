// Host-side check of the Stockham stage index logic in fft_core.h: runs the
// gather/scatter halves for every emulated thread of a tile and prints the
// transformed tile so the test can compare it with numpy.fft.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "fft_core.h"

using namespace opaw;
using namespace opaw::fft;

template <int N, int NS, int LINES, int THREADS, bool LFAST, int LSTR, int ESTR, int S>
void host_stages(cx<double>* lds, const cx<double>* tw) {
  if constexpr (NS < N) {
    constexpr int RAD = pick_radix(N / NS);
    using St = Stage<double, N, RAD, NS, LINES, THREADS, LFAST, LSTR, ESTR, S>;
    std::vector<std::array<std::array<cx<double>, RAD>, St::K>> regs(THREADS);
    for (int t = 0; t < THREADS; ++t) {
      cx<double> v[St::K][RAD];
      St::gather(lds, v, t);
      for (int q = 0; q < St::K; ++q)
        for (int r = 0; r < RAD; ++r) regs[t][q][r] = v[q][r];
    }
    for (int t = 0; t < THREADS; ++t) {
      cx<double> v[St::K][RAD];
      for (int q = 0; q < St::K; ++q)
        for (int r = 0; r < RAD; ++r) v[q][r] = regs[t][q][r];
      St::scatter(lds, v, tw, t);
    }
    host_stages<N, NS * RAD, LINES, THREADS, LFAST, LSTR, ESTR, S>(lds, tw);
  }
}

template <int N, int LINES, bool LFAST, int S>
void run() {
  constexpr int THREADS = 64;
  constexpr int LSTR = LFAST ? 1 : N;
  constexpr int ESTR = LFAST ? LINES : 1;
  std::vector<cx<double>> tw(N), lds(N * LINES);
  for (int m = 0; m < N; ++m) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * m / N;
    tw[m] = mk<double>((double)std::cos(a), (double)std::sin(a));
  }
  // deterministic input: element (l, j)
  for (int l = 0; l < LINES; ++l)
    for (int j = 0; j < N; ++j)
      lds[l * LSTR + j * ESTR] = mk<double>(std::sin(0.37 * (j + 1) + l), std::cos(0.11 * j * j + 0.5 * l));
  host_stages<N, 1, LINES, THREADS, LFAST, LSTR, ESTR, S>(lds.data(), tw.data());
  for (int l = 0; l < LINES; ++l)
    for (int j = 0; j < N; ++j)
      std::printf("%d %d %d %d %.17e %.17e\n", N, S, l, j, lds[l * LSTR + j * ESTR].x, lds[l * LSTR + j * ESTR].y);
}

// runtime-plan transforms (lengths outside the compiled list): every stage
// out of place, all butterflies of a stage on one emulated thread
void run_gen(int N, int LINES, bool LFAST, int S) {
  const int LSTR = LFAST ? 1 : N;
  const int ESTR = LFAST ? LINES : 1;
  const Plan P = make_plan(N);
  if (P.nr == 0) {
    std::printf("bad plan %d\n", N);
    std::exit(1);
  }
  std::vector<cx<double>> tw(N), a(N * LINES), b(N * LINES);
  for (int m = 0; m < N; ++m) {
    long double ang = -2.0L * 3.14159265358979323846264338327950288L * m / N;
    tw[m] = mk<double>((double)std::cos(ang), (double)std::sin(ang));
  }
  for (int l = 0; l < LINES; ++l)
    for (int j = 0; j < N; ++j)
      a[l * LSTR + j * ESTR] = mk<double>(std::sin(0.37 * (j + 1) + l), std::cos(0.11 * j * j + 0.5 * l));
  cx<double>* src = a.data();
  cx<double>* dst = b.data();
  int ns = 1;
  for (int s = 0; s < P.nr; ++s) {
    if (S < 0)
      gen_stage<double, -1>(src, dst, tw.data(), N, LINES, LFAST, LSTR, ESTR, P.rad[s], ns, 0, 1);
    else
      gen_stage<double, 1>(src, dst, tw.data(), N, LINES, LFAST, LSTR, ESTR, P.rad[s], ns, 0, 1);
    std::swap(src, dst);
    ns *= P.rad[s];
  }
  for (int l = 0; l < LINES; ++l)
    for (int j = 0; j < N; ++j)
      std::printf("%d %d %d %d %.17e %.17e\n", N, S, l, j, src[l * LSTR + j * ESTR].x, src[l * LSTR + j * ESTR].y);
}

// the xy-plane pass of the kinetic operator: NY rows of length NX (row
// stride LS, the padded LDS row) and then NX columns of length NY, as the
// plane kernel runs them with 512 threads; prints element (y, x)
template <int NX, int NY, int LS, int S>
void run_plane() {
  constexpr int THREADS = 512;
  std::vector<cx<double>> twx(NX), twy(NY), lds(NY * LS);
  for (int m = 0; m < NX; ++m) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * m / NX;
    twx[m] = mk<double>((double)std::cos(a), (double)std::sin(a));
  }
  for (int m = 0; m < NY; ++m) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * m / NY;
    twy[m] = mk<double>((double)std::cos(a), (double)std::sin(a));
  }
  for (int y = 0; y < NY; ++y)
    for (int x = 0; x < NX; ++x)
      lds[y * LS + x] = mk<double>(std::sin(0.37 * (x + 1) + 0.7 * y), std::cos(0.11 * x * x + 0.5 * y));
  if (S < 0) {
    host_stages<NX, 1, NY, THREADS, false, LS, 1, S>(lds.data(), twx.data());
    host_stages<NY, 1, NX, THREADS, true, 1, LS, S>(lds.data(), twy.data());
  } else {
    host_stages<NY, 1, NX, THREADS, true, 1, LS, S>(lds.data(), twy.data());
    host_stages<NX, 1, NY, THREADS, false, LS, 1, S>(lds.data(), twx.data());
  }
  for (int y = 0; y < NY; ++y)
    for (int x = 0; x < NX; ++x)
      std::printf("%d %d %d %d %d %.17e %.17e\n", NX, NY, S, y, x, lds[y * LS + x].x, lds[y * LS + x].y);
}

#include <array>
#include <cstring>
int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "plane") == 0) {
    run_plane<64, 48, 72, -1>();
    run_plane<80, 64, 88, 1>();
    run_plane<96, 64, 104, -1>();
    run_plane<128, 64, 136, 1>();
    run_plane<48, 48, 56, -1>();
    return 0;
  }
  // radices 7, 11, 13 and their mixes with 2, 3, 5 (e.g. the 56 x 52 x 32
  // naphthalene grid and its doubled 112 x 104 x 64 box)
  const int gen_sizes[] = {7, 14, 28, 56, 112, 224, 13, 26, 52, 104, 208, 11, 22, 44, 77, 91, 143, 36, 60, 100, 18, 27};
  int k = 0;
  for (int n : gen_sizes) {
    run_gen(n, 1 + k % 3, k % 2 == 0, (k % 4 < 2) ? -1 : 1);
    ++k;
  }
  run<8, 3, false, -1>();
  run<12, 5, true, -1>();
  run<16, 2, false, 1>();
  run<20, 4, true, 1>();
  run<48, 3, false, -1>();
  run<64, 4, true, -1>();
  run<96, 2, false, 1>();
  run<10, 6, false, -1>();
  run<5, 3, true, 1>();
  run<6, 2, false, -1>();
  run<128, 2, true, -1>();
  run<192, 1, false, 1>();
  run<256, 1, false, -1>();
  run<40, 2, true, -1>();
  run<80, 2, false, 1>();
  run<160, 1, true, -1>();
  return 0;
}
