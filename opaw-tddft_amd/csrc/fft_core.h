// LDS-resident Stockham FFT for a tile of lines, gfx950.
//
// A workgroup owns LINES lines of length N held in LDS.  Each Stockham stage
// reads RAD elements per butterfly into registers, barriers, applies the
// twiddles and the RAD-point DFT, and writes the butterfly back in
// auto-sorted order (in place, register staged, so the tile needs no second
// LDS buffer).  Radices 8/4/2/3/5 are chosen at compile time from N.
//
// Sign convention follows FFTW: DIR=-1 is FFTW_FORWARD (e^{-i}), DIR=+1 is
// FFTW_BACKWARD (e^{+i}); both are unnormalised.
#pragma once
#include "opaw_common.h"

namespace opaw {
namespace fft {

constexpr int pick_radix(int rem) {
  return (rem % 8 == 0 && rem != 16) ? 8
       : (rem % 4 == 0) ? 4
       : (rem % 2 == 0) ? 2
       : (rem % 3 == 0) ? 3
       : (rem % 5 == 0) ? 5 : 0;
}

constexpr bool supported(int n) {
  if (n < 2) return false;
  while (n > 1) {
    int r = pick_radix(n);
    if (r == 0) return false;
    n /= r;
  }
  return true;
}

// ---------------------------------------------------------------------------
// small DFTs on registers; S = DIR (-1 forward, +1 backward)
template <typename R, int S>
__host__ __device__ __forceinline__ void dft2(cx<R>* v) {
  cx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <typename R, int S>
__host__ __device__ __forceinline__ void dft3(cx<R>* v) {
  const R h = R(0.86602540378443864676372317075294);
  cx<R> s12 = v[1] + v[2];
  cx<R> d12 = v[1] - v[2];
  cx<R> t = mk(v[0].x - R(0.5) * s12.x, v[0].y - R(0.5) * s12.y);
  // u = S*h*(x1-x2); X1 = t + i u, X2 = t - i u
  R ux = R(S) * h * d12.x, uy = R(S) * h * d12.y;
  v[0] = v[0] + s12;
  v[1] = mk(t.x - uy, t.y + ux);
  v[2] = mk(t.x + uy, t.y - ux);
}

template <typename R, int S>
__host__ __device__ __forceinline__ void dft4(cx<R>* v) {
  cx<R> a = v[0] + v[2], b = v[0] - v[2];
  cx<R> c = v[1] + v[3], d = v[1] - v[3];
  cx<R> di = mul_si(d, S);  // S*i*(x1-x3)
  v[0] = a + c;
  v[1] = b + di;
  v[2] = a - c;
  v[3] = b - di;
}

template <typename R, int S>
__host__ __device__ __forceinline__ void dft5(cx<R>* v) {
  const R c1 = R(0.30901699437494742410229341718282);   // cos(2pi/5)
  const R c2 = R(-0.80901699437494742410229341718282);  // cos(4pi/5)
  const R s1 = R(0.95105651629515357211643933337938);   // sin(2pi/5)
  const R s2 = R(0.58778525229247312916870595463907);   // sin(4pi/5)
  cx<R> t1 = v[1] + v[4], t2 = v[2] + v[3];
  cx<R> t3 = v[1] - v[4], t4 = v[2] - v[3];
  cx<R> x0 = v[0];
  cx<R> a1 = mk(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
  cx<R> a2 = mk(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
  cx<R> b1 = mk(R(S) * (s1 * t3.x + s2 * t4.x), R(S) * (s1 * t3.y + s2 * t4.y));
  cx<R> b2 = mk(R(S) * (s2 * t3.x - s1 * t4.x), R(S) * (s2 * t3.y - s1 * t4.y));
  v[0] = x0 + t1 + t2;
  v[1] = mk(a1.x - b1.y, a1.y + b1.x);
  v[4] = mk(a1.x + b1.y, a1.y - b1.x);
  v[2] = mk(a2.x - b2.y, a2.y + b2.x);
  v[3] = mk(a2.x + b2.y, a2.y - b2.x);
}

template <typename R, int S>
__host__ __device__ __forceinline__ void dft8(cx<R>* v) {
  const R c = R(0.70710678118654752440084436210485);
  cx<R> a[4], b[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    a[n] = v[n] + v[n + 4];
    b[n] = v[n] - v[n + 4];
  }
  // b_n *= w^n, w = e^{S 2 pi i / 8}
  b[1] = cmul(b[1], mk(c, R(S) * c));
  b[2] = mul_si(b[2], S);
  b[3] = cmul(b[3], mk(-c, R(S) * c));
  dft4<R, S>(a);
  dft4<R, S>(b);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = a[k];
    v[2 * k + 1] = b[k];
  }
}

template <typename R, int RAD, int S>
__host__ __device__ __forceinline__ void dft(cx<R>* v) {
  if constexpr (RAD == 2) dft2<R, S>(v);
  else if constexpr (RAD == 3) dft3<R, S>(v);
  else if constexpr (RAD == 4) dft4<R, S>(v);
  else if constexpr (RAD == 5) dft5<R, S>(v);
  else if constexpr (RAD == 8) dft8<R, S>(v);
}

// ---------------------------------------------------------------------------
// One Stockham stage over the tile.  Element (line l, index j) lives at
// lds[l*LSTR + j*ESTR].  LFAST: consecutive threads walk lines first (use when
// LSTR == 1 so that lanes touch consecutive LDS words).  The stage is split
// into gather (tile -> registers) and scatter (twiddle, DFT, registers ->
// tile) so that a barrier separates them; both halves are host-callable so
// the index logic can be exercised on the CPU.
template <typename R, int N, int RAD, int NS, int LINES, int THREADS, bool LFAST,
          int LSTR, int ESTR, int S>
struct Stage {
  static constexpr int M = N / RAD;             // butterflies per line
  static constexpr int NB = LINES * M;          // butterflies per tile
  static constexpr int K = (NB + THREADS - 1) / THREADS;
  static constexpr int TWS = N / (NS * RAD);    // twiddle table stride

  __host__ __device__ static __forceinline__ void gather(const cx<R>* lds, cx<R> (&v)[K][RAD], int tid) {
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int t = tid + q * THREADS;
      if ((NB % THREADS == 0) || t < NB) {
        const int l = LFAST ? (t % LINES) : (t / M);
        const int j = LFAST ? (t / LINES) : (t % M);
#pragma unroll
        for (int r = 0; r < RAD; ++r) v[q][r] = lds[l * LSTR + (j + r * M) * ESTR];
      }
    }
  }

  __host__ __device__ static __forceinline__ void scatter(cx<R>* lds, cx<R> (&v)[K][RAD],
                                                          const cx<R>* tw, int tid) {
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int t = tid + q * THREADS;
      if ((NB % THREADS == 0) || t < NB) {
        const int l = LFAST ? (t % LINES) : (t / M);
        const int j = LFAST ? (t / LINES) : (t % M);
        const int k = j % NS;
        if constexpr (NS > 1) {
#pragma unroll
          for (int r = 1; r < RAD; ++r) {
            cx<R> w = tw[r * k * TWS];
            if constexpr (S > 0) w.y = -w.y;
            v[q][r] = cmul(v[q][r], w);
          }
        }
        dft<R, RAD, S>(v[q]);
        const int base = (j / NS) * NS * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) lds[l * LSTR + (base + r * NS) * ESTR] = v[q][r];
      }
    }
  }
};

#if defined(__HIPCC__)
template <typename R, int N, int NS, int LINES, int THREADS, bool LFAST, int LSTR, int ESTR, int S>
__device__ __forceinline__ void stages(cx<R>* lds, const cx<R>* tw) {
  if constexpr (NS < N) {
    constexpr int RAD = pick_radix(N / NS);
    static_assert(RAD != 0, "unsupported FFT length");
    using St = Stage<R, N, RAD, NS, LINES, THREADS, LFAST, LSTR, ESTR, S>;
    cx<R> v[St::K][RAD];
    St::gather(lds, v, threadIdx.x);
    __syncthreads();
    St::scatter(lds, v, tw, threadIdx.x);
    __syncthreads();
    stages<R, N, NS * RAD, LINES, THREADS, LFAST, LSTR, ESTR, S>(lds, tw);
  }
}

// Full length-N transform of every line of the tile.  Caller must have
// synchronised after filling the tile; returns synchronised.
template <typename R, int N, int LINES, int THREADS, bool LFAST, int LSTR, int ESTR, int S>
__device__ __forceinline__ void lines(cx<R>* lds, const cx<R>* tw) {
  stages<R, N, 1, LINES, THREADS, LFAST, LSTR, ESTR, S>(lds, tw);
}

// Copy the forward twiddle table exp(-2 pi i m / N), m < N, into LDS.
template <typename R>
__device__ __forceinline__ void load_twiddles(cx<R>* dst, const cx<R>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}
#endif  // __HIPCC__

}  // namespace fft
}  // namespace opaw
