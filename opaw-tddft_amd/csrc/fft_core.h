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

// ---------------------------------------------------------------------------
// Runtime-length transforms (lengths outside the compiled list): the radix
// plan is built on the host (radices 8, 4, 2, 3, 5, 7, 11, 13) and each stage
// runs out of place between two LDS buffers.  Radices without a hand-written
// butterfly use a direct DFT whose roots come from the twiddle table
// (exp(-2 pi i m / N) = tw[m]).
constexpr int kMaxPlan = 16;
struct Plan {
  int nr;
  int rad[kMaxPlan];
};

// radix plan of n, or nr = 0 when n has a prime factor above 13
inline Plan make_plan(int n) {
  Plan P{};
  P.nr = 0;
  if (n < 2) return P;
  int m = n;
  while (m > 1 && P.nr < kMaxPlan) {
    int r = 0;
    if (m % 8 == 0 && m != 16) r = 8;
    else if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else {
      const int odd[5] = {3, 5, 7, 11, 13};
      for (int i = 0; i < 5; ++i)
        if (m % odd[i] == 0) {
          r = odd[i];
          break;
        }
    }
    if (r == 0) {
      P.nr = 0;
      return P;
    }
    P.rad[P.nr++] = r;
    m /= r;
  }
  if (m != 1) P.nr = 0;
  return P;
}

template <typename R, int RAD, int S>
__host__ __device__ __forceinline__ void dft_direct(cx<R>* v, const cx<R>* tw, int N) {
  cx<R> o[RAD];
  const int step = N / RAD;
#pragma unroll
  for (int m = 0; m < RAD; ++m) {
    cx<R> acc = v[0];
#pragma unroll
    for (int q = 1; q < RAD; ++q) {
      cx<R> w = tw[((q * m) % RAD) * step];
      if constexpr (S > 0) w.y = -w.y;
      acc = acc + cmul(v[q], w);
    }
    o[m] = acc;
  }
#pragma unroll
  for (int m = 0; m < RAD; ++m) v[m] = o[m];
}

// one butterfly of a runtime stage: src -> dst (same index scheme as Stage)
template <typename R, int RAD, int S>
__host__ __device__ __forceinline__ void gen_butterfly(const cx<R>* src, cx<R>* dst, const cx<R>* tw, int N, int l,
                                                       int j, int M, int ns, int LSTR, int ESTR) {
  cx<R> v[RAD];
#pragma unroll
  for (int r = 0; r < RAD; ++r) v[r] = src[l * LSTR + (j + r * M) * ESTR];
  const int k = j % ns;
  if (ns > 1) {
    const int tws = N / (ns * RAD);
#pragma unroll
    for (int r = 1; r < RAD; ++r) {
      cx<R> w = tw[r * k * tws];
      if constexpr (S > 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  if constexpr (RAD == 2 || RAD == 3 || RAD == 4 || RAD == 5 || RAD == 8)
    dft<R, RAD, S>(v);
  else
    dft_direct<R, RAD, S>(v, tw, N);
  const int base = (j / ns) * ns * RAD + k;
#pragma unroll
  for (int r = 0; r < RAD; ++r) dst[l * LSTR + (base + r * ns) * ESTR] = v[r];
}

// all stages of a runtime plan for `lines` lines; element (l, j) of a buffer
// at l*LSTR + j*ESTR; tid / nthreads walk the butterflies.  Returns the
// buffer holding the result (a or b).
template <typename R, int S>
__host__ __device__ inline cx<R>* gen_stage(const cx<R>* src, cx<R>* dst, const cx<R>* tw, int N, int lines,
                                            bool lfast, int LSTR, int ESTR, int rad, int ns, int tid,
                                            int nthreads) {
  const int M = N / rad, NB = lines * M;
  for (int t = tid; t < NB; t += nthreads) {
    const int l = lfast ? (t % lines) : (t / M);
    const int j = lfast ? (t / lines) : (t % M);
    switch (rad) {
      case 2: gen_butterfly<R, 2, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      case 3: gen_butterfly<R, 3, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      case 4: gen_butterfly<R, 4, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      case 5: gen_butterfly<R, 5, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      case 7: gen_butterfly<R, 7, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      case 8: gen_butterfly<R, 8, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      case 11: gen_butterfly<R, 11, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
      default: gen_butterfly<R, 13, S>(src, dst, tw, N, l, j, M, ns, LSTR, ESTR); break;
    }
  }
  return dst;
}

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

template <typename R, int S>
__device__ __forceinline__ cx<R>* lines_gen(cx<R>* a, cx<R>* b, const cx<R>* tw, int N, int lines, bool lfast,
                                            int LSTR, int ESTR, const Plan& P) {
  int ns = 1;
  for (int s = 0; s < P.nr; ++s) {
    gen_stage<R, S>(a, b, tw, N, lines, lfast, LSTR, ESTR, P.rad[s], ns, threadIdx.x, blockDim.x);
    __syncthreads();
    cx<R>* t = a;
    a = b;
    b = t;
    ns *= P.rad[s];
  }
  return a;
}

// Copy the forward twiddle table exp(-2 pi i m / N), m < N, into LDS.
template <typename R>
__device__ __forceinline__ void load_twiddles(cx<R>* dst, const cx<R>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}
#endif  // __HIPCC__

}  // namespace fft
}  // namespace opaw
