// Kernels of the ground-state solver.
//   cheb_combine       the Chebyshev recurrence of chebyr_m (src/chebyr_m.f90)
//   residual_partials  |H phi_i - e_i phi_i|^2 by chunks (convergence judge)
//   real_dots, diis_predict, to_f32
//                      DIIS on V_KS (diis_general, src/diis.f90:12-153)
//   precond_*          the k-space preconditioner of the block solver: the
//                      Teter-Payne-Allan ratio K(x) times exp(g(x, y)), g a
//                      two-hidden-layer tanh MLP; forward, loss sums and the
//                      parameter gradient (backward) for its training
// Every reduction writes per-chunk partial sums that the host adds in a
// fixed order, so results do not depend on scheduling.
#include "opaw_gs.h"

namespace opaw {
int mlp_param_count(const MlpShape& s);
namespace {

constexpr int kTB = 256;

inline unsigned blocks_for(long long n) { return (unsigned)((n + kTB - 1) / kTB); }

__device__ __forceinline__ void chunk_bounds(long long n, int nchunk, int c, long long& lo, long long& hi) {
  const long long per = (n + nchunk - 1) / nchunk;
  lo = (long long)c * per;
  hi = lo + per < n ? lo + per : n;
}

template <typename R>
__global__ __launch_bounds__(kTB) void k_cheb(cx<R>* __restrict__ cc, const cx<R>* __restrict__ cb,
                                              const cx<R>* __restrict__ ca, long long n, double havg, double dh,
                                              int first) {
  const long long i = (long long)blockIdx.x * kTB + threadIdx.x;
  if (i >= n) return;
  const cx<R> h = cc[i], b = cb[i];
  double x = ((double)h.x - havg * (double)b.x) / dh;
  double y = ((double)h.y - havg * (double)b.y) / dh;
  if (!first) {
    const cx<R> a = ca[i];
    x = 2.0 * x - (double)a.x;
    y = 2.0 * y - (double)a.y;
  }
  cc[i] = mk<R>((R)x, (R)y);
}

// block = (chunk c, band i): one partial per block
template <typename R>
__global__ __launch_bounds__(kTB) void k_resid(const cx<R>* __restrict__ hp, const cx<R>* __restrict__ phi,
                                               const double* __restrict__ eig, long long N, int nb,
                                               double* __restrict__ part) {
  const int c = blockIdx.x, i = blockIdx.y;
  long long lo, hi;
  chunk_bounds(N, kGsChunks, c, lo, hi);
  const double e = eig[i];
  const cx<R>* h = hp + (size_t)i * N;
  const cx<R>* p = phi + (size_t)i * N;
  double s = 0.0;
  for (long long r = lo + threadIdx.x; r < hi; r += kTB) {
    const double dx = (double)h[r].x - e * (double)p[r].x;
    const double dy = (double)h[r].y - e * (double)p[r].y;
    s += dx * dx + dy * dy;
  }
  __shared__ double red[kTB];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)c * nb + i] = red[0];
}

// block = (chunk c, pair i, pair j)
__global__ __launch_bounds__(kTB) void k_real_dots(const double* __restrict__ a, long long N, int K,
                                                   double* __restrict__ part) {
  const int c = blockIdx.x, i = blockIdx.y, j = blockIdx.z;
  long long lo, hi;
  chunk_bounds(N, kGsChunks, c, lo, hi);
  const double* ai = a + (size_t)i * N;
  const double* aj = a + (size_t)j * N;
  double s = 0.0;
  for (long long r = lo + threadIdx.x; r < hi; r += kTB) s += ai[r] * aj[r];
  __shared__ double red[kTB];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[((size_t)c * K + i) * K + j] = red[0];
}

__global__ __launch_bounds__(kTB) void k_diis_predict(const double* __restrict__ bank, const double* __restrict__ err,
                                                      const double* __restrict__ coef, int K, double mix, long long N,
                                                      double* __restrict__ out) {
  const long long r = (long long)blockIdx.x * kTB + threadIdx.x;
  if (r >= N) return;
  double s = 0.0;
  for (int k = 0; k < K; ++k) s += (bank[(size_t)k * N + r] + mix * err[(size_t)k * N + r]) * coef[k];
  out[r] = s;
}

__global__ __launch_bounds__(kTB) void k_to_f32(const double* __restrict__ in, float* __restrict__ out, long long n) {
  const long long r = (long long)blockIdx.x * kTB + threadIdx.x;
  if (r < n) out[r] = (float)in[r];
}

__global__ __launch_bounds__(kTB) void k_scale_c(c64* __restrict__ a, long long n, double s) {
  const long long r = (long long)blockIdx.x * kTB + threadIdx.x;
  if (r < n) a[r] = mk(a[r].x * s, a[r].y * s);
}

__global__ __launch_bounds__(kTB) void k_axpy_c(const c64* __restrict__ a, const c64* __restrict__ b, double s,
                                                long long n, c64* __restrict__ out) {
  const long long r = (long long)blockIdx.x * kTB + threadIdx.x;
  if (r < n) out[r] = mk(a[r].x - s * b[r].x, a[r].y - s * b[r].y);
}

// ---------------------------------------------------------------------------
// the preconditioner multiplier and its MLP
constexpr int kMaxH = 16;

__device__ __forceinline__ double tpa(double x) {
  const double p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
  return p / (p + 16.0 * x * x * x * x);
}

struct Fwd {
  double u0, u1, h1[kMaxH], h2[kMaxH], g;
};

// params: W1 (nh x 2, row-major), b1 (nh), W2 (nh x nh), b2 (nh), w3 (nh), b3
__device__ __forceinline__ void mlp_forward(const double* __restrict__ P, int nh, double u0, double u1, Fwd& f) {
  f.u0 = u0;
  f.u1 = u1;
  const double* W1 = P;
  const double* b1 = W1 + 2 * nh;
  const double* W2 = b1 + nh;
  const double* b2 = W2 + nh * nh;
  const double* w3 = b2 + nh;
  const double b3 = w3[nh];
  for (int a = 0; a < nh; ++a) f.h1[a] = tanh(W1[2 * a] * u0 + W1[2 * a + 1] * u1 + b1[a]);
  double g = b3;
  for (int a = 0; a < nh; ++a) {
    double z = b2[a];
    for (int b = 0; b < nh; ++b) z += W2[a * nh + b] * f.h1[b];
    f.h2[a] = tanh(z);
    g += w3[a] * f.h2[a];
  }
  f.g = g;
}

__device__ __forceinline__ double multiplier(const double* __restrict__ P, int nh, double x, double y, Fwd& f) {
  double m = tpa(x);
  if (P) {
    mlp_forward(P, nh, log1p(x), y, f);
    m *= exp(f.g);
  } else {
    f.g = 0.0;
  }
  return m;
}

// w = M r in k space, with the FFT's 1/N: the multiplier depends on k only
// through ek(k), so it is evaluated once per (band, k shell):
// tab[i * nsh + s] = M_i(eks[s]) / N.
__global__ __launch_bounds__(kTB) void k_precond_table(const double* __restrict__ eks, int nsh, long long N, int nb,
                                                       const double* __restrict__ tkin,
                                                       const double* __restrict__ yfeat, const double* __restrict__ P,
                                                       int nh, double* __restrict__ tab) {
  const long long q = (long long)blockIdx.x * kTB + threadIdx.x;
  if (q >= (long long)nsh * nb) return;
  const int i = (int)(q / nsh);
  const int sh = (int)(q - (long long)i * nsh);
  Fwd f;
  tab[q] = multiplier(P, nh, eks[sh] * (double)N / tkin[i], yfeat[i], f) / (double)N;
}

// rk[i * N + k] *= tab[i * nsh + shell(k)]; one thread per k over all bands
__global__ __launch_bounds__(kTB) void k_precond_lookup(c64* __restrict__ rk, long long N, int nb,
                                                        const int* __restrict__ shell_of,
                                                        const double* __restrict__ tab, int nsh) {
  const long long k = (long long)blockIdx.x * kTB + threadIdx.x;
  if (k >= N) return;
  const int sh = shell_of[k];
  for (int i = 0; i < nb; ++i) {
    const double m = tab[(size_t)i * nsh + sh];
    const size_t q = (size_t)i * N + k;
    rk[q] = mk(rk[q].x * m, rk[q].y * m);
  }
}

// block = (chunk c, band i): ReA, ImA, B, Ee partials
__global__ __launch_bounds__(kTB) void k_loss_sums(const c64* __restrict__ rk, const c64* __restrict__ ekf, long long N,
                                                   int nb, const double* __restrict__ ek, const double* __restrict__ tkin,
                                                   const double* __restrict__ yfeat, const double* __restrict__ P, int nh,
                                                   double* __restrict__ part) {
  const int c = blockIdx.x, i = blockIdx.y;
  long long lo, hi;
  chunk_bounds(N, kGsChunks, c, lo, hi);
  double ar = 0.0, ai = 0.0, bb = 0.0, ee = 0.0;
  for (long long k = lo + threadIdx.x; k < hi; k += kTB) {
    Fwd f;
    const double m = multiplier(P, nh, ek[k] * (double)N / tkin[i], yfeat[i], f);
    const c64 r = rk[(size_t)i * N + k], e = ekf[(size_t)i * N + k];
    // conj(e) r
    const double cr = e.x * r.x + e.y * r.y, ci = e.x * r.y - e.y * r.x;
    ar += m * cr;
    ai += m * ci;
    bb += m * m * (r.x * r.x + r.y * r.y);
    ee += e.x * e.x + e.y * e.y;
  }
  __shared__ double red[4][kTB];
  red[0][threadIdx.x] = ar;
  red[1][threadIdx.x] = ai;
  red[2][threadIdx.x] = bb;
  red[3][threadIdx.x] = ee;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int t = 0; t < 4; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[((size_t)c * nb + i) * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// parameter gradient: samples q = i*N + k in tiles of kTile, per-sample
// backward into LDS, then one thread per parameter sums the tile
constexpr int kTile = 64;                  // LDS: 4 x 64 x 16 doubles + small = 34 KB
constexpr int kSlots = (16 * 16 + 5 * 16 + 1 + kTile - 1) / kTile;   // parameters per thread

__global__ __launch_bounds__(kTile) void k_precond_grad(const c64* __restrict__ rk, const c64* __restrict__ ekf,
                                                        long long N, int nb, const double* __restrict__ ek,
                                                        const double* __restrict__ tkin,
                                                        const double* __restrict__ yfeat, const double* __restrict__ P,
                                                        int nh, const double* __restrict__ coefA,
                                                        const double* __restrict__ coefB, int npar,
                                                        double* __restrict__ gpart, int nchunk) {
  __shared__ double s_u[kTile][2];
  __shared__ double s_h1[kTile][kMaxH];
  __shared__ double s_dz1[kTile][kMaxH];
  __shared__ double s_h2[kTile][kMaxH];
  __shared__ double s_dz2[kTile][kMaxH];
  __shared__ double s_dg[kTile];
  const long long total = N * nb;
  long long lo, hi;
  chunk_bounds(total, nchunk, blockIdx.x, lo, hi);
  const double* W2 = P + 3 * nh;
  const double* w3 = W2 + nh * nh + nh;
  // each thread owns parameters p = t + slot * kTile
  double acc[kSlots];
  for (int slot = 0; slot < kSlots; ++slot) acc[slot] = 0.0;
  for (long long t0 = lo; t0 < hi; t0 += kTile) {
    const long long q = t0 + threadIdx.x;
    if (q < hi) {
      const int i = (int)(q / N);
      const long long k = q - (long long)i * N;
      Fwd f;
      const double m = multiplier(P, nh, ek[k] * (double)N / tkin[i], yfeat[i], f);
      const c64 r = rk[q], e = ekf[q];
      const double cr = e.x * r.x + e.y * r.y, ci = e.x * r.y - e.y * r.x;
      // dL/dg = (Re(cA conj(e) r) + cB m |r|^2) m
      const double dg = (coefA[2 * i] * cr - coefA[2 * i + 1] * ci + coefB[i] * m * (r.x * r.x + r.y * r.y)) * m;
      s_u[threadIdx.x][0] = f.u0;
      s_u[threadIdx.x][1] = f.u1;
      s_dg[threadIdx.x] = dg;
      double dz2[kMaxH];
      for (int a = 0; a < nh; ++a) {
        dz2[a] = dg * w3[a] * (1.0 - f.h2[a] * f.h2[a]);
        s_h2[threadIdx.x][a] = f.h2[a];
        s_dz2[threadIdx.x][a] = dz2[a];
        s_h1[threadIdx.x][a] = f.h1[a];
      }
      for (int b = 0; b < nh; ++b) {
        double dh1 = 0.0;
        for (int a = 0; a < nh; ++a) dh1 += W2[a * nh + b] * dz2[a];
        s_dz1[threadIdx.x][b] = dh1 * (1.0 - f.h1[b] * f.h1[b]);
      }
    } else {
      s_u[threadIdx.x][0] = s_u[threadIdx.x][1] = 0.0;
      s_dg[threadIdx.x] = 0.0;
      for (int a = 0; a < nh; ++a) s_h2[threadIdx.x][a] = s_dz2[threadIdx.x][a] = s_h1[threadIdx.x][a] = s_dz1[threadIdx.x][a] = 0.0;
    }
    __syncthreads();
    const int ns = (int)(hi - t0 < kTile ? hi - t0 : kTile);
    for (int slot = 0; slot < kSlots; ++slot) {
      const int p = threadIdx.x + slot * kTile;
      if (p >= npar) break;
      double s = 0.0;
      if (p < 2 * nh) {                       // W1[a][c]
        const int a = p / 2, cc = p % 2;
        for (int t = 0; t < ns; ++t) s += s_dz1[t][a] * s_u[t][cc];
      } else if (p < 3 * nh) {                // b1
        const int a = p - 2 * nh;
        for (int t = 0; t < ns; ++t) s += s_dz1[t][a];
      } else if (p < 3 * nh + nh * nh) {      // W2[a][b]
        const int a = (p - 3 * nh) / nh, b = (p - 3 * nh) % nh;
        for (int t = 0; t < ns; ++t) s += s_dz2[t][a] * s_h1[t][b];
      } else if (p < 4 * nh + nh * nh) {      // b2
        const int a = p - 3 * nh - nh * nh;
        for (int t = 0; t < ns; ++t) s += s_dz2[t][a];
      } else if (p < 5 * nh + nh * nh) {      // w3
        const int a = p - 4 * nh - nh * nh;
        for (int t = 0; t < ns; ++t) s += s_dg[t] * s_h2[t][a];
      } else {                                // b3
        for (int t = 0; t < ns; ++t) s += s_dg[t];
      }
      acc[slot] += s;
    }
    __syncthreads();
  }
  for (int slot = 0; slot < kSlots; ++slot) {
    const int p = threadIdx.x + slot * kTile;
    if (p < npar) gpart[(size_t)blockIdx.x * npar + p] = acc[slot];
  }
}

// out_j = sum_{i<m} T[i + m j] in_i for j < nout (FP64 fields, T column-major)
// on the matrix cores: per workgroup 64 outputs x 64 points, four waves of
// 16 outputs each.  The inputs and the coefficients of
// 16 inputs at a time go through LDS (the coefficients with the input index
// fastest: a column of T is contiguous), the next 16 are loaded into
// registers meanwhile.  Complex products as four real 16x16x4 FP64 MFMAs:
// Re += Tr Ir - Ti Ii, Im += Tr Ii + Ti Ir.  Lane l of a product holds
// output row l/16 + 4e of point column l%16.
typedef double f64x4r __attribute__((ext_vector_type(4)));
constexpr int kRmP = 64, kRmJ = 64, kRmK = 16;
__global__ __launch_bounds__(256) void k_rotate_mfma(const c64* __restrict__ in, c64* __restrict__ out, long long N,
                                                      int m, int nout, const c64* __restrict__ T) {
  __shared__ c64 sI[kRmK][kRmP + 1];
  __shared__ c64 sT[kRmK][kRmJ + 1];
  const long long p0 = (long long)blockIdx.x * kRmP;
  const int j0 = blockIdx.y * kRmJ;
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int lr = lane % 16, lk = lane / 16;
  c64 vi[4], vt[4];
  auto load = [&](int i0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int slot = (int)threadIdx.x + 256 * q;
      const int ki = slot / kRmP, ci = slot % kRmP;     // inputs: points fastest
      const int kt = slot % kRmK, ct = slot / kRmK;     // coefficients: inputs fastest
      vi[q] = in[(long long)min(i0 + ki, m - 1) * N + min(p0 + ci, N - 1)];
      vt[q] = T[(long long)min(i0 + kt, m - 1) + (long long)m * min(j0 + ct, nout - 1)];
    }
  };
  f64x4r cr[4], ci4[4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
    for (int e = 0; e < 4; ++e) cr[v][e] = ci4[v][e] = 0.0;
  load(0);
  for (int i0 = 0; i0 < m; i0 += kRmK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int slot = (int)threadIdx.x + 256 * q;
      const int ki = slot / kRmP, ci = slot % kRmP;
      const int kt = slot % kRmK, ct = slot / kRmK;
      sI[ki][ci] = (i0 + ki < m && p0 + ci < N) ? vi[q] : mk(0.0, 0.0);
      sT[kt][ct] = (i0 + kt < m && j0 + ct < nout) ? vt[q] : mk(0.0, 0.0);
    }
    __syncthreads();
    if (i0 + kRmK < m) load(i0 + kRmK);
#pragma unroll
    for (int kk = 0; kk < kRmK / 4; ++kk) {
      const c64 a = sT[4 * kk + lk][16 * w + lr];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const c64 b = sI[4 * kk + lk][16 * v + lr];
        cr[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, cr[v], 0, 0, 0);
        cr[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, cr[v], 0, 0, 0);
        ci4[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, ci4[v], 0, 0, 0);
        ci4[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, ci4[v], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int j = j0 + 16 * w + lk + 4 * e;
      const long long p = p0 + 16 * v + lr;
      if (j < nout && p < N) out[(long long)j * N + p] = mk(cr[v][e], ci4[v][e]);
    }
}

// r_i = h_i - e_i phi_i
__global__ __launch_bounds__(kTB) void k_resid_vec(const c64* __restrict__ hp, const c64* __restrict__ phi,
                                                   const double* __restrict__ eig, long long N, int nb,
                                                   c64* __restrict__ r) {
  const long long q = (long long)blockIdx.x * kTB + threadIdx.x;
  if (q >= N * nb) return;
  const double e = eig[q / N];
  r[q] = mk(hp[q].x - e * phi[q].x, hp[q].y - e * phi[q].y);
}

// block = (chunk c, band i): sum ek |fk|^2 and sum |fk|^2 (k space)
__global__ __launch_bounds__(kTB) void k_kin(const c64* __restrict__ fk, const double* __restrict__ ek, long long N,
                                             int nb, double* __restrict__ part) {
  const int c = blockIdx.x, i = blockIdx.y;
  long long lo, hi;
  chunk_bounds(N, kGsChunks, c, lo, hi);
  double a = 0.0, b = 0.0;
  for (long long k = lo + threadIdx.x; k < hi; k += kTB) {
    const c64 v = fk[(size_t)i * N + k];
    const double w = v.x * v.x + v.y * v.y;
    a += ek[k] * w;
    b += w;
  }
  __shared__ double red[2][kTB];
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((size_t)c * nb + i) * 2] = red[0][0];
    part[((size_t)c * nb + i) * 2 + 1] = red[1][0];
  }
}

__global__ __launch_bounds__(kTB) void k_sub_real(const double* __restrict__ a, const double* __restrict__ b,
                                                  long long n, double* __restrict__ out) {
  const long long r = (long long)blockIdx.x * kTB + threadIdx.x;
  if (r < n) out[r] = a[r] - b[r];
}

// acc[p] += sum over rows b < nrow of part[b * npar + p]: one block per
// parameter, strided partial sums and a fixed-order tree
__global__ __launch_bounds__(kTB) void k_colsum_acc(const double* __restrict__ part, int nrow, int npar,
                                                    double* __restrict__ acc) {
  const int p = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nrow; b += kTB) s += part[(size_t)b * npar + p];
  __shared__ double red[kTB];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[p] += red[0];
}

// shell sums of one prepared pair: z[i][s] = sum over k of shell s of
// conj(e_i(k)) r_i(k), w2[i][s] = sum |r_i(k)|^2; the k of shell s are
// perm[off[s] .. off[s+1]); block = (shell s, band i)
__global__ __launch_bounds__(kTB) void k_shell_agg(const c64* __restrict__ rk, const c64* __restrict__ ekf,
                                                   long long N, const int* __restrict__ perm,
                                                   const int* __restrict__ off, int nsh, c64* __restrict__ z,
                                                   double* __restrict__ w2) {
  const int sh = blockIdx.x, i = blockIdx.y;
  const int a = off[sh], b = off[sh + 1];
  double zr = 0.0, zi = 0.0, ww = 0.0;
  for (int q = a + (int)threadIdx.x; q < b; q += kTB) {
    const long long k = perm[q];
    const c64 r = rk[(size_t)i * N + k], e = ekf[(size_t)i * N + k];
    zr += e.x * r.x + e.y * r.y;
    zi += e.x * r.y - e.y * r.x;
    ww += r.x * r.x + r.y * r.y;
  }
  __shared__ double red[3][kTB];
  red[0][threadIdx.x] = zr;
  red[1][threadIdx.x] = zi;
  red[2][threadIdx.x] = ww;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int t = 0; t < 3; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    z[(size_t)i * nsh + sh] = mk(red[0][0], red[1][0]);
    w2[(size_t)i * nsh + sh] = red[2][0];
  }
}

// per band: A = sum_s M_s z_s, B = sum_s M_s^2 w2_s (one block per band)
__global__ __launch_bounds__(kTB) void k_loss_shells(const c64* __restrict__ z, const double* __restrict__ w2, int nsh,
                                                     const double* __restrict__ eks, double N,
                                                     const double* __restrict__ tkin, const double* __restrict__ yfeat,
                                                     const double* __restrict__ P, int nh, double* __restrict__ out) {
  const int i = blockIdx.x;
  double ar = 0.0, ai = 0.0, bb = 0.0;
  for (int sh = threadIdx.x; sh < nsh; sh += kTB) {
    Fwd f;
    const double m = multiplier(P, nh, eks[sh] * N / tkin[i], yfeat[i], f);
    const c64 v = z[(size_t)i * nsh + sh];
    ar += m * v.x;
    ai += m * v.y;
    bb += m * m * w2[(size_t)i * nsh + sh];
  }
  __shared__ double red[3][kTB];
  red[0][threadIdx.x] = ar;
  red[1][threadIdx.x] = ai;
  red[2][threadIdx.x] = bb;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int t = 0; t < 3; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[(size_t)i * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// parameter gradient over the (band, shell) samples q = i * nsh + s, as
// k_precond_grad with conj(e) r -> z_s and |r|^2 -> w2_s
__global__ __launch_bounds__(kTile) void k_grad_shells(const c64* __restrict__ z, const double* __restrict__ w2,
                                                       int nsh, int nb, const double* __restrict__ eks, double N,
                                                       const double* __restrict__ tkin,
                                                       const double* __restrict__ yfeat, const double* __restrict__ P,
                                                       int nh, const double* __restrict__ coefA,
                                                       const double* __restrict__ coefB, int npar,
                                                       double* __restrict__ gpart, int nchunk) {
  __shared__ double s_u[kTile][2];
  __shared__ double s_h1[kTile][kMaxH];
  __shared__ double s_dz1[kTile][kMaxH];
  __shared__ double s_h2[kTile][kMaxH];
  __shared__ double s_dz2[kTile][kMaxH];
  __shared__ double s_dg[kTile];
  const long long total = (long long)nsh * nb;
  long long lo, hi;
  chunk_bounds(total, nchunk, blockIdx.x, lo, hi);
  const double* W2 = P + 3 * nh;
  const double* w3 = W2 + nh * nh + nh;
  double acc[kSlots];
  for (int slot = 0; slot < kSlots; ++slot) acc[slot] = 0.0;
  for (long long t0 = lo; t0 < hi; t0 += kTile) {
    const long long q = t0 + threadIdx.x;
    if (q < hi) {
      const int i = (int)(q / nsh);
      const int sh = (int)(q - (long long)i * nsh);
      Fwd f;
      const double m = multiplier(P, nh, eks[sh] * N / tkin[i], yfeat[i], f);
      const c64 v = z[q];
      const double dg = (coefA[2 * i] * v.x - coefA[2 * i + 1] * v.y + coefB[i] * m * w2[q]) * m;
      s_u[threadIdx.x][0] = f.u0;
      s_u[threadIdx.x][1] = f.u1;
      s_dg[threadIdx.x] = dg;
      double dz2[kMaxH];
      for (int a = 0; a < nh; ++a) {
        dz2[a] = dg * w3[a] * (1.0 - f.h2[a] * f.h2[a]);
        s_h2[threadIdx.x][a] = f.h2[a];
        s_dz2[threadIdx.x][a] = dz2[a];
        s_h1[threadIdx.x][a] = f.h1[a];
      }
      for (int b = 0; b < nh; ++b) {
        double dh1 = 0.0;
        for (int a = 0; a < nh; ++a) dh1 += W2[a * nh + b] * dz2[a];
        s_dz1[threadIdx.x][b] = dh1 * (1.0 - f.h1[b] * f.h1[b]);
      }
    }
    __syncthreads();
    const int ns = (int)(hi - t0 < kTile ? hi - t0 : kTile);
    for (int slot = 0; slot < kSlots; ++slot) {
      const int p = threadIdx.x + slot * kTile;
      if (p >= npar) break;
      double sm = 0.0;
      if (p < 2 * nh) {
        const int a = p / 2, cc = p % 2;
        for (int t = 0; t < ns; ++t) sm += s_dz1[t][a] * s_u[t][cc];
      } else if (p < 3 * nh) {
        const int a = p - 2 * nh;
        for (int t = 0; t < ns; ++t) sm += s_dz1[t][a];
      } else if (p < 3 * nh + nh * nh) {
        const int a = (p - 3 * nh) / nh, b = (p - 3 * nh) % nh;
        for (int t = 0; t < ns; ++t) sm += s_dz2[t][a] * s_h1[t][b];
      } else if (p < 4 * nh + nh * nh) {
        const int a = p - 3 * nh - nh * nh;
        for (int t = 0; t < ns; ++t) sm += s_dz2[t][a];
      } else if (p < 5 * nh + nh * nh) {
        const int a = p - 4 * nh - nh * nh;
        for (int t = 0; t < ns; ++t) sm += s_dg[t] * s_h2[t][a];
      } else {
        for (int t = 0; t < ns; ++t) sm += s_dg[t];
      }
      acc[slot] += sm;
    }
    __syncthreads();
  }
  for (int slot = 0; slot < kSlots; ++slot) {
    const int p = threadIdx.x + slot * kTile;
    if (p < npar) gpart[(size_t)blockIdx.x * npar + p] = acc[slot];
  }
}

}  // namespace

hipError_t shell_aggregate(hipStream_t st, const c64* rk, const c64* ekf, long long N, int nb, const int* perm,
                           const int* off, int nsh, c64* z, double* w2) {
  hipLaunchKernelGGL(k_shell_agg, dim3(nsh, nb), dim3(kTB), 0, st, rk, ekf, N, perm, off, nsh, z, w2);
  return hipGetLastError();
}

hipError_t loss_shells(hipStream_t st, const c64* z, const double* w2, int nsh, int nb, const double* eks, long long N,
                       const double* tkin, const double* yfeat, const double* params, MlpShape s, double* out) {
  hipLaunchKernelGGL(k_loss_shells, dim3(nb), dim3(kTB), 0, st, z, w2, nsh, eks, (double)N, tkin, yfeat, params, s.nh,
                     out);
  return hipGetLastError();
}

hipError_t grad_shells(hipStream_t st, const c64* z, const double* w2, int nsh, int nb, const double* eks, long long N,
                       const double* tkin, const double* yfeat, const double* params, MlpShape s, const double* coefA,
                       const double* coefB, double* gpart, int nchunk) {
  hipLaunchKernelGGL(k_grad_shells, dim3(nchunk), dim3(kTile), 0, st, z, w2, nsh, nb, eks, (double)N, tkin, yfeat,
                     params, s.nh, coefA, coefB, mlp_param_count(s), gpart, nchunk);
  return hipGetLastError();
}

hipError_t colsum_accumulate(hipStream_t st, const double* part, int nrow, int npar, double* acc) {
  hipLaunchKernelGGL(k_colsum_acc, dim3(npar), dim3(kTB), 0, st, part, nrow, npar, acc);
  return hipGetLastError();
}

hipError_t sub_real(hipStream_t st, const double* a, const double* b, long long n, double* out) {
  hipLaunchKernelGGL(k_sub_real, dim3(blocks_for(n)), dim3(kTB), 0, st, a, b, n, out);
  return hipGetLastError();
}

hipError_t rotate_rect(hipStream_t st, const c64* in, c64* out, long long N, int m, int nout, const c64* T) {
  if (m < 1 || nout < 1 || N < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_rotate_mfma, dim3((unsigned)((N + kRmP - 1) / kRmP), (nout + kRmJ - 1) / kRmJ), dim3(256), 0,
                     st, in, out, N, m, nout, T);
  return hipGetLastError();
}

hipError_t residual_vectors(hipStream_t st, const c64* hp, const c64* phi, const double* eig, long long N, int nb,
                            c64* r) {
  hipLaunchKernelGGL(k_resid_vec, dim3(blocks_for(N * nb)), dim3(kTB), 0, st, hp, phi, eig, N, nb, r);
  return hipGetLastError();
}

hipError_t kinetic_partials(hipStream_t st, const c64* fk, const double* ek, long long N, int nb, double* part) {
  hipLaunchKernelGGL(k_kin, dim3(kGsChunks, nb), dim3(kTB), 0, st, fk, ek, N, nb, part);
  return hipGetLastError();
}

template <typename R>
hipError_t cheb_combine(hipStream_t st, cx<R>* cc, const cx<R>* cb, const cx<R>* ca, long long n, double havg,
                        double dh, int first) {
  hipLaunchKernelGGL((k_cheb<R>), dim3(blocks_for(n)), dim3(kTB), 0, st, cc, cb, ca, n, havg, dh, first);
  return hipGetLastError();
}

template <typename R>
hipError_t residual_partials(hipStream_t st, const cx<R>* hp, const cx<R>* phi, const double* eig, long long N,
                             int nb, double* part) {
  hipLaunchKernelGGL((k_resid<R>), dim3(kGsChunks, nb), dim3(kTB), 0, st, hp, phi, eig, N, nb, part);
  return hipGetLastError();
}

hipError_t real_dots(hipStream_t st, const double* a, long long N, int K, double* part) {
  hipLaunchKernelGGL(k_real_dots, dim3(kGsChunks, K, K), dim3(kTB), 0, st, a, N, K, part);
  return hipGetLastError();
}

hipError_t diis_predict(hipStream_t st, const double* bank, const double* err, const double* coef, int K, double mix,
                        long long N, double* out) {
  hipLaunchKernelGGL(k_diis_predict, dim3(blocks_for(N)), dim3(kTB), 0, st, bank, err, coef, K, mix, N, out);
  return hipGetLastError();
}

hipError_t to_f32(hipStream_t st, const double* in, float* out, long long n) {
  hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(n)), dim3(kTB), 0, st, in, out, n);
  return hipGetLastError();
}

hipError_t scale_complex(hipStream_t st, c64* a, long long n, double s) {
  hipLaunchKernelGGL(k_scale_c, dim3(blocks_for(n)), dim3(kTB), 0, st, a, n, s);
  return hipGetLastError();
}

hipError_t axpy_complex(hipStream_t st, const c64* a, const c64* b, double s, long long n, c64* out) {
  hipLaunchKernelGGL(k_axpy_c, dim3(blocks_for(n)), dim3(kTB), 0, st, a, b, s, n, out);
  return hipGetLastError();
}

int mlp_param_count(const MlpShape& s) { return s.nh * s.nin + s.nh + s.nh * s.nh + s.nh + s.nh + 1; }

hipError_t precond_apply_shells(hipStream_t st, c64* rk, long long N, int nb, const int* shell_of, const double* eks,
                                int nsh, const double* tkin, const double* yfeat, const double* params, MlpShape s,
                                double* tab) {
  hipLaunchKernelGGL(k_precond_table, dim3(blocks_for((long long)nsh * nb)), dim3(kTB), 0, st, eks, nsh, N, nb, tkin,
                     yfeat, params, s.nh, tab);
  hipLaunchKernelGGL(k_precond_lookup, dim3(blocks_for(N)), dim3(kTB), 0, st, rk, N, nb, shell_of, tab, nsh);
  return hipGetLastError();
}

hipError_t precond_loss_sums(hipStream_t st, const c64* rk, const c64* ekf, long long N, int nb, const double* ek,
                             const double* tkin, const double* yfeat, const double* params, MlpShape s, double* part) {
  hipLaunchKernelGGL(k_loss_sums, dim3(kGsChunks, nb), dim3(kTB), 0, st, rk, ekf, N, nb, ek, tkin, yfeat, params, s.nh,
                     part);
  return hipGetLastError();
}

hipError_t precond_grad(hipStream_t st, const c64* rk, const c64* ekf, long long N, int nb, const double* ek,
                        const double* tkin, const double* yfeat, const double* params, MlpShape s,
                        const double* coefA, const double* coefB, double* gpart, int nchunk) {
  hipLaunchKernelGGL(k_precond_grad, dim3(nchunk), dim3(kTile), 0, st, rk, ekf, N, nb, ek, tkin, yfeat, params, s.nh,
                     coefA, coefB, mlp_param_count(s), gpart, nchunk);
  return hipGetLastError();
}

template hipError_t cheb_combine<double>(hipStream_t, cx<double>*, const cx<double>*, const cx<double>*, long long,
                                         double, double, int);
template hipError_t cheb_combine<float>(hipStream_t, cx<float>*, const cx<float>*, const cx<float>*, long long, double,
                                        double, int);
template hipError_t residual_partials<double>(hipStream_t, const cx<double>*, const cx<double>*, const double*,
                                              long long, int, double*);
template hipError_t residual_partials<float>(hipStream_t, const cx<float>*, const cx<float>*, const double*, long long,
                                             int, double*);

}  // namespace opaw
