// Batched 3D FFT passes for the OPAW hot path (gfx950).
//
// A 3D transform is three passes of 1D line FFTs held in LDS:
//   * x pass: a workgroup owns LINES whole x-lines (contiguous in HBM);
//   * y / z pass: a workgroup owns XT consecutive x-columns of one y- or
//     z-line family, read as XT-wide contiguous rows (coalesced).
// Every pass addresses its input and output through a "view" (offsets and
// leading dimensions), and loads / stores only a window along the transform
// axis; outside the load window the line is zero.  That lets the same kernels
// run the kinetic operator on the orbital grid and the Martyna-Tuckerman
// Hartree convolution on the doubled box with the zero padding pruned (the
// embedding of the density and the crop of V_H are folded into the first and
// last passes).
//
// Passes may run in place (in == out): every tile is read completely before
// the barrier that precedes its stores, and tiles do not overlap.
//
// Fused epilogues:
//   EPI_MULT  (y/z pass): forward FFT, multiply by a real table T(k) (ek for
//             the kinetic term, vk for Hartree), inverse FFT, store.
//   EPI_HPSI  (x pass):   out = IFFT result + vks(sys) * u (+ sum_a p_a wnl_a)
//             (h_phi, src/hphi.f90:14-21 with addvnl's add-back gathered)
// and a fused input correction: when cpre is given, the first pass loads
// u = in + cpre (the compact S^-1/2 correction of the union of boxes,
// src/snphi.f90), and the H.psi epilogue rebuilds the same u for its vks * u
// term and adds the compact nonlocal correction cnl.
#include "fft_core.h"
#include "opaw_fields.h"
#include "opaw_kernels.h"

namespace opaw {

struct View {
  long long fs;   // field stride (elements)
  int sx, sy, sz; // coordinate offsets added before addressing
  int ldx, ldy;   // leading dimensions
  // tables even in every k axis (the Hartree kernel vk): nonzero fold
  // lengths store only the octant 0..n/2 per axis, read at min(i, n - i)
  int fx = 0, fy = 0, fz = 0;
};

__device__ __forceinline__ long long vaddr(const View& v, int f, int x, int y, int z) {
  return (long long)f * v.fs + (long long)(x + v.sx) +
         (long long)v.ldx * ((long long)(y + v.sy) + (long long)v.ldy * (long long)(z + v.sz));
}

// address of a multiplier table entry, through the octant fold when set
__device__ __forceinline__ long long taddr(const View& v, int x, int y, int z) {
  if (v.fx) {
    x += v.sx;
    y += v.sy;
    z += v.sz;
    x = min(x, v.fx - x);
    y = min(y, v.fy - y);
    z = min(z, v.fz - z);
    return (long long)x + (long long)v.ldx * ((long long)y + (long long)v.ldy * (long long)z);
  }
  return vaddr(v, 0, x, y, z);
}

struct XLines {
  int y0, ly, z0, lz, nf;   // line set: y in [y0,y0+ly), z in [z0,z0+lz), f < nf
  int la, lb;               // load window along x
  int sa, sb;               // store window along x
};

struct SLines {
  int nx;                   // x extent of the line set (x in [0,nx))
  int o0, lo, nf;           // other coordinate in [o0,o0+lo), f < nf
  int la, lb, sa, sb;       // windows along the transform axis
};

enum { EPI_NONE = 0, EPI_MULT = 1, EPI_HPSI = 2, EPI_RK4F = 3, EPI_RK4 = 4 };

constexpr int kThreads = 256;

template <int N>
constexpr int x_lines() { return (1024 / N) > 0 ? (1024 / N) : 1; }

// x-tile line stride: lines of a multiple-of-16 length would all start on the
// same LDS bank group; a stride of N + 8 elements spreads a wave's lines over
// both halves of the 16 bank groups that one 16-byte element spans.
template <int N>
constexpr int x_stride() { return (N % 16 == 0) ? N + 8 : N; }

// ---------------------------------------------------------------------------
// x pass.  N > 0: a compiled length (static LDS, in-place register-staged
// stages).  N == 0: the runtime length n with its radix plan (dynamic LDS
// holding two tiles and the twiddles, out-of-place stages).
template <typename R, int N, int S, int EPI>
__global__ __launch_bounds__(kThreads) void k_fft_x(const cx<R>* in, View vin,
                                                     cx<R>* out, View vout, XLines L,
                                                     const cx<R>* __restrict__ twg,
                                                     const cx<R>* __restrict__ u,
                                                     const R* __restrict__ vks, int nf_per_sys,
                                                     const int* __restrict__ umap, const c64* __restrict__ cpre,
                                                     const c64* __restrict__ cnl, int U, int n_rt,
                                                     fft::Plan plan) {
  const int n = N > 0 ? N : n_rt;
  const int LINES = N > 0 ? x_lines<(N > 0 ? N : 1)>() : ((1024 / n) > 0 ? 1024 / n : 1);
  const int LS = N > 0 ? x_stride<(N > 0 ? N : 1)>() : ((n % 16 == 0) ? n + 8 : n);
  cx<R>* tile;
  cx<R>* tile2 = nullptr;
  cx<R>* tw;
  if constexpr (N > 0) {
    __shared__ cx<R> s_tile[x_lines<(N > 0 ? N : 1)>() * x_stride<(N > 0 ? N : 1)>()];
    __shared__ cx<R> s_tw[N > 0 ? N : 1];
    tile = s_tile;
    tw = s_tw;
  } else {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    tile = reinterpret_cast<cx<R>*>(dyn_lds);
    tile2 = tile + LINES * LS;
    tw = tile2 + LINES * LS;
  }
  fft::load_twiddles(tw, twg, n);
  const long long nlines = (long long)L.ly * L.lz * L.nf;
  const long long line0 = (long long)blockIdx.x * LINES;
  for (int idx = threadIdx.x; idx < LINES * n; idx += kThreads) {
    const int l = idx / n, x = idx % n;
    const long long gl = line0 + l;
    cx<R> val = mk<R>(0, 0);
    if (gl < nlines && x >= L.la && x < L.lb) {
      const int y = L.y0 + (int)(gl % L.ly);
      const long long r = gl / L.ly;
      const int z = L.z0 + (int)(r % L.lz);
      const int f = (int)(r / L.lz);
      val = in[vaddr(vin, f, x, y, z)];
      if (EPI != EPI_HPSI && cpre) {
        const int k = umap[vaddr(vin, 0, x, y, z)];
        if (k >= 0) {
          const c64 c = cpre[(long long)k * U + f];
          val = mk<R>((R)((double)val.x + c.x), (R)((double)val.y + c.y));
        }
      }
    }
    tile[l * LS + x] = val;
  }
  __syncthreads();
  const cx<R>* res = tile;
  if constexpr (N > 0)
    fft::lines<R, (N > 0 ? N : 1), x_lines<(N > 0 ? N : 1)>(), kThreads, false, x_stride<(N > 0 ? N : 1)>(), 1, S>(
        tile, tw);
  else
    res = fft::lines_gen<R, S>(tile, tile2, tw, n, LINES, false, LS, 1, plan);
  for (int idx = threadIdx.x; idx < LINES * n; idx += kThreads) {
    const int l = idx / n, x = idx % n;
    const long long gl = line0 + l;
    if (gl < nlines && x >= L.sa && x < L.sb) {
      const int y = L.y0 + (int)(gl % L.ly);
      const long long r = gl / L.ly;
      const int z = L.z0 + (int)(r % L.lz);
      const int f = (int)(r / L.lz);
      const long long o = vaddr(vout, f, x, y, z);
      cx<R> val = res[l * LS + x];
      if constexpr (EPI == EPI_HPSI) {
        const int sys = f / nf_per_sys;
        const long long pt = o - (long long)f * vout.fs;
        const double v = (double)vks[(long long)sys * vout.fs + pt];
        const cx<R> uu = u[o];
        double ur = (double)uu.x, ui = (double)uu.y;
        double ar = (double)val.x, ai = (double)val.y;
        if (cpre || cnl) {
          const int k = umap[pt];
          if (k >= 0) {
            if (cpre) {
              const c64 c = cpre[(long long)k * U + f];
              ur += c.x;
              ui += c.y;
            }
            if (cnl) {
              const c64 c = cnl[(long long)k * U + f];
              ar += c.x;
              ai += c.y;
            }
          }
        }
        val = mk<R>((R)(ar + v * ur), (R)(ai + v * ui));
      }
      out[o] = val;
    }
  }
}

// ---------------------------------------------------------------------------
// y (AXIS=1) or z (AXIS=2) pass over XT x-columns per workgroup (16: more,
// smaller workgroups measured faster than 32- or 64-wide tiles on the 64^3
// z pass)
constexpr int kSXT = 16;
// columns per workgroup of a compiled length: 32 for lines of at most 48
// points, so a workgroup still moves a useful amount of data
constexpr int sxt_of(int n) { return n <= 48 ? 2 * kSXT : kSXT; }

template <typename R, int N, int AXIS, int S, int EPI>
__global__ __launch_bounds__(kThreads) void k_fft_s(const cx<R>* in, View vin,
                                                     cx<R>* out, View vout, SLines L,
                                                     const cx<R>* __restrict__ twg,
                                                     const R* __restrict__ tab, View vtab, int n_rt, int xt_rt,
                                                     fft::Plan plan) {
  // N == 0: runtime length n_rt over xt_rt columns (dynamic LDS: two tiles
  // and the twiddles)
  const int n = N > 0 ? N : n_rt;
  constexpr int CXT = sxt_of(N > 0 ? N : 256);
  const int XT = N > 0 ? CXT : xt_rt;
  cx<R>* tile;
  cx<R>* tile2 = nullptr;
  cx<R>* tw;
  if constexpr (N > 0) {
    __shared__ cx<R> s_tile[CXT * (N > 0 ? N : 1)];
    __shared__ cx<R> s_tw[N > 0 ? N : 1];
    tile = s_tile;
    tw = s_tw;
  } else {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    tile = reinterpret_cast<cx<R>*>(dyn_lds);
    tile2 = tile + XT * n;
    tw = tile2 + XT * n;
  }
  fft::load_twiddles(tw, twg, n);
  const int xtiles = (L.nx + XT - 1) / XT;
  long long b = blockIdx.x;
  const int xt = (int)(b % xtiles);
  b /= xtiles;
  const int o = L.o0 + (int)(b % L.lo);
  const int f = (int)(b / L.lo);
  const int x0 = xt * XT;
  for (int idx = threadIdx.x; idx < XT * n; idx += kThreads) {
    const int c = idx % XT, j = idx / XT;
    const int x = x0 + c;
    cx<R> val = mk<R>(0, 0);
    if (x < L.nx && j >= L.la && j < L.lb) {
      const long long a = (AXIS == 1) ? vaddr(vin, f, x, j, o) : vaddr(vin, f, x, o, j);
      val = in[a];
    }
    tile[idx] = val;
  }
  __syncthreads();
  cx<R>* res = tile;
  if constexpr (N > 0)
    fft::lines<R, (N > 0 ? N : 1), CXT, kThreads, true, 1, CXT, S>(tile, tw);
  else
    res = fft::lines_gen<R, S>(tile, tile2, tw, n, XT, true, 1, XT, plan);
  if constexpr (EPI == EPI_MULT) {
    for (int idx = threadIdx.x; idx < XT * n; idx += kThreads) {
      const int c = idx % XT, j = idx / XT;
      const int x = x0 + c;
      if (x < L.nx) {
        const long long a = (AXIS == 1) ? taddr(vtab, x, j, o) : taddr(vtab, x, o, j);
        const R t = tab[a];
        res[idx] = mk(res[idx].x * t, res[idx].y * t);
      }
    }
    __syncthreads();
    if constexpr (N > 0)
      fft::lines<R, (N > 0 ? N : 1), CXT, kThreads, true, 1, CXT, -S>(tile, tw);
    else
      res = fft::lines_gen<R, -S>(res, res == tile ? tile2 : tile, tw, n, XT, true, 1, XT, plan);
  }
  for (int idx = threadIdx.x; idx < XT * n; idx += kThreads) {
    const int c = idx % XT, j = idx / XT;
    const int x = x0 + c;
    if (x < L.nx && j >= L.sa && j < L.sb) {
      const long long a = (AXIS == 1) ? vaddr(vout, f, x, j, o) : vaddr(vout, f, x, o, j);
      out[a] = res[idx];
    }
  }
}

// ---------------------------------------------------------------------------
// xy-plane pass: a workgroup holds one (field, z) plane in LDS and transforms
// its rows (x) and columns (y) in one visit to HBM, so the kinetic operator is
// plane-forward, z (fwd * ek inv), plane-inverse: three passes instead of five.
//   EPI_NONE, S=-1: u = in (+ cpre) -> FFT_xy -> out
//   EPI_HPSI, S=+1: in -> IFFT_xy -> out = result + vks * u + cnl,
//                   u = uin (+ cpre)
//   EPI_RK4F / EPI_RK4 (RK4 stage 1 / stages 2-3, src/rk4.f90:35-56): the
//                   same h, then k = -i h, acc = k (or acc += 2k) and
//                   out = p + c k (stage 1 transforms p itself: uin is p); h itself is stored only on the union of
//                   boxes, compactly in hU[k*U + f] (FP64), for the C
//                   projection whose add-back is applied to acc and out
//                   afterwards (it is linear in h)
constexpr int kPlaneThreads = 512;

template <typename R, int NX, int NY>
constexpr int plane_stride() {
  // pad rows by 8 elements (spreads a wave's rows over the LDS bank groups)
  // when the padded plane still fits the 160 KB of LDS
  return ((NX % 16 == 0) && (long)NY * (NX + 8) * (long)sizeof(cx<R>) + (NX + NY) * (long)sizeof(cx<R>) <= 163840)
             ? NX + 8
             : NX;
}

template <typename R, int NX, int NY, int S, int EPI>
__global__ __launch_bounds__(kPlaneThreads) void k_fft_plane(const cx<R>* in, cx<R>* out, int nz, long long N,
                                                              const cx<R>* __restrict__ twxg,
                                                              const cx<R>* __restrict__ twyg,
                                                              const cx<R>* uin,
                                                              const R* __restrict__ vks, int nf_per_sys,
                                                              const int* __restrict__ upt,
                                                              const int* __restrict__ uzr,
                                                              const c64* __restrict__ cpre,
                                                              const c64* __restrict__ cnl, int U,
                                                              PlaneRk4<R> rk) {
  // uin may alias out (RK4 stages 2-3 update y in place: each element is read
  // and written by the same thread), so neither is __restrict__
  constexpr bool HP = EPI == EPI_HPSI || EPI == EPI_RK4F || EPI == EPI_RK4;
  constexpr int LS = plane_stride<R, NX, NY>();
  constexpr int NP = NX * NY;
  constexpr int PER = (NP + kPlaneThreads - 1) / kPlaneThreads;
  constexpr bool EXACT = NP % kPlaneThreads == 0;
  // the epilogue's global operands are loaded before the transforms when
  // they fit in a few registers, so their latency hides behind the FFT
  constexpr bool EARLY = HP && PER * (3 * sizeof(R)) <= 96;
  __shared__ cx<R> tile[NY * LS];
  __shared__ cx<R> twx[NX];
  __shared__ cx<R> twy[NY];
  fft::load_twiddles(twx, twxg, NX);
  fft::load_twiddles(twy, twyg, NY);
  const int f = blockIdx.x / nz;
  const int z = blockIdx.x % nz;
  const long long plane0 = (long long)z * NP;            // offset of the plane in a field
  const cx<R>* src = in + (long long)f * N + plane0;
  // plane element idx of this thread's q-th slot, clamped for the loads
  auto slot = [](int q) { return (int)threadIdx.x + q * kPlaneThreads; };
  auto lidx = [&](int q) { return EXACT ? slot(q) : min(slot(q), NP - 1); };
  auto tix = [](int idx) { return (idx / NX) * LS + idx % NX; };
  const R* vk = nullptr;
  const cx<R>* us = nullptr;
  if constexpr (HP) {
    vk = vks + (long long)(f / nf_per_sys) * N + plane0;
    us = uin + (long long)f * N + plane0;
  }
  cx<R> uu[EARLY ? PER : 1];
  R vv[EARLY ? PER : 1];
  {
    cx<R> v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) v[q] = src[lidx(q)];
    if constexpr (EARLY) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        uu[q] = us[lidx(q)];
        vv[q] = vk[lidx(q)];
      }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q)
      if (EXACT || slot(q) < NP) tile[tix(slot(q))] = v[q];
  }
  // union points of this plane: a contiguous range (the union is sorted by
  // grid point); forward pass: u = in + cpre there
  const bool useu = cpre || cnl || (EPI == EPI_RK4F || EPI == EPI_RK4);
  const int k0 = useu ? uzr[z] : 0, k1 = useu ? uzr[z + 1] : 0;
  if (EPI == EPI_NONE && cpre) {
    __syncthreads();
    for (int k = k0 + (int)threadIdx.x; k < k1; k += kPlaneThreads) {
      const int loc = (int)(upt[k] - plane0);
      const c64 c = cpre[(long long)k * U + f];
      cx<R>& t = tile[tix(loc)];
      t = mk<R>((R)((double)t.x + c.x), (R)((double)t.y + c.y));
    }
  }
  __syncthreads();
  if constexpr (S < 0) {
    fft::lines<R, NX, NY, kPlaneThreads, false, LS, 1, S>(tile, twx);   // rows (x)
    fft::lines<R, NY, NX, kPlaneThreads, true, 1, LS, S>(tile, twy);    // columns (y)
  } else {
    fft::lines<R, NY, NX, kPlaneThreads, true, 1, LS, S>(tile, twy);
    fft::lines<R, NX, NY, kPlaneThreads, false, LS, 1, S>(tile, twx);
  }
  cx<R>* dst = out + (long long)f * N + plane0;
  if constexpr (HP) {
    // h = T u + vks u + cnl with u = uin + cpre: the union terms
    // vks * cpre + cnl go into the tile first, then the plain epilogue
    if (cpre || cnl) {
      for (int k = k0 + (int)threadIdx.x; k < k1; k += kPlaneThreads) {
        const int loc = (int)(upt[k] - plane0);
        double ar = 0.0, ai = 0.0;
        if (cpre) {
          const double w = (double)vk[loc];
          const c64 c = cpre[(long long)k * U + f];
          ar += w * c.x;
          ai += w * c.y;
        }
        if (cnl) {
          const c64 c = cnl[(long long)k * U + f];
          ar += c.x;
          ai += c.y;
        }
        cx<R>& t = tile[tix(loc)];
        t = mk<R>((R)((double)t.x + ar), (R)((double)t.y + ai));
      }
      __syncthreads();
    }
    if constexpr (EPI == EPI_HPSI) {
      if constexpr (!EARLY) {
        // the operands of a batch are requested together: out may alias
        // uin, so the compiler cannot move a load above an earlier store,
        // but every element is read and written by one thread, so reading a
        // batch before writing it is safe in place
        constexpr int HB = PER <= 24 ? PER : 8;
        for (int q0 = 0; q0 < PER; q0 += HB) {
          cx<R> ub[HB];
          R vb[HB];
#pragma unroll
          for (int s = 0; s < HB; ++s) {
            const int q = min(q0 + s, PER - 1);
            ub[s] = us[lidx(q)];
            vb[s] = vk[lidx(q)];
          }
#pragma unroll
          for (int s = 0; s < HB; ++s) {
            const int q = q0 + s;
            if (q < PER && (EXACT || slot(q) < NP)) {
              const cx<R> t = tile[tix(slot(q))];
              const double w = (double)vb[s];
              dst[slot(q)] = mk<R>((R)((double)t.x + w * (double)ub[s].x), (R)((double)t.y + w * (double)ub[s].y));
            }
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          if (EXACT || slot(q) < NP) {
            const cx<R> t = tile[tix(slot(q))];
            const double w = (double)vv[q];
            dst[slot(q)] = mk<R>((R)((double)t.x + w * (double)uu[q].x), (R)((double)t.y + w * (double)uu[q].y));
          }
        }
      }
    } else {
      // RK4 stage update in the epilogue
      const cx<R>* pb = rk.p + (long long)f * N + plane0;
      cx<R>* ab = rk.acc + (long long)f * N + plane0;
      for (int q0 = 0; q0 < PER; q0 += 4) {
        cx<R> u4[4], p4[4], a4[4];
        R v4[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int q = min(q0 + s, PER - 1);
          if constexpr (EARLY) {
            u4[s] = uu[q];
            v4[s] = vv[q];
          } else {
            u4[s] = us[lidx(q)];
            v4[s] = vk[lidx(q)];
          }
          // stage 1 transforms p itself (uin == p): no second read
          if constexpr (EPI == EPI_RK4F)
            p4[s] = u4[s];
          else
            p4[s] = pb[lidx(q)];
          if constexpr (EPI == EPI_RK4) a4[s] = ab[lidx(q)];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int q = q0 + s;
          if (q < PER && (EXACT || slot(q) < NP)) {
            cx<R>& tt = tile[tix(slot(q))];
            const double w = (double)v4[s];
            const cx<R> h = mk<R>((R)((double)tt.x + w * (double)u4[s].x), (R)((double)tt.y + w * (double)u4[s].y));
            tt = h;   // kept for the compact copy on the union of boxes
            const double kr = (double)h.y, ki = -(double)h.x;
            if constexpr (EPI == EPI_RK4F)
              ab[slot(q)] = mk<R>((R)kr, (R)ki);
            else
              ab[slot(q)] = mk<R>((R)((double)a4[s].x + 2.0 * kr), (R)((double)a4[s].y + 2.0 * ki));
            dst[slot(q)] = mk<R>((R)((double)p4[s].x + rk.c * kr), (R)((double)p4[s].y + rk.c * ki));
          }
        }
      }
      __syncthreads();
      for (int k = k0 + (int)threadIdx.x; k < k1; k += kPlaneThreads) {
        const cx<R> h = tile[tix((int)(upt[k] - plane0))];
        rk.hU[(long long)k * U + f] = mk((double)h.x, (double)h.y);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < PER; ++q)
      if (EXACT || slot(q) < NP) dst[slot(q)] = tile[tix(slot(q))];
  }
}

// (nx, ny) planes with a compiled plane kernel: the squares and the
// rectangles of the polyacene grids (64 x 48, 80 x 64, 96 x 64, 128 x 64);
// a pair runs on the plane path where its unpadded plane and twiddles fit
// the 160 KB of LDS (all in FP32; FP64 up to 96 x 96 and 128 x 64)
#define OPAW_PLANE_PAIRS(X)                                                                              \
  X(16, 16) X(24, 24) X(32, 32) X(48, 48) X(64, 64) X(80, 80) X(96, 96) X(128, 128) X(64, 48) X(80, 64) \
  X(96, 64) X(128, 64)

template <typename R, int NX, int NY>
constexpr bool plane_fits() {
  return (long)NY * NX * (long)sizeof(cx<R>) + (NX + NY) * (long)sizeof(cx<R>) <= 163840;
}

template <typename R>
bool plane_supported(int nx, int ny) {
#define OPAW_CASE(a, b) \
  if (nx == a && ny == b) return plane_fits<R, a, b>();
  OPAW_PLANE_PAIRS(OPAW_CASE)
#undef OPAW_CASE
  return false;
}
template bool plane_supported<double>(int, int);
template bool plane_supported<float>(int, int);

template <typename R, int S, int EPI>
static hipError_t launch_plane(int nx, int ny, hipStream_t st, const cx<R>* in, cx<R>* out, int nz, int nf,
                               long long N, const cx<R>* twx, const cx<R>* twy, const cx<R>* uin, const R* vks,
                               int nfps, const int* upt, const int* uzr, const c64* cpre, const c64* cnl, int U,
                               PlaneRk4<R> rk = PlaneRk4<R>{}) {
  dim3 grid((unsigned)((long long)nf * nz));
#define OPAW_CASE(a, b)                                                                                      \
  if (nx == a && ny == b) {                                                                                  \
    if constexpr (plane_fits<R, a, b>()) {                                                                   \
      hipLaunchKernelGGL((k_fft_plane<R, a, b, S, EPI>), grid, dim3(kPlaneThreads), 0, st, in, out, nz, N, twx, \
                         twy, uin, vks, nfps, upt, uzr, cpre, cnl, U, rk);                                   \
      return hipGetLastError();                                                                              \
    }                                                                                                        \
    return hipErrorInvalidValue;                                                                             \
  }
  OPAW_PLANE_PAIRS(OPAW_CASE)
#undef OPAW_CASE
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// host-side dispatch over the compiled line lengths
#define OPAW_FFT_SIZES(X) \
  X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) \
  X(128) X(160) X(192) X(256)

static bool fft_size_compiled(int n) {
  switch (n) {
#define OPAW_CASE(n) case n: return true;
    OPAW_FFT_SIZES(OPAW_CASE)
#undef OPAW_CASE
    default: return false;
  }
}

// runtime-length transforms: dynamic LDS of the x pass (two tiles of
// 1024/n lines plus twiddles) and of the y/z pass (two tiles of xt columns
// plus twiddles) stay within 64 KiB
static constexpr int kMaxRuntimeN = 256;

static int x_lines_rt(int n) {
  int l = 1024 / n;
  return l > 0 ? l : 1;
}

static int s_xt_rt(int n, int esize) {
  int xt = kSXT;
  while (xt > 1 && (size_t)(2 * xt + 1) * n * esize > 65536) xt /= 2;
  return xt;
}

bool fft_size_supported(int n) {
  if (fft_size_compiled(n)) return true;
  return n >= 2 && n <= kMaxRuntimeN && fft::make_plan(n).nr > 0;
}

template <typename R, int S, int EPI>
static hipError_t launch_x_n(int n, dim3 grid, hipStream_t st, const cx<R>* in, View vin, cx<R>* out,
                             View vout, XLines L, const cx<R>* tw, const cx<R>* u, const R* vks,
                             int nfps, const int* umap, const c64* cpre, const c64* cnl, int U) {
  fft::Plan none{};
  switch (n) {
#define OPAW_CASE(nn)                                                                   \
  case nn:                                                                              \
    hipLaunchKernelGGL((k_fft_x<R, nn, S, EPI>), grid, dim3(kThreads), 0, st, in, vin, \
                       out, vout, L, tw, u, vks, nfps, umap, cpre, cnl, U, nn, none);   \
    return hipGetLastError();
    OPAW_FFT_SIZES(OPAW_CASE)
#undef OPAW_CASE
    default: {
      const fft::Plan P = fft::make_plan(n);
      if (P.nr == 0 || n > kMaxRuntimeN) return hipErrorInvalidValue;
      const int lines = x_lines_rt(n);
      const int ls = (n % 16 == 0) ? n + 8 : n;
      const size_t lds = (size_t)(2 * lines * ls + n) * sizeof(cx<R>);
      hipLaunchKernelGGL((k_fft_x<R, 0, S, EPI>), grid, dim3(kThreads), lds, st, in, vin, out, vout, L, tw, u, vks,
                         nfps, umap, cpre, cnl, U, n, P);
      return hipGetLastError();
    }
  }
}

template <typename R, int AXIS, int S, int EPI>
static hipError_t launch_s_n(int n, dim3 grid, hipStream_t st, const cx<R>* in, View vin, cx<R>* out,
                             View vout, SLines L, const cx<R>* tw, const R* tab, View vtab) {
  fft::Plan none{};
  switch (n) {
#define OPAW_CASE(nn)                                                                        \
  case nn:                                                                                   \
    hipLaunchKernelGGL((k_fft_s<R, nn, AXIS, S, EPI>), grid, dim3(kThreads), 0, st, in, vin, \
                       out, vout, L, tw, tab, vtab, nn, sxt_of(nn), none);                   \
    return hipGetLastError();
    OPAW_FFT_SIZES(OPAW_CASE)
#undef OPAW_CASE
    default: {
      const fft::Plan P = fft::make_plan(n);
      if (P.nr == 0 || n > kMaxRuntimeN) return hipErrorInvalidValue;
      const int xt = s_xt_rt(n, (int)sizeof(cx<R>));
      const size_t lds = (size_t)(2 * xt * n + n) * sizeof(cx<R>);
      hipLaunchKernelGGL((k_fft_s<R, 0, AXIS, S, EPI>), grid, dim3(kThreads), lds, st, in, vin, out, vout, L, tw,
                         tab, vtab, n, xt, P);
      return hipGetLastError();
    }
  }
}

template <typename R>
hipError_t fft_x_pass(int cls, int n, int dir, int epi, hipStream_t st, const cx<R>* in, const View& vin,
                      cx<R>* out, const View& vout, const XLines& L, const cx<R>* tw,
                      const cx<R>* u, const R* vks, int nfps, const int* umap = nullptr,
                      const c64* cpre = nullptr, const c64* cnl = nullptr, int U = 0) {
  const long long nlines = (long long)L.ly * L.lz * L.nf;
  if (nlines <= 0) return hipSuccess;
  const int lines = x_lines_rt(n);
  dim3 grid((unsigned)((nlines + lines - 1) / lines));
  hipError_t e;
  prof_begin(st, cls);
  if (epi == EPI_HPSI)
    e = launch_x_n<R, 1, EPI_HPSI>(n, grid, st, in, vin, out, vout, L, tw, u, vks, nfps, umap, cpre, cnl, U);
  else if (dir < 0)
    e = launch_x_n<R, -1, EPI_NONE>(n, grid, st, in, vin, out, vout, L, tw, u, vks, nfps, umap, cpre, cnl, U);
  else
    e = launch_x_n<R, 1, EPI_NONE>(n, grid, st, in, vin, out, vout, L, tw, u, vks, nfps, umap, cpre, cnl, U);
  prof_end(st, cls);
  return e;
}

template <typename R>
hipError_t fft_s_pass(int cls, int n, int axis, int dir, int epi, hipStream_t st, const cx<R>* in,
                      const View& vin, cx<R>* out, const View& vout, const SLines& L,
                      const cx<R>* tw, const R* tab, const View& vtab) {
  const int xt = fft_size_compiled(n) ? sxt_of(n) : s_xt_rt(n, (int)sizeof(cx<R>));
  const long long ntile = (long long)((L.nx + xt - 1) / xt) * L.lo * L.nf;
  if (ntile <= 0) return hipSuccess;
  dim3 grid((unsigned)ntile);
  hipError_t e;
  prof_begin(st, cls);
  if (axis == 1) {
    if (epi == EPI_MULT) e = launch_s_n<R, 1, -1, EPI_MULT>(n, grid, st, in, vin, out, vout, L, tw, tab, vtab);
    else if (dir < 0) e = launch_s_n<R, 1, -1, EPI_NONE>(n, grid, st, in, vin, out, vout, L, tw, tab, vtab);
    else e = launch_s_n<R, 1, 1, EPI_NONE>(n, grid, st, in, vin, out, vout, L, tw, tab, vtab);
  } else {
    if (epi == EPI_MULT) e = launch_s_n<R, 2, -1, EPI_MULT>(n, grid, st, in, vin, out, vout, L, tw, tab, vtab);
    else if (dir < 0) e = launch_s_n<R, 2, -1, EPI_NONE>(n, grid, st, in, vin, out, vout, L, tw, tab, vtab);
    else e = launch_s_n<R, 2, 1, EPI_NONE>(n, grid, st, in, vin, out, vout, L, tw, tab, vtab);
  }
  prof_end(st, cls);
  return e;
}

// ---------------------------------------------------------------------------
// The kinetic z pass of 64-point lines held in registers: out = IFFT_z(ek *
// FFT_z(in)) along z for XT x-columns of one (field, y) per workgroup,
// 64 = 8 x 8.  Thread (l, j) loads z = j + 8m (m < 8) of column l; a radix-8
// DFT over m, the twiddle w64^(j k1) and one LDS exchange give thread
// (l, k1) the values X[k1 + 8 k2] of its column after a second radix-8 DFT;
// it multiplies them by ek, and the mirror sequence (radix-8 inverse over
// k2, twiddle w64^-(a k1), exchange, radix-8 inverse over k1) leaves thread
// (l, a) holding z = a + 8b, stored as it was loaded.  Two LDS exchanges
// per transform in place of the Stockham tile's stage round trips; the LDS
// is column-fastest, so every exchange is conflict-free.  In place is
// allowed (a workgroup reads its columns before it writes them).
template <typename R, int XT>
__global__ __launch_bounds__(XT * 8) void k_zpass64(const cx<R>* in, cx<R>* out, int nx, int ny, long long N,
                                                   const R* __restrict__ ek, const cx<R>* __restrict__ twg) {
  __shared__ cx<R> x[64 * XT];
  const int l = threadIdx.x % XT, j = threadIdx.x / XT;
  const int col = blockIdx.x * XT + l, y = blockIdx.y, f = blockIdx.z;
  const long long sz = (long long)nx * ny;
  const long long base = (long long)f * N + col + (long long)nx * y;
  cx<R> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = in[base + (j + 8 * m) * sz];
  fft::dft<R, 8, -1>(v);                                  // over m: index k1
#pragma unroll
  for (int k = 1; k < 8; ++k) v[k] = cmul(v[k], twg[(j * k) & 63]);
#pragma unroll
  for (int k = 0; k < 8; ++k) x[(j * 8 + k) * XT + l] = v[k];
  __syncthreads();
  const int k1 = j;
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = x[(i * 8 + k1) * XT + l];
  fft::dft<R, 8, -1>(v);                                  // over j: v[k2] = X[k1 + 8 k2]
  const R* e = ek + col + (long long)nx * y;
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) {
    const R t = e[(k1 + 8 * k2) * sz];
    v[k2] = mk(v[k2].x * t, v[k2].y * t);
  }
  fft::dft<R, 8, 1>(v);                                   // over k2: index a
#pragma unroll
  for (int a = 1; a < 8; ++a) {
    const cx<R> w = twg[(a * k1) & 63];
    v[a] = cmul(v[a], mk(w.x, -w.y));
  }
  __syncthreads();                                        // the exchange reads above are done
#pragma unroll
  for (int a = 0; a < 8; ++a) x[(k1 * 8 + a) * XT + l] = v[a];
  __syncthreads();
  const int a = j;
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = x[(i * 8 + a) * XT + l];
  fft::dft<R, 8, 1>(v);                                   // over k1: v[b] = out(z = a + 8 b)
#pragma unroll
  for (int b = 0; b < 8; ++b) out[base + (a + 8 * b) * sz] = v[b];
}

// 16-point DFT in registers with sign S (natural order in and out):
// n = 4 n1 + n2, k = k1 + 4 k2, four DFT4s over n1, the twiddles
// w16^(n2 k1), four DFT4s over n2
template <int S>
__device__ __forceinline__ void dft16(c64 (&v)[16]) {
  constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
  const c64 w[4][4] = {{mk(1.0, 0.0), mk(1.0, 0.0), mk(1.0, 0.0), mk(1.0, 0.0)},
                       {mk(1.0, 0.0), mk(c1, S * s1), mk(h, S * h), mk(s1, S * c1)},
                       {mk(1.0, 0.0), mk(h, S * h), mk(0.0, S * 1.0), mk(-h, S * h)},
                       {mk(1.0, 0.0), mk(s1, S * c1), mk(-h, S * h), mk(-c1, -S * s1)}};
  c64 t[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    c64 a[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    fft::dft<double, 4, S>(a);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) t[n2][k1] = cmul(a[k1], w[n2][k1]);
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    c64 b[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
    fft::dft<double, 4, S>(b);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
  }
}

// FP64 passes of 128-point lines in registers along y (AXIS 1) or z
// (AXIS 2), XT x-columns of one (other coordinate o) per workgroup: 128 =
// 16 x 8.  Thread (l, j) loads z = j + 8m (m < 16; zero outside the input
// window [la, lb)), takes a DFT16 over m, the twiddle w128^(S j k1) and one
// LDS exchange; a DFT8 over j for k1 = j and j + 8 leaves it X[k1 + 16 k2].
//   MODE 0: one transform of sign S; the outputs inside [sa, sb) are stored.
//   MODE 1: S = -1, then X * tab (read through vtab) and the inverse by the
//           mirror sequence (inverse DFT8 over k2, twiddle w128^-(jj k1),
//           exchange, inverse DFT16 over k1), which returns z = j + 8m to
//           thread (l, j); the values inside [sa, sb) are stored.
// In place is allowed: every line is read before the first barrier.
template <int AXIS, int MODE, int S, int XT>
__global__ __launch_bounds__(XT * 8) void k_lines128(const c64* in, c64* out, View v, SLines L,
                                                    const c64* __restrict__ twg, const double* __restrict__ tab,
                                                    View vtab) {
  __shared__ c64 x[128 * XT];
  const int l = threadIdx.x % XT, j = threadIdx.x / XT;
  const int col = blockIdx.x * XT + l;
  const int o = L.o0 + blockIdx.y, f = blockIdx.z;
  const bool on = col < L.nx;
  auto addr = [&](int z) -> long long { return AXIS == 1 ? vaddr(v, f, col, z, o) : vaddr(v, f, col, o, z); };
  c64 u[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int z = j + 8 * m;
    u[m] = (on && z >= L.la && z < L.lb) ? in[addr(z)] : mk(0.0, 0.0);
  }
  dft16<S>(u);                                            // over m: index k1
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) {
    const c64 w = twg[(j * k1) & 127];
    u[k1] = cmul(u[k1], mk(w.x, S * -w.y));               // w128^(S j k1); twg = e^(-2 pi i m/128)
  }
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) x[(j * 16 + k1) * XT + l] = u[k1];
  __syncthreads();
  c64 r[2][8];                                            // r[h][k2] = X[k1 + 16 k2], k1 = j + 8 h
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    const int k1 = j + 8 * hh;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[hh][i] = x[(i * 16 + k1) * XT + l];
    fft::dft<double, 8, S>(r[hh]);
  }
  if constexpr (MODE == 0) {
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        const int k = j + 8 * hh + 16 * k2;
        if (on && k >= L.sa && k < L.sb) out[addr(k)] = r[hh][k2];
      }
  } else {
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int k1 = j + 8 * hh;
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        const int k = k1 + 16 * k2;
        const double t = on ? tab[AXIS == 1 ? taddr(vtab, col, k, o) : taddr(vtab, col, o, k)] : 0.0;
        r[hh][k2] = mk(r[hh][k2].x * t, r[hh][k2].y * t);
      }
      fft::dft<double, 8, -S>(r[hh]);                     // over k2: index jj
#pragma unroll
      for (int jj = 1; jj < 8; ++jj) {
        const c64 w = twg[(jj * k1) & 127];
        r[hh][jj] = cmul(r[hh][jj], mk(w.x, S * w.y));    // w128^(-S jj k1)
      }
    }
    __syncthreads();                                      // the exchange reads above are done
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int k1 = j + 8 * hh;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) x[(k1 * 8 + jj) * XT + l] = r[hh][jj];
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) u[k1] = x[(k1 * 8 + j) * XT + l];
    dft16<-S>(u);                                         // over k1: u[m] = out(z = j + 8 m)
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int z = j + 8 * m;
      if (on && z >= L.sa && z < L.sb) out[addr(z)] = u[m];
    }
  }
}

// a 128-point line pass of k_lines128 over the line set L (nf fields)
template <int AXIS, int MODE, int S>
static hipError_t lines128(int cls, hipStream_t st, const c64* in, c64* out, View v, SLines L, const c64* tw,
                           const double* tab, View vtab) {
  constexpr int XT = 16;
  const dim3 grid((unsigned)((L.nx + XT - 1) / XT), (unsigned)L.lo, (unsigned)L.nf);
  prof_begin(st, cls);
  hipLaunchKernelGGL((k_lines128<AXIS, MODE, S, XT>), grid, dim3(XT * 8), 0, st, in, out, v, L, tw, tab, vtab);
  prof_end(st, cls);
  return hipGetLastError();
}

// Roots of unity exp(2 pi i m / N), m < N, evaluated at compile time (Taylor
// series after reduction to [-pi, pi]; 1e-16 accurate), for the composite
// register DFTs below.
constexpr double ct_sin_cos(double x, bool want_sin) {
  const double pi = 3.14159265358979323846;
  while (x > pi) x -= 2.0 * pi;
  while (x < -pi) x += 2.0 * pi;
  double sum = want_sin ? x : 1.0, term = sum;
  const double x2 = x * x;
  for (int k = 1; k < 26; ++k) {
    const double a = want_sin ? (2.0 * k) * (2.0 * k + 1.0) : (2.0 * k - 1.0) * (2.0 * k);
    term = -term * x2 / a;
    sum += term;
  }
  return sum;
}
template <int N>
struct Roots {
  double c[N], s[N];
  constexpr Roots() : c(), s() {
    for (int m = 0; m < N; ++m) {
      const double x = 2.0 * 3.14159265358979323846 * m / N;
      c[m] = ct_sin_cos(x, false);
      s[m] = ct_sin_cos(x, true);
    }
  }
};

// A*B-point DFT in registers with sign S (natural order in and out):
// n = B n1 + n2, k = k1 + A k2: B DFT_A over n1, the twiddles
// w_AB^(S n2 k1), A DFT_B over n2
template <typename R, int A, int B, int S>
__device__ __forceinline__ void dft_ab(cx<R>* v) {
  constexpr Roots<A * B> W{};
  cx<R> t[B][A];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
    cx<R> a[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) a[n1] = v[B * n1 + n2];
    fft::dft<R, A, S>(a);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      const int m = (n2 * k1) % (A * B);
      t[n2][k1] = m == 0 ? a[k1] : cmul(a[k1], mk((R)W.c[m], (R)(S * W.s[m])));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cx<R> b[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) b[n2] = t[n2][k1];
    fft::dft<R, B, S>(b);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = b[k2];
  }
}
// the P-point register DFT of the generic z pass (P = N / 8)
template <typename R, int P, int S>
__device__ __forceinline__ void dft_p(cx<R>* v) {
  if constexpr (P == 6) dft_ab<R, 3, 2, S>(v);
  else dft_ab<R, 4, 3, S>(v);                              // P == 12
}

// The kinetic z pass of N-point lines in registers for N = 8 P (P = 6, 12:
// nz = 48, 96), the scheme of k_zpass64 with the P-point DFT composite:
// thread (l, j) loads z = j + 8m (m < P), DFT_P over m, twiddle w_N^(j k1),
// exchange, DFT8 over j for k1 = j, j + 8 (< P), ek, and the mirror sequence.
template <typename R, int N, int XT>
__global__ __launch_bounds__(XT * 8) void k_zpass_gen(const cx<R>* in, cx<R>* out, int nx, int ny, long long NF,
                                                      const R* __restrict__ ek, const cx<R>* __restrict__ twg) {
  constexpr int P = N / 8, H = (P + 7) / 8;
  __shared__ cx<R> x[N * XT];
  const int l = threadIdx.x % XT, j = threadIdx.x / XT;
  const int col = blockIdx.x * XT + l, y = blockIdx.y, f = blockIdx.z;
  const long long sz = (long long)nx * ny;
  const long long base = (long long)f * NF + col + (long long)nx * y;
  cx<R> u[P];
#pragma unroll
  for (int m = 0; m < P; ++m) u[m] = in[base + (j + 8 * m) * sz];
  dft_p<R, P, -1>(u);                                      // over m: index k1
#pragma unroll
  for (int k = 1; k < P; ++k) u[k] = cmul(u[k], twg[(j * k) % N]);
#pragma unroll
  for (int k = 0; k < P; ++k) x[(j * P + k) * XT + l] = u[k];
  __syncthreads();
  cx<R> r[H][8];
  const R* e = ek + col + (long long)nx * y;
#pragma unroll
  for (int hh = 0; hh < H; ++hh) {
    const int k1 = j + 8 * hh;
    if (k1 < P) {
#pragma unroll
      for (int i = 0; i < 8; ++i) r[hh][i] = x[(i * P + k1) * XT + l];
      fft::dft<R, 8, -1>(r[hh]);                           // over j: r[k2] = X[k1 + P k2]
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        const R t = e[(k1 + P * k2) * sz];
        r[hh][k2] = mk(r[hh][k2].x * t, r[hh][k2].y * t);
      }
      fft::dft<R, 8, 1>(r[hh]);                            // over k2: index jj
#pragma unroll
      for (int jj = 1; jj < 8; ++jj) r[hh][jj] = cmulc(r[hh][jj], twg[(jj * k1) % N]);
    }
  }
  __syncthreads();                                         // the exchange reads above are done
#pragma unroll
  for (int hh = 0; hh < H; ++hh) {
    const int k1 = j + 8 * hh;
    if (k1 < P) {
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) x[(k1 * 8 + jj) * XT + l] = r[hh][jj];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) u[k1] = x[(k1 * 8 + j) * XT + l];
  dft_p<R, P, 1>(u);                                       // over k1: u[m] = out(z = j + 8 m)
#pragma unroll
  for (int m = 0; m < P; ++m) out[base + (j + 8 * m) * sz] = u[m];
}

// columns per workgroup of k_zpass_gen (0: not covered)
static int zpass_gen_xt(const GridGeom& g) {
  if (g.nz != 48 && g.nz != 96) return 0;
  if (g.nx % 32 == 0) return 32;
  if (g.nx % 16 == 0) return 16;
  return 0;
}

template <typename R>
static hipError_t zpass_gen(const GridGeom& g, hipStream_t st, const cx<R>* in, cx<R>* out, int nf, const R* ek,
                            const cx<R>* twz) {
  const int xt = zpass_gen_xt(g);
  const long long NF = (long long)g.nx * g.ny * g.nz;
  const dim3 grid((unsigned)(g.nx / xt), (unsigned)g.ny, (unsigned)nf);
  prof_begin(st, KC_FFT_Z);
  if (g.nz == 48) {
    if (xt == 32)
      hipLaunchKernelGGL((k_zpass_gen<R, 48, 32>), grid, dim3(256), 0, st, in, out, g.nx, g.ny, NF, ek, twz);
    else
      hipLaunchKernelGGL((k_zpass_gen<R, 48, 16>), grid, dim3(128), 0, st, in, out, g.nx, g.ny, NF, ek, twz);
  } else {
    if (xt == 32)
      hipLaunchKernelGGL((k_zpass_gen<R, 96, 32>), grid, dim3(256), 0, st, in, out, g.nx, g.ny, NF, ek, twz);
    else
      hipLaunchKernelGGL((k_zpass_gen<R, 96, 16>), grid, dim3(128), 0, st, in, out, g.nx, g.ny, NF, ek, twz);
  }
  prof_end(st, KC_FFT_Z);
  return hipGetLastError();
}

// columns per workgroup of k_zpass64 for an nx-wide grid (0: not covered)
static int zpass64_xt(const GridGeom& g) {
  if (g.nz != 64) return 0;
  if (g.nx % 32 == 0) return 32;
  if (g.nx % 16 == 0) return 16;
  return 0;
}

template <typename R>
static hipError_t zpass64(const GridGeom& g, hipStream_t st, const cx<R>* in, cx<R>* out, int nf, const R* ek,
                          const cx<R>* twz) {
  const int xt = zpass64_xt(g);
  const long long N = (long long)g.nx * g.ny * g.nz;
  const dim3 grid((unsigned)(g.nx / xt), (unsigned)g.ny, (unsigned)nf);
  prof_begin(st, KC_FFT_Z);
  if (xt == 32)
    hipLaunchKernelGGL((k_zpass64<R, 32>), grid, dim3(256), 0, st, in, out, g.nx, g.ny, N, ek, twz);
  else
    hipLaunchKernelGGL((k_zpass64<R, 16>), grid, dim3(128), 0, st, in, out, g.nx, g.ny, N, ek, twz);
  prof_end(st, KC_FFT_Z);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Kinetic + local potential (+ gathered projector terms):
//   out_f = IFFT(ek * FFT(u_f)) + vks(sys(f)) * u_f [+ cnl_f],
//   u_f = in_f [+ cpre_f]   (compact corrections on the union of boxes)
// (src/hphi.f90:14-21 and the add-back of :42-60).  tmp must hold nf fields.
template <typename R>
hipError_t apply_t_plus_v(const GridGeom& g, hipStream_t st, const cx<R>* in, cx<R>* out, cx<R>* tmp,
                          int nf, int nf_per_sys, const R* ek, const R* vks,
                          const cx<R>* twx, const cx<R>* twy, const cx<R>* twz,
                          const int* umap, const int* upt, const int* uzr, const c64* cpre, const c64* cnl,
                          int U, hipEvent_t before_inv, const PlaneRk4<R>* rk, hipStream_t st_inv,
                          hipEvent_t ev_mid) {
  const long long N = (long long)g.nx * g.ny * g.nz;
  View v{N, 0, 0, 0, g.nx, g.ny};
  XLines Lx{0, g.ny, 0, g.nz, nf, 0, g.nx, 0, g.nx};
  SLines Ly{g.nx, 0, g.nz, nf, 0, g.ny, 0, g.ny};
  SLines Lz{g.nx, 0, g.ny, nf, 0, g.nz, 0, g.nz};
  View vt{0, 0, 0, 0, g.nx, g.ny};
  hipError_t e;
  // the last pass runs on si: st, or st_inv once the passes before it (on st)
  // have ended (ev_mid).  The split is keyed on the event: a null st_inv is
  // the legacy default stream, a valid target.
  hipStream_t si = st;
  auto hand_over = [&]() -> hipError_t {
    if (!ev_mid) return hipSuccess;
    hipError_t r = hipEventRecord(ev_mid, st);
    if (r == hipSuccess) r = hipStreamWaitEvent(st_inv, ev_mid, 0);
    si = st_inv;
    return r;
  };
  if (plane_supported<R>(g.nx, g.ny)) {
    prof_begin(st, KC_FFT_X);
    e = launch_plane<R, -1, EPI_NONE>(g.nx, g.ny, st, in, tmp, g.nz, nf, N, twx, twy, nullptr, nullptr, nf_per_sys,
                                      upt, uzr, cpre, nullptr, U);
    prof_end(st, KC_FFT_X);
    if (e) return e;
    if (zpass64_xt(g) > 0) {
      if ((e = zpass64<R>(g, st, tmp, tmp, nf, ek, twz))) return e;
    } else if (zpass_gen_xt(g) > 0) {
      if ((e = zpass_gen<R>(g, st, tmp, tmp, nf, ek, twz))) return e;
    } else if ((e = fft_s_pass<R>(KC_FFT_Z, g.nz, 2, -1, EPI_MULT, st, tmp, v, tmp, v, Lz, twz, ek, vt))) {
      return e;
    }
    if ((e = hand_over())) return e;
    if (before_inv && (e = hipStreamWaitEvent(si, before_inv, 0))) return e;
    prof_begin(si, KC_FFT_XH);
    if (rk == nullptr)
      e = launch_plane<R, 1, EPI_HPSI>(g.nx, g.ny, si, tmp, out, g.nz, nf, N, twx, twy, in, vks, nf_per_sys, upt,
                                       uzr, cpre, cnl, U);
    else if (rk->first)
      e = launch_plane<R, 1, EPI_RK4F>(g.nx, g.ny, si, tmp, out, g.nz, nf, N, twx, twy, in, vks, nf_per_sys, upt,
                                       uzr, cpre, cnl, U, *rk);
    else
      e = launch_plane<R, 1, EPI_RK4>(g.nx, g.ny, si, tmp, out, g.nz, nf, N, twx, twy, in, vks, nf_per_sys, upt,
                                      uzr, cpre, cnl, U, *rk);
    prof_end(si, KC_FFT_XH);
    return e;
  }
  if (rk) return hipErrorInvalidValue;   // the RK4 epilogue exists on the plane path only
  if ((e = fft_x_pass<R>(KC_FFT_X, g.nx, -1, EPI_NONE, st, in, v, tmp, v, Lx, twx, nullptr, nullptr, 1, umap, cpre,
                         nullptr, U)))
    return e;
  if ((e = fft_s_pass<R>(KC_FFT_Y, g.ny, 1, -1, EPI_NONE, st, tmp, v, tmp, v, Ly, twy, nullptr, vt))) return e;
  if ((e = fft_s_pass<R>(KC_FFT_Z, g.nz, 2, -1, EPI_MULT, st, tmp, v, tmp, v, Lz, twz, ek, vt))) return e;
  if ((e = fft_s_pass<R>(KC_FFT_Y, g.ny, 1, 1, EPI_NONE, st, tmp, v, tmp, v, Ly, twy, nullptr, vt))) return e;
  if ((e = hand_over())) return e;
  if (before_inv && (e = hipStreamWaitEvent(si, before_inv, 0))) return e;
  return fft_x_pass<R>(KC_FFT_XH, g.nx, 1, EPI_HPSI, si, tmp, v, out, v, Lx, twx, in, vks, nf_per_sys, umap, cpre, cnl,
                       U);
}

// Hartree convolution on the (scale*nx, scale*ny, scale*nz) box:
//   out = crop( IFFT( vk * FFT( embed(rho) ) ) )
// for one complex field rho (two real densities packed as re + i im; the
// kernel vk is real and even so the two results do not mix).  work must hold
// the doubled box.  Zero lines of the padded box are never touched.
// the Hartree kernel vk is stored as its octant (NX/2+1, NY/2+1, NZ/2+1)
static View vk_octant_view(int NX, int NY, int NZ) {
  View v{0, 0, 0, 0, NX / 2 + 1, NY / 2 + 1};
  v.fx = NX;
  v.fy = NY;
  v.fz = NZ;
  return v;
}

hipError_t hartree_conv(const GridGeom& g, int scale, hipStream_t st, const c64* rho, c64* out,
                        c64* work, const double* vk, const c64* twx, const c64* twy,
                        const c64* twz) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int NX = scale * nx, NY = scale * ny, NZ = scale * nz;
  const int ox = NX / 2 - nx / 2, oy = NY / 2 - ny / 2, oz = NZ / 2 - nz / 2;
  View small{0, -ox, -oy, -oz, nx, ny};      // small-grid array addressed in big coordinates
  View big{0, 0, 0, 0, NX, NY};
  hipError_t e;
  // x forward: lines (y,z) inside the density support, x loaded from the small grid
  XLines Lx1{oy, ny, oz, nz, 1, ox, ox + nx, 0, NX};
  if ((e = fft_x_pass<double>(KC_HARTREE, NX, -1, EPI_NONE, st, rho, small, work, big, Lx1, twx, nullptr, nullptr, 1))) return e;
  // y forward: z inside the support, y loaded only where nonzero
  SLines Ly1{NX, oz, nz, 1, oy, oy + ny, 0, NY};
  if ((e = fft_s_pass<double>(KC_HARTREE, NY, 1, -1, EPI_NONE, st, work, big, work, big, Ly1, twy, nullptr, big))) return e;
  // z: forward, * vk, inverse; only the cropped z range is stored
  SLines Lz{NX, 0, NY, 1, oz, oz + nz, oz, oz + nz};
  const View voct = vk_octant_view(NX, NY, NZ);
  if ((e = fft_s_pass<double>(KC_HARTREE, NZ, 2, -1, EPI_MULT, st, work, big, work, big, Lz, twz, vk, voct))) return e;
  // y inverse: z in crop, store cropped y
  SLines Ly2{NX, oz, nz, 1, 0, NY, oy, oy + ny};
  if ((e = fft_s_pass<double>(KC_HARTREE, NY, 1, 1, EPI_NONE, st, work, big, work, big, Ly2, twy, nullptr, big))) return e;
  // x inverse: (y,z) in crop, store cropped x into the small grid
  XLines Lx2{oy, ny, oz, nz, 1, 0, NX, ox, ox + nx};
  return fft_x_pass<double>(KC_HARTREE, NX, 1, EPI_NONE, st, work, big, out, small, Lx2, twx, nullptr, nullptr, 1);
}

// ---------------------------------------------------------------------------
// Fused first pass of the rebuild's Hartree solve on the doubled box
// (NX = 2 nx, NY = 2 ny, NZ = 2 nz, the density embedded at offsets
// o = n / 2 along each axis).
//
// k_hart_fwd, one workgroup per (small z plane, y parity p): the
// renormalisation and pair packing of the plane (dens + compensation charge,
// times the factor to rnel), its embedding in ny rows of length NX, the x
// transform, and the y transform with the zero padding used exactly: with
// o = ny / 2,
//   Y[2j + p] = (-1)^j DFT_ny( x[m] w^(p (m + o)) )[j],  w = exp(-2 pi i / NY),
// so each parity is one ny-point transform of the rows held in LDS, written
// to the rows 2j + p of the doubled-box plane oz + z.  The z pass (forward,
// * vk, inverse) and the cropped y and x inverse passes are the plain line
// passes; V_KS follows as a separate kernel (an inverse fused with V_KS in
// one workgroup per plane measured slower than the two line passes: one
// LDS-filling workgroup per CU keeps too few loads in flight).
constexpr int kHT = 512;

template <int NX, int NY>
constexpr bool hart_fits() {
  // ny rows of NX + 8 plus the three twiddle tables
  return (long)(NY / 2) * (NX + 8) * 16 + (NX + NY + NY / 2) * 16 <= 163840;
}

template <int NX, int NY>
__global__ __launch_bounds__(kHT) void k_hart_fwd(HartreeRebuild H, int q, int nsys, int nz,
                                                   c64* __restrict__ work, const c64* __restrict__ twxg,
                                                   const c64* __restrict__ twyg) {
  constexpr int nx = NX / 2, ny = NY / 2, ox = nx / 2, oy = ny / 2;
  constexpr int LS = NX + 8;
  __shared__ c64 tile[ny * LS];
  __shared__ c64 twx[NX];
  __shared__ c64 twy[NY];
  __shared__ c64 twh[ny];
  __shared__ double red[kHT / 64][4];
  __shared__ double facs[2];
  const int z = blockIdx.x, p = blockIdx.y;
  const long long N = (long long)nx * ny * nz;
  fft::load_twiddles(twx, twxg, NX);
  fft::load_twiddles(twy, twyg, NY);
  for (int m = threadIdx.x; m < ny; m += kHT) twh[m] = twyg[2 * m];   // exp(-2 pi i m / ny)
  // renormalisation factors: every workgroup sums the partials in the same
  // order (lanes, then waves), so all hold the same factors
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < 2; ++k) {
    const int s = 2 * q + k;
    if (s >= nsys || !H.renorm) break;
    for (int c = threadIdx.x; c < H.nd; c += kHT) v[2 * k] += H.dpart[(long long)s * H.nd + c];
    for (int c = threadIdx.x; c < H.nn; c += kHT) v[2 * k + 1] += H.npart[(long long)s * H.nn + c];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int k = threadIdx.x, s = 2 * q + k;
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kHT / 64; ++w) {
      a += red[w][2 * k];
      b += red[w][2 * k + 1];
    }
    const double f = (H.renorm && s < nsys) ? H.rnel / (b * H.dv + a * H.dv) : 1.0;
    facs[k] = f;
    if (z == 0 && p == 0 && s < nsys) H.fac[s] = f;
  }
  // the embedded rows: zero outside the x window
  for (int e = threadIdx.x; e < ny * (NX - nx); e += kHT) {
    const int y = e / (NX - nx), c = e % (NX - nx);
    tile[y * LS + (c < ox ? c : c + nx)] = mk(0.0, 0.0);
  }
  // the plane's points, loads batched (sparse index first, then the values)
  constexpr int NP = nx * ny;
  constexpr int CNT = (NP + kHT - 1) / kHT;
  auto pt = [&](int u) { return min((int)threadIdx.x + u * kHT, NP - 1); };
  const long long r0 = (long long)z * NP;
  int kk[CNT];
#pragma unroll
  for (int u = 0; u < CNT; ++u) kk[u] = H.smap[r0 + pt(u)];
  double dd[2][CNT], hh[2][CNT];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int s = min(2 * q + t, nsys - 1);
#pragma unroll
    for (int u = 0; u < CNT; ++u) {
      dd[t][u] = H.dens[(long long)s * N + r0 + pt(u)];
      hh[t][u] = H.nhs[(long long)s * H.S + max(kk[u], 0)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < CNT; ++u) {
    const int e = (int)threadIdx.x + u * kHT;
    if (NP % kHT == 0 || e < NP) {
      double pr[2] = {0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 2; ++t)
        if (2 * q + t < nsys) pr[t] = dd[t][u] * facs[t] + (kk[u] >= 0 ? hh[t][u] : 0.0) * facs[t];
      tile[(e / nx) * LS + ox + e % nx] = mk(pr[0], pr[1]);
    }
  }
  __syncthreads();
  fft::lines<double, NX, ny, kHT, false, LS, 1, -1>(tile, twx);
  if (p) {
    for (int e = threadIdx.x; e < ny * NX; e += kHT) {
      const int y = e / NX, x = e % NX;
      tile[y * LS + x] = cmul(tile[y * LS + x], twy[y + oy]);
    }
    __syncthreads();
  }
  fft::lines<double, ny, NX, kHT, true, 1, LS, -1>(tile, twh);
  const int oz = nz / 2;
  c64* dst = work + (long long)(oz + z) * NX * NY;
  for (int e = threadIdx.x; e < ny * NX; e += kHT) {
    const int j = e / NX, x = e % NX;
    c64 t = tile[j * LS + x];
    if (j & 1) t = mk(-t.x, -t.y);
    dst[(long long)(2 * j + p) * NX + x] = t;
  }
}

// small-grid (nx, ny) planes with a compiled fused path
#define OPAW_HART_PAIRS(X) X(16, 16) X(32, 32) X(48, 48) X(64, 64) X(64, 48) X(48, 64)

bool hartree_fused_supported(const GridGeom& g, int scale) {
  if (scale != 2 || g.nx % 2 || g.ny % 2 || g.nz % 2) return false;
#define OPAW_CASE(a, b) \
  if (g.nx == a && g.ny == b) return hart_fits<2 * a, 2 * b>();
  OPAW_HART_PAIRS(OPAW_CASE)
#undef OPAW_CASE
  return false;
}

hipError_t hartree_fused(const GridGeom& g, hipStream_t st, const HartreeRebuild& H, int q, int nsys, c64* work,
                         c64* vh, const double* vk, const c64* twx, const c64* twy, const c64* twz) {
  if (!hartree_fused_supported(g, 2)) return hipErrorNotSupported;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int NX = 2 * nx, NY = 2 * ny, NZ = 2 * nz, oz = nz / 2;
  prof_begin(st, KC_HARTREE);
#define OPAW_CASE(a, b)                                                                                    \
  if (nx == a && ny == b)                                                                                  \
    hipLaunchKernelGGL((k_hart_fwd<2 * a, 2 * b>), dim3(nz, 2), dim3(kHT), 0, st, H, q, nsys, nz, work, twx, \
                       twy);
  OPAW_HART_PAIRS(OPAW_CASE)
#undef OPAW_CASE
  prof_end(st, KC_HARTREE);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  View small{0, -nx / 2, -ny / 2, -oz, nx, ny};
  View big{0, 0, 0, 0, NX, NY};
  SLines Lz{NX, 0, NY, 1, oz, oz + nz, oz, oz + nz};
  const View voct = vk_octant_view(NX, NY, NZ);
  SLines Ly2{NX, oz, nz, 1, 0, NY, ny / 2, ny / 2 + ny};
  if (NZ == 128 && NY == 128) {
    if ((e = lines128<2, 1, -1>(KC_HARTREE, st, work, work, big, Lz, twz, vk, voct))) return e;
    if ((e = lines128<1, 0, 1>(KC_HARTREE, st, work, work, big, Ly2, twy, nullptr, big))) return e;
  } else {
    if ((e = fft_s_pass<double>(KC_HARTREE, NZ, 2, -1, EPI_MULT, st, work, big, work, big, Lz, twz, vk, voct))) return e;
    if ((e = fft_s_pass<double>(KC_HARTREE, NY, 1, 1, EPI_NONE, st, work, big, work, big, Ly2, twy, nullptr, big)))
      return e;
  }
  XLines Lx2{ny / 2, ny, oz, nz, 1, 0, NX, nx / 2, nx / 2 + nx};
  return fft_x_pass<double>(KC_HARTREE, NX, 1, EPI_NONE, st, work, big, vh, small, Lx2, twx, nullptr, nullptr, 1);
}

template hipError_t apply_t_plus_v<double>(const GridGeom&, hipStream_t, const c64*, c64*, c64*, int, int,
                                           const double*, const double*, const c64*, const c64*, const c64*,
                                           const int*, const int*, const int*, const c64*, const c64*, int,
                                           hipEvent_t, const PlaneRk4<double>*, hipStream_t, hipEvent_t);
template hipError_t apply_t_plus_v<float>(const GridGeom&, hipStream_t, const c32*, c32*, c32*, int, int,
                                          const float*, const float*, const c32*, const c32*, const c32*,
                                          const int*, const int*, const int*, const c64*, const c64*, int,
                                          hipEvent_t, const PlaneRk4<float>*, hipStream_t, hipEvent_t);

// plain 3D transform of nf fields (used by tests and the Hartree-free paths)
template <typename R>
hipError_t fft3d(const GridGeom& g, hipStream_t st, int dir, const cx<R>* in, cx<R>* out, int nf,
                 const cx<R>* twx, const cx<R>* twy, const cx<R>* twz) {
  const long long N = (long long)g.nx * g.ny * g.nz;
  View v{N, 0, 0, 0, g.nx, g.ny};
  View vt{0, 0, 0, 0, g.nx, g.ny};
  XLines Lx{0, g.ny, 0, g.nz, nf, 0, g.nx, 0, g.nx};
  SLines Ly{g.nx, 0, g.nz, nf, 0, g.ny, 0, g.ny};
  SLines Lz{g.nx, 0, g.ny, nf, 0, g.nz, 0, g.nz};
  hipError_t e;
  if ((e = fft_x_pass<R>(KC_OTHER, g.nx, dir, EPI_NONE, st, in, v, out, v, Lx, twx, nullptr, nullptr, 1))) return e;
  if ((e = fft_s_pass<R>(KC_OTHER, g.ny, 1, dir, EPI_NONE, st, out, v, out, v, Ly, twy, nullptr, vt))) return e;
  return fft_s_pass<R>(KC_OTHER, g.nz, 2, dir, EPI_NONE, st, out, v, out, v, Lz, twz, nullptr, vt);
}
template hipError_t fft3d<double>(const GridGeom&, hipStream_t, int, const c64*, c64*, int, const c64*,
                                  const c64*, const c64*);
template hipError_t fft3d<float>(const GridGeom&, hipStream_t, int, const c32*, c32*, int, const c32*,
                                 const c32*, const c32*);

}  // namespace opaw
