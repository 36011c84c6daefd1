// Projector-box kernels of the OPAW hot path (gfx950).
//
//  box_project: ca(f, coff_a + i) = dv * sum_{box_a} u_f(r) p_{a,i}(r)
//               (src/proj.f90:1-44 and src/proj1.f90:1-46; projectors are real
//               at the gamma point so conj(p) = p).  One workgroup per
//               (atom, field); FP64 accumulation; wave64 shuffles then LDS.
//  box_add:     u_f(r) += sum_i p_{a,i}(r) w_{a,i} for every atom a whose box
//               holds r, with w = sn(n) * ca (src/snphi.f90:15-48) or
//               w = D ca (addvnl, src/hphi.f90:24-65).  One workgroup per
//               (z-plane, field) walks the atoms whose boxes cut that plane in
//               atom order, with a barrier between atoms, so overlapping boxes
//               (C-C neighbours) add deterministically without atomics.
// Both follow the periodic wrap mod(ix+nx-1,nx)+1 of the reference edge path.
#include "opaw_kernels.h"

namespace opaw {

constexpr int kBoxThreads = 256;

__device__ __forceinline__ int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <typename R>
__global__ __launch_bounds__(kBoxThreads) void k_box_project(const AtomDesc* __restrict__ atoms,
                                                             const double* __restrict__ tabs,
                                                             const cx<R>* __restrict__ u, long long N,
                                                             int nx, int ny, int nz, int ctot,
                                                             double dv, c64* __restrict__ ca) {
  const int a = blockIdx.x;
  const int f = blockIdx.y;
  const AtomDesc A = atoms[a];
  const double* tab = tabs + A.tab;
  const cx<R>* uf = u + (long long)f * N;
  double accr[kMaxMs], acci[kMaxMs];
#pragma unroll
  for (int i = 0; i < kMaxMs; ++i) accr[i] = acci[i] = 0.0;
  const int n1 = A.nr[0], n2 = A.nr[1];
  for (int p = threadIdx.x; p < A.npt; p += kBoxThreads) {
    const int lx = p % n1;
    const int r = p / n1;
    const int ly = r % n2;
    const int lz = r / n2;
    int gx = A.start[0] + lx, gy = A.start[1] + ly, gz = A.start[2] + lz;
    if (A.wrap) {
      gx = wrapi(gx, nx);
      gy = wrapi(gy, ny);
      gz = wrapi(gz, nz);
    }
    const cx<R> v = uf[(long long)gx + (long long)nx * ((long long)gy + (long long)ny * gz)];
    const double vr = (double)v.x, vi = (double)v.y;
#pragma unroll
    for (int i = 0; i < kMaxMs; ++i) {
      if (i < A.ms) {
        const double t = tab[(long long)i * A.npt + p];
        accr[i] += vr * t;
        acci[i] += vi * t;
      }
    }
  }
  __shared__ double red[kBoxThreads / 64][kMaxMs][2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxMs; ++i) {
    if (i < A.ms) {
      const double sr = wave_sum(accr[i]);
      const double si = wave_sum(acci[i]);
      if (lane == 0) {
        red[wave][i][0] = sr;
        red[wave][i][1] = si;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < A.ms) {
    const int i = threadIdx.x;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int w = 0; w < kBoxThreads / 64; ++w) {
      sr += red[w][i][0];
      si += red[w][i][1];
    }
    ca[(long long)f * ctot + A.coff + i] = mk(sr * dv, si * dv);
  }
}

template <typename R>
__global__ __launch_bounds__(kBoxThreads) void k_box_add(const AtomDesc* __restrict__ atoms,
                                                         const double* __restrict__ tabs,
                                                         const int* __restrict__ plane_ptr,
                                                         const PlaneEntry* __restrict__ plane,
                                                         cx<R>* __restrict__ out, long long N, int nx,
                                                         int ny, int nz, int ctot,
                                                         const c64* __restrict__ ca, int mode,
                                                         const double* __restrict__ sn,
                                                         const double* __restrict__ dij, int nf_per_sys,
                                                         int mtot) {
  const int z = blockIdx.x;
  const int f = blockIdx.y;
  cx<R>* of = out + (long long)f * N;
  __shared__ double wr[kMaxMs], wi[kMaxMs];
  const int e0 = plane_ptr[z], e1 = plane_ptr[z + 1];
  for (int e = e0; e < e1; ++e) {
    const PlaneEntry pe = plane[e];
    const AtomDesc A = atoms[pe.atom];
    const c64* c = ca + (long long)f * ctot + A.coff;
    if (threadIdx.x < A.ms) {
      const int i = threadIdx.x;
      double xr, xi;
      if (mode == 0) {
        const double s = sn[A.coff + i];
        xr = s * c[i].x;
        xi = s * c[i].y;
      } else {
        const int sys = f / nf_per_sys;
        const double* D = dij + (long long)sys * mtot + A.moff;
        xr = 0.0;
        xi = 0.0;
        for (int j = 0; j < A.ms; ++j) {
          const double d = D[i + A.ms * j];
          xr += d * c[j].x;
          xi += d * c[j].y;
        }
      }
      wr[i] = xr;
      wi[i] = xi;
    }
    __syncthreads();
    const int n1 = A.nr[0], n2 = A.nr[1];
    const int gz = A.wrap ? wrapi(A.start[2] + pe.lz, nz) : A.start[2] + pe.lz;
    const double* tab = tabs + A.tab + (long long)pe.lz * n1 * n2;
    for (int q = threadIdx.x; q < n1 * n2; q += kBoxThreads) {
      const int lx = q % n1, ly = q / n1;
      int gx = A.start[0] + lx, gy = A.start[1] + ly;
      if (A.wrap) {
        gx = wrapi(gx, nx);
        gy = wrapi(gy, ny);
      }
      double vr = 0.0, vi = 0.0;
      for (int i = 0; i < A.ms; ++i) {
        const double t = tab[(long long)i * A.npt + q];
        vr += t * wr[i];
        vi += t * wi[i];
      }
      const long long g = (long long)gx + (long long)nx * ((long long)gy + (long long)ny * gz);
      cx<R> o = of[g];
      o.x = (R)((double)o.x + vr);
      o.y = (R)((double)o.y + vi);
      of[g] = o;
    }
    __syncthreads();
  }
}

template <typename R>
hipError_t box_project(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf,
                       int which, double dv, c64* ca) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  const long long N = (long long)g.nx * g.ny * g.nz;
  prof_begin(st, KC_BOX_PROJ);
  hipLaunchKernelGGL((k_box_project<R>), dim3(bt.natom, nf), dim3(kBoxThreads), 0, st, bt.atoms,
                     which ? bt.p1 : bt.p, u, N, g.nx, g.ny, g.nz, bt.ctot, dv, ca);
  prof_end(st, KC_BOX_PROJ);
  return hipGetLastError();
}

template <typename R>
hipError_t box_add(const GridGeom& g, const BoxTables& bt, hipStream_t st, cx<R>* out, int nf,
                   int which, const c64* ca, int mode, double n, const double* sn, const double* dij,
                   int nf_per_sys, int mtot) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  const long long N = (long long)g.nx * g.ny * g.nz;
  (void)n;
  prof_begin(st, KC_BOX_ADD);
  hipLaunchKernelGGL((k_box_add<R>), dim3(g.nz, nf), dim3(kBoxThreads), 0, st, bt.atoms,
                     which ? bt.p1 : bt.p, bt.plane_ptr, bt.plane, out, N, g.nx, g.ny, g.nz, bt.ctot, ca,
                     mode, sn, dij, nf_per_sys, mtot);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

template hipError_t box_project<double>(const GridGeom&, const BoxTables&, hipStream_t, const c64*, int, int,
                                        double, c64*);
template hipError_t box_project<float>(const GridGeom&, const BoxTables&, hipStream_t, const c32*, int, int,
                                       double, c64*);
template hipError_t box_add<double>(const GridGeom&, const BoxTables&, hipStream_t, c64*, int, int,
                                    const c64*, int, double, const double*, const double*, int, int);
template hipError_t box_add<float>(const GridGeom&, const BoxTables&, hipStream_t, c32*, int, int,
                                   const c64*, int, double, const double*, const double*, int, int);

}  // namespace opaw
