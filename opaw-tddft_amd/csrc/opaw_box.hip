// Projector-box kernels of the OPAW hot path (gfx950).
//
// Projection  ca(f, coff_a + i) = dv * sum_{box_a} u_f(r) p_{a,i}(r)
// (src/proj.f90:1-44, src/proj1.f90:1-46; the tables are real at the gamma
// point so conj(p) = p).  A workgroup owns a 64-point chunk of one atom's box
// and 16 fields: 16 lanes walk the chunk's points (consecutive points ->
// consecutive lanes, coalesced box rows), the 16 lane groups are the fields,
// each lane accumulates ms complex sums, and the 16 point lanes are reduced
// with four xor-shuffles.  One partial per chunk is written; a finishing
// kernel sums the chunks of an atom with a fixed shuffle tree (deterministic)
// and a weights kernel forms
//     w = sn(n) * ca                      (sn_phi, src/snphi.f90:24-28)
//     w = D(sys) ca                       (addvnl, src/hphi.f90:43-47)
// With a compact correction cpre the input is u = in + cpre (S^-1/2 applied on
// the fly), so a projection of S^-1/2 psi never needs S^-1/2 psi in memory.
//
// Add-back: union_corr turns a weight set into the compact correction
// corr(k, f) over the union of boxes (opaw_gather.h); consumers add it.
#include "opaw_box.h"

namespace opaw {

constexpr int kBoxThreads = 256;
constexpr int kFieldsPerWG = 16;
constexpr int kPointLanes = kBoxThreads / kFieldsPerWG;   // 16

template <typename R, int MSP, bool PRE>
__global__ __launch_bounds__(kBoxThreads) void k_proj(const AtomDesc* __restrict__ atoms,
                                                      const int2* __restrict__ chunks, const int* __restrict__ bpt,
                                                      const int* __restrict__ bk, const double* __restrict__ tabs,
                                                      const cx<R>* __restrict__ u, long long N, int nf,
                                                      const c64* __restrict__ cpre, int U,
                                                      double* __restrict__ part) {
  const int ch = blockIdx.x;
  const int kl = threadIdx.x % kPointLanes;
  const int f = blockIdx.y * kFieldsPerWG + threadIdx.x / kPointLanes;
  const int2 cd = chunks[ch];                     // (atom, first point)
  const AtomDesc A = atoms[cd.x];
  double accr[MSP], acci[MSP];
#pragma unroll
  for (int i = 0; i < MSP; ++i) accr[i] = acci[i] = 0.0;
  const int kend = min(A.npt, cd.y + kProjChunk);
  if (f < nf) {
    const cx<R>* uf = u + (long long)f * N;
#pragma unroll 4
    for (int k = cd.y + kl; k < kend; k += kPointLanes) {
      const int pt = bpt[A.boff + k];
      const cx<R> v = uf[pt];
      double vr = (double)v.x, vi = (double)v.y;
      if constexpr (PRE) {
        const c64 c = cpre[(long long)bk[A.boff + k] * nf + f];
        vr += c.x;
        vi += c.y;
      }
      const double* t = tabs + A.tab + k;
#pragma unroll
      for (int i = 0; i < MSP; ++i) {
        if (i < A.ms) {
          const double tv = t[(long long)i * A.npt];
          accr[i] += vr * tv;
          acci[i] += vi * tv;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < MSP; ++i) {
#pragma unroll
    for (int off = kPointLanes / 2; off > 0; off >>= 1) {
      accr[i] += __shfl_xor(accr[i], off, 64);
      acci[i] += __shfl_xor(acci[i], off, 64);
    }
  }
  if (kl == 0 && f < nf) {
    double* pp = part + ((long long)ch * nf + f) * MSP * 2;
#pragma unroll
    for (int i = 0; i < MSP; ++i) {
      if (i < A.ms) {
        pp[2 * i] = accr[i];
        pp[2 * i + 1] = acci[i];
      }
    }
  }
}

// ca(f, c) = dv * sum over the chunks of c's atom: one wave per (field,
// coefficient), lanes over chunks, fixed shuffle tree (deterministic)
template <int MSP>
__global__ __launch_bounds__(256) void k_proj_finish(const int2* __restrict__ coef, const int2* __restrict__ achunk,
                                                     const double* __restrict__ part, int nf, int ctot, double dv,
                                                     c64* __restrict__ ca) {
  const int f = blockIdx.y;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= ctot) return;
  const int2 ai = coef[c];                       // (atom, i)
  const int2 cr = achunk[ai.x];                  // (first chunk, chunk count)
  double sr = 0.0, si = 0.0;
  for (int q = lane; q < cr.y; q += 64) {
    const double* pp = part + (((long long)(cr.x + q) * nf + f) * MSP + ai.y) * 2;
    sr += pp[0];
    si += pp[1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if (lane == 0) ca[(long long)f * ctot + c] = mk(sr * dv, si * dv);
}

// w = D(sys) ca or sn * ca from existing coefficients, one workgroup per field
__global__ __launch_bounds__(128) void k_weights(const AtomDesc* __restrict__ atoms, const int2* __restrict__ coef,
                                                 const c64* __restrict__ ca, int nf, int ctot, int mode,
                                                 const double* __restrict__ sn, const double* __restrict__ dij,
                                                 int nf_per_sys, int mtot, c64* __restrict__ w) {
  const int f = blockIdx.x;
  const c64* row = ca + (long long)f * ctot;
  for (int c = threadIdx.x; c < ctot; c += blockDim.x) {
    const int2 ai = coef[c];
    c64 out;
    if (mode == 1) {
      const double s = sn[c];
      out = mk(s * row[c].x, s * row[c].y);
    } else {
      const AtomDesc A = atoms[ai.x];
      const double* D = dij + (long long)(f / nf_per_sys) * mtot + A.moff;
      double xr = 0.0, xi = 0.0;
      for (int j = 0; j < A.ms; ++j) {
        const double d = D[ai.y + A.ms * j];
        xr += d * row[A.coff + j].x;
        xi += d * row[A.coff + j].y;
      }
      out = mk(xr, xi);
    }
    w[(long long)c * nf + f] = out;
  }
}

// corr(k, f) for every union point k: one wave per point, lanes = fields
// (weights are coefficient-major, w[c*nf + f], so lane loads coalesce and the
// table value is shared by the wave); the ms loads of an entry are issued
// together (unrolled to MSP with a guard)
template <int MSP>
__global__ __launch_bounds__(kBoxThreads) void k_union_corr(UnionTabs U, const double* __restrict__ tab,
                                                            const c64* __restrict__ w, int nf,
                                                            c64* __restrict__ corr) {
  const int k = blockIdx.x * (kBoxThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= U.U) return;
  const int e0 = U.uptr[k], e1 = U.uptr[k + 1];
  for (int f0 = 0; f0 < nf; f0 += 64) {
    const int f = f0 + lane;
    const bool on = f < nf;
    double xr = 0.0, xi = 0.0;
    for (int e = e0; e < e1; ++e) {
      const UEntry E = U.ue[e];
      const double* t = tab + E.toff;
      const c64* ww = w + (long long)E.coff * nf + (on ? f : 0);
      double tv[MSP];
      c64 wv[MSP];
#pragma unroll
      for (int i = 0; i < MSP; ++i) {
        const bool ok = i < E.ms;
        tv[i] = ok ? t[i] : 0.0;
        wv[i] = ok ? ww[(long long)i * nf] : mk(0.0, 0.0);
      }
#pragma unroll
      for (int i = 0; i < MSP; ++i) {
        xr += tv[i] * wv[i].x;
        xi += tv[i] * wv[i].y;
      }
    }
    if (on) corr[(long long)k * nf + f] = mk(xr, xi);
  }
}

// out_f(upt[k]) += corr(k, f)
template <typename R>
__global__ __launch_bounds__(kBoxThreads) void k_union_apply(const int* __restrict__ upt, int U,
                                                             cx<R>* __restrict__ out, long long N, int nf,
                                                             const c64* __restrict__ corr) {
  const int k = blockIdx.x * kBoxThreads + threadIdx.x;
  const int f = blockIdx.y;
  if (k >= U) return;
  const c64 c = corr[(long long)k * nf + f];
  const long long o = (long long)f * N + upt[k];
  cx<R> v = out[o];
  v.x = (R)((double)v.x + c.x);
  v.y = (R)((double)v.y + c.y);
  out[o] = v;
}

// ---------------------------------------------------------------------------
template <typename R, int MSP>
static void launch_proj(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf,
                        const double* tab, const c64* cpre, double* part) {
  const long long N = (long long)g.nx * g.ny * g.nz;
  dim3 grid(bt.nchunks, (nf + kFieldsPerWG - 1) / kFieldsPerWG);
  if (cpre)
    hipLaunchKernelGGL((k_proj<R, MSP, true>), grid, dim3(kBoxThreads), 0, st, bt.atoms, bt.chunks, bt.bpt, bt.bk,
                       tab, u, N, nf, cpre, bt.U.U, part);
  else
    hipLaunchKernelGGL((k_proj<R, MSP, false>), grid, dim3(kBoxThreads), 0, st, bt.atoms, bt.chunks, bt.bpt, bt.bk,
                       tab, u, N, nf, cpre, bt.U.U, part);
}

template <typename R>
hipError_t box_coeffs(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf, int which,
                      const c64* cpre, double dv, double* part, c64* ca, int mode, const double* sn,
                      const double* dij, int nf_per_sys, int mtot, c64* w) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  const double* tab = which ? bt.p1 : bt.p;
  prof_begin(st, KC_BOX_PROJ);
  if (bt.msp == 8)
    launch_proj<R, 8>(g, bt, st, u, nf, tab, cpre, part);
  else
    launch_proj<R, 16>(g, bt, st, u, nf, tab, cpre, part);
  prof_end(st, KC_BOX_PROJ);
  hipError_t e = hipGetLastError();
  if (e) return e;
  prof_begin(st, KC_BOX_ADD);
  const dim3 fg((bt.ctot + 3) / 4, nf);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_proj_finish<8>, fg, dim3(256), 0, st, bt.coef, bt.achunk, part, nf, bt.ctot, dv, ca);
  else
    hipLaunchKernelGGL(k_proj_finish<16>, fg, dim3(256), 0, st, bt.coef, bt.achunk, part, nf, bt.ctot, dv, ca);
  prof_end(st, KC_BOX_ADD);
  if ((e = hipGetLastError())) return e;
  if (mode == 0) return hipSuccess;
  return box_weights(bt, st, ca, nf, mode, sn, dij, nf_per_sys, mtot, w);
}

hipError_t box_weights(const BoxTables& bt, hipStream_t st, const c64* ca, int nf, int mode, const double* sn,
                       const double* dij, int nf_per_sys, int mtot, c64* w) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  prof_begin(st, KC_BOX_ADD);
  hipLaunchKernelGGL(k_weights, dim3(nf), dim3(128), 0, st, bt.atoms, bt.coef, ca, nf, bt.ctot, mode, sn, dij,
                     nf_per_sys, mtot, w);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

hipError_t union_corr(const BoxTables& bt, hipStream_t st, int nf, int which, const c64* w, c64* corr) {
  if (bt.natom == 0 || nf == 0 || bt.U.U == 0) return hipSuccess;
  prof_begin(st, KC_BOX_ADD);
  const int per = kBoxThreads / 64;
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_union_corr<8>, dim3((bt.U.U + per - 1) / per), dim3(kBoxThreads), 0, st, bt.U,
                       which ? bt.U.p1 : bt.U.p, w, nf, corr);
  else
    hipLaunchKernelGGL(k_union_corr<16>, dim3((bt.U.U + per - 1) / per), dim3(kBoxThreads), 0, st, bt.U,
                       which ? bt.U.p1 : bt.U.p, w, nf, corr);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

template <typename R>
hipError_t union_apply(const GridGeom& g, const BoxTables& bt, hipStream_t st, cx<R>* out, int nf, const c64* corr) {
  if (bt.natom == 0 || nf == 0 || bt.U.U == 0) return hipSuccess;
  const long long N = (long long)g.nx * g.ny * g.nz;
  prof_begin(st, KC_BOX_ADD);
  hipLaunchKernelGGL((k_union_apply<R>), dim3((bt.U.U + kBoxThreads - 1) / kBoxThreads, nf), dim3(kBoxThreads), 0,
                     st, bt.U.upt, bt.U.U, out, N, nf, corr);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

template hipError_t box_coeffs<double>(const GridGeom&, const BoxTables&, hipStream_t, const c64*, int, int,
                                       const c64*, double, double*, c64*, int, const double*, const double*, int, int,
                                       c64*);
template hipError_t box_coeffs<float>(const GridGeom&, const BoxTables&, hipStream_t, const c32*, int, int,
                                      const c64*, double, double*, c64*, int, const double*, const double*, int, int,
                                      c64*);
template hipError_t union_apply<double>(const GridGeom&, const BoxTables&, hipStream_t, c64*, int, const c64*);
template hipError_t union_apply<float>(const GridGeom&, const BoxTables&, hipStream_t, c32*, int, const c64*);

}  // namespace opaw
