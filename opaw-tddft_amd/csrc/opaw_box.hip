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

// SRC: 0 = the grid fields u, 1 = u + cpre (compact correction),
// 2 = the compact values cpre alone (fields stored only on the union of
// boxes, cpre[k*nf + f]; u is not read), 4 = by blockIdx.z: z = 0, 1 the
// grid fields u with tabs0 / tabs1, z = 2 the compact values cpre with tabs0
// into part2 (three projections in one launch)
template <typename R, int MSP, int SRC>
__global__ __launch_bounds__(kBoxThreads) void k_proj(const AtomDesc* __restrict__ atoms,
                                                      const int2* __restrict__ chunks, const int* __restrict__ bpt,
                                                      const int* __restrict__ bk, const double* __restrict__ tabs0,
                                                      const double* __restrict__ tabs1,
                                                      const cx<R>* __restrict__ u, long long N, int nf,
                                                      const c64* __restrict__ cpre, int /*table sets, = gridDim.z*/,
                                                      double* __restrict__ part0, double* __restrict__ part1,
                                                      double* __restrict__ part2) {
  // blockIdx.z selects the table set (two projections of the same u in one
  // launch; SRC 4: a third one of the compact values)
  const double* __restrict__ tabs = blockIdx.z == 1 ? tabs1 : tabs0;
  double* __restrict__ part = blockIdx.z == 0 ? part0 : (blockIdx.z == 1 ? part1 : part2);
  // where the values come from (uniform per workgroup)
  const bool gsrc = SRC == 4 ? blockIdx.z < 2 : SRC < 2;
  const bool csrc = SRC == 4 ? blockIdx.z == 2 : SRC != 0;
  const int ch = blockIdx.x;
  const int kl = threadIdx.x % kPointLanes;
  const int f = blockIdx.y * kFieldsPerWG + threadIdx.x / kPointLanes;
  const int2 cd = chunks[ch];                     // (atom, first point)
  const AtomDesc A = atoms[cd.x];
  double accr[MSP], acci[MSP];
#pragma unroll
  for (int i = 0; i < MSP; ++i) accr[i] = acci[i] = 0.0;
  const int kend = min(A.npt, cd.y + kProjChunk);
  constexpr int PPT = kProjChunk / kPointLanes;   // points per lane
  // Every load is unconditional (point index clamped into the chunk,
  // projector index clamped to the atom's ms, field clamped to nf - 1) so the
  // compiler issues them together; masked points contribute zero and
  // projectors i >= ms and fields f >= nf are never written.
  const int fc = min(f, nf - 1);
  const cx<R>* uf = gsrc ? u + (long long)fc * N : nullptr;
  int kk[PPT];
  bool ok[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int k = cd.y + kl + q * kPointLanes;
    ok[q] = k < kend;
    kk[q] = ok[q] ? k : cd.y;
  }
  int pt[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) pt[q] = gsrc ? bpt[A.boff + kk[q]] : 0;
  double tv[PPT][MSP];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const double* t = tabs + A.tab + kk[q];
#pragma unroll
    for (int i = 0; i < MSP; ++i) tv[q][i] = t[(long long)(i < A.ms ? i : 0) * A.npt];
  }
  double vr[PPT], vi[PPT];
  if (gsrc) {
    cx<R> v[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) v[q] = uf[pt[q]];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      vr[q] = ok[q] ? (double)v[q].x : 0.0;
      vi[q] = ok[q] ? (double)v[q].y : 0.0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < PPT; ++q) vr[q] = vi[q] = 0.0;
  }
  if (csrc) {
    int bq[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) bq[q] = bk[A.boff + kk[q]];   // every box point is a union point
    c64 cv[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) cv[q] = cpre[(long long)bq[q] * nf + fc];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      if (ok[q]) {
        vr[q] += cv[q].x;
        vi[q] += cv[q].y;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
#pragma unroll
    for (int i = 0; i < MSP; ++i) {
      accr[i] += vr[q] * tv[q][i];
      acci[i] += vi[q] * tv[q][i];
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

// ca(f, coff_a + i) = dv * sum over the atom's chunks (one wave per
// coefficient, lanes over chunks, fixed shuffle tree: deterministic), then the
// weights of the atom (mode 1: sn * ca, mode 2: D(sys) ca, 0: none).  One
// workgroup per (atom, field).
template <int MSP>
__global__ __launch_bounds__(256) void k_proj_finish(const AtomDesc* __restrict__ atoms,
                                                     const int2* __restrict__ achunk,
                                                     const double* __restrict__ part0, int nf, int ctot, double dv,
                                                     c64* __restrict__ ca0, int mode0, const double* __restrict__ sn,
                                                     const double* __restrict__ dij, int nf_per_sys, int mtot,
                                                     c64* __restrict__ w0, const double* __restrict__ part1,
                                                     c64* __restrict__ ca1, const double* __restrict__ part2,
                                                     c64* __restrict__ ca2, c64* __restrict__ w2) {
  // blockIdx.z = 1: the second projection of a dual launch (coefficients
  // only); z = 2: a third one with the sn weights (mode 1) into w2
  const int z = blockIdx.z;
  const double* __restrict__ part = z == 0 ? part0 : (z == 1 ? part1 : part2);
  c64* __restrict__ ca = z == 0 ? ca0 : (z == 1 ? ca1 : ca2);
  const int mode = z == 0 ? mode0 : (z == 1 ? 0 : 1);
  c64* __restrict__ w = z == 2 ? w2 : w0;
  const int a = blockIdx.x, f = blockIdx.y;
  const AtomDesc A = atoms[a];
  const int2 cr = achunk[a];                     // (first chunk, chunk count)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ c64 row[kMaxMs];
  for (int i = wave; i < A.ms; i += 4) {
    double sr = 0.0, si = 0.0;
    for (int q = lane; q < cr.y; q += 64) {
      const double* pp = part + (((long long)(cr.x + q) * nf + f) * MSP + i) * 2;
      sr += pp[0];
      si += pp[1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_xor(sr, off, 64);
      si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) {
      const c64 v = mk(sr * dv, si * dv);
      ca[(long long)f * ctot + A.coff + i] = v;
      row[i] = v;
    }
  }
  if (mode == 0) return;
  __syncthreads();
  const int i = threadIdx.x;
  if (i < A.ms) {
    c64 out;
    if (mode == 1) {
      const double s = sn[A.coff + i];
      out = mk(s * row[i].x, s * row[i].y);
    } else {
      const double* D = dij + (long long)(f / nf_per_sys) * mtot + A.moff;
      double xr = 0.0, xi = 0.0;
      for (int j = 0; j < A.ms; ++j) {
        const double d = D[i + A.ms * j];
        xr += d * row[j].x;
        xi += d * row[j].y;
      }
      out = mk(xr, xi);
    }
    w[(long long)(A.coff + i) * nf + f] = out;
  }
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

// corr(k, f) for every union point k: one wave per point, lanes = fields.
// The point's atom entries are flattened into (coefficient, table value)
// pairs in atom order, read in batches of 8 so their loads and the weight
// loads (coefficient-major, w[c*nf + f]: coalesced over lanes) are in flight
// together.
__global__ __launch_bounds__(kBoxThreads) void k_union_corr(const int* __restrict__ pptr,
                                                            const int* __restrict__ pc,
                                                            const double* __restrict__ pv, int U,
                                                            const c64* __restrict__ w, int nf,
                                                            c64* __restrict__ corr) {
  const int k = blockIdx.x * (kBoxThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= U) return;
  const int p0 = pptr[k], p1 = pptr[k + 1];
  for (int f0 = 0; f0 < nf; f0 += 64) {
    const int f = f0 + lane;
    const bool on = f < nf;
    const c64* wf = w + (on ? f : 0);
    double xr = 0.0, xi = 0.0;
    for (int p = p0; p < p1; p += 8) {
      int c[8];
      double t[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const bool ok = p + q < p1;
        c[q] = ok ? pc[p + q] : 0;
        t[q] = ok ? pv[p + q] : 0.0;
      }
      c64 wv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) wv[q] = wf[(long long)c[q] * nf];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        xr += t[q] * wv[q].x;
        xi += t[q] * wv[q].y;
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
                        const double* tab, const c64* cpre, double* part, const double* tab1 = nullptr,
                        double* part1 = nullptr, double* part2 = nullptr) {
  const long long N = (long long)g.nx * g.ny * g.nz;
  dim3 grid(bt.nchunks, (nf + kFieldsPerWG - 1) / kFieldsPerWG, part2 ? 3 : (tab1 ? 2 : 1));
  if (part2)
    hipLaunchKernelGGL((k_proj<R, MSP, 4>), grid, dim3(kBoxThreads), 0, st, bt.atoms, bt.chunks, bt.bpt, bt.bk,
                       tab, tab1, u, N, nf, cpre, bt.U.U, part, part1, part2);
  else if (cpre && !u)
    hipLaunchKernelGGL((k_proj<R, MSP, 2>), grid, dim3(kBoxThreads), 0, st, bt.atoms, bt.chunks, bt.bpt, bt.bk,
                       tab, tab1, u, N, nf, cpre, bt.U.U, part, part1, nullptr);
  else if (cpre)
    hipLaunchKernelGGL((k_proj<R, MSP, 1>), grid, dim3(kBoxThreads), 0, st, bt.atoms, bt.chunks, bt.bpt, bt.bk,
                       tab, tab1, u, N, nf, cpre, bt.U.U, part, part1, nullptr);
  else
    hipLaunchKernelGGL((k_proj<R, MSP, 0>), grid, dim3(kBoxThreads), 0, st, bt.atoms, bt.chunks, bt.bpt, bt.bk,
                       tab, tab1, u, N, nf, cpre, bt.U.U, part, part1, nullptr);
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
  const dim3 fg(bt.natom, nf);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_proj_finish<8>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, ca, mode,
                       sn, dij, nf_per_sys, mtot, w, nullptr, nullptr, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL(k_proj_finish<16>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, ca, mode,
                       sn, dij, nf_per_sys, mtot, w, nullptr, nullptr, nullptr, nullptr, nullptr);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// Dual projection of one set of fields: cA = dv P1^T u with weights
// wA = sn cA, and cB = dv P^T u, in one projection launch and one finish
// launch (blockIdx.z selects the table set).  part must hold two partial sets.
template <typename R>
hipError_t box_coeffs_dual(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf,
                           double dv, double* part, c64* cA, const double* sn, c64* wA, c64* cB) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  double* part1 = part + (size_t)bt.nchunks * nf * bt.msp * 2;
  prof_begin(st, KC_BOX_PROJ);
  if (bt.msp == 8)
    launch_proj<R, 8>(g, bt, st, u, nf, bt.p1, nullptr, part, bt.p, part1);
  else
    launch_proj<R, 16>(g, bt, st, u, nf, bt.p1, nullptr, part, bt.p, part1);
  prof_end(st, KC_BOX_PROJ);
  hipError_t e = hipGetLastError();
  if (e) return e;
  prof_begin(st, KC_BOX_ADD);
  const dim3 fg(bt.natom, nf, 2);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_proj_finish<8>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, cA, 1, sn,
                       nullptr, nf, 0, wA, part1, cB, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL(k_proj_finish<16>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, cA, 1, sn,
                       nullptr, nf, 0, wA, part1, cB, nullptr, nullptr, nullptr);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// cB += O wA for every (atom a, field f): the projections of S^-1/2 u from
// those of u, since P_a^T S^-1/2 u dv = P_a^T u dv + sum_b O_ab wA_b with the
// fixed overlaps O_ab(i, j) = dv sum_r p_ai(r) p1_bj(r) over the common points
// of the two boxes (host-built, FP64, grid order).  Per atom the nonzero
// columns are flattened: column l is coefficient oc[optr[a] + l] and
// O(i, l) = ov[ovoff[a] + i * L_a + l].  One wave per (atom, field): lanes
// over columns, a fixed shuffle tree over the lanes (deterministic).
//
// With wC (the RK4 stage's outer-S^-1/2 weights) the column weights are
// wA + a1 (-i) wC: cB of y + a1 (-i) P1 wC from cB of y, as the stage's
// union add-back changes y by a1 (-i) P1 wC.
template <int MSP>
__global__ __launch_bounds__(64) void k_bfix(const AtomDesc* __restrict__ atoms, const int* __restrict__ optr,
                                             const int* __restrict__ ovoff, const int* __restrict__ oc,
                                             const double* __restrict__ ov, const c64* __restrict__ wA, int nf,
                                             int ctot, c64* __restrict__ cB, const c64* __restrict__ wC, double a1) {
  const int a = blockIdx.x, f = blockIdx.y, lane = threadIdx.x;
  const AtomDesc A = atoms[a];
  const int l0 = optr[a], La = optr[a + 1] - l0;
  const double* O = ov + ovoff[a];
  double xr[MSP], xi[MSP];
#pragma unroll
  for (int i = 0; i < MSP; ++i) xr[i] = xi[i] = 0.0;
  for (int lb = 0; lb < La; lb += 64) {
    const int l = min(lb + lane, La - 1);
    const bool on = lb + lane < La;
    const int c = oc[l0 + l];
    double o[MSP];
#pragma unroll
    for (int i = 0; i < MSP; ++i) o[i] = O[(long long)(i < A.ms ? i : 0) * La + l];
    c64 w = wA[(long long)c * nf + f];
    if (wC) {
      const c64 v = wC[(long long)c * nf + f];
      w = mk(w.x + a1 * v.y, w.y - a1 * v.x);
    }
    const double wr = on ? w.x : 0.0, wi = on ? w.y : 0.0;
#pragma unroll
    for (int i = 0; i < MSP; ++i) {
      xr[i] += o[i] * wr;
      xi[i] += o[i] * wi;
    }
  }
#pragma unroll
  for (int i = 0; i < MSP; ++i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      xr[i] += __shfl_xor(xr[i], off, 64);
      xi[i] += __shfl_xor(xi[i], off, 64);
    }
  }
#pragma unroll
  for (int i = 0; i < MSP; ++i) {
    if (lane == i && i < A.ms) {
      c64& b = cB[(long long)f * ctot + A.coff + i];
      b = mk(b.x + xr[i], b.y + xi[i]);
    }
  }
}

hipError_t box_bfix(const BoxTables& bt, hipStream_t st, int nf, const c64* wA, c64* cB, const c64* wC, double a1) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  prof_begin(st, KC_BOX_ADD);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_bfix<8>, dim3(bt.natom, nf), dim3(64), 0, st, bt.atoms, bt.optr, bt.ovoff, bt.oc, bt.ov, wA,
                       nf, bt.ctot, cB, wC, a1);
  else
    hipLaunchKernelGGL(k_bfix<16>, dim3(bt.natom, nf), dim3(64), 0, st, bt.atoms, bt.optr, bt.ovoff, bt.oc, bt.ov, wA,
                       nf, bt.ctot, cB, wC, a1);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// cA += a1 (-i) O11 wC, then wA = sn cA, for every (atom a, field f): the
// P1 coefficients of y + a1 (-i) P1 wC from those of y, with the fixed p1-p1
// box overlaps O11_ab(i, j) = dv sum_r p1_ai(r) p1_bj(r) (same columns as
// k_bfix's O).  One wave per (atom, field), fixed shuffle tree.
template <int MSP>
__global__ __launch_bounds__(64) void k_afix(const AtomDesc* __restrict__ atoms, const int* __restrict__ optr,
                                             const int* __restrict__ ovoff, const int* __restrict__ oc,
                                             const double* __restrict__ ov1, const c64* __restrict__ wC, int nf,
                                             int ctot, double a1, const double* __restrict__ sn,
                                             c64* __restrict__ cA, c64* __restrict__ wA) {
  const int a = blockIdx.x, f = blockIdx.y, lane = threadIdx.x;
  const AtomDesc A = atoms[a];
  const int l0 = optr[a], La = optr[a + 1] - l0;
  const double* O = ov1 + ovoff[a];
  double xr[MSP], xi[MSP];
#pragma unroll
  for (int i = 0; i < MSP; ++i) xr[i] = xi[i] = 0.0;
  for (int lb = 0; lb < La; lb += 64) {
    const int l = min(lb + lane, La - 1);
    const bool on = lb + lane < La;
    const int c = oc[l0 + l];
    double o[MSP];
#pragma unroll
    for (int i = 0; i < MSP; ++i) o[i] = O[(long long)(i < A.ms ? i : 0) * La + l];
    const c64 w = wC[(long long)c * nf + f];
    const double wr = on ? w.x : 0.0, wi = on ? w.y : 0.0;
#pragma unroll
    for (int i = 0; i < MSP; ++i) {
      xr[i] += o[i] * wr;
      xi[i] += o[i] * wi;
    }
  }
#pragma unroll
  for (int i = 0; i < MSP; ++i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      xr[i] += __shfl_xor(xr[i], off, 64);
      xi[i] += __shfl_xor(xi[i], off, 64);
    }
  }
#pragma unroll
  for (int i = 0; i < MSP; ++i) {
    if (lane == i && i < A.ms) {
      const int c = A.coff + i;
      c64& v = cA[(long long)f * ctot + c];
      v = mk(v.x + a1 * xi[i], v.y - a1 * xr[i]);
      const double s = sn[c];
      wA[(long long)c * nf + f] = mk(s * v.x, s * v.y);
    }
  }
}

hipError_t box_afix(const BoxTables& bt, hipStream_t st, int nf, const c64* wC, double a1, const double* sn, c64* cA,
                    c64* wA) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  prof_begin(st, KC_BOX_ADD);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_afix<8>, dim3(bt.natom, nf), dim3(64), 0, st, bt.atoms, bt.optr, bt.ovoff, bt.oc, bt.ov1, wC,
                       nf, bt.ctot, a1, sn, cA, wA);
  else
    hipLaunchKernelGGL(k_afix<16>, dim3(bt.natom, nf), dim3(64), 0, st, bt.atoms, bt.optr, bt.ovoff, bt.oc, bt.ov1, wC,
                       nf, bt.ctot, a1, sn, cA, wA);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// The three projections after an RK4 stage's inverse kinetic pass in one
// launch: A = P1^T y and B = P^T y of the grid fields y, and C = P1^T h of
// the compact h (hU), then their coefficients and the weights wC = sn C.
template <typename R>
hipError_t box_coeffs_triple(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* y, const c64* hU,
                             int nf, double dv, double* part, c64* cA, c64* cB, c64* cC, const double* sn, c64* wC) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  double* part1 = part + (size_t)bt.nchunks * nf * bt.msp * 2;
  double* part2 = part1 + (size_t)bt.nchunks * nf * bt.msp * 2;
  prof_begin(st, KC_BOX_PROJ);
  if (bt.msp == 8)
    launch_proj<R, 8>(g, bt, st, y, nf, bt.p1, hU, part, bt.p, part1, part2);
  else
    launch_proj<R, 16>(g, bt, st, y, nf, bt.p1, hU, part, bt.p, part1, part2);
  prof_end(st, KC_BOX_PROJ);
  hipError_t e = hipGetLastError();
  if (e) return e;
  prof_begin(st, KC_BOX_ADD);
  const dim3 fg(bt.natom, nf, 3);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_proj_finish<8>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, cA, 0, sn,
                       nullptr, nf, 0, nullptr, part1, cB, part2, cC, wC);
  else
    hipLaunchKernelGGL(k_proj_finish<16>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, cA, 0, sn,
                       nullptr, nf, 0, nullptr, part1, cB, part2, cC, wC);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}
template hipError_t box_coeffs_triple<double>(const GridGeom&, const BoxTables&, hipStream_t, const c64*, const c64*,
                                              int, double, double*, c64*, c64*, c64*, const double*, c64*);
template hipError_t box_coeffs_triple<float>(const GridGeom&, const BoxTables&, hipStream_t, const c32*, const c64*,
                                             int, double, double*, c64*, c64*, c64*, const double*, c64*);

hipError_t box_weights(const BoxTables& bt, hipStream_t st, const c64* ca, int nf, int mode, const double* sn,
                       const double* dij, int nf_per_sys, int mtot, c64* w) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  prof_begin(st, KC_BOX_ADD);
  hipLaunchKernelGGL(k_weights, dim3(nf), dim3(128), 0, st, bt.atoms, bt.coef, ca, nf, bt.ctot, mode, sn, dij,
                     nf_per_sys, mtot, w);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// The same corrections with the weights staged in LDS.  Every union point
// reads ~20 weight rows, so the global kernel above re-reads the (small)
// weight matrix from L2 about 20 times per launch.  Here a workgroup owns a
// block of consecutive union points (host-built: at most kUnionBlockRows
// points and kUnionBlockPairs pairs) and FG fields: it copies the weights of
// those fields for every coefficient (ctot x FG) and the block's pairs into
// LDS with coalesced loads, then every sum reads LDS only.  Lanes are
// (point, field); the pair order per point is the one of k_union_corr.
//
// APPLY: instead of storing corr, add it into the grid fields out
// (out_f(upt[k]) += corr(k, f)): the block's results are staged in LDS and
// re-read point-fastest, so the read-modify-writes walk consecutive grid
// points of one field.
template <int FG, typename R, int MODE>
__global__ __launch_bounds__(kBoxThreads) void k_union_corr_lds(const int2* __restrict__ ublk,
                                                                const int* __restrict__ pptr,
                                                                const int* __restrict__ pc,
                                                                const int* __restrict__ ucptr,
                                                                const int* __restrict__ ucol,
                                                                const double* __restrict__ pv,
                                                                const c64* __restrict__ w, int nf, int ctot,
                                                                c64* __restrict__ corr, const int* __restrict__ upt,
                                                                cx<R>* __restrict__ out, long long N,
                                                                cx<R>* __restrict__ out2, double a1, double a2,
                                                                const c64* __restrict__ w2) {
  // MODE 0: corr(k, f) = sum; 1: out_f(upt[k]) += sum; 2 (RK4 stage
  // add-back of the outer S^-1/2, linear in h): with kc = -i sum,
  // out_f += a1 kc and out2_f += a2 kc at upt[k]; 3: MODE 2 and, from the
  // second weight set w2 staged beside w, corr(k, f) = its sum
  constexpr bool APPLY = MODE != 0;
  constexpr int NSET = MODE == 3 ? 2 : 1;
  extern __shared__ __align__(16) double smem[];
  c64* wl = reinterpret_cast<c64*>(smem);                 // [ctot][FG]
  c64* wl2 = wl + (long long)ctot * FG;                   // [ctot][FG], MODE 3
  double* tl = smem + 2 * NSET * (long long)ctot * FG;    // pair values
  int* cl = reinterpret_cast<int*>(tl + kUnionBlockPairs);  // pair coefficients
  int* rl = cl + kUnionBlockPairs;                        // row starts (block-relative)
  c64* res = reinterpret_cast<c64*>(smem + 2 * NSET * (long long)ctot * FG + (kUnionBlockPairs * 12 + (kUnionBlockRows + 1) * 4 + 15) / 16 * 2);
  const int2 blk = ublk[blockIdx.x];                      // (first point, point count)
  const int f0 = blockIdx.y * FG;
  const int q0 = pptr[blk.x];
  const int np = pptr[blk.x + blk.y] - q0;                // >= 1: every union point has a pair
  // staging of the block's own coefficient rows (ctot bounds the list),
  // every thread issuing its loads in batches (clamped indices); pc holds
  // indices into the list
  const int c0 = ucptr[blockIdx.x];
  const int nc = ucptr[blockIdx.x + 1] - c0;
  const int nw = nc * FG;
  for (int b0 = 0; b0 < nw; b0 += kBoxThreads * 8) {
    c64 t[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int q = min(b0 + s * kBoxThreads + (int)threadIdx.x, nw - 1);
      t[s] = w[(long long)ucol[c0 + q / FG] * nf + min(f0 + q % FG, nf - 1)];
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int q = b0 + s * kBoxThreads + (int)threadIdx.x;
      if (q < nw) wl[q] = (f0 + q % FG < nf) ? t[s] : mk(0.0, 0.0);
    }
    if constexpr (NSET == 2) {
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int q = min(b0 + s * kBoxThreads + (int)threadIdx.x, nw - 1);
        t[s] = w2[(long long)ucol[c0 + q / FG] * nf + min(f0 + q % FG, nf - 1)];
      }
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int q = b0 + s * kBoxThreads + (int)threadIdx.x;
        if (q < nw) wl2[q] = (f0 + q % FG < nf) ? t[s] : mk(0.0, 0.0);
      }
    }
  }
  {
    constexpr int PB = kUnionBlockPairs / kBoxThreads;
    double tv[PB];
    int cv[PB];
#pragma unroll
    for (int s = 0; s < PB; ++s) {
      const int q = min(s * kBoxThreads + (int)threadIdx.x, np - 1);
      tv[s] = pv[q0 + q];
      cv[s] = pc[q0 + q];
    }
#pragma unroll
    for (int s = 0; s < PB; ++s) {
      const int q = s * kBoxThreads + (int)threadIdx.x;
      if (q < np) {
        tl[q] = tv[s];
        cl[q] = cv[s];
      }
    }
  }
  if ((int)threadIdx.x <= blk.y) rl[threadIdx.x] = pptr[blk.x + threadIdx.x] - q0;
  __syncthreads();
  constexpr int PL = kBoxThreads / FG;
  const int j = threadIdx.x % FG;
  const int f = f0 + j;
  for (int r = threadIdx.x / FG; r < blk.y; r += PL) {
    const int a = rl[r], e = rl[r + 1];
    double xr = 0.0, xi = 0.0, yr = 0.0, yi = 0.0;
    int p = a;
    for (; p + 4 <= e; p += 4) {
      int cc[4];
      double tt[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        cc[s] = cl[p + s];
        tt[s] = tl[p + s];
      }
      c64 wv[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) wv[s] = wl[cc[s] * FG + j];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        xr += tt[s] * wv[s].x;
        xi += tt[s] * wv[s].y;
      }
      if constexpr (NSET == 2) {
#pragma unroll
        for (int s = 0; s < 4; ++s) wv[s] = wl2[cc[s] * FG + j];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          yr += tt[s] * wv[s].x;
          yi += tt[s] * wv[s].y;
        }
      }
    }
    for (; p < e; ++p) {
      const c64 wv = wl[cl[p] * FG + j];
      const double t = tl[p];
      xr += t * wv.x;
      xi += t * wv.y;
      if constexpr (NSET == 2) {
        const c64 wv2 = wl2[cl[p] * FG + j];
        yr += t * wv2.x;
        yi += t * wv2.y;
      }
    }
    if constexpr (APPLY)
      res[r * FG + j] = mk(xr, xi);
    else if (f < nf)
      corr[(long long)(blk.x + r) * nf + f] = mk(xr, xi);
    if constexpr (NSET == 2)
      if (f < nf) corr[(long long)(blk.x + r) * nf + f] = mk(yr, yi);
  }
  if constexpr (APPLY) {
    __syncthreads();
    for (int q = threadIdx.x; q < blk.y * FG; q += kBoxThreads) {
      const int r = q % blk.y, jj = q / blk.y;
      if (f0 + jj < nf) {
        const long long gi = (long long)(f0 + jj) * N + upt[blk.x + r];
        const c64 v = res[r * FG + jj];
        if constexpr (MODE == 1) {
          cx<R>& o = out[gi];
          o = mk<R>((R)((double)o.x + v.x), (R)((double)o.y + v.y));
        } else {
          const double kr = v.y, ki = -v.x;
          cx<R>& o = out[gi];
          o = mk<R>((R)((double)o.x + a1 * kr), (R)((double)o.y + a1 * ki));
          cx<R>& o2 = out2[gi];
          o2 = mk<R>((R)((double)o2.x + a2 * kr), (R)((double)o2.y + a2 * ki));
        }
      }
    }
  }
}

// LDS of the kernel for coefficient lists of at most ctot rows and nset
// weight sets
static size_t union_lds_bytes(int ctot, int fg, bool apply, int nset = 1) {
  size_t s = (size_t)nset * ctot * fg * 16 +
             ((size_t)kUnionBlockPairs * 12 + (kUnionBlockRows + 1) * 4 + 15) / 16 * 16;
  if (apply) s += (size_t)kUnionBlockRows * fg * 16;
  return s;
}

// fields per workgroup of the LDS kernel for ctot coefficients, within 64 KiB
// of LDS (0: use the global kernel)
static int union_lds_fields(int ctot, bool apply, int nset = 1) {
  for (int fg : {16, 8, 4})
    if (union_lds_bytes(ctot, fg, apply, nset) <= 65536) return fg;
  return 0;
}

template <typename R, int MODE>
static void launch_union_lds(const BoxTables& bt, hipStream_t st, int fg, int nf, const double* pv, const c64* w,
                             c64* corr, cx<R>* out, long long N, cx<R>* out2 = nullptr, double a1 = 0.0,
                             double a2 = 0.0, const c64* w2 = nullptr) {
  const size_t lds = union_lds_bytes(bt.ucmax, fg, MODE != 0, MODE == 3 ? 2 : 1);
  const dim3 grid(bt.nublk, (nf + fg - 1) / fg);
  if (fg == 16)
    hipLaunchKernelGGL((k_union_corr_lds<16, R, MODE>), grid, dim3(kBoxThreads), lds, st, bt.ublk, bt.pptr, bt.pcl,
                       bt.ucptr, bt.ucol, pv, w, nf, bt.ucmax, corr, bt.U.upt, out, N, out2, a1, a2, w2);
  else if (fg == 8)
    hipLaunchKernelGGL((k_union_corr_lds<8, R, MODE>), grid, dim3(kBoxThreads), lds, st, bt.ublk, bt.pptr, bt.pcl,
                       bt.ucptr, bt.ucol, pv, w, nf, bt.ucmax, corr, bt.U.upt, out, N, out2, a1, a2, w2);
  else
    hipLaunchKernelGGL((k_union_corr_lds<4, R, MODE>), grid, dim3(kBoxThreads), lds, st, bt.ublk, bt.pptr, bt.pcl,
                       bt.ucptr, bt.ucol, pv, w, nf, bt.ucmax, corr, bt.U.upt, out, N, out2, a1, a2, w2);
}

hipError_t union_corr(const BoxTables& bt, hipStream_t st, int nf, int which, const c64* w, c64* corr) {
  if (bt.natom == 0 || nf == 0 || bt.U.U == 0) return hipSuccess;
  prof_begin(st, KC_BOX_ADD);
  const double* pv = which ? bt.pv1 : bt.pv;
  const int fg = bt.nublk > 0 ? union_lds_fields(bt.ucmax, false) : 0;
  if (fg > 0) {
    launch_union_lds<double, 0>(bt, st, fg, nf, pv, w, corr, nullptr, 0);
  } else {
    const int per = kBoxThreads / 64;
    hipLaunchKernelGGL(k_union_corr, dim3((bt.U.U + per - 1) / per), dim3(kBoxThreads), 0, st, bt.pptr, bt.pc, pv,
                       bt.U.U, w, nf, corr);
  }
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

template <typename R>
hipError_t union_corr_apply(const GridGeom& g, const BoxTables& bt, hipStream_t st, int nf, int which,
                            const c64* w, c64* corr, cx<R>* out) {
  if (bt.natom == 0 || nf == 0 || bt.U.U == 0) return hipSuccess;
  const int fg = bt.nublk > 0 ? union_lds_fields(bt.ucmax, true) : 0;
  if (fg == 0) {
    hipError_t e = union_corr(bt, st, nf, which, w, corr);
    if (e) return e;
    return union_apply<R>(g, bt, st, out, nf, corr);
  }
  prof_begin(st, KC_BOX_ADD);
  launch_union_lds<R, 1>(bt, st, fg, nf, which ? bt.pv1 : bt.pv, w, corr, out, (long long)g.nx * g.ny * g.nz);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// RK4 stage add-back: with kc = -i corr(k, f) of the weights w, y += a1 kc
// and acc += a2 kc on the union of boxes (fallback without union blocks:
// corr then two scattered passes)
template <typename R>
__global__ __launch_bounds__(kBoxThreads) void k_union_rk4(const int* __restrict__ upt, int U, cx<R>* __restrict__ y,
                                                           cx<R>* __restrict__ acc, long long N, int nf,
                                                           const c64* __restrict__ corr, double a1, double a2) {
  const int k = blockIdx.x * kBoxThreads + threadIdx.x;
  const int f = blockIdx.y;
  if (k >= U) return;
  const c64 v = corr[(long long)k * nf + f];
  const double kr = v.y, ki = -v.x;
  const long long gi = (long long)f * N + upt[k];
  const cx<R> a = y[gi];
  y[gi] = mk<R>((R)((double)a.x + a1 * kr), (R)((double)a.y + a1 * ki));
  const cx<R> b = acc[gi];
  acc[gi] = mk<R>((R)((double)b.x + a2 * kr), (R)((double)b.y + a2 * ki));
}

template <typename R>
hipError_t union_rk4_apply(const GridGeom& g, const BoxTables& bt, hipStream_t st, int nf, const c64* w, c64* corr,
                           cx<R>* y, cx<R>* acc, double a1, double a2) {
  if (bt.natom == 0 || nf == 0 || bt.U.U == 0) return hipSuccess;
  const long long N = (long long)g.nx * g.ny * g.nz;
  const int fg = bt.nublk > 0 ? union_lds_fields(bt.ucmax, true) : 0;
  if (fg == 0) {
    hipError_t e = union_corr(bt, st, nf, 1, w, corr);
    if (e) return e;
    prof_begin(st, KC_BOX_ADD);
    hipLaunchKernelGGL((k_union_rk4<R>), dim3((bt.U.U + kBoxThreads - 1) / kBoxThreads, nf), dim3(kBoxThreads), 0,
                       st, bt.U.upt, bt.U.U, y, acc, N, nf, corr, a1, a2);
    prof_end(st, KC_BOX_ADD);
    return hipGetLastError();
  }
  prof_begin(st, KC_BOX_ADD);
  launch_union_lds<R, 2>(bt, st, fg, nf, bt.pv1, w, corr, y, N, acc, a1, a2);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// RK4 stage add-back of wC and corrA = P1 wA in one LDS pass (fallback
// without a fitting LDS layout: the two separate passes)
template <typename R>
hipError_t union_rk4_both(const GridGeom& g, const BoxTables& bt, hipStream_t st, int nf, const c64* wC,
                          const c64* wA, c64* corrC, c64* corrA, cx<R>* y, cx<R>* acc, double a1, double a2) {
  if (bt.natom == 0 || nf == 0 || bt.U.U == 0) return hipSuccess;
  const long long N = (long long)g.nx * g.ny * g.nz;
  const int fg = bt.nublk > 0 ? union_lds_fields(bt.ucmax, true, 2) : 0;
  if (fg == 0) {
    hipError_t e = union_rk4_apply<R>(g, bt, st, nf, wC, corrC, y, acc, a1, a2);
    if (e) return e;
    return union_corr(bt, st, nf, 1, wA, corrA);
  }
  prof_begin(st, KC_BOX_ADD);
  launch_union_lds<R, 3>(bt, st, fg, nf, bt.pv1, wC, corrA, y, N, acc, a1, a2, wA);
  prof_end(st, KC_BOX_ADD);
  return hipGetLastError();
}

// coefficients (P1, weights sn) of fields stored only on the union of boxes
// (vals[k*nf + f])
hipError_t box_coeffs_compact(const GridGeom& g, const BoxTables& bt, hipStream_t st, const c64* vals, int nf,
                              double dv, double* part, c64* ca, const double* sn, c64* w) {
  if (bt.natom == 0 || nf == 0) return hipSuccess;
  prof_begin(st, KC_BOX_PROJ);
  if (bt.msp == 8)
    launch_proj<double, 8>(g, bt, st, nullptr, nf, bt.p1, vals, part);
  else
    launch_proj<double, 16>(g, bt, st, nullptr, nf, bt.p1, vals, part);
  prof_end(st, KC_BOX_PROJ);
  hipError_t e = hipGetLastError();
  if (e) return e;
  prof_begin(st, KC_BOX_ADD);
  const dim3 fg(bt.natom, nf);
  if (bt.msp == 8)
    hipLaunchKernelGGL(k_proj_finish<8>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, ca, 1, sn,
                       nullptr, nf, 0, w, nullptr, nullptr, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL(k_proj_finish<16>, fg, dim3(256), 0, st, bt.atoms, bt.achunk, part, nf, bt.ctot, dv, ca, 1, sn,
                       nullptr, nf, 0, w, nullptr, nullptr, nullptr, nullptr, nullptr);
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
template hipError_t box_coeffs_dual<double>(const GridGeom&, const BoxTables&, hipStream_t, const c64*, int, double,
                                            double*, c64*, const double*, c64*, c64*);
template hipError_t box_coeffs_dual<float>(const GridGeom&, const BoxTables&, hipStream_t, const c32*, int, double,
                                           double*, c64*, const double*, c64*, c64*);
template hipError_t union_rk4_apply<double>(const GridGeom&, const BoxTables&, hipStream_t, int, const c64*, c64*,
                                            c64*, c64*, double, double);
template hipError_t union_rk4_apply<float>(const GridGeom&, const BoxTables&, hipStream_t, int, const c64*, c64*,
                                           c32*, c32*, double, double);
template hipError_t union_rk4_both<double>(const GridGeom&, const BoxTables&, hipStream_t, int, const c64*,
                                            const c64*, c64*, c64*, c64*, c64*, double, double);
template hipError_t union_rk4_both<float>(const GridGeom&, const BoxTables&, hipStream_t, int, const c64*,
                                           const c64*, c64*, c64*, c32*, c32*, double, double);
template hipError_t union_corr_apply<double>(const GridGeom&, const BoxTables&, hipStream_t, int, int, const c64*,
                                             c64*, c64*);
template hipError_t union_corr_apply<float>(const GridGeom&, const BoxTables&, hipStream_t, int, int, const c64*,
                                            c64*, c32*);

}  // namespace opaw
