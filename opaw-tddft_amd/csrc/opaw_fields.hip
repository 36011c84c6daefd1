// Grid-field and small-matrix kernels of the OPAW TDDFT hot path (gfx950).
//
// All reductions are deterministic: a fixed number of chunks per row, each
// reduced by one workgroup (wave64 shuffles + LDS), then summed in chunk
// order by a finishing kernel.  Densities, rho_ij, D_ij, norms and dipoles
// accumulate in FP64 for both precisions.
#include "opaw_fields.h"

#include <type_traits>

namespace opaw {

constexpr int kT = 256;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// block-wide sum of NV values per thread (result valid in thread 0)
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[kT / 64][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double s = wsum(v[i]);
    if (lane == 0) red[w][i] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int k = 0; k < kT / 64; ++k) s += red[k][i];
      v[i] = s;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void chunk_range(long long n, int nchunk, int c, long long& b, long long& e) {
  const long long per = (n + nchunk - 1) / nchunk;
  b = (long long)c * per;
  e = b + per < n ? b + per : n;
}

// ---------------------------------------------------------------------------
// Streaming field kernels run on 2D grids: blockIdx.y is the field (or
// system), blockIdx.x covers the grid points, kV points per thread.
constexpr int kV = 4;
// orbitals whose loads a thread issues together in the per-point orbital sums
constexpr int kOrbBatch = 8;

// RK4 stage update (src/rk4.f90:35-56 and :22), k = -i hs with hs the
// complete shs output (its outer S^-1/2 add-back already in h):
//   stage 1: acc = k;        y = p + c k
//   stage 2,3: acc += 2k;    y = p + c k
// Loads are unconditional (indices clamped) so a thread's kV loads are in
// flight together; only the stores are guarded.
template <typename R, bool FIRST>
__global__ __launch_bounds__(kT) void k_rk4_update(int N, const cx<R>* __restrict__ h, const cx<R>* __restrict__ p,
                                                    cx<R>* __restrict__ acc, cx<R>* __restrict__ y, double c) {
  const long long off = (long long)blockIdx.y * N;
  const int base = blockIdx.x * (kT * kV) + threadIdx.x;
  cx<R> hh[kV], pp[kV], aa[kV];
#pragma unroll
  for (int j = 0; j < kV; ++j) {
    const long long i = off + min(base + j * kT, N - 1);
    hh[j] = h[i];
    pp[j] = p[i];
    if constexpr (!FIRST) aa[j] = acc[i];
  }
#pragma unroll
  for (int j = 0; j < kV; ++j) {
    const int r = base + j * kT;
    if (r >= N) continue;
    const double kr = (double)hh[j].y, ki = -(double)hh[j].x;
    if constexpr (FIRST)
      acc[off + r] = mk<R>((R)kr, (R)ki);
    else
      acc[off + r] = mk<R>((R)((double)aa[j].x + 2.0 * kr), (R)((double)aa[j].y + 2.0 * ki));
    y[off + r] = mk<R>((R)((double)pp[j].x + c * kr), (R)((double)pp[j].y + c * ki));
  }
}

// RK4 stage 4 (src/rk4.f90:22): p = p + dt/6 (acc + k), k = -i h, fused with
// the per-field partial sums of |p|^2 for the norm restore (norm1,
// src/rk4.f90:92-102).  Grid (nchunk, nf); the chunking of k_norm_partial.
template <typename R>
__global__ __launch_bounds__(kT) void k_rk4_final(long long N, int nchunk, const cx<R>* __restrict__ h,
                                                   cx<R>* __restrict__ p, const cx<R>* __restrict__ acc,
                                                   double dt6, double* __restrict__ part) {
  const int ch = blockIdx.x, f = blockIdx.y;
  long long lo, hi;
  chunk_range(N, nchunk, ch, lo, hi);
  const long long off = (long long)f * N;
  double v[1] = {0.0};
  for (long long b0 = lo; b0 < hi; b0 += (long long)kT * kV) {
    cx<R> hh[kV], pp[kV], aa[kV];
#pragma unroll
    for (int j = 0; j < kV; ++j) {
      const long long r = min(b0 + threadIdx.x + (long long)j * kT, hi - 1);
      hh[j] = h[off + r];
      pp[j] = p[off + r];
      aa[j] = acc[off + r];
    }
#pragma unroll
    for (int j = 0; j < kV; ++j) {
      const long long r = b0 + threadIdx.x + (long long)j * kT;
      if (r >= hi) continue;
      const double kr = (double)hh[j].y, ki = -(double)hh[j].x;
      const cx<R> q = mk<R>((R)((double)pp[j].x + dt6 * ((double)aa[j].x + kr)),
                            (R)((double)pp[j].y + dt6 * ((double)aa[j].y + ki)));
      p[off + r] = q;
      v[0] += (double)q.x * (double)q.x + (double)q.y * (double)q.y;
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[(long long)f * nchunk + ch] = v[0];
}

// per-field partial sums of |psi|^2
template <typename R>
__global__ __launch_bounds__(kT) void k_norm_partial(const cx<R>* __restrict__ u, long long N, int nchunk,
                                                      double* __restrict__ part) {
  const int c = blockIdx.x, f = blockIdx.y;
  long long b, e;
  chunk_range(N, nchunk, c, b, e);
  const cx<R>* uf = u + (long long)f * N;
  double v[1] = {0.0};
  for (long long i = b + threadIdx.x; i < e; i += kT) {
    const double x = uf[i].x, y = uf[i].y;
    v[0] += x * x + y * y;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[(long long)f * nchunk + c] = v[0];
}

// out[r] = scale * sum_c part[r*nchunk + c]  (then sqrt if want_sqrt); one
// workgroup per row, fixed reduction tree
__global__ __launch_bounds__(kT) void k_finish_rows(const double* __restrict__ part, int /*rows: one per workgroup*/, int nchunk,
                                                     double scale, int want_sqrt, double* __restrict__ out) {
  const int r = blockIdx.x;
  double v[1] = {0.0};
  for (int c = threadIdx.x; c < nchunk; c += kT) v[0] += part[(long long)r * nchunk + c];
  block_sum<1>(v);
  if (threadIdx.x == 0) {
    const double s = v[0] * scale;
    out[r] = want_sqrt ? sqrt(s) : s;
  }
}

// psi_f *= num[f] / den[f]   (src/rk4.f90:92-102, norm1)
template <typename R>
__global__ __launch_bounds__(kT) void k_scale_fields(cx<R>* __restrict__ u, int N, const double* __restrict__ num,
                                                      const double* __restrict__ den) {
  const int f = blockIdx.y;
  const double s = num[f] / den[f];
  cx<R>* uf = u + (long long)f * N;
  const int base = blockIdx.x * (kT * kV) + threadIdx.x;
  cx<R> v[kV];
#pragma unroll
  for (int j = 0; j < kV; ++j) v[j] = uf[min(base + j * kT, N - 1)];
#pragma unroll
  for (int j = 0; j < kV; ++j) {
    const int r = base + j * kT;
    if (r < N) uf[r] = mk<R>((R)((double)v[j].x * s), (R)((double)v[j].y * s));
  }
}

// The norm restore fused with the dipole of the restored orbitals
// (src/rk4.f90:92-102 then src/tddft_plot.f90:40-55): psi_f *= num[f]/den[f]
// for the nb orbitals of system s, and the partial sums of r * 2 sum|psi|^2.
// Grid (nchunk, nsys), the chunking of k_dipole_partial.
template <typename R>
__global__ __launch_bounds__(kT) void k_scale_dipole(cx<R>* __restrict__ u, int nx, int ny, int nz, int nb,
                                                      int nchunk, double dx, double dy, double dz,
                                                      const double* __restrict__ num,
                                                      const double* __restrict__ den,
                                                      double* __restrict__ part) {
  const int ch = blockIdx.x, s = blockIdx.y;
  const long long N = (long long)nx * ny * nz;
  long long lo, hi;
  chunk_range(N, nchunk, ch, lo, hi);
  extern __shared__ double fac[];
  for (int o = threadIdx.x; o < nb; o += kT) fac[o] = num[s * nb + o] / den[s * nb + o];
  __syncthreads();
  double v[3] = {0.0, 0.0, 0.0};
  cx<R>* us = u + (long long)s * nb * N;
  for (long long i = lo + threadIdx.x; i < hi; i += kT) {
    double rho = 0.0;
    for (int o0 = 0; o0 < nb; o0 += kOrbBatch) {
      cx<R> w[kOrbBatch];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) w[q] = us[(long long)min(o0 + q, nb - 1) * N + i];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) {
        if (o0 + q < nb) {
          const double s = fac[o0 + q];
          const cx<R> v = mk<R>((R)((double)w[q].x * s), (R)((double)w[q].y * s));
          us[(long long)(o0 + q) * N + i] = v;
          rho += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        }
      }
    }
    rho *= 2.0;
    const int ii = (int)i;
    const int ix = ii % nx, t = ii / nx;
    const int iy = t % ny, iz = t / ny;
    v[0] += rho * ((ix - nx / 2) * dx);
    v[1] += rho * ((iy - ny / 2) * dy);
    v[2] += rho * ((iz - nz / 2) * dz);
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[((long long)s * 3 + k) * nchunk + ch] = v[k];
}

// dens[s][r] = 2 sum_{o<nb} |psi~_{s,o}(r)|^2 (src/tddft_dens.f90:65-76) with
// psi~ = psi + cpre (compact S^-1/2 correction) when cpre is set (on the fly,
// src/ncpaw_make_hamiltonian.f90:228-235)
template <typename R>
__global__ __launch_bounds__(kT) void k_density(const cx<R>* __restrict__ u, int N, int nb,
                                                 double* __restrict__ dens, const int* __restrict__ umap,
                                                 const c64* __restrict__ cpre, int U) {
  const int s = blockIdx.y;
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= N) return;
  const cx<R>* us = u + (long long)s * nb * N + r;
  const int k = cpre ? umap[r] : -1;
  double acc = 0.0;
  for (int o0 = 0; o0 < nb; o0 += kOrbBatch) {
    cx<R> v[kOrbBatch];
#pragma unroll
    for (int q = 0; q < kOrbBatch; ++q) v[q] = us[(long long)min(o0 + q, nb - 1) * N];
    double vr[kOrbBatch], vi[kOrbBatch];
#pragma unroll
    for (int q = 0; q < kOrbBatch; ++q) {
      vr[q] = (double)v[q].x;
      vi[q] = (double)v[q].y;
    }
    if (k >= 0) {
      c64 c[kOrbBatch];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) c[q] = cpre[(long long)k * U + s * nb + min(o0 + q, nb - 1)];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) {
        vr[q] += c[q].x;
        vi[q] += c[q].y;
      }
    }
#pragma unroll
    for (int q = 0; q < kOrbBatch; ++q)
      if (o0 + q < nb) acc += vr[q] * vr[q] + vi[q] * vi[q];
  }
  dens[(long long)s * N + r] = 2.0 * acc;
}

// compensation charge (src/nhat.f90:13-28):
//   nhat(r) = sum_a sum_L g_a(r, L) wl_a(L),  wl_a(L) = sum_ij rho_ij qijlm(L,i,j)
// wl per (system, atom, L) first, then one thread per grid point gathers the
// sphere entries of that point in atom order (no atomics).
__global__ __launch_bounds__(64) void k_nhat_w(const AtomDesc* __restrict__ atoms, int natom,
                                                const double* __restrict__ rhoij, int mtot,
                                                const double* __restrict__ qtab, int nl2, int /*nsys: one per blockIdx.y*/,
                                                double* __restrict__ wl) {
  const int a = blockIdx.x, s = blockIdx.y;
  const AtomDesc A = atoms[a];
  const int m2 = A.ms * A.ms;
  __shared__ double rho[kMaxMs * kMaxMs];
  const double* rg = rhoij + (long long)s * mtot + A.moff;
  for (int q = threadIdx.x; q < m2; q += 64) rho[q] = rg[q];
  __syncthreads();
  const double* qq = qtab + A.toff_q;
  for (int L = threadIdx.x; L < nl2; L += 64) {
    double w = 0.0;
#pragma unroll 8
    for (int q = 0; q < m2; ++q) w += rho[q] * qq[L + nl2 * q];
    wl[((long long)s * natom + a) * nl2 + L] = w;
  }
}

__global__ __launch_bounds__(kT) void k_nhat_grid(SphereTabs S, const double* __restrict__ sph_g,
                                                   const double* __restrict__ wl, int natom, int nl2, int N,
                                                   double* __restrict__ nhat) {
  const int s = blockIdx.y;
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= N) return;
  const int k = S.smap[r];
  double v = 0.0;
  if (k >= 0) {
    for (int e = S.sptr[k]; e < S.sptr[k + 1]; ++e) {
      const int2 en = S.se[e];                  // (atom, g offset of L = 0)
      const int a = en.x;
      const int stride = S.sn[a];
      const double* w = wl + ((long long)s * natom + a) * nl2;
      for (int L = 0; L < nl2; ++L) v += sph_g[en.y + (long long)L * stride] * w[L];
    }
  }
  nhat[(long long)s * N + r] = v;
}

// partial sums of two real fields per system (nhat and dens) for the rnel
// renormalisation (src/nhat.f90:31-41)
__global__ __launch_bounds__(kT) void k_sum2_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                      long long N, int nchunk, double* __restrict__ part) {
  const int c = blockIdx.x, s = blockIdx.y;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  double v[2] = {0.0, 0.0};
  for (long long i = lo + threadIdx.x; i < hi; i += kT) {
    v[0] += a[(long long)s * N + i];
    v[1] += b[(long long)s * N + i];
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    part[((long long)s * 2 + 0) * nchunk + c] = v[0];
    part[((long long)s * 2 + 1) * nchunk + c] = v[1];
  }
}

// fac[s] = rnel / (dv*sum nhat + dv*sum dens); one workgroup per system
__global__ __launch_bounds__(kT) void k_renorm_factor(const double* __restrict__ part, int /*nsys: one per workgroup*/, int nchunk,
                                                       double dv, double rnel, double* __restrict__ fac) {
  const int s = blockIdx.x;
  double v[2] = {0.0, 0.0};
  for (int c = threadIdx.x; c < nchunk; c += kT) {
    v[0] += part[((long long)s * 2 + 0) * nchunk + c];
    v[1] += part[((long long)s * 2 + 1) * nchunk + c];
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) fac[s] = rnel / (v[0] * dv + v[1] * dv);
}

// dens, nhat *= fac (when fac != null); rho pair for the Hartree solve:
//   pair[q] = (dens[2q]+nhat[2q]) + i (dens[2q+1]+nhat[2q+1])
__global__ __launch_bounds__(kT) void k_scale_pack(double* __restrict__ dens, double* __restrict__ nhat,
                                                    const double* __restrict__ fac, int N, int nsys,
                                                    c64* __restrict__ pair) {
  const int q = blockIdx.y;
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= N) return;
  double v[2] = {0.0, 0.0};
  for (int k = 0; k < 2; ++k) {
    const int s = 2 * q + k;
    if (s < nsys) {
      const long long j = (long long)s * N + r;
      double d = dens[j], h = nhat[j];
      if (fac) {
        d *= fac[s];
        h *= fac[s];
        dens[j] = d;
        nhat[j] = h;
      }
      v[k] = d + h;
    }
  }
  pair[(long long)q * N + r] = mk(v[0], v[1]);
}

// vxc[s] = LDA(dens[s] + ncoret); vh[s] from the packed Hartree output
// (minus its mean when periodic); vks[s] = vh + vxc + vloc
// (src/tddft_pot.f90:130-149)
__global__ __launch_bounds__(kT) void k_vks(double* __restrict__ dens, const double* __restrict__ ncoret,
                                             const c64* __restrict__ vhpair, const double* __restrict__ vhmean,
                                             const double* __restrict__ vloc, int N,
                                             double* __restrict__ vxc, double* __restrict__ vh,
                                             double* __restrict__ vks, float* __restrict__ vksf,
                                             const double* __restrict__ fac) {
  const int s = blockIdx.y;
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= N) return;
  const long long i = (long long)s * N + r;
  double dn = dens[i];
  if (fac) {
    dn *= fac[s];
    dens[i] = dn;
  }
  const double x = lda_vxc(dn + ncoret[r]);
  const c64 p = vhpair[(long long)(s / 2) * N + r];
  double h = (s & 1) ? p.y : p.x;
  if (vhmean) h -= vhmean[s];
  const double v = h + x + vloc[r];
  vxc[i] = x;
  vh[i] = h;
  vks[i] = v;
  if (vksf) vksf[i] = (float)v;
}

// partial sums of the packed Hartree output (re -> system 2q, im -> 2q+1)
__global__ __launch_bounds__(kT) void k_pair_sum_partial(const c64* __restrict__ pair, long long N, int nchunk,
                                                          double* __restrict__ part) {
  const int c = blockIdx.x, q = blockIdx.y;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  double v[2] = {0.0, 0.0};
  for (long long i = lo + threadIdx.x; i < hi; i += kT) {
    const c64 p = pair[(long long)q * N + i];
    v[0] += p.x;
    v[1] += p.y;
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    part[((long long)q * 2 + 0) * nchunk + c] = v[0];
    part[((long long)q * 2 + 1) * nchunk + c] = v[1];
  }
}

// ---------------------------------------------------------------------------
// rho_ij = 2 sum_o Re(conj(c_i) c_j)  (src/tddft_rhoij.f90:88-114)
__global__ void k_rhoij(const AtomDesc* __restrict__ atoms, const c64* __restrict__ ca, int ctot, int nb,
                        int mtot, double* __restrict__ rhoij) {
  const int a = blockIdx.x, s = blockIdx.y;
  const AtomDesc A = atoms[a];
  for (int q = threadIdx.x; q < A.ms * A.ms; q += blockDim.x) {
    const int i = q % A.ms, j = q / A.ms;
    double acc = 0.0;
    for (int o = 0; o < nb; ++o) {
      const c64* c = ca + (long long)(s * nb + o) * ctot + A.coff;
      acc += 2.0 * (c[i].x * c[j].x + c[i].y * c[j].y);
    }
    rhoij[(long long)s * mtot + A.moff + q] = acc;
  }
}

// D_ij (src/get_dij.f90:1-156, pawdij src/libpaw/4_m_pawdij.F90:168-1198):
// dij0 + dijhartree (eijkl . rho_sel, pawdijhartree) + dijxc (from the
// on-site kernel, when the type has libpaw tables) + dijhat, the grid term
// sum_L q_ij^L dv sum_sph (vks - vxc) g_L (pawdijhat,
// src/libpaw/4_m_pawdij.F90:2227-2514) over the atom's D^hat sphere, then
// transfer_dij's packed upper-triangle symmetric fill and +-param_dij_max
// clip.  One workgroup per (atom, system); rho_sel, q_ij^L and the sphere
// moments are staged in LDS so the contractions run from LDS.
__global__ __launch_bounds__(kT) void k_dij(const AtomDesc* __restrict__ atoms, const double* __restrict__ rhoij,
                                             int mtot, const double* __restrict__ qtab,
                                             const double* __restrict__ d0tab, const double* __restrict__ etab,
                                             const int* __restrict__ sph_idx,
                                             const double* __restrict__ sph_g, const double* __restrict__ vks,
                                             const double* __restrict__ vxc, int subtract_vxc, long long N,
                                             double dv, const double* __restrict__ dxc, double* __restrict__ dij,
                                             DijWeights wt) {
  const int a = blockIdx.x, s = blockIdx.y;
  const AtomDesc A = atoms[a];
  const int ms = A.ms, m2 = ms * ms, nl2 = A.dh_nl;
  __shared__ double red[kT / 64][kMaxNl2];
  __shared__ double prod[kMaxNl2];
  __shared__ double rho[kMaxMs * kMaxMs];
  __shared__ double qq[kMaxNl2 * kMaxMs * kMaxMs];
  __shared__ double M[kMaxMs * kMaxMs];
  const double* rg = rhoij + (long long)s * mtot + A.moff;
  for (int q = threadIdx.x; q < m2; q += kT) {
    const double rv = rg[q];
    rho[q] = fabs(rv) > kRhoijSelect ? rv : 0.0;   // src/get_dij.f90:80-88 selection
  }
  // loads are issued in batches with clamped indices (no load behind a
  // branch), so a thread's loads are in flight together
  const int nq = nl2 * m2;
  for (int q0 = 0; q0 < nq; q0 += kT * 4) {
    double t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) t[u] = qtab[A.toff_qh + min(q0 + u * kT + (int)threadIdx.x, nq - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + u * kT + (int)threadIdx.x;
      if (q < nq) qq[q] = t[u];
    }
  }
  double acc[kMaxNl2];
#pragma unroll
  for (int L = 0; L < kMaxNl2; ++L) acc[L] = 0.0;
  for (int p = threadIdx.x; p < A.dh_n; p += kT) {
    const long long r = sph_idx[A.dh_off + p];
    double g[kMaxNl2];
#pragma unroll
    for (int L = 0; L < kMaxNl2; ++L)
      g[L] = sph_g[(long long)A.dh_goff + (long long)min(L, nl2 - 1) * A.dh_n + p];
    const double vk = vks[(long long)s * N + r];
    const double vx = vxc[(long long)s * N + r];
    const double pot = subtract_vxc ? vk - vx : vk;
#pragma unroll
    for (int L = 0; L < kMaxNl2; ++L) acc[L] += pot * g[L];
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int L = 0; L < kMaxNl2; ++L) {
    if (L < nl2) {
      const double t = wsum(acc[L]);
      if (lane == 0) red[w][L] = t;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nl2) {
    double t = 0.0;
    for (int k = 0; k < kT / 64; ++k) t += red[k][threadIdx.x];
    prod[threadIdx.x] = t * dv;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < m2; q += kT) {
    const int i = q % ms, j = q / ms;
    double d = d0tab[A.toff_d0 + q];
    const double* e = etab + A.toff_e + (long long)q;   // E(i,j,k,l) at i + ms(j + ms(k + ms l))
    for (int kl0 = 0; kl0 < m2; kl0 += 16) {
      double ev[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) ev[u] = e[(long long)min(kl0 + u, m2 - 1) * m2];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (kl0 + u < m2) d += ev[u] * rho[kl0 + u];
    }
    for (int L = 0; L < nl2; ++L) d += qq[L + nl2 * (i + ms * j)] * prod[L];
    if (A.ot >= 0) d += dxc[(long long)s * mtot + A.moff + q];
    M[q] = d;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < m2; q += kT) {
    const int i = q % ms, j = q / ms;
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    double d = M[lo + ms * hi];
    if (fabs(d) > kDijMax) d = d / fabs(d) * kDijMax;
    dij[(long long)s * mtot + A.moff + q] = d;
    rho[q] = d;   // the final D of the atom, for the weights below
  }
  if (!wt.w) return;
  __syncthreads();
  // nonlocal weights of the system's fields: w(coff+i, f) = sum_j D_ij
  // cB(f, coff+j), in the order of the separate weights kernel (bit-identical)
  for (int q = threadIdx.x; q < wt.nb * ms; q += kT) {
    const int o = q % wt.nb, i = q / wt.nb;
    const int f = wt.f0 + s * wt.nb + o;
    const c64* row = wt.cb + (long long)f * wt.ctot + A.coff;
    double xr = 0.0, xi = 0.0;
    for (int j = 0; j < ms; ++j) {
      const double d = rho[i + ms * j];
      xr += d * row[j].x;
      xi += d * row[j].y;
    }
    wt.w[(long long)(A.coff + i) * wt.nf + f] = mk(xr, xi);
  }
}

// ---------------------------------------------------------------------------
// dipole partial sums (src/tddft_plot.f90:40-55, h_type=1 density 2 sum|phi|^2)
template <typename R>
__global__ __launch_bounds__(kT) void k_dipole_partial(const cx<R>* __restrict__ u, int nx, int ny, int nz,
                                                        int nb, int nchunk, double dx, double dy, double dz,
                                                        double* __restrict__ part) {
  const int c = blockIdx.x, s = blockIdx.y;
  const long long N = (long long)nx * ny * nz;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  double v[3] = {0.0, 0.0, 0.0};
  const cx<R>* us = u + (long long)s * nb * N;
  for (long long i = lo + threadIdx.x; i < hi; i += kT) {
    double rho = 0.0;
    for (int o0 = 0; o0 < nb; o0 += kOrbBatch) {
      cx<R> w[kOrbBatch];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) w[q] = us[(long long)min(o0 + q, nb - 1) * N + i];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q)
        if (o0 + q < nb) rho += (double)w[q].x * (double)w[q].x + (double)w[q].y * (double)w[q].y;
    }
    rho *= 2.0;
    const int ii = (int)i;
    const int ix = ii % nx, t = ii / nx;
    const int iy = t % ny, iz = t / ny;
    v[0] += rho * ((ix - nx / 2) * dx);
    v[1] += rho * ((iy - ny / 2) * dy);
    v[2] += rho * ((iz - nz / 2) * dz);
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[((long long)s * 3 + k) * nchunk + c] = v[k];
}

// phi_pert = phi * exp(-i sm r_pol)  (src/tddft_excite.f90:8-15)
template <typename R>
__global__ __launch_bounds__(kT) void k_kick(const cx<R>* __restrict__ src, cx<R>* __restrict__ dst, int nx,
                                              int ny, int nz, int nb, int ipol, double d, double sm) {
  const long long N = (long long)nx * ny * nz;
  const long long n = N * nb;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) {
    const long long r = i % N;
    int k;
    int nn;
    if (ipol == 0) {
      k = (int)(r % nx);
      nn = nx;
    } else if (ipol == 1) {
      k = (int)((r / nx) % ny);
      nn = ny;
    } else {
      k = (int)(r / ((long long)nx * ny));
      nn = nz;
    }
    const double th = sm * ((k - nn / 2) * d);
    double sn, cs;
    sincos(th, &sn, &cs);
    const cx<R> v = src[i];
    const double vr = v.x, vi = v.y;
    // (vr + i vi)(cs - i sn)
    dst[i] = mk<R>((R)(vr * cs + vi * sn), (R)(vi * cs - vr * sn));
  }
}

// ---------------------------------------------------------------------------
// Gram partials: G[c][i][j] = sum_{r in chunk c} conj(a_i(r)) b_j(r)
// 16x16 tile of (i, j) per workgroup; points staged through LDS.
template <typename R>
__global__ __launch_bounds__(256) void k_gram_partial(const cx<R>* __restrict__ a, const cx<R>* __restrict__ b,
                                                       long long N, int nb, int nchunk, c64* __restrict__ part) {
  const int ti = blockIdx.x, tj = blockIdx.y, c = blockIdx.z;
  const int li = threadIdx.x % 16, lj = threadIdx.x / 16;
  const int i = ti * 16 + li, j = tj * 16 + lj;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  __shared__ c64 sa[16][17], sb[16][17];
  double gr = 0.0, gi = 0.0;
  for (long long r0 = lo; r0 < hi; r0 += 16) {
    // load 16 points x 16 orbitals of a and b (orbital = lj row, point = li col)
    {
      const int o = ti * 16 + lj;
      const long long r = r0 + li;
      c64 v = mk(0.0, 0.0);
      if (o < nb && r < hi) {
        const cx<R> w = a[(long long)o * N + r];
        v = mk((double)w.x, (double)w.y);
      }
      sa[lj][li] = v;
      const int o2 = tj * 16 + lj;
      c64 v2 = mk(0.0, 0.0);
      if (o2 < nb && r < hi) {
        const cx<R> w = b[(long long)o2 * N + r];
        v2 = mk((double)w.x, (double)w.y);
      }
      sb[lj][li] = v2;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const c64 x = sa[li][k], y = sb[lj][k];
      gr += x.x * y.x + x.y * y.y;
      gi += x.x * y.y - x.y * y.x;
    }
    __syncthreads();
  }
  if (i < nb && j < nb) part[((long long)c * nb + i) * nb + j] = mk(gr, gi);
}

// Gram partials with the whole m x m block in one workgroup per point chunk
// (m <= 16 RB), so every element of a and b is read once:
// G[c][i][j] = sum_{r in chunk c} conj(a_i(r)) b_j(r).  Thread (tx, ty) of a
// 16 x 16 grid owns i = tx + 16 ri, j = ty + 16 rj (ri, rj < RB).  Points go
// through LDS 16 at a time; the next 16 are loaded into registers while the
// current ones are summed.
template <typename R, int RB>
__global__ __launch_bounds__(256) void k_gram_block(const cx<R>* __restrict__ a, const cx<R>* __restrict__ b,
                                                     long long N, int m, int nchunk, c64* __restrict__ part) {
  constexpr int P = 16;
  constexpr int M = 16 * RB;
  __shared__ c64 sa[M][P + 1], sb[M][P + 1];
  const int c = blockIdx.x;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  // loader slots: RB elements of a and of b per thread, element (o, p) with
  // o = slot / P, p = slot % P (16 consecutive points of one orbital per
  // 16 lanes); orbital and point are clamped, out-of-range values zeroed
  cx<R> va[RB], vb[RB];
  auto load = [&](long long r0) {
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      const int slot = (int)threadIdx.x + 256 * q;
      const int o = slot / P, p = slot % P;
      const long long r = min(r0 + p, hi - 1);
      const int oc = min(o, m - 1);
      va[q] = a[(long long)oc * N + r];
      vb[q] = b[(long long)oc * N + r];
    }
  };
  double gr[RB][RB], gi[RB][RB];
#pragma unroll
  for (int x = 0; x < RB; ++x)
#pragma unroll
    for (int y = 0; y < RB; ++y) gr[x][y] = gi[x][y] = 0.0;
  if (lo < hi) load(lo);
  for (long long r0 = lo; r0 < hi; r0 += P) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      const int slot = (int)threadIdx.x + 256 * q;
      const int o = slot / P, p = slot % P;
      const bool ok = o < m && r0 + p < hi;
      sa[o][p] = ok ? mk((double)va[q].x, (double)va[q].y) : mk(0.0, 0.0);
      sb[o][p] = ok ? mk((double)vb[q].x, (double)vb[q].y) : mk(0.0, 0.0);
    }
    __syncthreads();
    if (r0 + P < hi) load(r0 + P);
#pragma unroll
    for (int k = 0; k < P; ++k) {
      c64 x[RB], y[RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        x[q] = sa[tx + 16 * q][k];
        y[q] = sb[ty + 16 * q][k];
      }
#pragma unroll
      for (int u = 0; u < RB; ++u)
#pragma unroll
        for (int v = 0; v < RB; ++v) {
          gr[u][v] += x[u].x * y[v].x + x[u].y * y[v].y;
          gi[u][v] += x[u].x * y[v].y - x[u].y * y[v].x;
        }
    }
  }
#pragma unroll
  for (int u = 0; u < RB; ++u)
#pragma unroll
    for (int v = 0; v < RB; ++v) {
      const int i = tx + 16 * u, j = ty + 16 * v;
      if (i < m && j < m) part[((long long)c * m + i) * m + j] = mk(gr[u][v], gi[u][v]);
    }
}

// FP64 Gram of up to 16 RB fields on the matrix cores, same partials as
// k_gram_block.  With G = A^H B, Re G = Ar^T Br + Ai^T Bi and
// Im G = Ar^T Bi - Ai^T Br: four v_mfma_f64_16x16x4f64 per 16 x 16 tile and
// 4 points.  Wave w owns rows 16w..16w+15 of the block (RB tiles); 16
// points of every field go through LDS per step (16 consecutive points of
// one field per 16 lanes), the next 16 are loaded into registers meanwhile.
// Operand layout of the 16x16x4 f64 product: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; lane l receives
// D[l / 16 + 4 i][l % 16] in element i.
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int RB>
__global__ __launch_bounds__(64 * RB) void k_gram_mfma(const c64* __restrict__ a, const c64* __restrict__ b,
                                                       long long N, int m, int nchunk, c64* __restrict__ part) {
  constexpr int P = 16;
  constexpr int M = 16 * RB;
  constexpr int NT = 64 * RB;
  constexpr int SL = M * P / NT;   // loader slots per thread: 4
  __shared__ c64 sa[M][P + 1], sb[M][P + 1];
  const int c = blockIdx.x;
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int lr = lane % 16, lk = lane / 16;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  c64 va[SL], vb[SL];
  auto load = [&](long long r0) {
#pragma unroll
    for (int q = 0; q < SL; ++q) {
      const int slot = (int)threadIdx.x + NT * q;
      const int o = slot / P, p = slot % P;
      const long long r = min(r0 + p, hi - 1);
      const int oc = min(o, m - 1);
      va[q] = a[(long long)oc * N + r];
      vb[q] = b[(long long)oc * N + r];
    }
  };
  f64x4 gr[RB], gi[RB];
#pragma unroll
  for (int v = 0; v < RB; ++v)
    for (int i = 0; i < 4; ++i) gr[v][i] = gi[v][i] = 0.0;
  if (lo < hi) load(lo);
  for (long long r0 = lo; r0 < hi; r0 += P) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SL; ++q) {
      const int slot = (int)threadIdx.x + NT * q;
      const int o = slot / P, p = slot % P;
      const bool ok = o < m && r0 + p < hi;
      sa[o][p] = ok ? va[q] : mk(0.0, 0.0);
      sb[o][p] = ok ? vb[q] : mk(0.0, 0.0);
    }
    __syncthreads();
    if (r0 + P < hi) load(r0 + P);
#pragma unroll
    for (int kk = 0; kk < P / 4; ++kk) {
      const c64 x = sa[16 * w + lr][4 * kk + lk];
#pragma unroll
      for (int v = 0; v < RB; ++v) {
        const c64 y = sb[16 * v + lr][4 * kk + lk];
        gr[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, gr[v], 0, 0, 0);
        gr[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.y, gr[v], 0, 0, 0);
        gi[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, gi[v], 0, 0, 0);
        gi[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, y.x, gi[v], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < RB; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = 16 * w + lk + 4 * e, j = 16 * v + lr;
      if (i < m && j < m) part[((long long)c * m + i) * m + j] = mk(gr[v][e], gi[v][e]);
    }
}

// More than 64 FP64 fields: the same product on 64 x 64 tiles of G, one
// workgroup (four waves, each 16 rows of the tile) per (tile pair, point
// chunk).  Tile pairs vary fastest in the workgroup index, so the tiles of
// one chunk run together and share the chunk's operands through the caches.
// sym (a == b): only the tiles I <= J; an off-diagonal tile also writes its
// conjugate transpose, G_ji = conj(G_ij), which is the same sum of the same
// products.
__global__ __launch_bounds__(256) void k_gram_mfma_tiles(const c64* __restrict__ a, const c64* __restrict__ b,
                                                          long long N, int m, int T, int ntp, int sym, int nchunk,
                                                          c64* __restrict__ part) {
  constexpr int P = 16, M = 64, NT = 256;
  constexpr int SL = M * P / NT;   // loader slots per thread: 4
  __shared__ c64 sa[M][P + 1], sb[M][P + 1];
  const int tp = blockIdx.x % ntp;
  const int c = blockIdx.x / ntp;
  int I, J;
  if (sym) {
    int k = tp;
    I = 0;
    while (k >= T - I) {
      k -= T - I;
      ++I;
    }
    J = I + k;
  } else {
    I = tp / T;
    J = tp % T;
  }
  const int ia0 = M * I, jb0 = M * J;
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int lr = lane % 16, lk = lane / 16;
  long long lo, hi;
  chunk_range(N, nchunk, c, lo, hi);
  c64 va[SL], vb[SL];
  auto load = [&](long long r0) {
#pragma unroll
    for (int q = 0; q < SL; ++q) {
      const int slot = (int)threadIdx.x + NT * q;
      const int o = slot / P, p = slot % P;
      const long long r = min(r0 + p, hi - 1);
      va[q] = a[(long long)min(ia0 + o, m - 1) * N + r];
      vb[q] = b[(long long)min(jb0 + o, m - 1) * N + r];
    }
  };
  f64x4 gr[4], gi[4];
#pragma unroll
  for (int v = 0; v < 4; ++v)
    for (int i = 0; i < 4; ++i) gr[v][i] = gi[v][i] = 0.0;
  if (lo < hi) load(lo);
  for (long long r0 = lo; r0 < hi; r0 += P) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SL; ++q) {
      const int slot = (int)threadIdx.x + NT * q;
      const int o = slot / P, p = slot % P;
      const bool inr = r0 + p < hi;
      sa[o][p] = (inr && ia0 + o < m) ? va[q] : mk(0.0, 0.0);
      sb[o][p] = (inr && jb0 + o < m) ? vb[q] : mk(0.0, 0.0);
    }
    __syncthreads();
    if (r0 + P < hi) load(r0 + P);
#pragma unroll
    for (int kk = 0; kk < P / 4; ++kk) {
      const c64 x = sa[16 * w + lr][4 * kk + lk];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const c64 y = sb[16 * v + lr][4 * kk + lk];
        gr[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.x, gr[v], 0, 0, 0);
        gr[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, y.y, gr[v], 0, 0, 0);
        gi[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, y.y, gi[v], 0, 0, 0);
        gi[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.y, y.x, gi[v], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = ia0 + 16 * w + lk + 4 * e, j = jb0 + 16 * v + lr;
      if (i < m && j < m) {
        part[((long long)c * m + i) * m + j] = mk(gr[v][e], gi[v][e]);
        if (sym && I != J) part[((long long)c * m + j) * m + i] = mk(gr[v][e], -gi[v][e]);
      }
    }
}

// g[q] = dv * sum over chunks c of part[c * nb * nb + q].  A workgroup takes
// kGfQ consecutive entries with kGfL lanes each: lane l sums chunks l,
// l + kGfL, ... (the kGfQ entries of one chunk are adjacent in memory), and
// the lane sums are added in LDS in a fixed tree, so the order is fixed.
constexpr int kGfQ = 8, kGfL = 32;
__global__ __launch_bounds__(kGfQ* kGfL) void k_gram_finish(const c64* __restrict__ part, int nb, int nchunk,
                                                             double dv, c64* __restrict__ g) {
  const int ql = threadIdx.x % kGfQ, l = threadIdx.x / kGfQ;
  const int nq = nb * nb;
  const int q = blockIdx.x * kGfQ + ql;
  double sr = 0.0, si = 0.0;
  if (q < nq)
    for (int c = l; c < nchunk; c += kGfL) {
      const c64 v = part[(long long)c * nq + q];
      sr += v.x;
      si += v.y;
    }
  __shared__ double rr[kGfQ * kGfL], ri[kGfQ * kGfL];
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int o = kGfL / 2; o > 0; o >>= 1) {
    if (l < o) {
      rr[threadIdx.x] += rr[threadIdx.x + o * kGfQ];
      ri[threadIdx.x] += ri[threadIdx.x + o * kGfQ];
    }
    __syncthreads();
  }
  if (l == 0 && q < nq) g[q] = mk(rr[ql] * dv, ri[ql] * dv);
}

// out_j = sum_i T[i + nb*j] in_i   (nb x nb complex FP64 T, column-major)
template <typename R>
__global__ __launch_bounds__(kT) void k_rotate(const cx<R>* __restrict__ in, cx<R>* __restrict__ out, long long N,
                                                int nb, const c64* __restrict__ T) {
  constexpr int JT = 8;
  const int j0 = blockIdx.y * JT;
  const long long r = (long long)blockIdx.x * kT + threadIdx.x;
  if (r >= N) return;
  double ar[JT], ai[JT];
#pragma unroll
  for (int q = 0; q < JT; ++q) ar[q] = ai[q] = 0.0;
  for (int i = 0; i < nb; ++i) {
    const cx<R> w = in[(long long)i * N + r];
    const double wr = w.x, wi = w.y;
#pragma unroll
    for (int q = 0; q < JT; ++q) {
      if (j0 + q < nb) {
        const c64 t = T[i + (long long)nb * (j0 + q)];
        ar[q] += wr * t.x - wi * t.y;
        ai[q] += wr * t.y + wi * t.x;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < JT; ++q)
    if (j0 + q < nb) out[(long long)(j0 + q) * N + r] = mk<R>((R)ar[q], (R)ai[q]);
}

template <typename R>
__global__ __launch_bounds__(kT) void k_convert(const cx<R>* __restrict__ in, c64* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT)
    out[i] = mk((double)in[i].x, (double)in[i].y);
}

__global__ __launch_bounds__(kT) void k_fill(double* __restrict__ p, long long n, double v) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) p[i] = v;
}

// ---------------------------------------------------------------------------
// Fused rebuild kernels (one system group of the per-stage rebuild).
//
// density with per-block partial sums: block (c, s) covers points
// [c * kDensBlock, (c+1) * kDensBlock) of system s and writes
// dpart[s * gridDim.x + c] = sum of its dens values
constexpr int kDensBlock = kT;

template <typename R>
__global__ __launch_bounds__(kT) void k_density_sum(const cx<R>* __restrict__ u, int N, int nb,
                                                     double* __restrict__ dens, const int* __restrict__ umap,
                                                     const c64* __restrict__ cpre, int U,
                                                     double* __restrict__ dpart) {
  const int s = blockIdx.y;
  double tot[1] = {0.0};
  {
    const int r = blockIdx.x * kDensBlock + threadIdx.x;
    const int rc = min(r, N - 1);
    const cx<R>* us = u + (long long)s * nb * N + rc;
    const int k = cpre ? umap[rc] : -1;
    double acc = 0.0;
    for (int o0 = 0; o0 < nb; o0 += kOrbBatch) {
      cx<R> v[kOrbBatch];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) v[q] = us[(long long)min(o0 + q, nb - 1) * N];
      double vr[kOrbBatch], vi[kOrbBatch];
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q) {
        vr[q] = (double)v[q].x;
        vi[q] = (double)v[q].y;
      }
      if (k >= 0) {
        c64 c[kOrbBatch];
#pragma unroll
        for (int q = 0; q < kOrbBatch; ++q) c[q] = cpre[(long long)k * U + s * nb + min(o0 + q, nb - 1)];
#pragma unroll
        for (int q = 0; q < kOrbBatch; ++q) {
          vr[q] += c[q].x;
          vi[q] += c[q].y;
        }
      }
#pragma unroll
      for (int q = 0; q < kOrbBatch; ++q)
        if (o0 + q < nb) acc += vr[q] * vr[q] + vi[q] * vi[q];
    }
    if (r < N) {
      dens[(long long)s * N + r] = 2.0 * acc;
      tot[0] = 2.0 * acc;
    }
  }
  block_sum<1>(tot);
  if (threadIdx.x == 0) dpart[(long long)s * gridDim.x + blockIdx.x] = tot[0];
}

// rho_ij (src/tddft_rhoij.f90:75-115) and the compensation weights
// wl(L) = sum_ij rho_ij qijlm(L,i,j) (src/nhat.f90:13-28) of one (atom,
// system) in one workgroup
__global__ __launch_bounds__(64) void k_rhoij_nhatw(const AtomDesc* __restrict__ atoms, const c64* __restrict__ ca,
                                                    int ctot, int nb, int mtot, double* __restrict__ rhoij,
                                                    const double* __restrict__ qtab, int nl2, int natom,
                                                    double* __restrict__ wl) {
  const int a = blockIdx.x, s = blockIdx.y;
  const AtomDesc A = atoms[a];
  const int m2 = A.ms * A.ms, nq = nl2 * m2;
  __shared__ double rho[kMaxMs * kMaxMs];
  __shared__ double qq[kMaxNl2 * kMaxMs * kMaxMs];
  // qijlm of the type into LDS (batched loads, clamped indices)
  for (int t0 = 0; t0 < nq; t0 += 64 * 4) {
    double t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) t[u] = qtab[A.toff_q + min(t0 + u * 64 + (int)threadIdx.x, nq - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (t0 + u * 64 + (int)threadIdx.x < nq) qq[t0 + u * 64 + threadIdx.x] = t[u];
  }
  for (int q = threadIdx.x; q < m2; q += 64) {
    const int i = q % A.ms, j = q / A.ms;
    double acc = 0.0;
    for (int o0 = 0; o0 < nb; o0 += kOrbBatch) {
      c64 ci[kOrbBatch], cj[kOrbBatch];
#pragma unroll
      for (int u = 0; u < kOrbBatch; ++u) {
        const c64* c = ca + (long long)(s * nb + min(o0 + u, nb - 1)) * ctot + A.coff;
        ci[u] = c[i];
        cj[u] = c[j];
      }
#pragma unroll
      for (int u = 0; u < kOrbBatch; ++u)
        if (o0 + u < nb) acc += 2.0 * (ci[u].x * cj[u].x + ci[u].y * cj[u].y);
    }
    rho[q] = acc;
    rhoij[(long long)s * mtot + A.moff + q] = acc;
  }
  __syncthreads();
  for (int L = threadIdx.x; L < nl2; L += 64) {
    double w = 0.0;
    for (int q = 0; q < m2; ++q) w += rho[q] * qq[L + nl2 * q];
    wl[((long long)s * natom + a) * nl2 + L] = w;
  }
}

// nhat on the union of the compensation spheres only (it is zero elsewhere):
// nhs[s * S + k] = sum over the atoms of point k of sum_L g wl, with per-block
// partial sums npart[s * gridDim.x + blk]
template <int NLM>
__global__ __launch_bounds__(64) void k_nhat_sparse(SphereTabs T, int S, const double* __restrict__ sph_g,
                                                    const double* __restrict__ wl, int natom, int nl2,
                                                    double* __restrict__ nhs, double* __restrict__ npart) {
  const int s = blockIdx.y;
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int kc = min(k, S - 1);
  double v = 0.0;
  const int e0 = T.sptr[kc], e1 = T.sptr[kc + 1];
  for (int e = e0; e < e1; ++e) {
    const int2 en = T.se[e];
    const int stride = T.sn[en.x];
    const double* w = wl + ((long long)s * natom + en.x) * nl2;
    double g[NLM], ww[NLM];
#pragma unroll
    for (int L = 0; L < NLM; ++L) {
      const int l = min(L, nl2 - 1);
      g[L] = sph_g[en.y + (long long)l * stride];
      ww[L] = w[l];
    }
#pragma unroll
    for (int L = 0; L < NLM; ++L)
      if (L < nl2) v += g[L] * ww[L];
  }
  if (k < S) nhs[(long long)s * S + k] = v;
  else v = 0.0;
  // 64-thread block: one wave, fixed shuffle tree
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (threadIdx.x == 0) npart[(long long)s * gridDim.x + blockIdx.x] = v;
}

// rnel renormalisation (src/nhat.f90:31-41) fused with the Hartree packing:
// every workgroup forms fac[s] = rnel / (dv sum dens + dv sum nhat) from the
// partial sums in the same fixed order (so all agree bit for bit), then
// dens *= fac and pair[q] = fac (dens + nhat)[2q] + i fac (dens + nhat)[2q+1]
// with nhat read from the sparse sphere values
__global__ __launch_bounds__(kT) void k_renorm_pack(double* __restrict__ dens, const int* __restrict__ smap,
                                                    const double* __restrict__ nhs, int S,
                                                    const double* __restrict__ dpart, int nd,
                                                    const double* __restrict__ npart, int nn, double dv,
                                                    double rnel, int renorm, int N, int nsys,
                                                    c64* __restrict__ pair, double* __restrict__ fac_out) {
  const int q = blockIdx.y;
  __shared__ double facs[2];
  for (int k = 0; k < 2; ++k) {
    const int s = 2 * q + k;
    if (s >= nsys) break;
    double v[2] = {0.0, 0.0};
    for (int c = threadIdx.x; c < nd; c += kT) v[0] += dpart[(long long)s * nd + c];
    for (int c = threadIdx.x; c < nn; c += kT) v[1] += npart[(long long)s * nn + c];
    block_sum<2>(v);
    if (threadIdx.x == 0) {
      const double f = renorm ? rnel / (v[1] * dv + v[0] * dv) : 1.0;
      facs[k] = f;
      if (blockIdx.x == 0) fac_out[s] = f;
    }
  }
  __syncthreads();
  const int base = blockIdx.x * (kT * kV) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kV; ++j) {
    const int r = base + j * kT;
    if (r >= N) continue;
    const int k = smap[r];
    double v[2] = {0.0, 0.0};
    for (int t = 0; t < 2; ++t) {
      const int s = 2 * q + t;
      if (s < nsys) {
        const long long i = (long long)s * N + r;
        const double d = dens[i] * facs[t];
        const double h = (k >= 0 ? nhs[(long long)s * S + k] : 0.0) * facs[t];
        dens[i] = d;
        v[t] = d + h;
      }
    }
    pair[(long long)q * N + r] = mk(v[0], v[1]);
  }
}

// full nhat field of one system from the sparse values: nhat = fac * nhs on
// the sphere union, 0 elsewhere
__global__ __launch_bounds__(kT) void k_nhat_expand(const int* __restrict__ smap, const double* __restrict__ nhs,
                                                    const double* __restrict__ fac, int N,
                                                    double* __restrict__ nhat) {
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= N) return;
  const int k = smap[r];
  nhat[r] = k >= 0 ? fac[0] * nhs[k] : 0.0;
}

// ---------------------------------------------------------------------------
// host wrappers
static unsigned grid_for(long long n) {
  long long b = (n + kT - 1) / kT;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static dim3 grid_rows(long long N, int rows, int per_thread) {
  return dim3((unsigned)((N + (long long)kT * per_thread - 1) / ((long long)kT * per_thread)), (unsigned)rows);
}

template <typename R>
hipError_t rk4_update(hipStream_t st, int stage, long long N, int nf, const cx<R>* h, cx<R>* p, cx<R>* acc,
                      cx<R>* y, double c, double dt, double* part, double dv, double* norms) {
  if (stage == 1) {
    hipLaunchKernelGGL((k_rk4_update<R, true>), grid_rows(N, nf, kV), dim3(kT), 0, st, (int)N, h, p, acc, y, c);
  } else if (stage < 4) {
    hipLaunchKernelGGL((k_rk4_update<R, false>), grid_rows(N, nf, kV), dim3(kT), 0, st, (int)N, h, p, acc, y, c);
  } else {
    hipLaunchKernelGGL((k_rk4_final<R>), dim3(kChunks, nf), dim3(kT), 0, st, N, kChunks, h, p, acc, dt / 6.0, part);
    hipLaunchKernelGGL(k_finish_rows, dim3(nf), dim3(kT), 0, st, part, nf, kChunks, dv, 1, norms);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t field_norms(hipStream_t st, const cx<R>* u, long long N, int nf, double dv, double* part, double* out) {
  hipLaunchKernelGGL((k_norm_partial<R>), dim3(kChunks, nf), dim3(kT), 0, st, u, N, kChunks, part);
  hipLaunchKernelGGL(k_finish_rows, dim3(nf), dim3(kT), 0, st, part, nf, kChunks, dv, 1, out);
  return hipGetLastError();
}

template <typename R>
hipError_t scale_fields(hipStream_t st, cx<R>* u, long long N, int nf, const double* num, const double* den) {
  hipLaunchKernelGGL((k_scale_fields<R>), grid_rows(N, nf, kV), dim3(kT), 0, st, u, (int)N, num, den);
  return hipGetLastError();
}

template <typename R>
hipError_t scale_dipole(hipStream_t st, cx<R>* u, int nx, int ny, int nz, int nsys, int nb, double dx, double dy,
                        double dz, const double* num, const double* den, double* part, double* out) {
  hipLaunchKernelGGL((k_scale_dipole<R>), dim3(kChunks, nsys), dim3(kT), nb * sizeof(double), st, u, nx, ny, nz, nb,
                     kChunks, dx, dy, dz, num, den, part);
  hipLaunchKernelGGL(k_finish_rows, dim3(3 * nsys), dim3(kT), 0, st, part, 3 * nsys, kChunks, dx * dy * dz, 0, out);
  return hipGetLastError();
}

template <typename R>
hipError_t density_sum(hipStream_t st, const cx<R>* u, long long N, int nsys, int nb, double* dens,
                       const int* umap, const c64* cpre, int U, double* dpart, int* nd) {
  const int nc = (int)((N + kDensBlock - 1) / kDensBlock);
  *nd = nc;
  hipLaunchKernelGGL((k_density_sum<R>), dim3(nc, nsys), dim3(kT), 0, st, u, (int)N, nb, dens, umap, cpre, U,
                     dpart);
  return hipGetLastError();
}

hipError_t rhoij_nhatw(hipStream_t st, const AtomDesc* atoms, int natom, const c64* ca, int ctot, int nsys, int nb,
                       int mtot, double* rho, const double* qtab, int nl2, double* wl) {
  if (natom == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rhoij_nhatw, dim3(natom, nsys), dim3(64), 0, st, atoms, ca, ctot, nb, mtot, rho, qtab, nl2,
                     natom, wl);
  return hipGetLastError();
}

hipError_t nhat_sparse(hipStream_t st, const SphereTabs& T, int S, const double* sph_g, const double* wl, int natom,
                       int nl2, int nsys, double* nhs, double* npart, int* nn) {
  const int nc = (S + 63) / 64;
  *nn = nc;
  if (S == 0) return hipSuccess;
  if (nl2 <= 9)
    hipLaunchKernelGGL(k_nhat_sparse<9>, dim3(nc, nsys), dim3(64), 0, st, T, S, sph_g, wl, natom, nl2, nhs, npart);
  else if (nl2 <= 16)
    hipLaunchKernelGGL(k_nhat_sparse<16>, dim3(nc, nsys), dim3(64), 0, st, T, S, sph_g, wl, natom, nl2, nhs, npart);
  else
    hipLaunchKernelGGL(k_nhat_sparse<kMaxNl2>, dim3(nc, nsys), dim3(64), 0, st, T, S, sph_g, wl, natom, nl2, nhs,
                       npart);
  return hipGetLastError();
}

hipError_t renorm_pack(hipStream_t st, double* dens, const int* smap, const double* nhs, int S, const double* dpart,
                       int nd, const double* npart, int nn, double dv, double rnel, int renorm, long long N, int nsys,
                       c64* pair, double* fac) {
  hipLaunchKernelGGL(k_renorm_pack, grid_rows(N, (nsys + 1) / 2, kV), dim3(kT), 0, st, dens, smap, nhs, S, dpart,
                     nd, npart, nn, dv, rnel, renorm, (int)N, nsys, pair, fac);
  return hipGetLastError();
}

hipError_t nhat_expand(hipStream_t st, const int* smap, const double* nhs, const double* fac, long long N,
                       double* nhat) {
  hipLaunchKernelGGL(k_nhat_expand, dim3((unsigned)((N + kT - 1) / kT)), dim3(kT), 0, st, smap, nhs, fac, (int)N,
                     nhat);
  return hipGetLastError();
}

template <typename R>
hipError_t density(hipStream_t st, const cx<R>* u, long long N, int nsys, int nb, double* dens, const int* umap,
                   const c64* cpre, int U) {
  hipLaunchKernelGGL((k_density<R>), grid_rows(N, nsys, 1), dim3(kT), 0, st, u, (int)N, nb, dens, umap, cpre, U);
  return hipGetLastError();
}

hipError_t fill(hipStream_t st, double* p, long long n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kT), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t nhat(hipStream_t st, const AtomDesc* atoms, int natom, const double* rhoij, int mtot,
                const double* qtab, int nl2, const SphereTabs& S, const double* sph_g, long long N, int nsys,
                double* wl, double* nh) {
  if (natom > 0) {
    hipLaunchKernelGGL(k_nhat_w, dim3(natom, nsys), dim3(64), 0, st, atoms, natom, rhoij, mtot, qtab, nl2, nsys,
                       wl);
  }
  hipLaunchKernelGGL(k_nhat_grid, grid_rows(N, nsys, 1), dim3(kT), 0, st, S, sph_g, wl, natom, nl2, (int)N, nh);
  return hipGetLastError();
}

hipError_t renorm_factor(hipStream_t st, const double* nh, const double* dens, long long N, int nsys, double dv,
                         double rnel, double* part, double* fac) {
  hipLaunchKernelGGL(k_sum2_partial, dim3(kChunks, nsys), dim3(kT), 0, st, nh, dens, N, kChunks, part);
  hipLaunchKernelGGL(k_renorm_factor, dim3(nsys), dim3(kT), 0, st, part, nsys, kChunks, dv, rnel, fac);
  return hipGetLastError();
}

hipError_t scale_pack(hipStream_t st, double* dens, double* nh, const double* fac, long long N, int nsys,
                      c64* pair) {
  hipLaunchKernelGGL(k_scale_pack, grid_rows(N, (nsys + 1) / 2, 1), dim3(kT), 0, st, dens, nh, fac, (int)N, nsys,
                     pair);
  return hipGetLastError();
}

hipError_t pair_means(hipStream_t st, const c64* pair, long long N, int nsys, double* part, double* means) {
  const int npair = (nsys + 1) / 2;
  hipLaunchKernelGGL(k_pair_sum_partial, dim3(kChunks, npair), dim3(kT), 0, st, pair, N, kChunks, part);
  hipLaunchKernelGGL(k_finish_rows, dim3(2 * npair), dim3(kT), 0, st, part, 2 * npair, kChunks, 1.0 / (double)N, 0,
                     means);
  return hipGetLastError();
}

// *flag = 1 when a and b differ in any of n words
__global__ __launch_bounds__(kT) void k_words_differ(const unsigned* __restrict__ a, const unsigned* __restrict__ b,
                                                      long long n, int* __restrict__ flag) {
  bool d = false;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT)
    d = d || a[i] != b[i];
  if (d) *flag = 1;
}

hipError_t words_differ(hipStream_t st, const unsigned* a, const unsigned* b, long long n, int* flag) {
  if (n <= 0) return hipSuccess;
  const long long blocks = std::min<long long>((n + kT - 1) / kT, 4096);
  hipLaunchKernelGGL(k_words_differ, dim3((unsigned)blocks), dim3(kT), 0, st, a, b, n, flag);
  return hipGetLastError();
}

hipError_t vks_build(hipStream_t st, double* dens, const double* ncoret, const c64* vhpair,
                     const double* vhmean, const double* vloc, long long N, int nsys, double* vxc, double* vh,
                     double* vks, float* vksf, const double* fac) {
  hipLaunchKernelGGL(k_vks, grid_rows(N, nsys, 1), dim3(kT), 0, st, dens, ncoret, vhpair, vhmean, vloc, (int)N,
                     vxc, vh, vks, vksf, fac);
  return hipGetLastError();
}

hipError_t rhoij(hipStream_t st, const AtomDesc* atoms, int natom, const c64* ca, int ctot, int nsys, int nb,
                 int mtot, double* rho) {
  if (natom == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rhoij, dim3(natom, nsys), dim3(64), 0, st, atoms, ca, ctot, nb, mtot, rho);
  return hipGetLastError();
}

hipError_t dij(hipStream_t st, const AtomDesc* atoms, int natom, const double* rho, int mtot, const double* qtab,
               const double* d0tab, const double* etab, const int* sph_idx, const double* sph_g,
               const double* vks, const double* vxc, int subtract_vxc, long long N, int nsys, double dv,
               const double* dxc, double* d, const DijWeights* wt) {
  if (natom == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dij, dim3(natom, nsys), dim3(kT), 0, st, atoms, rho, mtot, qtab, d0tab, etab, sph_idx,
                     sph_g, vks, vxc, subtract_vxc, N, dv, dxc, d, wt ? *wt : DijWeights{});
  return hipGetLastError();
}

template <typename R>
hipError_t dipole(hipStream_t st, const cx<R>* u, int nx, int ny, int nz, int nsys, int nb, double dx, double dy,
                  double dz, double* part, double* out) {
  hipLaunchKernelGGL((k_dipole_partial<R>), dim3(kChunks, nsys), dim3(kT), 0, st, u, nx, ny, nz, nb, kChunks, dx, dy,
                     dz, part);
  hipLaunchKernelGGL(k_finish_rows, dim3(3 * nsys), dim3(kT), 0, st, part, 3 * nsys, kChunks, dx * dy * dz, 0, out);
  return hipGetLastError();
}

template <typename R>
hipError_t kick(hipStream_t st, const cx<R>* src, cx<R>* dst, int nx, int ny, int nz, int nb, int ipol, double d,
                double sm) {
  hipLaunchKernelGGL((k_kick<R>), dim3(grid_for((long long)nx * ny * nz * nb)), dim3(kT), 0, st, src, dst, nx, ny,
                     nz, nb, ipol, d, sm);
  return hipGetLastError();
}

template <typename R>
hipError_t gram(hipStream_t st, const cx<R>* a, const cx<R>* b, long long N, int nb, double dv, c64* part, c64* g) {
  const int t = (nb + 15) / 16;
  if (t > 4) {
    if constexpr (std::is_same<R, double>::value) {
      const int T = (nb + 63) / 64;
      const int sym = (const void*)a == (const void*)b;
      const int ntp = sym ? T * (T + 1) / 2 : T * T;
      hipLaunchKernelGGL(k_gram_mfma_tiles, dim3(ntp * kGramChunks), dim3(256), 0, st, a, b, N, nb, T, ntp, sym,
                         kGramChunks, part);
    } else {
      hipLaunchKernelGGL((k_gram_partial<R>), dim3(t, t, kGramChunks), dim3(256), 0, st, a, b, N, nb, kGramChunks,
                         part);
    }
    hipLaunchKernelGGL(k_gram_finish, dim3((nb * nb + kGfQ - 1) / kGfQ), dim3(kGfQ * kGfL), 0, st, part, nb,
                       kGramChunks, dv, g);
    return hipGetLastError();
  }
  const dim3 grid(kGramBlockChunks);
  if constexpr (std::is_same<R, double>::value) {
    if (t == 1)
      hipLaunchKernelGGL((k_gram_mfma<1>), grid, dim3(64), 0, st, a, b, N, nb, kGramBlockChunks, part);
    else if (t == 2)
      hipLaunchKernelGGL((k_gram_mfma<2>), grid, dim3(128), 0, st, a, b, N, nb, kGramBlockChunks, part);
    else if (t == 3)
      hipLaunchKernelGGL((k_gram_mfma<3>), grid, dim3(192), 0, st, a, b, N, nb, kGramBlockChunks, part);
    else
      hipLaunchKernelGGL((k_gram_mfma<4>), grid, dim3(256), 0, st, a, b, N, nb, kGramBlockChunks, part);
  } else if (t == 1)
    hipLaunchKernelGGL((k_gram_block<R, 1>), grid, dim3(256), 0, st, a, b, N, nb, kGramBlockChunks, part);
  else if (t == 2)
    hipLaunchKernelGGL((k_gram_block<R, 2>), grid, dim3(256), 0, st, a, b, N, nb, kGramBlockChunks, part);
  else if (t == 3)
    hipLaunchKernelGGL((k_gram_block<R, 3>), grid, dim3(256), 0, st, a, b, N, nb, kGramBlockChunks, part);
  else
    hipLaunchKernelGGL((k_gram_block<R, 4>), grid, dim3(256), 0, st, a, b, N, nb, kGramBlockChunks, part);
  hipLaunchKernelGGL(k_gram_finish, dim3((nb * nb + kGfQ - 1) / kGfQ), dim3(kGfQ * kGfL), 0, st, part, nb,
                     kGramBlockChunks, dv, g);
  return hipGetLastError();
}

template <typename R>
hipError_t rotate(hipStream_t st, const cx<R>* in, cx<R>* out, long long N, int nb, const c64* T) {
  hipLaunchKernelGGL((k_rotate<R>), dim3((unsigned)((N + kT - 1) / kT), (nb + 7) / 8), dim3(kT), 0, st, in, out, N,
                     nb, T);
  return hipGetLastError();
}

template <typename R>
hipError_t convert_to_c64(hipStream_t st, const cx<R>* in, c64* out, long long n) {
  hipLaunchKernelGGL((k_convert<R>), dim3(grid_for(n)), dim3(kT), 0, st, in, out, n);
  return hipGetLastError();
}

#define OPAW_INST(R)                                                                                            \
  template hipError_t rk4_update<R>(hipStream_t, int, long long, int, const cx<R>*, cx<R>*, cx<R>*, cx<R>*,     \
                                    double, double, double*, double, double*);                                  \
  template hipError_t scale_dipole<R>(hipStream_t, cx<R>*, int, int, int, int, int, double, double, double,     \
                                      const double*, const double*, double*, double*);                          \
  template hipError_t field_norms<R>(hipStream_t, const cx<R>*, long long, int, double, double*, double*);      \
  template hipError_t scale_fields<R>(hipStream_t, cx<R>*, long long, int, const double*, const double*);       \
  template hipError_t density<R>(hipStream_t, const cx<R>*, long long, int, int, double*, const int*,          \
                                 const c64*, int);                                                              \
  template hipError_t density_sum<R>(hipStream_t, const cx<R>*, long long, int, int, double*, const int*,      \
                                     const c64*, int, double*, int*);                                           \
  template hipError_t dipole<R>(hipStream_t, const cx<R>*, int, int, int, int, int, double, double, double,     \
                                double*, double*);                                                              \
  template hipError_t kick<R>(hipStream_t, const cx<R>*, cx<R>*, int, int, int, int, int, double, double);      \
  template hipError_t gram<R>(hipStream_t, const cx<R>*, const cx<R>*, long long, int, double, c64*, c64*);     \
  template hipError_t rotate<R>(hipStream_t, const cx<R>*, cx<R>*, long long, int, const c64*);                 \
  template hipError_t convert_to_c64<R>(hipStream_t, const cx<R>*, c64*, long long);
OPAW_INST(double)
OPAW_INST(float)

}  // namespace opaw
