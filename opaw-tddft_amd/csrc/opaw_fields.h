// Declarations of the grid-field / small-matrix kernels (opaw_fields.hip).
#pragma once
#include "opaw_gather.h"
#include "opaw_kernels.h"

namespace opaw {

constexpr int kChunks = 256;          // fixed chunk count of deterministic reductions
constexpr int kGramChunks = 64;         // Gram partials of the tiled kernel (more than 64 fields)
constexpr int kGramBlockChunks = 512;   // Gram partials of the whole-block kernel (up to 64 fields)
// partial-sum entries the Gram of m fields needs
inline size_t gram_partials(int m) {
  return (size_t)(m <= 64 ? kGramBlockChunks : kGramChunks) * (size_t)m * (size_t)m;
}
constexpr double kXcDensThreshold = 1e-15;  // per-point LDA density cutoff
constexpr double kRhoijSelect = 1e-10;      // src/get_dij.f90:82
constexpr double kDijMax = 500.0;           // src/0_parameter.f90 param_dij_max

// LDA: Slater exchange + Perdew-Wang 1992 correlation, unpolarised
// (libxc ids 1 and 12 as called from src/tddft_pot.f90:163-177)
__device__ __forceinline__ double lda_vxc(double n) {
  if (!(n > kXcDensThreshold)) return 0.0;
  const double pi = 3.14159265358979323846;
  const double cx = cbrt(3.0 / pi);
  const double n13 = cbrt(n);
  const double vx = -cx * n13;
  const double rs = cbrt(3.0 / (4.0 * pi * n));
  const double A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const double srs = sqrt(rs);
  const double q0 = -2.0 * A * (1.0 + a1 * rs);
  const double q1 = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  const double dq1 = A * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
  const double lg = log1p(1.0 / q1);
  const double ec = q0 * lg;
  const double dec = -2.0 * A * a1 * lg - q0 * dq1 / (q1 * q1 + q1);
  return vx + ec - rs / 3.0 * dec;
}

// union of the compensation spheres: point k owns entries [sptr[k],sptr[k+1])
// = (atom, offset of its g value for L=0); g stride of atom a is sn[a]
struct SphereTabs {
  const int* smap;   // N: grid point -> sphere-union index or -1
  const int* sptr;   // S+1
  const int2* se;    // entries
  const int* sn;     // per-atom sphere point count (g stride)
};

// stage 4 also writes the per-field norms (dv sum|p|^2, then sqrt) to norms
template <typename R>
hipError_t rk4_update(hipStream_t st, int stage, long long N, int nf, const cx<R>* h, cx<R>* p, cx<R>* acc,
                      cx<R>* y, double c, double dt, double* part, double dv, double* norms);
template <typename R>
hipError_t field_norms(hipStream_t st, const cx<R>* u, long long N, int nf, double dv, double* part, double* out);
template <typename R>
hipError_t scale_fields(hipStream_t st, cx<R>* u, long long N, int nf, const double* num, const double* den);
// scale_fields fused with the dipole of the scaled orbitals (out: 3 per system)
template <typename R>
hipError_t scale_dipole(hipStream_t st, cx<R>* u, int nx, int ny, int nz, int nsys, int nb, double dx, double dy,
                        double dz, const double* num, const double* den, double* part, double* out);
template <typename R>
hipError_t density(hipStream_t st, const cx<R>* u, long long N, int nsys, int nb, double* dens, const int* umap,
                   const c64* cpre, int U);
// density with per-block partial sums (nd partials per system)
template <typename R>
hipError_t density_sum(hipStream_t st, const cx<R>* u, long long N, int nsys, int nb, double* dens,
                       const int* umap, const c64* cpre, int U, double* dpart, int* nd);
// rho_ij and the compensation weights wl in one launch
hipError_t rhoij_nhatw(hipStream_t st, const AtomDesc* atoms, int natom, const c64* ca, int ctot, int nsys, int nb,
                       int mtot, double* rho, const double* qtab, int nl2, double* wl);
// nhat on the sphere union (S points per system) with nn partial sums per system
hipError_t nhat_sparse(hipStream_t st, const SphereTabs& T, int S, const double* sph_g, const double* wl, int natom,
                       int nl2, int nsys, double* nhs, double* npart, int* nn);
// rnel factor from the partials, dens *= fac, Hartree pair of fac (dens + nhat)
hipError_t renorm_pack(hipStream_t st, double* dens, const int* smap, const double* nhs, int S, const double* dpart,
                       int nd, const double* npart, int nn, double dv, double rnel, int renorm, long long N, int nsys,
                       c64* pair, double* fac);
// full nhat field of one system from the sparse values
hipError_t nhat_expand(hipStream_t st, const int* smap, const double* nhs, const double* fac, long long N,
                       double* nhat);
hipError_t fill(hipStream_t st, double* p, long long n, double v);
hipError_t nhat(hipStream_t st, const AtomDesc* atoms, int natom, const double* rhoij, int mtot,
                const double* qtab, int nl2, const SphereTabs& S, const double* sph_g, long long N, int nsys,
                double* wl, double* nh);
hipError_t renorm_factor(hipStream_t st, const double* nh, const double* dens, long long N, int nsys, double dv,
                         double rnel, double* part, double* fac);
hipError_t scale_pack(hipStream_t st, double* dens, double* nh, const double* fac, long long N, int nsys,
                      c64* pair);
hipError_t pair_means(hipStream_t st, const c64* pair, long long N, int nsys, double* part, double* means);
// *flag (device, cleared by the caller) = 1 when a and b differ in any of n words
hipError_t words_differ(hipStream_t st, const unsigned* a, const unsigned* b, long long n, int* flag);
// with fac: dens_s is first scaled by fac_s and stored (the renormalisation
// whose factors the fused Hartree pass computed)
hipError_t vks_build(hipStream_t st, double* dens, const double* ncoret, const c64* vhpair,
                     const double* vhmean, const double* vloc, long long N, int nsys, double* vxc, double* vh,
                     double* vks, float* vksf, const double* fac = nullptr);
hipError_t rhoij(hipStream_t st, const AtomDesc* atoms, int natom, const c64* ca, int ctot, int nsys, int nb,
                 int mtot, double* rho);
// with wt: also the nonlocal weights w(c, f) = sum_j D_ij cB(f, c_j) of every
// field of the rebuilt systems (what the weights kernel computes from D)
struct DijWeights {
  const c64* cb;   // coefficients, field-major cb[f * ctot + c]
  c64* w;          // weights, coefficient-major w[c * nf + f]; null: none
  int ctot, nf, nb, f0;   // f0: first field of the rebuild's first system
};
// D^hat: the sphere of atom a is points [dh_off, dh_off + dh_n) of sph_idx /
// sph_g (g stored per atom, moment-major) with dh_nl moments, and q_ij^L at
// toff_qh of qtab; dxc (per system, stride mtot) is added for atoms with
// on-site tables (ot >= 0)
hipError_t dij(hipStream_t st, const AtomDesc* atoms, int natom, const double* rho, int mtot, const double* qtab,
               const double* d0tab, const double* etab, const int* sph_idx, const double* sph_g,
               const double* vks, const double* vxc, int subtract_vxc, long long N, int nsys, double dv,
               const double* dxc, double* d, const DijWeights* wt = nullptr);
template <typename R>
hipError_t dipole(hipStream_t st, const cx<R>* u, int nx, int ny, int nz, int nsys, int nb, double dx, double dy,
                  double dz, double* part, double* out);
template <typename R>
hipError_t kick(hipStream_t st, const cx<R>* src, cx<R>* dst, int nx, int ny, int nz, int nb, int ipol, double d,
                double sm);
template <typename R>
hipError_t gram(hipStream_t st, const cx<R>* a, const cx<R>* b, long long N, int nb, double dv, c64* part, c64* g);
template <typename R>
hipError_t rotate(hipStream_t st, const cx<R>* in, cx<R>* out, long long N, int nb, const c64* T);
template <typename R>
hipError_t convert_to_c64(hipStream_t st, const cx<R>* in, c64* out, long long n);

}  // namespace opaw
