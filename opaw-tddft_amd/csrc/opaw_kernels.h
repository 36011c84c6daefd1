// Host-side launch wrappers shared between the translation units.
#pragma once
#include "opaw_common.h"

namespace opaw {

struct GridGeom {
  int nx, ny, nz;
};

// Kernel-family timing hooks (HIP events on the launch stream), implemented
// by the context; no-ops unless profiling is enabled on the calling thread.
enum KClass { KC_FFT_X = 0, KC_FFT_Y = 1, KC_FFT_Z = 2, KC_FFT_XH = 3, KC_HARTREE = 4, KC_BOX_PROJ = 5,
              KC_BOX_ADD = 6, KC_RK4 = 7, KC_DENS = 8, KC_POT = 9, KC_OTHER = 10 };
void prof_begin(hipStream_t st, int cls);
void prof_end(hipStream_t st, int cls);

bool fft_size_supported(int n);

// RK4 stage fused into the last kinetic pass (plane path only): with it,
// apply_t_plus_v writes y = p + c k and acc (= k, or += 2k when !first),
// k = -i h, instead of h, and h on the union of boxes into hU[k*U + f].
// With first, the transformed input must be p itself (p is not read again).
template <typename R>
struct PlaneRk4 {
  const cx<R>* p = nullptr;
  cx<R>* acc = nullptr;
  c64* hU = nullptr;
  double c = 0.0;
  int first = 0;
};

// whether the kinetic operator of an nx x ny grid runs on the xy-plane path
template <typename R>
bool plane_supported(int nx, int ny);

// out_f = IFFT(ek FFT(u_f)) + vks(sys) u_f [+ cnl_f], where u_f = in_f [+ cpre_f]
// (compact corrections over the union of boxes stored as corr[k*U + f]: U is
// the row stride, i.e. the field count of the batch; umap maps a grid point to
// its union index k or -1, upt lists the union points in grid order and the
// union points of plane z are [uzr[z], uzr[z+1])).  cnl is read only by the
// last (inverse) pass, which first waits for before_inv when it is set, so
// cnl may still be in production on another stream during the first passes.
// With ev_mid set, the passes before the last run on st and the last pass on
// st_inv (which may be the null stream), after them: ev_mid is recorded on st
// and st_inv waits for it.
template <typename R>
hipError_t apply_t_plus_v(const GridGeom& g, hipStream_t st, const cx<R>* in, cx<R>* out, cx<R>* tmp,
                          int nf, int nf_per_sys, const R* ek, const R* vks,
                          const cx<R>* twx, const cx<R>* twy, const cx<R>* twz,
                          const int* umap, const int* upt, const int* uzr, const c64* cpre, const c64* cnl,
                          int U, hipEvent_t before_inv, const PlaneRk4<R>* rk = nullptr,
                          hipStream_t st_inv = nullptr, hipEvent_t ev_mid = nullptr);

template <typename R>
hipError_t fft3d(const GridGeom& g, hipStream_t st, int dir, const cx<R>* in, cx<R>* out, int nf,
                 const cx<R>* twx, const cx<R>* twy, const cx<R>* twz);

hipError_t hartree_conv(const GridGeom& g, int scale, hipStream_t st, const c64* rho, c64* out,
                        c64* work, const double* vk, const c64* twx, const c64* twy,
                        const c64* twz);

// The rebuild's renormalisation and Martyna-Tuckerman Hartree solve for
// doubled boxes whose (2 nx) x ny tile fits LDS, the renormalisation and pair
// packing fused into the first transform pass.  Pair q = systems 2q, 2q+1:
// fac_s = rnel / (dv (sum dpart_s + sum npart_s)) when renorm, else 1, is
// stored in fac; rho_s = fac_s (dens_s + nhs_s[smap]) (dens is read, not
// written: vks_build stores fac_s dens_s); vh[q] receives V_H of
// rho_2q + i rho_2q+1 on the small grid.
struct HartreeRebuild {
  const double* dens;      // nsys x N
  const int* smap;         // N: point -> sparse compensation index or -1
  const double* nhs;       // nsys x S sparse compensation charge
  int S;
  const double* dpart;     // nsys x nd partial sums of dens
  int nd;
  const double* npart;     // nsys x nn partial sums of the compensation charge
  int nn;
  double dv, rnel;
  int renorm;
  double* fac;             // nsys renormalisation factors (out)
};
bool hartree_fused_supported(const GridGeom& g, int scale);
hipError_t hartree_fused(const GridGeom& g, hipStream_t st, const HartreeRebuild& H, int q, int nsys, c64* work,
                         c64* vh, const double* vk, const c64* twx, const c64* twy, const c64* twz);

}  // namespace opaw
