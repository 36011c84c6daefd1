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

template <typename R>
hipError_t apply_t_plus_v(const GridGeom& g, hipStream_t st, const cx<R>* u, cx<R>* out, cx<R>* tmp,
                          int nf, int nf_per_sys, const R* ek, const R* vks,
                          const cx<R>* twx, const cx<R>* twy, const cx<R>* twz);

template <typename R>
hipError_t fft3d(const GridGeom& g, hipStream_t st, int dir, const cx<R>* in, cx<R>* out, int nf,
                 const cx<R>* twx, const cx<R>* twy, const cx<R>* twz);

hipError_t hartree_conv(const GridGeom& g, int scale, hipStream_t st, const c64* rho, c64* out,
                        c64* work, const double* vk, const c64* twx, const c64* twy,
                        const c64* twz);

// --- projector-box work (opaw_box.hip) -------------------------------------
struct BoxTables {
  const AtomDesc* atoms;     // device
  int natom;
  int ctot;                  // sum of ms over atoms
  const double* p;           // projector tables (all atoms, reference layout)
  const double* p1;          // orthonormal projector tables
  const float* pf;           // FP32 copies (mixed precision) or null
  const float* p1f;
  const double* s;           // per-atom overlap eigenvalues, packed by coff
  const int* plane_ptr;      // nz+1
  const PlaneEntry* plane;   // plane entries
};

// ca[f][coff+i] = dv * sum_box u_f * tab_i  (tab = p or p1), FP64 accumulation
template <typename R>
hipError_t box_project(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf,
                       int which, double dv, c64* ca);

// out_f[box] += sum_i tab_i * w_i with w = M c:
//   mode 0: w_i = sn_i c_i     (sn = (1+s)^n - 1 packed by coff, or sinv)
//   mode 1: w = D(sys) c       (D per system, packed at moff, ms*ms each)
template <typename R>
hipError_t box_add(const GridGeom& g, const BoxTables& bt, hipStream_t st, cx<R>* out, int nf,
                   int which, const c64* ca, int mode, double n, const double* sn, const double* dij,
                   int nf_per_sys, int mtot);

}  // namespace opaw
