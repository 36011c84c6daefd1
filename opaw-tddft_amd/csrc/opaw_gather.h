// Deterministic projector-box corrections in compact form (gfx950).
//
// The union of all atom boxes is a sorted list of grid points; point k owns
// the entries [uptr[k], uptr[k+1]) -- one per atom whose box holds the point,
// in atom order.  Entry e carries the atom's coefficient offset, its ms and
// the offset of the point's ms table values in a union-major copy of the
// projector tables.  For a weight set w (nf x ctot) the correction of field f
// at union point k is
//     corr(f, k) = sum_e sum_{i<ms_e} tab[toff_e + i] * w_f[coff_e + i]
// -- the reference's per-atom add-back (src/snphi.f90:29-47,
// src/hphi.f90:42-60) summed in atom order, without atomics.  It is computed
// once per weight set into a compact array corr[k*nf + f]; streaming kernels
// add corr(umap[r], f) at grid points inside the union.
#pragma once
#include "opaw_common.h"

namespace opaw {

struct UEntry {
  int coff;   // coefficient offset of the atom (sum of ms before it)
  int ms;     // projectors of the atom
  int toff;   // offset of this point's table values in the union tables
  int atom;
};

struct UnionTabs {
  const int* umap;      // N: grid point -> union index, -1 outside every box
  const int* uptr;      // U+1
  const UEntry* ue;     // entries
  const int* upt;       // U: grid point of union index k
  const double* p;      // union-major projector table values (local_p3d)
  const double* p1;     // union-major orthonormal projector values (local_p3d1)
  int U;
};

// compact corrections read by the streaming kernels
struct Corr {
  const int* umap;      // N: grid point -> union index or -1
  const c64* c;         // corr[k*nf + f]; null = none
  int U;
};

__device__ __forceinline__ c64 gather_corr(const UnionTabs& U, const double* __restrict__ tab, int k,
                                           const c64* __restrict__ w) {
  double xr = 0.0, xi = 0.0;
  const int e1 = U.uptr[k + 1];
  for (int e = U.uptr[k]; e < e1; ++e) {
    const UEntry E = U.ue[e];
    const double* t = tab + E.toff;
    const c64* ww = w + E.coff;
    for (int i = 0; i < E.ms; ++i) {
      const double tv = t[i];
      const c64 wv = ww[i];
      xr += tv * wv.x;
      xi += tv * wv.y;
    }
  }
  return mk(xr, xi);
}

}  // namespace opaw
