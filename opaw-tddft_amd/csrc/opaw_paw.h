// Set-up kernels shared with the context (opaw_paw.hip): the local potential
// of vloc_tuma_prep (src/vloc/vloc_prep.f90, src/vloc/make_cvkb.f90).
#pragma once
#include "opaw_common.h"

namespace opaw {

// cv[p] = v(|r_p|) of one atom type on the doubled box, r_p =
// ((ix - nxb/2) dx, ...), from the natural spline (x, y, y2) of the radial
// potential on x = max(rr, 0.1 dr1); v = y[0] below rr[0], charge / r above
// rr[nr-1] (make_cvkb, src/vloc/make_cvkb.f90:30-45)
hipError_t vloc_sample(hipStream_t st, const double* rr, const double* x, const double* y, const double* y2, int nr,
                       double charge, double dr1, int nxb, int nyb, int nzb, double dx, double dy, double dz,
                       c64* cv);
// cvn(k) += dv exp(-i r0.k) form(k) cvkb(k), form the structure factor of
// the natom atoms at coords (shift_cvkb, add_formfactor)
hipError_t vloc_accumulate(hipStream_t st, const c64* cvkb, c64* cvn, int nxb, int nyb, int nzb, double dx,
                           double dy, double dz, const double* coords, int natom);
// vloc(ix,iy,iz) = Re cvn(central box) / (nxb nyb nzb dv) (c_to_v_compress)
hipError_t vloc_crop(hipStream_t st, const c64* cvn, int nxb, int nyb, int nzb, int nx, int ny, int nz, double scale,
                     double* vloc);

}  // namespace opaw
