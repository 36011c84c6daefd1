// On-site (libpaw) part of D_ij: per-type radial tables and the kernel that
// turns rho_ij into Dij^xc at every Hamiltonian rebuild.
#pragma once
#include "opaw_kernels.h"

namespace opaw {

constexpr int kOnsiteMaxLm = 9;    // density moments up to l = 2 (partial waves l <= 1)
constexpr int kOnsiteMaxNij = 10;  // radial pairs of up to 4 partial waves

// One dataset's radial tables, offsets into the concatenated double table
// (otab) and int table (oint).  Radial arrays have nr points; pp / tpp are
// pair-major (nij, nr); gnt is (lm, ms, ms); kln is (ms, ms).
struct OnsiteType {
  int nr;          // pawtab%mesh_size
  int lm;          // lcut_size^2 moments
  int nij;         // radial pairs (basis_size (basis_size + 1) / 2)
  int usetcore;    // pseudo core density present
  long long rad, w, inv2, pp, tpp, core, tcore, gnt;
  int kln;
};

// Dij^xc of every (atom, system) whose type has tables (AtomDesc::ot >= 0):
// pawdensities (opt_dens = 1), pawxcm (pawxcdev = 1, ABINIT ixc = 7 LDA)
// for the all-electron and pseudo moments, pawdijxcm.  dxc receives the
// (ms, ms) matrices at the atoms' moff, per system (stride mtot); scratch
// holds onsite_scratch_doubles(nr_max) doubles per (atom, system).
hipError_t onsite_xc(hipStream_t st, const AtomDesc* atoms, int natom, const OnsiteType* ot, const double* otab,
                     const int* oint, const double* rhoij, int mtot, int nsys, double* scratch, int nr_max,
                     double* dxc);

// Scratch doubles of one (atom, system) for types of at most nr_max points.
size_t onsite_scratch_doubles(int nr_max);

}  // namespace opaw
