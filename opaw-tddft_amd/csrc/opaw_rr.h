// Rayleigh-Ritz on the device (opaw_rr.hip): the lowest eigenpairs of the
// subspace pencil H c = S c w of the block solver, with rocSOLVER's
// divide-and-conquer Hermitian eigensolver and rocBLAS products, so that the
// Gram matrices never leave HBM.
#pragma once
#include "opaw_common.h"

namespace opaw {

// Device work of an m-dimensional pencil, in bytes (the matrices, the
// eigenvalues and the flags; rocSOLVER's own workspace lives in the handle).
size_t rr_work_bytes(int m);
// flags[0] = kept rank, flags[1] / flags[2] = rocSOLVER info of the overlap
// and of the projected problem (0: converged), inside the work area
int* rr_flags(void* work, int m);

// S and H: row-major m x m Gram matrices (S[a*m+b] = <y_a|y_b>, H[a*m+b] =
// <y_a|H y_b>), left unchanged.  The overlap is scaled to a unit diagonal,
// its eigenvectors with eigenvalue <= floor * largest are dropped (the
// canonical orthogonalisation), the projected Hamiltonian is diagonalised
// and the lowest nout eigenpairs returned: C column-major m x nout (C[i + m
// j]), w[0..nout).  Everything is enqueued on st; the kept rank and the
// solver's info land in rr_flags(work, m), which the caller reads after its
// next synchronisation (rank < nout or info != 0: the result is not valid).
// *blas is a rocBLAS handle created on first use (release with
// rr_release).  Returns 0, or a rocBLAS status (> 0) / HIP error (< 0).
int rr_lowest(void** blas, hipStream_t st, int m, int nout, double floor, const c64* S, const c64* H, void* work,
              c64* C, double* w);
void rr_release(void* blas);

}  // namespace opaw
