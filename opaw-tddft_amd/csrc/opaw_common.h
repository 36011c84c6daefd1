// Shared definitions for the OPAW-TDDFT HIP library (gfx950).
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
// host-only build (CPU checks of index logic): no HIP headers
#ifndef __host__
#define __host__
#define __device__
#define __forceinline__ inline
#endif
#endif
#include <cstdint>
#include <cstddef>

namespace opaw {

// Interleaved complex of FP64 or FP32; layout-compatible with Fortran
// complex*16 / complex*8.
template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) {
  cx<R> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return mk(a.x * s, a.y * s); }
// multiply by s*i, s = +-1
template <typename R>
__host__ __device__ __forceinline__ cx<R> mul_si(cx<R> a, int s) {
  return s > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

using c64 = cx<double>;
using c32 = cx<float>;

enum Precision { PREC_F64 = 0, PREC_MIXED = 1 };

// Per-atom descriptor kept in device memory.  Box tables are stored in the
// reference layout local_p3d(nr1,nr2,nr3,ms): projector-major, x fastest.
struct AtomDesc {
  int start[3];    // 0-based ir_start (may be negative / overflow when periodic)
  int nr[3];       // nrough
  int ms;          // mstates
  int itype;
  int npt;         // nr1*nr2*nr3
  int wrap;        // box crosses the grid edge (reference `edge` flag)
  long long tab;   // offset of this atom's table in the P / P1 arrays (doubles)
  int coff;        // offset of this atom's coefficients (sum of ms before it)
  int moff;        // offset of this atom's ms*ms matrices (dij, rhoij)
  int sph_off;     // offset into sphere point arrays
  int sph_n;       // number of sphere points
  int toff_q;      // offset of the type's qijlm table (nl2*ms*ms)
  int toff_d0;     // offset of the type's dij0 table (ms*ms)
  int toff_e;      // offset of the type's eijkl table (ms^4)
  int boff;        // offset of this atom's box points in the per-box-point tables
  int ot;          // on-site (libpaw) table set of the type, -1: none (model on-site D_ij)
  // D^hat sphere: points [dh_off, dh_off + dh_n) of the D^hat point arrays,
  // their g values (dh_nl moments, moment-major) from dh_goff; the type's
  // q_ij^L table (dh_nl, ms, ms) at toff_qh
  int dh_off, dh_n, dh_nl, dh_goff, toff_qh;
};

constexpr int kMaxMs = 16;
constexpr int kMaxNl2 = 25;

}  // namespace opaw
