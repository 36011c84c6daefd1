// Projector-box work: tables, projection, weights and gathered add-back.
#pragma once
#include "opaw_gather.h"
#include "opaw_kernels.h"

namespace opaw {

constexpr int kProjChunk = 64;
// union-correction blocks: at most this many points and (coefficient, value)
// pairs per block
constexpr int kUnionBlockRows = 64;
constexpr int kUnionBlockPairs = 1024;    // box points per projection workgroup

struct BoxTables {
  const AtomDesc* atoms;     // device
  int natom;
  int ctot;                  // sum of ms over atoms
  int msp;                   // padded ms of the projection kernel (8 or 16)
  int nchunks;               // projection chunks over all atoms
  const int2* chunks;        // (atom, first point) per chunk
  const int2* achunk;        // (first chunk, count) per atom
  const int2* coef;          // (atom, i) per coefficient
  const double* p;           // per-atom projector tables (reference layout)
  const double* p1;          // per-atom orthonormal projector tables
  const int* bpt;            // grid point of every box point (atoms concatenated, boff)
  const int* bk;             // union index of every box point
  const int* pptr;           // U+1: (coefficient, value) pair range of union point k
  const int* pc;             // pair coefficient index
  const double* pv;          // pair value, projector table p
  const double* pv1;         // pair value, projector table p1
  const int* optr;           // natom+1: overlap columns of atom a (k_bfix)
  const int* ovoff;          // natom: offset of atom a's overlap values
  const int* oc;             // overlap column -> coefficient index
  const double* ov;          // overlap values O(i, l), row-major per atom
  const double* ov1;         // the same columns for the p1-p1 overlaps O11(i, l)
  const int2* ublk;          // union point blocks (first point, count) for the LDS kernel
  int nublk;                 // 0: no blocks (global kernel)
  const int* ucptr;          // nublk+1: coefficient list of block b is ucol[ucptr[b] .. ucptr[b+1])
  const int* ucol;           // coefficients referenced by each block (sorted)
  const int* pcl;            // pair coefficient as an index into its block's list
  int ucmax;                 // longest block coefficient list
  UnionTabs U;               // union-of-boxes gather tables
};

// Projection of nf fields (of in + cpre when cpre, a compact correction, is
// given), chunk partials in `part`, coefficients in ca (nf x ctot), and
// weights w (mode 1: sn * ca, mode 2: D(sys) ca, mode 0: none).
template <typename R>
hipError_t box_coeffs(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf, int which,
                      const c64* cpre, double dv, double* part, c64* ca, int mode, const double* sn,
                      const double* dij, int nf_per_sys, int mtot, c64* w);

// compact correction corr(k, f) (row k of U rows, field stride 1, row stride
// nf) = sum over the atom entries of union point k of tab * w, where the
// weights are coefficient-major, w[c*nf + f]
hipError_t union_corr(const BoxTables& bt, hipStream_t st, int nf, int which, const c64* w, c64* corr);

hipError_t box_weights(const BoxTables& bt, hipStream_t st, const c64* ca, int nf, int mode, const double* sn,
                       const double* dij, int nf_per_sys, int mtot, c64* w);

// out_f(upt[k]) += corr(k, f), in place
template <typename R>
hipError_t union_apply(const GridGeom& g, const BoxTables& bt, hipStream_t st, cx<R>* out, int nf, const c64* corr);
// cA = dv P1^T u, wA = sn cA and cB = dv P^T u in one projection launch
template <typename R>
hipError_t box_coeffs_dual(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* u, int nf,
                           double dv, double* part, c64* cA, const double* sn, c64* wA, c64* cB);
// RK4 stage add-back of the weights w (p1 tables): kc = -i corr,
// y += a1 kc, acc += a2 kc on the union of boxes
template <typename R>
hipError_t union_rk4_apply(const GridGeom& g, const BoxTables& bt, hipStream_t st, int nf, const c64* w, c64* corr,
                           cx<R>* y, cx<R>* acc, double a1, double a2);
// P1 coefficients and sn weights of fields stored only on the union of boxes
hipError_t box_coeffs_compact(const GridGeom& g, const BoxTables& bt, hipStream_t st, const c64* vals, int nf,
                              double dv, double* part, c64* ca, const double* sn, c64* w);
// cB += O wA: projections of S^-1/2 u from those of u (fixed box overlaps);
// with wC, cB += O (wA + a1 (-i) wC)
hipError_t box_bfix(const BoxTables& bt, hipStream_t st, int nf, const c64* wA, c64* cB, const c64* wC = nullptr,
                    double a1 = 0.0);
// cA += a1 (-i) O11 wC, wA = sn cA (P1 coefficients after an RK4 stage's
// union add-back, from those before it)
hipError_t box_afix(const BoxTables& bt, hipStream_t st, int nf, const c64* wC, double a1, const double* sn, c64* cA,
                    c64* wA);
// cA = dv P1^T y, cB = dv P^T y and cC = dv P1^T h (h compact, on the union
// of boxes) with wC = sn cC, in one projection launch and one finish launch;
// part holds three partial sets
template <typename R>
hipError_t box_coeffs_triple(const GridGeom& g, const BoxTables& bt, hipStream_t st, const cx<R>* y, const c64* hU,
                             int nf, double dv, double* part, c64* cA, c64* cB, c64* cC, const double* sn, c64* wC);
// RK4 stage add-back and the next stage's compact correction in one pass over
// the union of boxes: with kc = -i P1 wC, y += a1 kc and acc += a2 kc, and
// corrA = P1 wA
template <typename R>
hipError_t union_rk4_both(const GridGeom& g, const BoxTables& bt, hipStream_t st, int nf, const c64* wC,
                          const c64* wA, c64* corrC, c64* corrA, cx<R>* y, cx<R>* acc, double a1, double a2);
// union_corr added straight into the grid fields out (corr is scratch for the
// fallback path); the per-field order of additions is that of union_apply
template <typename R>
hipError_t union_corr_apply(const GridGeom& g, const BoxTables& bt, hipStream_t st, int nf, int which,
                            const c64* w, c64* corr, cx<R>* out);

}  // namespace opaw
