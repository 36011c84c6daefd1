// On-site exchange-correlation part of D_ij (libpaw pawdenpot + pawdij with
// pawxcdev = 1), computed on the GPU at every Hamiltonian rebuild so the time
// loop keeps D_ij on the device.
//
// Reference chain (src/get_dij.f90:14-30): set_rhoij selects |rho_ij| > 1e-10;
// pawdensities (src/libpaw/m_paw_denpot.F90:1216-1574) builds the lm moments
// n1_lm(r) = sum_ij rho_ij dltij G(lm, ij) phi_i phi_j / r^2 and the pseudo
// ones from tphi; pawxcm (src/libpaw/2_m_pawxc.F90:3763-4513) expands the
// LDA potential to first order in the non-spherical moments:
//   v_00 = sqrt(4pi) v(n_s) + v''(n_s) sum_{lm>0} n_lm^2 / (2 sqrt(4pi)),
//   v_lm = v'(n_s) n_lm,
// with n_s = (n_00 / sqrt(4pi) + core) / 2 the spin-up spherical density
// (clamped at 1e-14), v = xcpbe option 1 (ABINIT's ixc = 7 PW92,
// src/libpaw/0_xc_m_xcpbe.F90:881-945) and v', v'' by finite differences at
// n_s (1 +- 1e-4), evaluated from a Taylor jet of v; pawdijxcm (src/libpaw/4_m_pawdij.F90:1959-2178)
// integrates sum_lm G(lm, ij) int (v_lm phi_i phi_j - vt_lm tphi_i tphi_j)
// with simp_gen's weights.
//
// The radial mesh is cut into chunks of 128 points, one workgroup per
// (chunk, atom, system), so a rebuild fills the device with a few hundred
// workgroups: per point one thread evaluates the all-electron moments and
// potentials and another the pseudo ones, the weighted potential moments and
// the chunk's pair functions go to LDS, and one thread per (pair, lm) sums
// the chunk point by point.  A second kernel adds the chunk partials in chunk
// order, so results are bit-identical run to run, and applies the rare
// low-density interpolation of pawxcm (points with n < xc_denpos bracketed by
// denser points) from the spherical values kept in scratch.
#include "opaw_onsite.h"

namespace opaw {

namespace {

constexpr int kPts = 128;                             // radial points per chunk (workgroup)
constexpr int kOT = 2 * kPts;                          // all-electron + pseudo thread per point
constexpr int kFin = 128;
constexpr int kNP = kOnsiteMaxNij * kOnsiteMaxLm;     // (pair, moment) integrals
static_assert(kNP <= kPts, "one thread per (pair, lm) integral in each half");
constexpr double kSqfpi = 3.5449077018110318;         // sqrt(4 pi)
constexpr double kDelta = 1.0e-4;                      // pawxcm delta
constexpr double kRhoMin = 1.0e-14;                    // pawxcm rho_min
constexpr double kDenpos = 1.0e-14;                    // xc_denpos of src/get_dij.f90:20
constexpr double kRhoijSel = 1.0e-10;                  // src/get_dij.f90:82
// xcpbe (option 1) constants
constexpr double kTwoM13 = 0.7937005259840998;         // 2^(-1/3)
constexpr double kThreeFourthDivPi = 0.238732414637843;
constexpr double kSixPi2_13 = 3.8977770897207535;      // (6 pi^2)^(1/3)
constexpr double kRsfac = 0.6203504908994000;
constexpr double kSqRsfac = 0.7876233178997433;
constexpr double kSqRsfacInv = 1.269642451250142;
constexpr double kA = 0.031091, kA1 = 0.21370, kB1 = 7.5957, kB2 = 3.5876, kB3 = 1.6382, kB4 = 0.49294;

// pawxcm's differences of v at rho (1 +- delta) from a Taylor jet of v in s,
// v(rho (1 + s)) = sum_k c_k s^k:
//   d1 = (v(rho(1+d)) - v(rho(1-d))) / (4 d rho)             = (c1 + c3 d^2) / (2 rho)
//   d2 = (v(rho(1+d)) + v(rho(1-d)) - 2 v(rho)) / (4 d^2 rho^2) = (c2 + c4 d^2) / (2 rho^2)
// up to the c5 d^4 and c6 d^4 terms (below 1e-15 of the sums); the
// difference quotients themselves would multiply a one-ulp difference in v
// by 1/d^2 = 1e8.  c3 d^2 and c4 d^2 are ~1e-8 of the sums, so c3 and c4 come
// from an FP32 jet (their 1e-7 rounding lands at 1e-15) and c0..c2 from an
// FP64 one.  Every power of rho in xcpbe is a monomial in (1 + s), whose
// coefficients are binomial; the logarithm and the quotients use the
// recurrences of Taylor-mode differentiation.
template <typename T, int K>
struct Jet {
  T c[K];
};

// a (1 + s)^p
template <typename T, int K>
__device__ __forceinline__ Jet<T, K> jmono(T a, T p) {
  Jet<T, K> r;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    r.c[k] = a;
    a = a * (p - (T)k) / (T)(k + 1);
  }
  return r;
}
template <typename T, int K>
__device__ __forceinline__ void jaxpy(Jet<T, K>& y, T a, const Jet<T, K>& x) {
#pragma unroll
  for (int k = 0; k < K; ++k) y.c[k] += a * x.c[k];
}
template <typename T, int K>
__device__ __forceinline__ Jet<T, K> jmul(const Jet<T, K>& a, const Jet<T, K>& b) {
  Jet<T, K> r;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T acc = 0;
#pragma unroll
    for (int i = 0; i <= k; ++i) acc += a.c[i] * b.c[k - i];
    r.c[k] = acc;
  }
  return r;
}
// a / b
template <typename T, int K>
__device__ __forceinline__ Jet<T, K> jdiv(const Jet<T, K>& a, const Jet<T, K>& b) {
  Jet<T, K> r;
  const T i0 = (T)1 / b.c[0];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T acc = a.c[k];
#pragma unroll
    for (int i = 1; i <= k; ++i) acc -= b.c[i] * r.c[k - i];
    r.c[k] = acc * i0;
  }
  return r;
}
// log a for k >= 1 (c[0] is set by the caller)
template <typename T, int K>
__device__ __forceinline__ Jet<T, K> jlog_tail(const Jet<T, K>& a) {
  Jet<T, K> r;
  r.c[0] = 0;
  const T i0 = (T)1 / a.c[0];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    T acc = a.c[k];
#pragma unroll
    for (int i = 1; i < k; ++i) acc -= ((T)i / (T)k) * r.c[i] * a.c[k - i];
    r.c[k] = acc * i0;
  }
  return r;
}

__device__ __forceinline__ double dev_cbrt(double x) { return cbrt(x); }
__device__ __forceinline__ float dev_cbrt(float x) { return cbrtf(x); }
__device__ __forceinline__ double dev_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dev_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dev_log(double x) { return log(x); }
__device__ __forceinline__ float dev_log(float x) { return logf(x); }

// xcpbe option 1, nspden = 1, as a jet in s: the potential of the spin-up
// density rho_up (1 + s).  With t = (2 rho_up)^-1/3 and x = t^1/2:
// rs ~ t (1+s)^-1/3, sqrt(rs) ~ x (1+s)^-1/6, rs^-1/2 ~ 2 rho_up t^2 x
// (1+s)^1/6 and the exchange ~ (1+s)^1/3.  lg = -log(q1^2 / (q1^2 + q1)) =
// log(1 + 1/q1); its value follows xcpbe's own operations.
template <typename T, int K>
__device__ __forceinline__ Jet<T, K> lda_v_jet(T rho_up) {
  using J = Jet<T, K>;
  const T rum13 = (T)1 / dev_cbrt(rho_up);
  const T t = (T)kTwoM13 * rum13;
  const T x = dev_sqrt(t);
  const T rhoto6 = (T)2 * rho_up * t * t * x;
  const T rs = (T)kRsfac * t, sqr_rs = (T)kSqRsfac * x, rsm1_2 = (T)kSqRsfacInv * rhoto6;
  const T A = (T)kA;
  // q0 = -2A (1 + A1 rs)
  J q0 = jmono<T, K>((T)-2 * A * (T)kA1 * rs, (T)(-1.0 / 3.0));
  q0.c[0] += (T)-2 * A;
  // q1 = 2A (B1 sqrt(rs) + B2 rs + B3 rs^3/2 + B4 rs^2)
  J q1 = jmono<T, K>((T)2 * A * (T)kB1 * sqr_rs, (T)(-1.0 / 6.0));
  jaxpy(q1, (T)1, jmono<T, K>((T)2 * A * (T)kB2 * rs, (T)(-1.0 / 3.0)));
  jaxpy(q1, (T)1, jmono<T, K>((T)2 * A * (T)kB3 * rs * sqr_rs, (T)-0.5));
  jaxpy(q1, (T)1, jmono<T, K>((T)2 * A * (T)kB4 * rs * rs, (T)(-2.0 / 3.0)));
  // q1p = A (B1 rs^-1/2 + 2 B2 + 3 B3 sqrt(rs) + 4 B4 rs)
  J q1p = jmono<T, K>(A * (T)kB1 * rsm1_2, (T)(1.0 / 6.0));
  q1p.c[0] += (T)2 * A * (T)kB2;
  jaxpy(q1p, (T)1, jmono<T, K>((T)3 * A * (T)kB3 * sqr_rs, (T)(-1.0 / 6.0)));
  jaxpy(q1p, (T)1, jmono<T, K>((T)4 * A * (T)kB4 * rs, (T)(-1.0 / 3.0)));
  // 1/q1, w = 1 + 1/q1 = (q1 + 1)/q1, den = 1/(q1^2 + q1) = (1/q1)^2 / w
  const J one = jmono<T, K>((T)1, (T)0);
  const J iq1 = jdiv(one, q1);
  J w = iq1;
  w.c[0] += (T)1;
  const J den = jdiv(jmul(iq1, iq1), w);
  const T q10 = q1.c[0];
  J lg = jlog_tail(w);
  lg.c[0] = -dev_log(q10 * q10 * ((T)1 / (q10 * q10 + q10)));
  // v = 4/3 ex + q0 lg - rs/3 (-2 A A1 lg - q0 q1p den)
  J v = jmono<T, K>((T)(4.0 / 3.0) * ((T)(-kThreeFourthDivPi * kSixPi2_13) * rum13 * rum13 * rho_up),
                    (T)(1.0 / 3.0));
  jaxpy(v, (T)1, jmul(q0, lg));
  J decrs = jmul(jmul(q0, q1p), den);
  jaxpy(decrs, (T)2 * A * (T)kA1, lg);
  jaxpy(v, (T)(1.0 / 3.0), jmul(jmono<T, K>(rs, (T)(-1.0 / 3.0)), decrs));
  return v;
}

// pawxcm potential moments at one point from the density moments n (lm < L)
// and the core density; returns the total spherical density (rho_updn(r,1))
__device__ __forceinline__ double xc_moments(const double (&n)[kOnsiteMaxLm], int L, double core,
                                             double (&v)[kOnsiteMaxLm]) {
  const double r0 = n[0] + kSqfpi * core;
  double rs = r0 * (0.5 / kSqfpi);
  if (rs < kDenpos) rs = kDenpos;                      // mkdenpos
  const Jet<double, 3> c = lda_v_jet<double, 3>(rs);
  const Jet<float, 5> ct = lda_v_jet<float, 5>((float)rs);
  const double vi = c.c[0];
  const double dd = kDelta * kDelta;
  const bool on = rs > kRhoMin;
  const double d1 = on ? (c.c[1] + (double)ct.c[3] * dd) / (2.0 * rs) : 0.0;
  const double d2 = on ? (c.c[2] + (double)ct.c[4] * dd) / (2.0 * rs * rs) : 0.0;
  double v1sum = 0.0;
#pragma unroll
  for (int lm = 1; lm < kOnsiteMaxLm; ++lm)
    if (lm < L) v1sum += n[lm] * n[lm];
  v[0] = vi * kSqfpi + v1sum * d2 * (0.5 / kSqfpi);
#pragma unroll
  for (int lm = 1; lm < kOnsiteMaxLm; ++lm) v[lm] = lm < L ? d1 * n[lm] : 0.0;
  return r0;
}

// moments of the (pseudo) density at radial point r >= 1 from the pair
// coefficients C (nij x L)
__device__ __forceinline__ void moments_at(const double* __restrict__ C, const double* __restrict__ pp, int nr,
                                           int P, int L, double inv2, int r, double (&n)[kOnsiteMaxLm]) {
#pragma unroll
  for (int lm = 0; lm < kOnsiteMaxLm; ++lm) n[lm] = 0.0;
  for (int k = 0; k < P; ++k) {
    const double f = pp[(long long)k * nr + r] * inv2;
#pragma unroll
    for (int lm = 0; lm < kOnsiteMaxLm; ++lm)
      if (lm < L) n[lm] += C[k * kOnsiteMaxLm + lm] * f;
  }
}

__device__ __forceinline__ void moments(const double* __restrict__ C, const double* __restrict__ pp,
                                        const double* __restrict__ inv2, int nr, int P, int L, int r,
                                        double (&n)[kOnsiteMaxLm]) {
  if (r > 0) {
    moments_at(C, pp, nr, P, L, inv2[r], r, n);
    return;
  }
  // r = 0: pawrad_deducer0, f(1) = f(4) + 3 (f(2) - f(3))
  double a[kOnsiteMaxLm], b[kOnsiteMaxLm], c[kOnsiteMaxLm];
  moments_at(C, pp, nr, P, L, inv2[1], 1, a);
  moments_at(C, pp, nr, P, L, inv2[2], 2, b);
  moments_at(C, pp, nr, P, L, inv2[3], 3, c);
#pragma unroll
  for (int lm = 0; lm < kOnsiteMaxLm; ++lm) n[lm] = c[lm] + 3.0 * (a[lm] - b[lm]);
}

// Scratch of one (atom, system): r0, v0 (all-electron), r0t, v0t (pseudo)
// at every radial point, then per chunk the kNP partial integrals and the
// chunk's low-density flag.
__host__ __device__ constexpr long long scratch_per(int nr_max) {
  return 4LL * nr_max + (long long)((nr_max + kPts - 1) / kPts) * (kNP + 1);
}

// Grid (chunk, atom, system): threads [0, kPts) evaluate the all-electron
// moments and potentials at the chunk's points, threads [kPts, 2 kPts) the
// pseudo ones; then thread q < kNP of each half sums its (pair, lm) integral
// over the chunk from LDS, and the chunk's partial (all-electron minus
// pseudo) goes to scratch.
__global__ __launch_bounds__(kOT) void k_onsite_points(const AtomDesc* __restrict__ atoms,
                                                        const OnsiteType* __restrict__ ots,
                                                        const double* __restrict__ otab,
                                                        const int* __restrict__ oint,
                                                        const double* __restrict__ rhoij, int mtot,
                                                        double* __restrict__ scratch, int nr_max) {
  const int c = blockIdx.x, a = blockIdx.y, s = blockIdx.z;
  const AtomDesc A = atoms[a];
  if (A.ot < 0) return;                                // uniform over the workgroup
  const OnsiteType T = ots[A.ot];
  const int nr = T.nr;
  const int c0 = c * kPts;
  if (c0 >= nr) return;                                // uniform: this type has fewer chunks
  const int ms = A.ms, m2 = ms * ms, L = T.lm, P = T.nij;
  const double* gnt = otab + T.gnt;
  const int* kln = oint + T.kln;
  __shared__ double rho[kMaxMs * kMaxMs];
  __shared__ double G[kOnsiteMaxLm * kMaxMs * kMaxMs];
  __shared__ int K[kMaxMs * kMaxMs];
  __shared__ double C[kNP];
  __shared__ double V[2][kPts][kOnsiteMaxLm];          // w v_lm, all-electron and pseudo
  __shared__ double Pw[2][kOnsiteMaxNij][kPts + 1];    // pair functions of the chunk (padded rows)
  __shared__ double red[2][kNP];
  __shared__ int flag;
  const double* rg = rhoij + (long long)s * mtot + A.moff;
  for (int q = threadIdx.x; q < m2; q += kOT) {
    const double v = rg[q];
    rho[q] = fabs(v) > kRhoijSel ? v : 0.0;            // set_rhoij selection
    K[q] = kln[q];
  }
  for (int q = threadIdx.x; q < L * m2; q += kOT) G[q] = gnt[q];
  if (threadIdx.x == 0) flag = 0;
  __syncthreads();
  // pair coefficients C(k, lm) = sum over (i, j) with kln(i, j) = k of
  // rho_ij G(lm, i, j): both triangles of rho, i.e. dltij times the packed
  // element, as pawdensities accumulates it
  for (int q = threadIdx.x; q < kNP; q += kOT) {
    const int k = q / kOnsiteMaxLm, lm = q % kOnsiteMaxLm;
    double acc = 0.0;
    if (k < P && lm < L)
      for (int ij = 0; ij < m2; ++ij) acc += K[ij] == k ? rho[ij] * G[lm * m2 + ij] : 0.0;
    C[q] = acc;
  }
  __syncthreads();
  const int h = threadIdx.x / kPts, i = threadIdx.x % kPts, r = c0 + i;
  const double* pw = otab + (h == 0 ? T.pp : T.tpp);
  double* sc = scratch + (long long)(s * gridDim.y + a) * scratch_per(nr_max);
  const double tolr = kDenpos * (1.0 + 1e-6);
  if (r < nr) {
    double n[kOnsiteMaxLm], v[kOnsiteMaxLm], f[kOnsiteMaxNij];
    // the point's pair functions, all loads in flight together
#pragma unroll
    for (int k = 0; k < kOnsiteMaxNij; ++k) f[k] = pw[(long long)min(k, P - 1) * nr + r];
    if (r > 0) {
      const double i2 = otab[T.inv2 + r];
#pragma unroll
      for (int lm = 0; lm < kOnsiteMaxLm; ++lm) n[lm] = 0.0;
#pragma unroll
      for (int k = 0; k < kOnsiteMaxNij; ++k) {
        if (k < P) {
          const double fk = f[k] * i2;
#pragma unroll
          for (int lm = 0; lm < kOnsiteMaxLm; ++lm)
            if (lm < L) n[lm] += C[k * kOnsiteMaxLm + lm] * fk;
        }
      }
    } else {
      moments(C, pw, otab + T.inv2, nr, P, L, r, n);
    }
    const double cr = h == 0 ? otab[T.core + r] : (T.usetcore ? otab[T.tcore + r] : 0.0);
    const double r0 = xc_moments(n, L, cr, v);
    sc[2 * h * nr_max + r] = r0;
    sc[(2 * h + 1) * nr_max + r] = v[0];
    const double wr = otab[T.w + r];
#pragma unroll
    for (int lm = 0; lm < kOnsiteMaxLm; ++lm) V[h][i][lm] = wr * v[lm];
#pragma unroll
    for (int k = 0; k < kOnsiteMaxNij; ++k) Pw[h][k][i] = k < P ? f[k] : 0.0;
    if (r0 < tolr) flag = 1;
  } else {
#pragma unroll
    for (int lm = 0; lm < kOnsiteMaxLm; ++lm) V[h][i][lm] = 0.0;
    for (int k = 0; k < kOnsiteMaxNij; ++k) Pw[h][k][i] = 0.0;
  }
  __syncthreads();
  if (i < kNP) {
    const int k = i / kOnsiteMaxLm, lm = i % kOnsiteMaxLm;
    const int len = min(kPts, nr - c0);
    double acc = 0.0;
    for (int j = 0; j < len; ++j) acc += V[h][j][lm] * Pw[h][k][j];
    red[h][i] = acc;
  }
  __syncthreads();
  double* part = sc + 4LL * nr_max + (long long)c * (kNP + 1);
  if (threadIdx.x < kNP) part[threadIdx.x] = red[0][threadIdx.x] - red[1][threadIdx.x];
  if (threadIdx.x == 0) part[kNP] = flag ? 1.0 : 0.0;
}

// Grid (atom, system): the chunk partials added in chunk order, pawxcm's
// low-density interpolation where a chunk flagged it, and
// Dij^xc(i, j) = sum_lm G(lm, i, j) I(kln(i, j), lm).
__global__ __launch_bounds__(kFin) void k_onsite_finish(const AtomDesc* __restrict__ atoms,
                                                         const OnsiteType* __restrict__ ots,
                                                         const double* __restrict__ otab,
                                                         const int* __restrict__ oint, int mtot,
                                                         const double* __restrict__ scratch, int nr_max,
                                                         double* __restrict__ dxc) {
  const int a = blockIdx.x, s = blockIdx.y;
  const AtomDesc A = atoms[a];
  if (A.ot < 0) return;                                // uniform over the workgroup
  const OnsiteType T = ots[A.ot];
  const int ms = A.ms, m2 = ms * ms, L = T.lm, P = T.nij, nr = T.nr;
  const int nch = (nr + kPts - 1) / kPts;
  const double* gnt = otab + T.gnt;
  const int* kln = oint + T.kln;
  const double* sc = scratch + (long long)(s * gridDim.x + a) * scratch_per(nr_max);
  const double* part = sc + 4LL * nr_max;
  __shared__ double I[kNP];
  __shared__ double G[kOnsiteMaxLm * kMaxMs * kMaxMs];
  __shared__ int K[kMaxMs * kMaxMs];
  __shared__ int flag;
  if (threadIdx.x == 0) flag = 0;
  for (int q = threadIdx.x; q < m2; q += kFin) K[q] = kln[q];
  for (int q = threadIdx.x; q < L * m2; q += kFin) G[q] = gnt[q];
  __syncthreads();
  if (threadIdx.x < nch && part[(long long)threadIdx.x * (kNP + 1) + kNP] != 0.0) flag = 1;
  for (int q = threadIdx.x; q < kNP; q += kFin) {
    double t = 0.0;
    for (int cc = 0; cc < nch; ++cc) t += part[(long long)cc * (kNP + 1) + q];
    I[q] = t;
  }
  __syncthreads();
  if (flag && threadIdx.x == 0) {
    // pawxcm's interpolation of v_00 across points with rho < tol_rho that
    // lie between denser points: the correction of those points' share of
    // the l = 0 integrals, for the all-electron (h = 0) and pseudo (h = 1)
    // moments
    const double* rad = otab + T.rad;
    const double* w = otab + T.w;
    const double tolr = kDenpos * (1.0 + 1e-6);
    for (int h = 0; h < 2; ++h) {
      const double* r0v = sc + 2 * h * nr_max;
      const double* v0 = sc + (2 * h + 1) * nr_max;
      const double sign = h == 0 ? 1.0 : -1.0;
      const double* pw = otab + (h == 0 ? T.pp : T.tpp);
      int ir1 = 0, ir2 = 0;
      for (int ir = 1; ir <= nr; ++ir) {
        if (r0v[ir - 1] < tolr) {
          if (ir1 == 0) ir1 = ir - 1;
          ir2 = ir + 1;
        } else if (ir1 > 0) {
          if (ir1 > 1 || ir2 < nr) {
            const double fct = (v0[ir2 - 1] - v0[ir1 - 1]) / (rad[ir2 - 1] - rad[ir1 - 1]);
            for (int jr = ir1 + 1; jr < ir2; ++jr) {
              const double vn = v0[ir1 - 1] + fct * (rad[jr - 1] - rad[ir1 - 1]);
              const double dvv = (vn - v0[jr - 1]) * w[jr - 1] * sign;
              for (int k = 0; k < P; ++k) I[k * kOnsiteMaxLm] += dvv * pw[(long long)k * nr + jr - 1];
            }
          }
          ir1 = ir2 = 0;
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < m2; q += kFin) {
    const int k = K[q];
    double d = 0.0;
    for (int lm = 0; lm < L; ++lm) d += G[lm * m2 + q] * I[k * kOnsiteMaxLm + lm];
    dxc[(long long)s * mtot + A.moff + q] = d;
  }
}

}  // namespace

hipError_t onsite_xc(hipStream_t st, const AtomDesc* atoms, int natom, const OnsiteType* ot, const double* otab,
                     const int* oint, const double* rhoij, int mtot, int nsys, double* scratch, int nr_max,
                     double* dxc) {
  if (natom == 0 || !ot) return hipSuccess;
  const int nch = (nr_max + kPts - 1) / kPts;
  if (nch > kFin) return hipErrorInvalidValue;         // the finish kernel reads one flag per thread
  hipLaunchKernelGGL(k_onsite_points, dim3(nch, natom, nsys), dim3(kOT), 0, st, atoms, ot, otab, oint, rhoij, mtot,
                     scratch, nr_max);
  hipLaunchKernelGGL(k_onsite_finish, dim3(natom, nsys), dim3(kFin), 0, st, atoms, ot, otab, oint, mtot, scratch,
                     nr_max, dxc);
  return hipGetLastError();
}

size_t onsite_scratch_doubles(int nr_max) { return (size_t)scratch_per(nr_max); }

}  // namespace opaw
