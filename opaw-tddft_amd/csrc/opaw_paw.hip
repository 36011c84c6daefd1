// Grid-sized part of the PAW set-up (prep_paw, src/prepare_paw.f90):
//   opaw_paw_box     get_local_p3d (src/p3d.f90:2-114): projectors sampled on
//                    the fine grid around an atom and contracted onto the
//                    rough box with the spline B matrices of get_b
//   opaw_paw_sphere  map_localgrid + get_localg3d (src/map_localgrid.f90,
//                    src/g3d.f90): grid points within rcut of an atom, in
//                    grid order, and the shape functions g_L Y_L on them
// Both run once per atom at set-up.  The fine-grid sampling writes one value
// per (m-state, fine point) and the contraction is three passes of one
// small matrix along one axis each; the sphere is a stream compaction over
// the grid (block counts, host scan, ordered fill with wave ballots).
#include <cmath>
#include <vector>

#include "opaw_common.h"
#include "opaw_hip.h"

namespace opaw {
int io_fail(int code, const char* fmt, ...);
}

namespace {

// Fortran default-real literal of src/ylm.f90:13
constexpr double kPiYlm = (double)3.14159265359f;

// real spherical harmonics of src/ylm.f90:4-107 (l <= 4)
__device__ double ylm_dev(int l, int m, double xx, double yy, double zz) {
  const double pi = kPiYlm;
  const double p1 = 1.0 / (2.0 * sqrt(pi));
  switch (l) {
    case 0:
      return p1;
    case 1: {
      const double p2 = p1 * sqrt(3.0);
      return m == -1 ? p2 * yy : (m == 0 ? p2 * zz : p2 * xx);
    }
    case 2: {
      const double p3 = sqrt(15.0) / (2.0 * sqrt(pi));
      switch (m) {
        case -2: return p3 * xx * yy;
        case -1: return p3 * yy * zz;
        case 0: return p3 * (2.0 * zz * zz - xx * xx - yy * yy) / (2.0 * sqrt(3.0));
        case 1: return p3 * xx * zz;
        default: return p3 * (xx * xx - yy * yy) / 2.0;
      }
    }
    case 3:
      switch (m) {
        case -3: return 1.0 / 4.0 * sqrt(35.0 / 2.0 / pi) * yy * (3.0 * xx * xx - yy * yy);
        case -2: return 1.0 / 2.0 * sqrt(105.0 / pi) * xx * yy * zz;
        case -1: return 1.0 / 4.0 * sqrt(21.0 / 2.0 / pi) * yy * (4.0 * zz * zz - xx * xx - yy * yy);
        case 0: return 1.0 / 4.0 * sqrt(7.0 / pi) * zz * (2.0 * zz * zz - 3.0 * xx * xx - 3.0 * yy * yy);
        case 1: return 1.0 / 4.0 * sqrt(21.0 / 2.0 / pi) * xx * (4.0 * zz * zz - xx * xx - yy * yy);
        case 2: return 1.0 / 4.0 * sqrt(105.0 / pi) * (xx * xx - yy * yy) * zz;
        default: return 1.0 / 4.0 * sqrt(35.0 / 2.0 / pi) * xx * (xx * xx - 3.0 * yy * yy);
      }
    default:
      switch (m) {
        case -4: return 3.0 / 4.0 * sqrt(35.0 / pi) * xx * yy * (xx * xx - yy * yy);
        case -3: return 3.0 / 4.0 * sqrt(35.0 / 2.0 / pi) * (3.0 * xx * xx - yy * yy) * yy * zz;
        case -2: return 3.0 / 4.0 * sqrt(5.0 / pi) * xx * yy * (7.0 * zz * zz - 1.0);
        case -1: return 3.0 / 4.0 * sqrt(5.0 / 2.0 / pi) * yy * zz * (7.0 * zz * zz - 3.0);
        case 0: return 3.0 / 16.0 / sqrt(pi) * (35.0 * zz * zz * zz * zz - 30.0 * zz * zz + 3.0);
        case 1: return 3.0 / 4.0 * sqrt(5.0 / 2.0 / pi) * xx * zz * (7.0 * zz * zz - 3.0);
        case 2: return 3.0 / 8.0 * sqrt(5.0 / pi) * (xx * xx - yy * yy) * (7.0 * zz * zz - 1.0);
        case 3: return 3.0 / 4.0 * sqrt(35.0 / 2.0 / pi) * (xx * xx - 3.0 * yy * yy) * xx * zz;
        default:
          return 3.0 / 16.0 * sqrt(35.0 / pi) *
                 (xx * xx * (xx * xx - 3.0 * yy * yy) - yy * yy * (3.0 * xx * xx - yy * yy));
      }
  }
}

// minbyabs of src/p3d.f90:230-246: the image of smallest magnitude
__device__ __forceinline__ double min_image(double x, double c, double h) {
  double v = x - c;
  const double a = x + 2.0 * h - c, b = x - 2.0 * h - c;
  if (fabs(a) < fabs(v)) v = a;
  if (fabs(b) < fabs(v)) v = b;
  return v;
}

// src/interpolate.f90: rr[i] <= t < rr[j] by bisection (0-based), weights
__device__ __forceinline__ void bracket(const double* __restrict__ rr, int nr, double t, int& i, int& j, double& wi,
                                       double& wj) {
  i = 0;
  j = nr - 1;
  while (j - i > 1) {
    const int jm = (i + j + 2) / 2 - 1;   // (i+j)/2 on 1-based indices
    if (t < rr[jm])
      j = jm;
    else
      i = jm;
  }
  const double den = rr[j] - rr[i];
  wi = (rr[j] - t) / den;
  wj = (t - rr[i]) / den;
}

// fine-grid projectors: out[s*nf + p] (p = a + nf1 (b + nf2 c)) for the
// ms states of rad and then the ms states of rad1
__global__ __launch_bounds__(256) void k_box_fine(const double* __restrict__ rr, int nr,
                                                   const double* __restrict__ rad, const double* __restrict__ rad1,
                                                   const int* __restrict__ lm, int ms, int nf1, int nf2, int nf3,
                                                   double dxf, double dyf, double dzf, double x0, double y0,
                                                   double z0, double cx, double cy, double cz, double hx, double hy,
                                                   double hz, double* __restrict__ out) {
  const long long nf = (long long)nf1 * nf2 * nf3;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nf) return;
  const int a = (int)(p % nf1);
  const int b = (int)((p / nf1) % nf2);
  const int c = (int)(p / ((long long)nf1 * nf2));
  const double x = x0 + a * dxf, y = y0 + b * dyf, z = z0 + c * dzf;
  const double xc = min_image(x, cx, hx), yc = min_image(y, cy, hy), zc = min_image(z, cz, hz);
  const double r = fmax(sqrt(xc * xc + yc * yc + zc * zc), 1e-8);
  const double xx = xc / r, yy = yc / r, zz = zc / r;
  int i, j;
  double wi, wj;
  bracket(rr, nr, r, i, j, wi, wj);
  for (int s = 0; s < ms; ++s) {
    const double y_lm = ylm_dev(lm[2 * s], lm[2 * s + 1], xx, yy, zz);
    const double* rs = rad + (size_t)s * nr;
    const double* r1 = rad1 + (size_t)s * nr;
    out[(size_t)s * nf + p] = (rs[i] * wi + rs[j] * wj) * y_lm;
    out[(size_t)(ms + s) * nf + p] = (r1[i] * wi + r1[j] * wj) * y_lm;
  }
}

// contraction along one axis: in (d0, d1, d2, S) x fastest -> out with the
// axis length nf replaced by nrgh; B is Fortran b(nrgh, nf)
__global__ __launch_bounds__(256) void k_box_contract(const double* __restrict__ in, double* __restrict__ out,
                                                       const double* __restrict__ B, int axis, int d0, int d1,
                                                       int d2, int S, int nrgh) {
  int o[3] = {d0, d1, d2};
  const int nf = o[axis];
  o[axis] = nrgh;
  const long long total = (long long)o[0] * o[1] * o[2] * S;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int i0 = (int)(q % o[0]);
  const int i1 = (int)((q / o[0]) % o[1]);
  const int i2 = (int)((q / ((long long)o[0] * o[1])) % o[2]);
  const int s = (int)(q / ((long long)o[0] * o[1] * o[2]));
  int idx[3] = {i0, i1, i2};
  const int rgh = idx[axis];
  long long stride = 1;
  if (axis >= 1) stride *= d0;
  if (axis >= 2) stride *= d1;
  idx[axis] = 0;
  const double* src = in + (size_t)s * d0 * d1 * d2 + idx[0] + (long long)d0 * (idx[1] + (long long)d1 * idx[2]);
  double acc = 0.0;
  for (int f = 0; f < nf; ++f) acc = fma(B[rgh + (size_t)f * nrgh], src[(long long)f * stride], acc);
  out[q] = acc;
}

// sphere membership as in map_localgrid's get_r (src/map_localgrid.f90:84-92)
__device__ __forceinline__ bool in_sphere(long long p, int nx, int ny, double dx, double dy, double dz, double cx,
                                          double cy, double cz, double hx, double hy, double hz, double rcut) {
  const int ix = (int)(p % nx), iy = (int)((p / nx) % ny), iz = (int)(p / ((long long)nx * ny));
  const double x = -hx + ix * dx, y = -hy + iy * dy, z = -hz + iz * dz;
  const double xc = fmin(fmin(fabs(x - cx), fabs(x + hx * 2.0 - cx)), fabs(x - hx * 2.0 - cx));
  const double yc = fmin(fmin(fabs(y - cy), fabs(y + hy * 2.0 - cy)), fabs(y - hy * 2.0 - cy));
  const double zc = fmin(fmin(fabs(z - cz), fabs(z + hz * 2.0 - cz)), fabs(z - hz * 2.0 - cz));
  return sqrt(xc * xc + yc * yc + zc * zc) < rcut;
}

constexpr int kSphBlock = 256;

__global__ __launch_bounds__(kSphBlock) void k_sph_count(long long n, int nx, int ny, double dx, double dy, double dz,
                                                         double cx, double cy, double cz, double hx, double hy,
                                                         double hz, double rcut, int* __restrict__ counts) {
  __shared__ int wcount[kSphBlock / 64];
  const long long p = (long long)blockIdx.x * kSphBlock + threadIdx.x;
  const bool in = p < n && in_sphere(p, nx, ny, dx, dy, dz, cx, cy, cz, hx, hy, hz, rcut);
  const unsigned long long m = __ballot(in);
  if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kSphBlock / 64; ++w) t += wcount[w];
    counts[blockIdx.x] = t;
  }
}

// ordered fill: block offsets from the host scan, wave offsets from LDS,
// lane offsets from the ballot.  grid(ng,3) 1-based and g3d(ng,nl2) are
// Fortran layouts (local_grid, local_g3d of src/3_atom.f90:24-26)
__global__ __launch_bounds__(kSphBlock) void k_sph_fill(long long n, int nx, int ny, double dx, double dy, double dz,
                                                        double cx, double cy, double cz, double hx, double hy,
                                                        double hz, double rcut, double rcomp,
                                                        const double* __restrict__ rr, int nr,
                                                        const double* __restrict__ gl, int nrc, int nl2,
                                                        const int* __restrict__ offsets, int ng,
                                                        int* __restrict__ grid, double* __restrict__ g3d) {
  __shared__ int wcount[kSphBlock / 64];
  const long long p = (long long)blockIdx.x * kSphBlock + threadIdx.x;
  const bool in = p < n && in_sphere(p, nx, ny, dx, dy, dz, cx, cy, cz, hx, hy, hz, rcut);
  const unsigned long long m = __ballot(in);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wcount[wave] = __popcll(m);
  __syncthreads();
  if (!in) return;
  int off = offsets[blockIdx.x];
  for (int w = 0; w < wave; ++w) off += wcount[w];
  off += __popcll(m & ((1ull << lane) - 1ull));
  if (off >= ng) return;   // the host sized ng from the counts: never taken
  const int ix = (int)(p % nx), iy = (int)((p / nx) % ny), iz = (int)(p / ((long long)nx * ny));
  grid[off] = ix + 1;
  grid[off + ng] = iy + 1;
  grid[off + 2 * (size_t)ng] = iz + 1;
  // get_localg3d's get_xyz / get_radialpart / get_angularpart (src/g3d.f90:44-117)
  const double x = -hx + ix * dx, y = -hy + iy * dy, z = -hz + iz * dz;
  const double xc = min_image(x, cx, hx), yc = min_image(y, cy, hy), zc = min_image(z, cz, hz);
  const double r = fmax(sqrt(xc * xc + yc * yc + zc * zc), 1e-8);
  const double xx = xc / r, yy = yc / r, zz = zc / r;
  if (r > rcomp) {
    for (int L = 0; L < nl2; ++L) g3d[off + (size_t)L * ng] = 0.0;
    return;
  }
  int i, j;
  double wi, wj;
  bracket(rr, nr, r, i, j, wi, wj);
  i = min(i, nrc - 1);
  j = min(j, nrc - 1);
  for (int L = 0; L < nl2; ++L) {
    const int l = (int)floor(sqrt((double)L));
    const int mm = L - l * l - l;
    const double* g = gl + (size_t)l * nrc;
    g3d[off + (size_t)L * ng] = (g[i] * wi + g[j] * wj) * ylm_dev(l, mm, xx, yy, zz);
  }
}

struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

#define PCHK(expr)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return opaw::io_fail(OPAW_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(Dev& d, size_t count) {
  hipError_t e = hipMalloc(&d.p, count * sizeof(T));
  if (e != hipSuccess) return opaw::io_fail(OPAW_ERR_ALLOC, "paw set-up buffer of %zu bytes", count * sizeof(T));
  return OPAW_OK;
}

}  // namespace

extern "C" int opaw_paw_box(int nr, const double* rr, int ms, const double* rad, const double* rad1, const int* lm,
                            const int* nrough, const int* nfine, const double* dxf, const double* corner,
                            const double* coord, const double* half, const double* bx, const double* by,
                            const double* bz, double* p3d, double* p3d1) {
  using opaw::io_fail;
  if (!rr || !rad || !rad1 || !lm || !nrough || !nfine || !dxf || !corner || !coord || !half || !bx || !by || !bz ||
      !p3d || !p3d1)
    return io_fail(OPAW_ERR_ARG, "null argument");
  if (nr < 2 || ms < 1) return io_fail(OPAW_ERR_ARG, "nr >= 2 and ms >= 1 required");
  for (int k = 0; k < 3; ++k)
    if (nrough[k] < 2 || nfine[k] < 2 || nfine[k] > 4096 || nrough[k] > 4096)
      return io_fail(OPAW_ERR_ARG, "bad box shape");
  for (int s = 0; s < ms; ++s)
    if (lm[2 * s] < 0 || lm[2 * s] > 4 || lm[2 * s + 1] < -lm[2 * s] || lm[2 * s + 1] > lm[2 * s])
      return io_fail(OPAW_ERR_ARG, "bad (l, m) of state %d", s + 1);
  const long long nf = (long long)nfine[0] * nfine[1] * nfine[2];
  const int S = 2 * ms;
  Dev drr, drad, drad1, dlm, dbx, dby, dbz, da, db;
  int rc;
  if ((rc = dalloc<double>(drr, nr)) || (rc = dalloc<double>(drad, (size_t)ms * nr)) ||
      (rc = dalloc<double>(drad1, (size_t)ms * nr)) || (rc = dalloc<int>(dlm, 2 * (size_t)ms)) ||
      (rc = dalloc<double>(dbx, (size_t)nrough[0] * nfine[0])) ||
      (rc = dalloc<double>(dby, (size_t)nrough[1] * nfine[1])) ||
      (rc = dalloc<double>(dbz, (size_t)nrough[2] * nfine[2])) || (rc = dalloc<double>(da, (size_t)S * nf)) ||
      (rc = dalloc<double>(db, (size_t)S * nrough[0] * nfine[1] * nfine[2])))
    return rc;
  PCHK(hipMemcpy(drr.p, rr, nr * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(drad.p, rad, (size_t)ms * nr * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(drad1.p, rad1, (size_t)ms * nr * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(dlm.p, lm, 2 * (size_t)ms * sizeof(int), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(dbx.p, bx, (size_t)nrough[0] * nfine[0] * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(dby.p, by, (size_t)nrough[1] * nfine[1] * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(dbz.p, bz, (size_t)nrough[2] * nfine[2] * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_box_fine, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, 0, (const double*)drr.p, nr,
                     (const double*)drad.p, (const double*)drad1.p, (const int*)dlm.p, ms, nfine[0], nfine[1],
                     nfine[2], dxf[0], dxf[1], dxf[2], corner[0], corner[1], corner[2], coord[0], coord[1], coord[2],
                     half[0], half[1], half[2], (double*)da.p);
  PCHK(hipGetLastError());
  // x: (nf1,nf2,nf3) -> (nr1,nf2,nf3); y: -> (nr1,nr2,nf3); z: -> (nr1,nr2,nr3)
  const int d[3][3] = {{nfine[0], nfine[1], nfine[2]}, {nrough[0], nfine[1], nfine[2]},
                       {nrough[0], nrough[1], nfine[2]}};
  const Dev* bm[3] = {&dbx, &dby, &dbz};
  double* bufs[2] = {(double*)da.p, (double*)db.p};
  for (int axis = 0; axis < 3; ++axis) {
    long long tot = (long long)S * nrough[axis];
    for (int k = 0; k < 3; ++k)
      if (k != axis) tot *= d[axis][k];
    hipLaunchKernelGGL(k_box_contract, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0,
                       (const double*)bufs[axis & 1], bufs[(axis + 1) & 1], (const double*)bm[axis]->p, axis,
                       d[axis][0], d[axis][1], d[axis][2], S, nrough[axis]);
    PCHK(hipGetLastError());
  }
  // after three passes the result is in bufs[1] (x: a->b, y: b->a, z: a->b)
  const size_t nb = (size_t)nrough[0] * nrough[1] * nrough[2];
  PCHK(hipMemcpy(p3d, bufs[1], (size_t)ms * nb * sizeof(double), hipMemcpyDeviceToHost));
  PCHK(hipMemcpy(p3d1, bufs[1] + (size_t)ms * nb, (size_t)ms * nb * sizeof(double), hipMemcpyDeviceToHost));
  return OPAW_OK;
}

extern "C" int opaw_paw_sphere(int nx, int ny, int nz, double dx, double dy, double dz, const double* coord,
                               double rcut, double rcomp, int nr, const double* rr, int nl2, int nrc,
                               const double* gl, int max_points, int* ngrid, int* local_grid, double* local_g3d) {
  using opaw::io_fail;
  if (!coord || !rr || !gl || !ngrid) return io_fail(OPAW_ERR_ARG, "null argument");
  if (nx < 1 || ny < 1 || nz < 1 || nr < 2 || nrc < 2 || nrc > nr || nl2 < 1 || nl2 > 25)
    return io_fail(OPAW_ERR_ARG, "bad sphere arguments");
  const int nlg = (int)std::lround(std::sqrt((double)nl2));
  if (nlg * nlg != nl2) return io_fail(OPAW_ERR_ARG, "nl2 must be (2 nl - 1)^2");
  const long long n = (long long)nx * ny * nz;
  const long long nblk = (n + kSphBlock - 1) / kSphBlock;
  if (nblk > (1LL << 31)) return io_fail(OPAW_ERR_ARG, "grid too large");
  const double hx = nx * dx / 2.0, hy = ny * dy / 2.0, hz = nz * dz / 2.0;
  Dev dcounts;
  int rc;
  if ((rc = dalloc<int>(dcounts, (size_t)nblk))) return rc;
  hipLaunchKernelGGL(k_sph_count, dim3((unsigned)nblk), dim3(kSphBlock), 0, 0, n, nx, ny, dx, dy, dz, coord[0],
                     coord[1], coord[2], hx, hy, hz, rcut, (int*)dcounts.p);
  PCHK(hipGetLastError());
  std::vector<int> counts((size_t)nblk);
  PCHK(hipMemcpy(counts.data(), dcounts.p, (size_t)nblk * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> offsets((size_t)nblk);
  long long total = 0;
  for (long long b = 0; b < nblk; ++b) {
    offsets[(size_t)b] = (int)total;
    total += counts[(size_t)b];
  }
  *ngrid = (int)total;
  if (max_points <= 0) return OPAW_OK;   // size query
  if (total > max_points) return io_fail(OPAW_ERR_ARG, "sphere has %lld points, room for %d", total, max_points);
  if (!local_grid || !local_g3d) return io_fail(OPAW_ERR_ARG, "null output");
  if (total == 0) return OPAW_OK;
  Dev drr, dgl, doff, dgrid, dg3d;
  if ((rc = dalloc<double>(drr, nr)) || (rc = dalloc<double>(dgl, (size_t)nlg * nrc)) ||
      (rc = dalloc<int>(doff, (size_t)nblk)) || (rc = dalloc<int>(dgrid, 3 * (size_t)total)) ||
      (rc = dalloc<double>(dg3d, (size_t)nl2 * total)))
    return rc;
  PCHK(hipMemcpy(drr.p, rr, nr * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(dgl.p, gl, (size_t)nlg * nrc * sizeof(double), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(doff.p, offsets.data(), (size_t)nblk * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sph_fill, dim3((unsigned)nblk), dim3(kSphBlock), 0, 0, n, nx, ny, dx, dy, dz, coord[0],
                     coord[1], coord[2], hx, hy, hz, rcut, rcomp, (const double*)drr.p, nr, (const double*)dgl.p, nrc,
                     nl2, (const int*)doff.p, (int)total, (int*)dgrid.p, (double*)dg3d.p);
  PCHK(hipGetLastError());
  PCHK(hipMemcpy(local_grid, dgrid.p, 3 * (size_t)total * sizeof(int), hipMemcpyDeviceToHost));
  PCHK(hipMemcpy(local_g3d, dg3d.p, (size_t)nl2 * total * sizeof(double), hipMemcpyDeviceToHost));
  return OPAW_OK;
}

// ---------------------------------------------------------------------------
// local potential of vloc_tuma_prep (nonperiodic, Martyna-Tuckerman box)
#include "opaw_paw.h"

namespace opaw {
namespace {

__global__ __launch_bounds__(256) void k_vloc_sample(const double* __restrict__ rr, const double* __restrict__ xs,
                                                     const double* __restrict__ ys, const double* __restrict__ y2s,
                                                     int nr, double charge, double dr1, int nxb, int nyb, int nzb,
                                                     double dx, double dy, double dz, c64* __restrict__ cv) {
  const long long n = (long long)nxb * nyb * nzb;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % nxb), iy = (int)((p / nxb) % nyb), iz = (int)(p / ((long long)nxb * nyb));
  const double x = (ix - nxb / 2) * dx, y = (iy - nyb / 2) * dy, z = (iz - nzb / 2) * dz;
  const double rs = fmax(1e-10, sqrt(x * x + y * y + z * z));
  const double rl = fmax(rs, 0.1 * dr1);
  double v;
  if (rs < rr[0]) {
    v = ys[0];
  } else if (rs > rr[nr - 1]) {
    v = charge / rs;
  } else {
    // SPLINT_dn (src/lib/library_NRF_splint.f:6-33)
    int klo = 0, khi = nr - 1;
    while (khi - klo > 1) {
      const int k = (khi + klo) / 2;
      if (xs[k] > rl)
        khi = k;
      else
        klo = k;
    }
    const double h = xs[khi] - xs[klo];
    const double a = (xs[khi] - rl) / h, b = (rl - xs[klo]) / h;
    v = a * ys[klo] + b * ys[khi] + ((a * a * a - a) * y2s[klo] + (b * b * b - b) * y2s[khi]) * (h * h) / 6.0;
  }
  cv[p] = mk(v, 0.0);
}

__device__ __forceinline__ double fold_k(int i, int n, double d) {
  const double pi = 3.14159265358979323846;
  double k = (double)i * (2.0 * pi / ((double)n * d));
  if (k > pi / d) k -= 2.0 * pi / d;
  return k;
}

__global__ __launch_bounds__(256) void k_vloc_accumulate(const c64* __restrict__ cvkb, c64* __restrict__ cvn, int nxb,
                                                         int nyb, int nzb, double dx, double dy, double dz,
                                                         const double* __restrict__ coords, int natom) {
  const long long n = (long long)nxb * nyb * nzb;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % nxb), iy = (int)((p / nxb) % nyb), iz = (int)(p / ((long long)nxb * nyb));
  const double kx = fold_k(ix, nxb, dx), ky = fold_k(iy, nyb, dy), kz = fold_k(iz, nzb, dz);
  // shift_cvkb: exp(-i r0.k), r0 = -(n/2) d
  const double r0x = -(double)nxb / 2.0 * dx, r0y = -(double)nyb / 2.0 * dy, r0z = -(double)nzb / 2.0 * dz;
  double s, c;
  sincos(-(r0x * kx + r0y * ky + r0z * kz), &s, &c);
  // add_formfactor: sum over atoms of exp(-i (R - rstart).k), zero on the
  // Nyquist planes, rstart = -(n d)/2
  double fr = 0.0, fi = 0.0;
  if (ix != nxb / 2 && iy != nyb / 2 && iz != nzb / 2) {
    const double sx = -0.5 * dx * nxb, sy = -0.5 * dy * nyb, sz = -0.5 * dz * nzb;
    for (int a = 0; a < natom; ++a) {
      double cs, sn;
      sincos(-((coords[3 * a] - sx) * kx), &sn, &cs);
      double cs2, sn2;
      sincos(-((coords[3 * a + 1] - sy) * ky + (coords[3 * a + 2] - sz) * kz), &sn2, &cs2);
      fr += cs * cs2 - sn * sn2;
      fi += cs * sn2 + sn * cs2;
    }
  }
  const double dv = dx * dy * dz;
  const c64 v = cvkb[p];
  // (v dv) (c + i s) (fr + i fi)
  const double ar = (v.x * c - v.y * s) * dv, ai = (v.x * s + v.y * c) * dv;
  const c64 o = cvn[p];
  cvn[p] = mk(o.x + ar * fr - ai * fi, o.y + ar * fi + ai * fr);
}

__global__ __launch_bounds__(256) void k_vloc_crop(const c64* __restrict__ cvn, int nxb, int nyb, int nzb, int nx,
                                                   int ny, int nz, double scale, double* __restrict__ vloc) {
  const long long n = (long long)nx * ny * nz;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ix = (int)(p % nx), iy = (int)((p / nx) % ny), iz = (int)(p / ((long long)nx * ny));
  const int bx = nxb / 2 - nx / 2 + ix, by = nyb / 2 - ny / 2 + iy, bz = nzb / 2 - nz / 2 + iz;
  vloc[p] = cvn[(size_t)bx + (size_t)nxb * ((size_t)by + (size_t)nyb * bz)].x * scale;
}

}  // namespace

hipError_t vloc_sample(hipStream_t st, const double* rr, const double* x, const double* y, const double* y2, int nr,
                       double charge, double dr1, int nxb, int nyb, int nzb, double dx, double dy, double dz,
                       c64* cv) {
  const long long n = (long long)nxb * nyb * nzb;
  hipLaunchKernelGGL(k_vloc_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rr, x, y, y2, nr, charge, dr1,
                     nxb, nyb, nzb, dx, dy, dz, cv);
  return hipGetLastError();
}

hipError_t vloc_accumulate(hipStream_t st, const c64* cvkb, c64* cvn, int nxb, int nyb, int nzb, double dx,
                           double dy, double dz, const double* coords, int natom) {
  const long long n = (long long)nxb * nyb * nzb;
  hipLaunchKernelGGL(k_vloc_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cvkb, cvn, nxb, nyb, nzb,
                     dx, dy, dz, coords, natom);
  return hipGetLastError();
}

hipError_t vloc_crop(hipStream_t st, const c64* cvn, int nxb, int nyb, int nzb, int nx, int ny, int nz, double scale,
                     double* vloc) {
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_vloc_crop, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cvn, nxb, nyb, nzb, nx, ny, nz,
                     scale, vloc);
  return hipGetLastError();
}

}  // namespace opaw
