// Rayleigh-Ritz of the block solver on the device.  The pencil is small
// (m = 2 or 3 times the band count) and latency-bound; the gain is that the
// Gram matrices, the eigenvectors and the rotation coefficients stay in HBM
// and the solver's host round trip disappears.
//   rr_prep    unit-diagonal scaling D S D of the Hermitian part of S and the
//              Hermitian part of H, both column-major for rocSOLVER
//   zheevd     eigenpairs (l_k, u_k) of D S D
//   rr_basis   Xd[:, k] = D u_k / sqrt(l_k) for l_k > floor * l_max, else 0;
//              the kept rank
//   zgemm x2   Hp = Xd^H H Xd
//   rr_pad     dropped directions (zero rows and columns of Hp) get a
//              diagonal above the Frobenius norm of Hp, so they sort after
//              every kept eigenvalue
//   zheevd     eigenpairs (w, v) of Hp
//   zgemm      C = Xd v[:, :nout]
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "opaw_rr.h"

namespace opaw {
namespace {

constexpr int kRrT = 256;

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

struct RrLayout {
  size_t as, ah, t, hp, lam, wv, e, dsc, flags, total;
  explicit RrLayout(int m) {
    const size_t mm = align256((size_t)m * m * sizeof(c64));
    const size_t vm = align256((size_t)m * sizeof(double));
    as = 0;
    ah = as + mm;
    t = ah + mm;
    hp = t + mm;
    lam = hp + mm;
    wv = lam + vm;
    e = wv + vm;
    dsc = e + vm;
    flags = dsc + vm;
    total = flags + 256;
  }
};

__device__ __forceinline__ double unit_scale(const c64* S, int m, int i) {
  const double s = S[(size_t)i * m + i].x;
  return s > 0.0 ? 1.0 / sqrt(s) : 0.0;
}

// thread (a, b): column-major AS[a + m b] = d_a d_b (S(a,b) + conj S(b,a)) / 2
// and AH[a + m b] = (H(a,b) + conj H(b,a)) / 2 from the row-major inputs
__global__ __launch_bounds__(kRrT) void k_rr_prep(const c64* __restrict__ S, const c64* __restrict__ H, int m,
                                                  c64* __restrict__ AS, c64* __restrict__ AH,
                                                  double* __restrict__ dsc) {
  const int a = blockIdx.x * 16 + (threadIdx.x & 15);
  const int b = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (a >= m || b >= m) return;
  const double da = unit_scale(S, m, a), db = unit_scale(S, m, b);
  const c64 sab = S[(size_t)a * m + b], sba = S[(size_t)b * m + a];
  const c64 hab = H[(size_t)a * m + b], hba = H[(size_t)b * m + a];
  const double f = 0.5 * da * db;
  AS[a + (size_t)m * b] = mk((sab.x + sba.x) * f, (sab.y - sba.y) * f);
  AH[a + (size_t)m * b] = mk(0.5 * (hab.x + hba.x), 0.5 * (hab.y - hba.y));
  if (b == 0) dsc[a] = da;
}

__device__ __forceinline__ bool rr_kept(const double* lam, int m, int k, double floor) {
  const double top = lam[m - 1];
  return top > 0.0 && lam[k] > floor * top;
}

// block k: column k of the eigenvectors U (in place) becomes D u_k / sqrt(l_k)
// or zero; block 0 also counts the kept directions
__global__ __launch_bounds__(kRrT) void k_rr_basis(c64* __restrict__ U, const double* __restrict__ lam,
                                                   const double* __restrict__ dsc, int m, double floor,
                                                   int* __restrict__ flags) {
  const int k = blockIdx.x;
  const bool keep = rr_kept(lam, m, k, floor);
  const double s = keep ? 1.0 / sqrt(lam[k]) : 0.0;
  c64* col = U + (size_t)m * k;
  for (int i = threadIdx.x; i < m; i += kRrT) col[i] = cscale(col[i], s * dsc[i]);
  if (k != 0) return;
  __shared__ int cnt[kRrT];
  int n = 0;
  for (int j = threadIdx.x; j < m; j += kRrT) n += rr_kept(lam, m, j, floor) ? 1 : 0;
  cnt[threadIdx.x] = n;
  __syncthreads();
  for (int w = kRrT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) cnt[threadIdx.x] += cnt[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) flags[0] = cnt[0];
}

// one block: Hp(k, k) = |Hp|_F + 1 for every dropped k (its row and column
// are zero, Xd's column being zero)
__global__ __launch_bounds__(kRrT) void k_rr_pad(c64* __restrict__ Hp, const double* __restrict__ lam, int m,
                                                 double floor) {
  __shared__ double red[kRrT];
  double s = 0.0;
  const size_t mm = (size_t)m * m;
  for (size_t q = threadIdx.x; q < mm; q += kRrT) s += Hp[q].x * Hp[q].x + Hp[q].y * Hp[q].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRrT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double big = sqrt(red[0]) + 1.0;
  for (int k = threadIdx.x; k < m; k += kRrT)
    if (!rr_kept(lam, m, k, floor)) Hp[k + (size_t)m * k] = mk(big, 0.0);
}

#define RRB(expr)                                      \
  do {                                                 \
    rocblas_status s_ = (expr);                        \
    if (s_ != rocblas_status_success) return (int)s_;  \
  } while (0)
#define RRH(expr)                                      \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return -(int)e_;             \
  } while (0)

}  // namespace

size_t rr_work_bytes(int m) { return RrLayout(m).total; }

int* rr_flags(void* work, int m) { return (int*)((char*)work + RrLayout(m).flags); }

int rr_lowest(void** blas, hipStream_t st, int m, int nout, double floor, const c64* S, const c64* H, void* work,
              c64* C, double* w) {
  if (m < 1 || nout < 1 || nout > m) return (int)rocblas_status_invalid_size;
  if (!*blas) {
    rocblas_handle h = nullptr;
    RRB(rocblas_create_handle(&h));
    *blas = h;
  }
  rocblas_handle h = (rocblas_handle)*blas;
  RRB(rocblas_set_stream(h, st));
  RRB(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
  const RrLayout L(m);
  char* base = (char*)work;
  c64* AS = (c64*)(base + L.as);
  c64* AH = (c64*)(base + L.ah);
  c64* T = (c64*)(base + L.t);
  c64* Hp = (c64*)(base + L.hp);
  double* lam = (double*)(base + L.lam);
  double* wv = (double*)(base + L.wv);
  double* E = (double*)(base + L.e);
  double* dsc = (double*)(base + L.dsc);
  int* flags = (int*)(base + L.flags);
  const unsigned nt = (unsigned)((m + 15) / 16);
  k_rr_prep<<<dim3(nt, nt), kRrT, 0, st>>>(S, H, m, AS, AH, dsc);
  RRH(hipGetLastError());
  auto* as = (rocblas_double_complex*)AS;
  auto* ah = (rocblas_double_complex*)AH;
  auto* tt = (rocblas_double_complex*)T;
  auto* hp = (rocblas_double_complex*)Hp;
  RRB(rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, m, as, m, lam, E, flags + 1));
  k_rr_basis<<<m, kRrT, 0, st>>>(AS, lam, dsc, m, floor, flags);
  RRH(hipGetLastError());
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
  RRB(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, m, m, m, &one, ah, m, as, m, &zero, tt, m));
  RRB(rocblas_zgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, m, m, m, &one, as, m, tt, m,
                    &zero, hp, m));
  k_rr_pad<<<1, kRrT, 0, st>>>(Hp, lam, m, floor);
  RRH(hipGetLastError());
  RRB(rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, m, hp, m, wv, E, flags + 2));
  RRB(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, m, nout, m, &one, as, m, hp, m, &zero,
                    (rocblas_double_complex*)C, m));
  RRH(hipMemcpyAsync(w, wv, (size_t)nout * sizeof(double), hipMemcpyDeviceToDevice, st));
  return 0;
}

void rr_release(void* blas) {
  if (blas) (void)rocblas_destroy_handle((rocblas_handle)blas);
}

}  // namespace opaw
