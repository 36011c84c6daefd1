// Kernels of the ground-state solver (opaw_gs.hip): Chebyshev filter
// recurrence, residual norms, DIIS on a real field, the k-space
// preconditioner and its MLP correction.
#pragma once
#include "opaw_common.h"

namespace opaw {

// Chebyshev step of chebyr_m (src/chebyr_m.f90:30-36) on n complex values:
// first: cc = (cc - havg cb) / dh; else cc = 2 (cc - havg cb) / dh - ca
template <typename R>
hipError_t cheb_combine(hipStream_t st, cx<R>* cc, const cx<R>* cb, const cx<R>* ca, long long n, double havg,
                        double dh, int first);

// part[c * nb + i] = sum over chunk c of |hp_i - e_i phi_i|^2 (kGsChunks chunks)
constexpr int kGsChunks = 64;
template <typename R>
hipError_t residual_partials(hipStream_t st, const cx<R>* hp, const cx<R>* phi, const double* eig, long long N,
                             int nb, double* part);

// part[(c * K + i) * K + j] = sum over chunk c of a_i a_j (real fields)
hipError_t real_dots(hipStream_t st, const double* a, long long N, int K, double* part);
// out = sum_k coef_k (bank_k + mix err_k)
hipError_t diis_predict(hipStream_t st, const double* bank, const double* err, const double* coef, int K, double mix,
                        long long N, double* out);
// acc[p] += sum over rows of part (nrow x npar), in a fixed order
hipError_t colsum_accumulate(hipStream_t st, const double* part, int nrow, int npar, double* acc);
hipError_t sub_real(hipStream_t st, const double* a, const double* b, long long n, double* out);
hipError_t to_f32(hipStream_t st, const double* in, float* out, long long n);
hipError_t scale_complex(hipStream_t st, c64* a, long long n, double s);
// out = a - s b (complex)
hipError_t axpy_complex(hipStream_t st, const c64* a, const c64* b, double s, long long n, c64* out);

// out_j = sum_{i<m} T[i + m j] in_i for j < nout (rectangular rotation)
hipError_t rotate_rect(hipStream_t st, const c64* in, c64* out, long long N, int m, int nout, const c64* T);
// r_i = hp_i - eig_i phi_i for nb fields
hipError_t residual_vectors(hipStream_t st, const c64* hp, const c64* phi, const double* eig, long long N, int nb,
                            c64* r);
// part[(c * nb + i) * 2 + {0, 1}] = sums over chunk c of ek |f_i(k)|^2 and |f_i(k)|^2
hipError_t kinetic_partials(hipStream_t st, const c64* fk, const double* ek, long long N, int nb, double* part);

// k-space preconditioner: for each band i and k point,
//   M = K(x) exp(g(x, y_i)),  x = ek(k) N / t_i,  K the Teter-Payne-Allan
// polynomial ratio, g the MLP (zero when params is null); w = M r in k space
struct MlpShape {
  int nin = 2, nh = 16;   // 2 -> nh -> nh -> 1, tanh hidden layers
};
int mlp_param_count(const MlpShape& s);
// w = M r applied in place to rk with M evaluated once per (band, k shell)
// into tab (nb x nsh); shell_of[k] is the shell of k point k, eks[s] its ek
hipError_t precond_apply_shells(hipStream_t st, c64* rk, long long N, int nb, const int* shell_of, const double* eks,
                                int nsh, const double* tkin, const double* yfeat, const double* params, MlpShape s,
                                double* tab);
// training: for each band, A_i = sum conj(e) M r, B_i = sum M^2 |r|^2 (k space),
// loss_i = 1 - |A_i|^2 / (|e_i|^2 B_i); grads of sum_i loss_i w.r.t. params
hipError_t precond_loss_sums(hipStream_t st, const c64* rk, const c64* ekf, long long N, int nb, const double* ek,
                             const double* tkin, const double* yfeat, const double* params, MlpShape s, double* part);
hipError_t precond_grad(hipStream_t st, const c64* rk, const c64* ekf, long long N, int nb, const double* ek,
                        const double* tkin, const double* yfeat, const double* params, MlpShape s,
                        const double* coefA, const double* coefB, double* gpart, int nchunk);

// training on k shells (sets of k with equal ek): per pair and band, the
// shell sums z_s = sum conj(e) r and w2_s = sum |r|^2 (perm lists the k of
// shell s in [off[s], off[s+1])); the loss sums (A, B per band) and the
// parameter gradient over the (band, shell) samples
hipError_t shell_aggregate(hipStream_t st, const c64* rk, const c64* ekf, long long N, int nb, const int* perm,
                           const int* off, int nsh, c64* z, double* w2);
hipError_t loss_shells(hipStream_t st, const c64* z, const double* w2, int nsh, int nb, const double* eks, long long N,
                       const double* tkin, const double* yfeat, const double* params, MlpShape s, double* out);
hipError_t grad_shells(hipStream_t st, const c64* z, const double* w2, int nsh, int nb, const double* eks, long long N,
                       const double* tkin, const double* yfeat, const double* params, MlpShape s, const double* coefA,
                       const double* coefB, double* gpart, int nchunk);

}  // namespace opaw
