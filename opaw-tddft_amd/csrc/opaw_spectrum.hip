// Absorption spectrum of dipole time series: a damped Fourier transform of
// many series at many frequencies.  The reference writes the dipole series
// (d_dip.dat, src/tddft_plot.f90:100-105) and leaves the spectrum to
// external scripts (README.md:7-11); this is that post-processing step.
//
//   alpha_s(w_k) = -dt * sum_n c_n f(t_n) d_s(t_n) exp(i w_k t_n)
//   t_n = n dt, c_n = 1 - [n = 0]/2 - [n = nt-1]/2 (trapezoid)
//   f = 1 | exp(-t/tau) | exp(-t^2 / (2 tau^2))
//   S_s(w) = (2 w / pi) Im alpha_s(w)
//
// The minus sign: the reference kicks phi -> phi exp(-i sm r)
// (src/tddft_excite.f90:11-12), a potential sm r delta(t) acting on the
// electrons, and d_dip = (dip_pert - dip)/sm is the dipole of the electron
// density (src/tddft_plot.f90:40-57), so the polarizability is -d_dip(t)
// and S > 0 at absorption peaks, with sum rule integral S dw = N_electrons
// for a local Hamiltonian.
//
// One 64-lane wave per (series, frequency): the weighted series g = c f d is
// formed once (k_spec_weight) and read by every frequency from L2; each lane
// sums a strided subset of samples with FP64 sincos and the wave reduces by
// butterfly shuffles (fixed order: results do not depend on the launch).
#include <cmath>
#include <mutex>
#include <vector>

#include "opaw_common.h"
#include "opaw_hip.h"

namespace opaw {
int io_fail(int code, const char* fmt, ...);
}

namespace {

constexpr int kSpecWaves = 4;   // frequencies per 256-thread block

__global__ __launch_bounds__(256) void k_spec_weight(const double* __restrict__ d, double* __restrict__ g,
                                                      int nt, long long total, double dt, int kind,
                                                      double tau) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int n = (int)(i % nt);
  const double t = n * dt;
  double f = d[i];
  if (kind == 1) {
    f *= exp(-t / tau);
  } else if (kind == 2) {
    const double u = t / tau;
    f *= exp(-0.5 * u * u);
  }
  f *= 1.0 - (n == 0 ? 0.5 : 0.0) - (n == nt - 1 ? 0.5 : 0.0);   // trapezoid (0 for a single sample)
  g[i] = f;
}

__global__ __launch_bounds__(256) void k_spec_ft(const double* __restrict__ g, int nt, double dt, double w0,
                                                  double dw, int nw, double* __restrict__ alpha,
                                                  double* __restrict__ strength) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * kSpecWaves + (threadIdx.x >> 6);
  const int s = blockIdx.y;
  if (k >= nw) return;   // uniform per wave
  const double w = w0 + k * dw;
  const double* gs = g + (size_t)s * nt;
  double re = 0.0, im = 0.0;
  for (int n = lane; n < nt; n += 64) {
    double sn, cs;
    sincos(w * (n * dt), &sn, &cs);
    const double v = gs[n];
    re = fma(v, cs, re);
    im = fma(v, sn, im);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    re += __shfl_xor(re, o, 64);
    im += __shfl_xor(im, o, 64);
  }
  if (lane == 0) {
    const size_t o = (size_t)s * nw + k;
    re *= -dt;
    im *= -dt;
    alpha[2 * o] = re;
    alpha[2 * o + 1] = im;
    strength[o] = 2.0 * w / M_PI * im;
  }
}

struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

// Work space of opaw_spectrum_noalloc: the weighted series g of up to
// kSpecScratchMax samples (nsig * nt), in a file-scope device array so the
// entry point allocates nothing.  Each device holds its own copy for the
// life of the process; its device address is looked up once per device
// (hipGetSymbolAddress) and kept in scratch_ptr.
constexpr long long kSpecScratchMax = 1LL << 21;   // 2 Mi doubles, 16 MiB per device
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kSpecScratchMax];
static double* scratch_ptr[kMaxDevices] = {};
static std::mutex scratch_mu;

// the device address of this device's scratch array (nullptr and a message
// on failure)
int scratch_of_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return opaw::io_fail(OPAW_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return opaw::io_fail(OPAW_ERR_HIP, "device %d outside 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return opaw::io_fail(OPAW_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    scratch_ptr[dev] = (double*)p;
  }
  *out = scratch_ptr[dev];
  return OPAW_OK;
}

// a pointer the kernels may use: device (or managed) memory of the current device
bool is_device_pointer(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();   // a host pointer leaves an error behind
    return false;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == dev;
}

}  // namespace

extern "C" int opaw_spectrum(const double* d, int nsig, int nt, double dt, int damp_kind, double tau, double w0,
                             double dw, int nw, double* alpha, double* strength) {
  using opaw::io_fail;
  if (!d) return io_fail(OPAW_ERR_ARG, "null dipole series");
  if (nsig < 1 || nsig > 65535) return io_fail(OPAW_ERR_ARG, "nsig must be 1..65535");
  if (nt < 1 || nw < 1) return io_fail(OPAW_ERR_ARG, "nt and nw must be positive");
  if (!(dt > 0.0)) return io_fail(OPAW_ERR_ARG, "dt must be positive");
  if (damp_kind < 0 || damp_kind > 2) return io_fail(OPAW_ERR_ARG, "damp_kind must be 0, 1 or 2");
  if (damp_kind != 0 && !(tau > 0.0)) return io_fail(OPAW_ERR_ARG, "tau must be positive");
  const long long total = (long long)nsig * nt;
  const long long nout = (long long)nsig * nw;
  if (total > (1LL << 31) || nout > (1LL << 31)) return io_fail(OPAW_ERR_ARG, "spectrum too large");
  DevMem dd, dg, da, ds;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
    return e == hipSuccess;
  };
  if (!chk(hipMalloc(&dd.p, total * sizeof(double))) || !chk(hipMalloc(&dg.p, total * sizeof(double))) ||
      !chk(hipMalloc(&da.p, 2 * nout * sizeof(double))) || !chk(hipMalloc(&ds.p, nout * sizeof(double))))
    return io_fail(OPAW_ERR_ALLOC, "spectrum buffers: %s", hipGetErrorString(e));
  if (!chk(hipMemcpy(dd.p, d, total * sizeof(double), hipMemcpyHostToDevice)))
    return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  const unsigned gb = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(k_spec_weight, dim3(gb), dim3(256), 0, 0, (const double*)dd.p, (double*)dg.p, nt, total, dt,
                     damp_kind, tau);
  if (!chk(hipGetLastError())) return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  hipLaunchKernelGGL(k_spec_ft, dim3((unsigned)((nw + kSpecWaves - 1) / kSpecWaves), (unsigned)nsig), dim3(256), 0,
                     0, (const double*)dg.p, nt, dt, w0, dw, nw, (double*)da.p, (double*)ds.p);
  if (!chk(hipGetLastError())) return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  if (alpha && !chk(hipMemcpy(alpha, da.p, 2 * nout * sizeof(double), hipMemcpyDeviceToHost)))
    return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  if (strength && !chk(hipMemcpy(strength, ds.p, nout * sizeof(double), hipMemcpyDeviceToHost)))
    return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  if (!chk(hipDeviceSynchronize())) return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  return OPAW_OK;
}

// The same spectrum on caller device buffers, with no allocation: the
// weighted series lives in this library's static scratch array (at most
// kSpecScratchMax = 2^21 samples, nsig * nt), and both kernels are enqueued
// on `stream` (nullptr: the null stream) without a host synchronisation.
// d (nsig * nt), alpha (2 * nsig * nw) and strength (nsig * nw) must be
// device memory of the current device.  Calls that overlap on one device
// share the scratch array and race.
extern "C" int opaw_spectrum_noalloc(const double* d, int nsig, int nt, double dt, int damp_kind, double tau,
                                     double w0, double dw, int nw, double* alpha, double* strength, void* stream) {
  using opaw::io_fail;
  if (nsig < 1 || nsig > 65535) return io_fail(OPAW_ERR_ARG, "nsig must be 1..65535");
  if (nt < 1 || nw < 1) return io_fail(OPAW_ERR_ARG, "nt and nw must be positive");
  if (!(dt > 0.0)) return io_fail(OPAW_ERR_ARG, "dt must be positive");
  if (damp_kind < 0 || damp_kind > 2) return io_fail(OPAW_ERR_ARG, "damp_kind must be 0, 1 or 2");
  if (damp_kind != 0 && !(tau > 0.0)) return io_fail(OPAW_ERR_ARG, "tau must be positive");
  const long long total = (long long)nsig * nt;
  const long long nout = (long long)nsig * nw;
  if (total > kSpecScratchMax)
    return io_fail(OPAW_ERR_ARG, "nsig * nt = %lld exceeds the scratch of %lld samples", total, kSpecScratchMax);
  if (nout > (1LL << 31)) return io_fail(OPAW_ERR_ARG, "spectrum too large");
  if (!is_device_pointer(d) || !is_device_pointer(alpha) || !is_device_pointer(strength))
    return io_fail(OPAW_ERR_ARG, "d, alpha and strength must be device memory of the current device");
  double* g = nullptr;
  const int r = scratch_of_current_device(&g);
  if (r != OPAW_OK) return r;
  hipStream_t st = (hipStream_t)stream;
  const unsigned gb = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(k_spec_weight, dim3(gb), dim3(256), 0, st, d, g, nt, total, dt, damp_kind, tau);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  hipLaunchKernelGGL(k_spec_ft, dim3((unsigned)((nw + kSpecWaves - 1) / kSpecWaves), (unsigned)nsig), dim3(256), 0,
                     st, (const double*)g, nt, dt, w0, dw, nw, alpha, strength);
  e = hipGetLastError();
  if (e != hipSuccess) return io_fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
  return OPAW_OK;
}
