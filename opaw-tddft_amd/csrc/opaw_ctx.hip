// C ABI of the OPAW-TDDFT propagation hot path: context, resident state and
// the orchestration of the kernels (one HIP stream, no host synchronisation
// inside a time step).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "opaw_box.h"
#include "opaw_fields.h"
#include "opaw_gs.h"
#include "opaw_hip.h"
#include "opaw_linalg.h"
#include "opaw_onsite.h"
#include "opaw_paw.h"
#include "opaw_rr.h"

using namespace opaw;

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HCHK(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(OPAW_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define RCHK(expr)            \
  do {                        \
    int r_ = (expr);          \
    if (r_ != OPAW_OK) return r_; \
  } while (0)

// ---------------------------------------------------------------------------
// profiler: HIP event pairs around launches of each kernel family
struct Prof {
  bool on = false;
  unsigned mask = 0;
  struct Rec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  int open_cls = -1;
  double total[OPAW_K_COUNT] = {0};
  long long count[OPAW_K_COUNT] = {0};

  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    // timing-only events: no system-scope fence, so recording one does not
    // write back and invalidate the caches between the kernels it brackets
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
  }
  void harvest() {
    for (auto& r : recs) {
      float ms = 0.f;
      if (r.a && r.b && hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        total[r.cls] += ms;
        count[r.cls] += 1;
      }
      if (r.a) pool.push_back(r.a);
      if (r.b) pool.push_back(r.b);
    }
    recs.clear();
  }
  void destroy() {
    harvest();
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

static thread_local Prof* g_prof = nullptr;

namespace opaw {
void prof_begin(hipStream_t st, int cls) {
  Prof* p = g_prof;
  if (!p || !p->on || !((p->mask >> cls) & 1u)) return;
  if (p->recs.size() > 20000) p->harvest();
  Prof::Rec r{cls, p->take(), nullptr};
  if (r.a) (void)hipEventRecord(r.a, st);
  p->recs.push_back(r);
}
void prof_end(hipStream_t st, int cls) {
  Prof* p = g_prof;
  if (!p || !p->on || p->recs.empty()) return;
  Prof::Rec& r = p->recs.back();
  if (r.cls != cls || r.b) return;
  r.b = p->take();
  if (r.b) (void)hipEventRecord(r.b, st);
}
}  // namespace opaw

// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct opaw_ctx {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0, dz = 0;
  int periodic = 0, prec = 0, nsys = 0, nb = 0;
  long long N = 0;
  int scale = 2;
  GridGeom g{};
  GridGeom gbig{};
  hipStream_t st = nullptr;
  // side stream (non-blocking, default priority) for work that overlaps the
  // main chain inside one entry point; forked from and joined back into st
  // with the two events, so callers only ever see st
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // time loop: the rebuild's potential chain stays on st2 for the next shs
  // stage (pot_pending), whose forward passes wait only for corrA (ev_corr);
  // ev_nhat hands the main stream's rho_ij / nhat to that chain
  hipEvent_t ev_corr = nullptr, ev_nhat = nullptr;
  hipEvent_t ev_onsite = nullptr;   // the on-site XC kernel (st2) has ended
  hipEvent_t ev_mid = nullptr;   // the forward kinetic passes (st2) have ended
  bool pot_pending = false;
  double* mpart = nullptr;   // partials of the periodic potential means (side stream)
  // own non-blocking stream of the multi-context time loop (created on first
  // use), so the molecules of one batch run side by side
  hipStream_t bst = nullptr;
  Prof prof;
  std::vector<void*> allocs;
  long long bytes = 0;

  std::map<int, c64*> tw64;
  std::map<int, c32*> tw32;

  double* ek = nullptr;
  float* ekf = nullptr;
  bool ek_ready = false;

  // host staging of the tables
  int natom = 0, ntype = 0, nl2 = 0, ctot = 0, mtot = 0, sphtot = 0;
  std::vector<int> atom_type, type_ms, type_nr, atom_start;
  std::vector<std::vector<double>> hp, hp1, hs, hsinv;
  std::vector<std::vector<int>> hsph_idx;
  std::vector<std::vector<double>> hsph_g;
  std::vector<std::vector<double>> hq, hd0, he;
  bool tables_ready = false;
  std::vector<AtomDesc> hatoms;

  // device tables
  AtomDesc* datoms = nullptr;
  int msp = 8, nchunks = 0, U = 0, S = 0;
  int2 *chunks = nullptr, *achunk = nullptr, *coef = nullptr;
  int *umap = nullptr, *uptr = nullptr, *upt = nullptr, *bpt = nullptr, *bk = nullptr;
  int *pptr = nullptr, *pc = nullptr;
  int* uzr = nullptr;        // nz+1: union points of plane z are [uzr[z], uzr[z+1])
  int2* ublk = nullptr;
  int nublk = 0;
  int *ucptr = nullptr, *ucol = nullptr, *pcl = nullptr;
  int ucmax = 0;
  int *optr = nullptr, *ovoff = nullptr, *oc = nullptr;   // box overlap tables (k_bfix)
  double* ov = nullptr;
  double* ov1 = nullptr;   // p1-p1 overlaps on the same columns
  // rebuild: nhat kept on the sphere union (nhs, nsys x S, before the rnel
  // factor); the full nhat field of system s is stale while nhat_lazy[s]
  double* nhs = nullptr;
  double* rpart = nullptr;   // density / nhat partial sums of the rebuild
  std::vector<char> nhat_lazy;
  double *pv = nullptr, *pv1 = nullptr;
  UEntry* uent = nullptr;
  double *utp = nullptr, *utp1 = nullptr;
  int *smap = nullptr, *sptr = nullptr, *sphn = nullptr;
  int2* sent = nullptr;
  double* wl = nullptr;
  bool prep_valid = false;   // cA/wA/cB hold the coefficients of the resident orbitals
  // an RK4 stage left cA/wA/cB (B not yet fixed: st2) and corrA of these
  // fields, for the pipelined rebuild that follows it
  const void* coef_ready = nullptr;
  double *P = nullptr, *P1 = nullptr;
  double* sn[4] = {nullptr, nullptr, nullptr, nullptr};  // n = -1, -1/2, +1/2, +1
  double* sn_custom = nullptr;
  int* sph_idx = nullptr;
  double* sph_g = nullptr;
  double *qtab = nullptr, *d0tab = nullptr, *etab = nullptr;
  // libpaw on-site tables per type (opaw_set_onsite; nr == 0: the type keeps
  // the model on-site D_ij) and libpaw D^hat spheres per atom
  // (opaw_set_dij_sphere; empty: D^hat over the compensation sphere)
  struct OnsiteHost {
    int nr = 0, lm = 0, nij = 0, usetcore = 0;
    std::vector<double> rad, w, inv2, pp, tpp, core, tcore, gnt, qd;
    std::vector<int> kln;
  };
  std::vector<OnsiteHost> hon;
  std::vector<std::vector<int>> hdh_idx;
  std::vector<std::vector<double>> hdh_g;
  OnsiteType* otypes = nullptr;
  double* otab = nullptr;
  int* oint = nullptr;
  int* dh_idx = nullptr;
  double* dh_g = nullptr;
  double* dxc = nullptr;     // nsys x mtot: Dij^xc of the last rebuild
  double* oscr = nullptr;    // on-site kernel scratch
  int onr_max = 0;
  bool any_onsite = false;

  // potentials
  double *vloc = nullptr, *ncoret = nullptr;
  double rnel = 0;
  int usexcnhat = 0;
  bool pot_ready = false;
  double* vk = nullptr;
  bool vk_ready = false;

  // per-system state
  double *vks = nullptr, *dij = nullptr, *rhoij = nullptr, *dens = nullptr, *nhat = nullptr, *vxc = nullptr,
         *vh = nullptr;
  float* vksf = nullptr;
  double* fac = nullptr;
  double* nrm0 = nullptr;
  double* nrm1 = nullptr;
  double* means = nullptr;
  double* part = nullptr;
  double* dipbuf = nullptr;
  int dipcap = 0;
  double* dipstep = nullptr;  // 3 nsys: the dipole slot of a replayed time step

  // orbitals + work
  void* phi = nullptr;
  DevBuf wy, wacc, wh, wu, wtmp;
  // projection coefficients / weights of the three projection roles:
  // A = p1 of the input, B = p of S^-1/2 input, C = p1 of H S^-1/2 input
  DevBuf cA, wA, cB, wB, cC, wC, ppart;
  // compact corrections over the union of boxes (nf x U): S^-1/2 of the
  // input (A), nonlocal add-back (NL), S^-1/2 of H S^-1/2 input (C)
  DevBuf corrA, corrNL, corrC;
  DevBuf hU;   // h on the union of boxes (fused RK4 stages)
  DevBuf hin, hout, hca, hflag;   // staging of the host-array entry points
  double dt = 0.0;                // tddft_mod dt for the host-array rk4_prop_shs
  int coef_cap = 0;
  c64* pair = nullptr;
  c64* hwork = nullptr;
  c64* gpart = nullptr;
  c64* gmat = nullptr;
  c64* tmat = nullptr;

  // orbitals 1..nocc of a system make its density and rho_ij (default nb)
  int nocc = 0;
  // ground-state solver: DIIS on V_KS per system, block-solver buffers and
  // the recorded (residual, error) pairs that train the preconditioner
  int diis_L = 4;
  double diis_mix = 0.15;
  struct Diis {
    int K = 0, Kb = 0, stage = 0;
    double *bank = nullptr, *err = nullptr;
  };
  std::vector<Diis> diis;
  DevBuf gsY, gsHY, gsK, gsT, gsPart, gsG, gsM, gsSmall, gsSmallC;
  struct Snap {
    DevBuf phi, r;     // phi_t and r_t; after prepare: r_t and e_t in k space
    DevBuf ty;         // after prepare: t then y on the device
    DevBuf agg;        // after prepare: shell sums z (nb x nsh complex), then w2 (nb x nsh)
    std::vector<double> t, y, ee;   // ee: per band sum |e(k)|^2
    int sys = 0, nb = 0;
    bool ready = false;
  };
  std::vector<Snap> snaps;
  bool snap_keep = true;   // keep a prepared pair's k-space fields (opaw_precond_pair); else only its shell sums
  // k shells (sets of k points with equal ek) for the training sums and the
  // preconditioner table
  DevBuf shPerm, shOff, shEk;
  DevBuf shOf;        // shell of each k point
  void* pin = nullptr;   // pinned host staging of the block solver's small transfers
  size_t pin_bytes = 0;
  DevBuf gsTab;       // preconditioner multipliers per (band, shell)
  DevBuf gsRR;        // the device Rayleigh-Ritz work (matrices, eigenvalues, flags)
  void* blas = nullptr;   // rocBLAS handle of the device Rayleigh-Ritz, created on first use
  // the block solver's pencils of dimension >= rr_dev_min are solved on the
  // device (rocSOLVER), smaller ones on the host, where they are cheaper:
  // zheevd costs about 1 ms even at m = 48, the host 0.25 ms; they cross near
  // m = 144 (2 x 3.8 ms on the device against about 9 ms on the host at 160)
  int rr_dev_min = 144;
  int nsh = 0;

  size_t esize() const { return prec == OPAW_PREC_F64 ? sizeof(c64) : sizeof(c32); }
};

static int dalloc(opaw_ctx* c, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(OPAW_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  c->allocs.push_back(*p);
  c->bytes += (long long)bytes;
  return OPAW_OK;
}

template <typename T>
static int dalloc_t(opaw_ctx* c, T** p, size_t count) {
  void* v = nullptr;
  int r = dalloc(c, &v, count * sizeof(T));
  *p = (T*)v;
  return r;
}

static void dfree(opaw_ctx* c, void* p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

static int ensure(opaw_ctx* c, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return OPAW_OK;
  dfree(c, b.p);
  b.p = nullptr;
  b.bytes = 0;
  RCHK(dalloc(c, &b.p, bytes));
  b.bytes = bytes;
  return OPAW_OK;
}

struct ProfScope {
  Prof* prev;
  explicit ProfScope(opaw_ctx* c) : prev(g_prof) { g_prof = &c->prof; }
  ~ProfScope() { g_prof = prev; }
};

#define CTX_CHECK(c) \
  if (!(c)) return fail(OPAW_ERR_ARG, "null context"); \
  ProfScope prof_scope_(c)

// ---------------------------------------------------------------------------
// twiddles exp(-2 pi i m / n)
static int twiddles(opaw_ctx* c, int n) {
  if (c->tw64.count(n)) return OPAW_OK;
  std::vector<c64> h64(n);
  std::vector<c32> h32(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    long double a = -2.0L * pi * (long double)m / (long double)n;
    h64[m] = mk<double>((double)cosl(a), (double)sinl(a));
    h32[m] = mk<float>((float)cosl(a), (float)sinl(a));
  }
  c64* d64;
  c32* d32;
  RCHK(dalloc_t(c, &d64, n));
  RCHK(dalloc_t(c, &d32, n));
  HCHK(hipMemcpy(d64, h64.data(), n * sizeof(c64), hipMemcpyHostToDevice));
  HCHK(hipMemcpy(d32, h32.data(), n * sizeof(c32), hipMemcpyHostToDevice));
  c->tw64[n] = d64;
  c->tw32[n] = d32;
  return OPAW_OK;
}

template <typename R>
struct TW;
template <>
struct TW<double> {
  static const c64* get(opaw_ctx* c, int n) { return c->tw64[n]; }
};
template <>
struct TW<float> {
  static const c32* get(opaw_ctx* c, int n) { return c->tw32[n]; }
};

// ---------------------------------------------------------------------------
extern "C" {

const char* opaw_last_error(void) { return g_err.c_str(); }
int opaw_version(void) { return 10000; }
int opaw_fft_length_supported(int n) { return fft_size_supported(n) ? 1 : 0; }
int opaw_plane_path(int nx, int ny, int precision) {
  if (precision == OPAW_PREC_F64) return plane_supported<double>(nx, ny) ? 1 : 0;
  return plane_supported<float>(nx, ny) ? 1 : 0;
}

int opaw_create(opaw_ctx** out, int nx, int ny, int nz, double dx, double dy, double dz, int periodic,
                int precision, int nsys, int nb) {
  if (!out) return fail(OPAW_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (nx <= 0 || ny <= 0 || nz <= 0 || dx <= 0 || dy <= 0 || dz <= 0)
    return fail(OPAW_ERR_ARG, "invalid grid %d %d %d / %g %g %g", nx, ny, nz, dx, dy, dz);
  if (nsys <= 0 || nb <= 0) return fail(OPAW_ERR_ARG, "nsys and nb must be positive");
  if (precision != OPAW_PREC_F64 && precision != OPAW_PREC_MIXED) return fail(OPAW_ERR_ARG, "bad precision");
  const int scale = periodic ? 1 : 2;
  for (int n : {nx, ny, nz, scale * nx, scale * ny, scale * nz})
    if (!fft_size_supported(n)) return fail(OPAW_ERR_UNSUPPORTED, "no compiled FFT length %d", n);
  if (scale == 2 && (nx % 2 || ny % 2 || nz % 2))
    return fail(OPAW_ERR_ARG, "Martyna-Tuckerman doubling needs even grid dimensions");
  opaw_ctx* c = new opaw_ctx();
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->dx = dx;
  c->dy = dy;
  c->dz = dz;
  c->periodic = periodic ? 1 : 0;
  c->prec = precision;
  c->nsys = nsys;
  c->nb = nb;
  c->nocc = nb;
  c->diis.resize(nsys);
  c->N = (long long)nx * ny * nz;
  c->scale = scale;
  c->g = GridGeom{nx, ny, nz};
  c->gbig = GridGeom{scale * nx, scale * ny, scale * nz};
  int r = OPAW_OK;
  for (int n : {nx, ny, nz, scale * nx, scale * ny, scale * nz})
    if (r == OPAW_OK) r = twiddles(c, n);
  const long long N = c->N;
  const size_t nf = (size_t)nsys * nb;
  if (r == OPAW_OK) r = dalloc_t(c, &c->ek, N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->ekf, N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->vks, nsys * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->vksf, nsys * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->dens, nsys * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->nhat, nsys * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->vxc, nsys * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->vh, nsys * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->vloc, N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->ncoret, N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->fac, nsys);
  if (r == OPAW_OK) r = dalloc_t(c, &c->nrm0, nf);
  if (r == OPAW_OK) r = dalloc_t(c, &c->nrm1, nf);
  if (r == OPAW_OK) r = dalloc_t(c, &c->means, nsys + 1);
  if (r == OPAW_OK) r = dalloc_t(c, &c->part, (size_t)kChunks * std::max<size_t>(nf, 4 * (size_t)nsys + 4));
  if (r == OPAW_OK) r = dalloc_t(c, &c->mpart, (size_t)kChunks * ((size_t)nsys + 2));
  if (r == OPAW_OK) r = dalloc(c, &c->phi, nf * N * c->esize());
  if (r == OPAW_OK) r = dalloc_t(c, &c->pair, (size_t)((nsys + 1) / 2) * N);
  if (r == OPAW_OK) r = dalloc_t(c, &c->hwork, (size_t)c->gbig.nx * c->gbig.ny * c->gbig.nz);
  if (r == OPAW_OK)
    r = dalloc_t(c, &c->vk, (size_t)(c->gbig.nx / 2 + 1) * (c->gbig.ny / 2 + 1) * (c->gbig.nz / 2 + 1));
  if (r == OPAW_OK) r = dalloc_t(c, &c->gpart, gram_partials(nb));
  if (r == OPAW_OK) r = dalloc_t(c, &c->gmat, (size_t)nb * nb);
  if (r == OPAW_OK) r = dalloc_t(c, &c->tmat, (size_t)nb * nb);
  if (r == OPAW_OK) r = (hipMemset(c->phi, 0, nf * N * c->esize()) == hipSuccess) ? OPAW_OK : fail(OPAW_ERR_HIP, "memset");
  if (r == OPAW_OK && (hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming | hipEventReleaseToDevice) !=
                           hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming | hipEventReleaseToDevice) !=
                           hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_corr, hipEventDisableTiming | hipEventReleaseToDevice) !=
                           hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_nhat, hipEventDisableTiming | hipEventReleaseToDevice) !=
                           hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_mid, hipEventDisableTiming | hipEventReleaseToDevice) !=
                           hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_onsite, hipEventDisableTiming | hipEventReleaseToDevice) !=
                           hipSuccess))
    r = fail(OPAW_ERR_HIP, "side stream / events");
  if (r != OPAW_OK) {
    opaw_destroy(c);
    return r;
  }
  *out = c;
  return OPAW_OK;
}

int opaw_destroy(opaw_ctx* c) {
  if (!c) return OPAW_OK;
  (void)hipDeviceSynchronize();
  c->prof.destroy();
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_corr) (void)hipEventDestroy(c->ev_corr);
  if (c->ev_nhat) (void)hipEventDestroy(c->ev_nhat);
  if (c->ev_onsite) (void)hipEventDestroy(c->ev_onsite);
  if (c->ev_mid) (void)hipEventDestroy(c->ev_mid);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->bst) (void)hipStreamDestroy(c->bst);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->pin) (void)hipHostFree(c->pin);
  rr_release(c->blas);
  delete c;
  return OPAW_OK;
}

int opaw_set_stream(opaw_ctx* c, void* stream) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  c->st = (hipStream_t)stream;
  return OPAW_OK;
}

int opaw_synchronize(opaw_ctx* c) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

long long opaw_device_bytes(opaw_ctx* c) { return c ? c->bytes : 0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// small device kernels local to the context
__global__ void k_ek_prep(double* ek, float* ekf, int nx, int ny, int nz, double dx, double dy, double dz,
                          double ekcut) {
  const long long N = (long long)nx * ny * nz;
  const double pi = acos(-1.0);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % nx);
    const int iy = (int)((i / nx) % ny);
    const int iz = (int)(i / ((long long)nx * ny));
    double kx = ix * (2.0 * pi / (nx * dx));
    double ky = iy * (2.0 * pi / (ny * dy));
    double kz = iz * (2.0 * pi / (nz * dz));
    if (kx > pi / dx) kx = kx - 2.0 * pi / dx;
    if (ky > pi / dy) ky = ky - 2.0 * pi / dy;
    if (kz > pi / dz) kz = kz - 2.0 * pi / dz;
    double e = (kx * kx + ky * ky + kz * kz) / 2.0;
    if (e > ekcut) e = ekcut;
    e = e / ((double)nx * (double)ny * (double)nz);
    ek[i] = e;
    ekf[i] = (float)e;
  }
}

__global__ void k_vk_real(c64* cvk, int nx, int ny, int nz, double dx, double dy, double dz, double aa, int periodic) {
  const long long N = (long long)nx * ny * nz;
  const double dv = dx * dy * dz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    double v = 0.0;
    if (!periodic) {
      const int ix = (int)(i % nx);
      const int iy = (int)((i / nx) % ny);
      const int iz = (int)(i / ((long long)nx * ny));
      const double x = -nx / 2.0 * dx + ix * dx;
      const double y = -ny / 2.0 * dy + iy * dy;
      const double z = -nz / 2.0 * dz + iz * dz;
      double r2 = x * x + y * y + z * z;
      if (r2 < 1e-26) r2 = 1e-26;
      const double rp = sqrt(r2);
      v = erf(aa * rp) / rp * dv;
    }
    cvk[i] = mk(v, 0.0);
  }
}

__global__ void k_vk_finish(const c64* cvk, double* vk, int nx, int ny, int nz, double dx, double dy, double dz,
                            double aa, int periodic) {
  const long long N = (long long)nx * ny * nz;
  const double pi = acos(-1.0);
  const double r0x = -nx / 2.0 * dx, r0y = -ny / 2.0 * dy, r0z = -nz / 2.0 * dz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % nx);
    const int iy = (int)((i / nx) % ny);
    const int iz = (int)(i / ((long long)nx * ny));
    double kx = ix * (2.0 * pi / (nx * dx));
    double ky = iy * (2.0 * pi / (ny * dy));
    double kz = iz * (2.0 * pi / (nz * dz));
    if (kx > pi / dx) kx = kx - 2.0 * pi / dx;
    if (ky > pi / dy) ky = ky - 2.0 * pi / dy;
    if (kz > pi / dz) kz = kz - 2.0 * pi / dz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    double re;
    if (!periodic) {
      const double ph = r0x * kx + r0y * ky + r0z * kz;
      double sp, cp;
      sincos(ph, &sp, &cp);
      const c64 v = cvk[i];
      re = v.x * cp - v.y * sp;
      const double k2s = k2 / (4.0 * aa * aa);
      double cterm;
      if (k2s > 1e-8)
        cterm = 4.0 * pi / k2 * (1.0 - exp(-k2s));
      else
        cterm = 4.0 * pi / (4.0 * aa * aa) * (1.0 - 0.5 * k2s + k2s * k2s / 6.0);
      re += cterm;
    } else {
      re = (k2 > 1e-8) ? 4.0 * pi / k2 : 0.0;
    }
    // vk is even in every k axis: only the octant 0..n/2 per axis is kept,
    // (nx/2+1, ny/2+1, nz/2+1), and read at min(i, n - i)
    if (ix <= nx / 2 && iy <= ny / 2 && iz <= nz / 2)
      vk[ix + (long long)(nx / 2 + 1) * (iy + (long long)(ny / 2 + 1) * iz)] =
          re / ((double)nx * (double)ny * (double)nz);
  }
}

__global__ void k_copy_d2f(const double* a, float* b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

static unsigned blocks_for(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(b, 65536));
}

extern "C" {

int opaw_ek_prep(opaw_ctx* c, double ekcut) {
  CTX_CHECK(c);
  hipLaunchKernelGGL(k_ek_prep, dim3(blocks_for(c->N)), dim3(256), 0, c->st, c->ek, c->ekf, c->nx, c->ny, c->nz,
                     c->dx, c->dy, c->dz, ekcut);
  HCHK(hipGetLastError());
  c->ek_ready = true;
  return OPAW_OK;
}

int opaw_set_ek(opaw_ctx* c, const double* ek) {
  CTX_CHECK(c);
  if (!ek) return fail(OPAW_ERR_ARG, "null ek");
  HCHK(hipMemcpy(c->ek, ek, c->N * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_copy_d2f, dim3(blocks_for(c->N)), dim3(256), 0, c->st, c->ek, c->ekf, c->N);
  HCHK(hipGetLastError());
  HCHK(hipStreamSynchronize(c->st));
  c->ek_ready = true;
  return OPAW_OK;
}

int opaw_set_atoms(opaw_ctx* c, int natom, const int* atom_map, const int* ir_start, int ntype, const int* nrough,
                   const int* mstates, int nl2) {
  CTX_CHECK(c);
  if (natom < 0 || ntype < 0 || (natom > 0 && (!atom_map || !ir_start || !nrough || !mstates)))
    return fail(OPAW_ERR_ARG, "bad atom arguments");
  if (nl2 < 1 || nl2 > kMaxNl2) return fail(OPAW_ERR_ARG, "nl2=%d outside 1..%d", nl2, kMaxNl2);
  c->natom = natom;
  c->ntype = ntype;
  c->nl2 = nl2;
  c->atom_type.assign(natom, 0);
  c->atom_start.assign(3 * natom, 0);
  c->type_ms.assign(ntype, 0);
  c->type_nr.assign(3 * ntype, 0);
  const int dims[3] = {c->nx, c->ny, c->nz};
  for (int t = 0; t < ntype; ++t) {
    c->type_ms[t] = mstates[t];
    if (mstates[t] < 1 || mstates[t] > kMaxMs) return fail(OPAW_ERR_ARG, "mstates(%d)=%d outside 1..%d", t + 1, mstates[t], kMaxMs);
    for (int k = 0; k < 3; ++k) {
      c->type_nr[3 * t + k] = nrough[3 * t + k];
      // src/read_paw.f90:254-262: nrough may not exceed the grid
      if (nrough[3 * t + k] < 1 || nrough[3 * t + k] > dims[k])
        return fail(OPAW_ERR_ARG, "nrough(%d,%d)=%d outside 1..%d", k + 1, t + 1, nrough[3 * t + k], dims[k]);
    }
  }
  for (int a = 0; a < natom; ++a) {
    const int t = atom_map[a] - 1;
    if (t < 0 || t >= ntype) return fail(OPAW_ERR_ARG, "atom_map(%d)=%d outside 1..%d", a + 1, atom_map[a], ntype);
    c->atom_type[a] = t;
    bool outside = false;
    for (int k = 0; k < 3; ++k) {
      const int s0 = ir_start[3 * a + k] - 1;
      c->atom_start[3 * a + k] = s0;
      const int nr = c->type_nr[3 * t + k];
      if (s0 < 0 || s0 + nr > dims[k]) outside = true;
      if (s0 < -dims[k] + 1 || s0 >= dims[k]) return fail(OPAW_ERR_ARG, "ir_start of atom %d wraps more than once", a + 1);
    }
    // src/get_irs.f90:30-39
    if (outside && !c->periodic)
      return fail(OPAW_ERR_ARG, "non periodic, local grid of atom %d cannot go outside the box", a + 1);
  }
  c->hp.assign(natom, {});
  c->hp1.assign(natom, {});
  c->hs.assign(natom, {});
  c->hsinv.assign(natom, {});
  c->hsph_idx.assign(natom, {});
  c->hsph_g.assign(natom, {});
  c->hq.assign(ntype, {});
  c->hd0.assign(ntype, {});
  c->he.assign(ntype, {});
  c->hon.assign(ntype, {});
  c->hdh_idx.assign(natom, {});
  c->hdh_g.assign(natom, {});
  c->tables_ready = false;
  return OPAW_OK;
}

static int atom_ms(opaw_ctx* c, int a) { return c->type_ms[c->atom_type[a]]; }
static int atom_npt(opaw_ctx* c, int a) {
  const int t = c->atom_type[a];
  return c->type_nr[3 * t] * c->type_nr[3 * t + 1] * c->type_nr[3 * t + 2];
}

int opaw_set_projectors(opaw_ctx* c, int ia, const double* p3d, const double* p3d1, const double* s,
                        const double* sinv) {
  CTX_CHECK(c);
  const int a = ia - 1;
  if (a < 0 || a >= c->natom) return fail(OPAW_ERR_ARG, "atom index %d outside 1..%d", ia, c->natom);
  if (!p3d || !p3d1 || !s || !sinv) return fail(OPAW_ERR_ARG, "null projector table");
  const int ms = atom_ms(c, a), npt = atom_npt(c, a);
  c->hp[a].assign(p3d, p3d + (size_t)ms * npt);
  c->hp1[a].assign(p3d1, p3d1 + (size_t)ms * npt);
  c->hs[a].assign(s, s + ms);
  c->hsinv[a].assign(sinv, sinv + ms);
  c->tables_ready = false;
  return OPAW_OK;
}

int opaw_set_sphere(opaw_ctx* c, int ia, int ngrid, const int* local_grid, const double* local_g3d) {
  CTX_CHECK(c);
  const int a = ia - 1;
  if (a < 0 || a >= c->natom) return fail(OPAW_ERR_ARG, "atom index %d outside 1..%d", ia, c->natom);
  if (ngrid < 0 || (ngrid > 0 && (!local_grid || !local_g3d))) return fail(OPAW_ERR_ARG, "bad sphere arguments");
  std::vector<int> idx(ngrid);
  for (int p = 0; p < ngrid; ++p) {
    const int ix = local_grid[p] - 1, iy = local_grid[p + ngrid] - 1, iz = local_grid[p + 2 * ngrid] - 1;
    if (ix < 0 || ix >= c->nx || iy < 0 || iy >= c->ny || iz < 0 || iz >= c->nz)
      return fail(OPAW_ERR_ARG, "local_grid point %d of atom %d outside the grid", p + 1, ia);
    idx[p] = ix + c->nx * (iy + c->ny * iz);
  }
  c->hsph_idx[a] = idx;
  c->hsph_g[a].assign(local_g3d, local_g3d + (size_t)ngrid * c->nl2);
  c->tables_ready = false;
  return OPAW_OK;
}

int opaw_set_type(opaw_ctx* c, int it, const double* qijlm, const double* dij0, const double* eijkl) {
  CTX_CHECK(c);
  const int t = it - 1;
  if (t < 0 || t >= c->ntype) return fail(OPAW_ERR_ARG, "type index %d outside 1..%d", it, c->ntype);
  const int ms = c->type_ms[t];
  const size_t m2 = (size_t)ms * ms;
  c->hq[t].assign((size_t)c->nl2 * m2, 0.0);
  c->hd0[t].assign(m2, 0.0);
  c->he[t].assign(m2 * m2, 0.0);
  if (qijlm) std::copy(qijlm, qijlm + (size_t)c->nl2 * m2, c->hq[t].begin());
  if (dij0) std::copy(dij0, dij0 + m2, c->hd0[t].begin());
  if (eijkl) std::copy(eijkl, eijkl + m2 * m2, c->he[t].begin());
  c->tables_ready = false;
  return OPAW_OK;
}

int opaw_set_onsite(opaw_ctx* c, int it, int nr, int lm_size, int nij, int usetcore, const double* rad,
                    const double* w, const double* phiphj, const double* tphitphj, const double* coredens,
                    const double* tcoredens, const double* gnt, const int* kln, const double* qijl) {
  CTX_CHECK(c);
  const int t = it - 1;
  if (t < 0 || t >= c->ntype) return fail(OPAW_ERR_ARG, "type index %d outside 1..%d", it, c->ntype);
  if (nr < 4) return fail(OPAW_ERR_ARG, "on-site mesh of %d points", nr);
  if (lm_size < 1 || lm_size > kOnsiteMaxLm)
    return fail(OPAW_ERR_UNSUPPORTED, "on-site moments %d outside 1..%d (partial waves with l <= 1)", lm_size,
                kOnsiteMaxLm);
  if (nij < 1 || nij > kOnsiteMaxNij)
    return fail(OPAW_ERR_UNSUPPORTED, "radial pairs %d outside 1..%d", nij, kOnsiteMaxNij);
  if (!rad || !w || !phiphj || !tphitphj || !coredens || !gnt || !kln || !qijl || (usetcore && !tcoredens))
    return fail(OPAW_ERR_ARG, "null on-site table");
  const int ms = c->type_ms[t];
  const size_t m2 = (size_t)ms * ms;
  for (size_t q = 0; q < m2; ++q)
    if (kln[q] < 0 || kln[q] >= nij) return fail(OPAW_ERR_ARG, "radial pair index %d outside 0..%d", kln[q], nij - 1);
  if (rad[0] != 0.0) return fail(OPAW_ERR_ARG, "the on-site mesh must start at r = 0");
  auto& h = c->hon[t];
  h.nr = nr;
  h.lm = lm_size;
  h.nij = nij;
  h.usetcore = usetcore ? 1 : 0;
  h.rad.assign(rad, rad + nr);
  h.w.assign(w, w + nr);
  h.inv2.assign(nr, 0.0);
  for (int r = 1; r < nr; ++r) h.inv2[r] = 1.0 / (rad[r] * rad[r]);
  h.pp.assign(phiphj, phiphj + (size_t)nr * nij);
  h.tpp.assign(tphitphj, tphitphj + (size_t)nr * nij);
  h.core.assign(coredens, coredens + nr);
  h.tcore.assign(nr, 0.0);
  if (usetcore) h.tcore.assign(tcoredens, tcoredens + nr);
  h.gnt.assign(gnt, gnt + (size_t)lm_size * m2);
  h.kln.assign(kln, kln + m2);
  h.qd.assign(qijl, qijl + (size_t)lm_size * m2);
  c->tables_ready = false;
  return OPAW_OK;
}

int opaw_set_dij_sphere(opaw_ctx* c, int ia, int nfgd, const int* ifftsph, const double* gylm) {
  CTX_CHECK(c);
  const int a = ia - 1;
  if (a < 0 || a >= c->natom) return fail(OPAW_ERR_ARG, "atom index %d outside 1..%d", ia, c->natom);
  const auto& h = c->hon[c->atom_type[a]];
  if (h.nr == 0) return fail(OPAW_ERR_STATE, "the type of atom %d has no on-site tables (opaw_set_onsite)", ia);
  if (nfgd < 0 || (nfgd > 0 && (!ifftsph || !gylm))) return fail(OPAW_ERR_ARG, "bad D^hat sphere arguments");
  std::vector<int> idx(nfgd);
  for (int p = 0; p < nfgd; ++p) {
    const long long f = (long long)ifftsph[p] - 1;
    if (f < 0 || f >= c->N) return fail(OPAW_ERR_ARG, "ifftsph point %d of atom %d outside the grid", p + 1, ia);
    idx[p] = (int)f;
  }
  c->hdh_idx[a] = idx;
  c->hdh_g[a].assign(gylm, gylm + (size_t)nfgd * h.lm);
  c->tables_ready = false;
  return OPAW_OK;
}

}  // extern "C"

template <typename T>
static int upload(opaw_ctx* c, T** dst, const std::vector<T>& v) {
  RCHK(dalloc_t(c, dst, std::max<size_t>(v.size(), 1)));
  if (!v.empty()) HCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return OPAW_OK;
}

extern "C" {

int opaw_finalize_tables(opaw_ctx* c) {
  CTX_CHECK(c);
  const int natom = c->natom;
  for (int a = 0; a < natom; ++a)
    if (c->hp[a].empty()) return fail(OPAW_ERR_STATE, "projectors of atom %d not set", a + 1);
  // free previous device tables
  for (void* p : {(void*)c->datoms, (void*)c->P, (void*)c->P1, (void*)c->sn[0], (void*)c->sn[1], (void*)c->sn[2],
                  (void*)c->sn[3], (void*)c->sn_custom, (void*)c->sph_idx,
                  (void*)c->sph_g, (void*)c->qtab, (void*)c->d0tab, (void*)c->etab, (void*)c->dij, (void*)c->rhoij,
                  (void*)c->chunks, (void*)c->achunk, (void*)c->coef, (void*)c->umap, (void*)c->uptr, (void*)c->upt,
                  (void*)c->uent, (void*)c->utp, (void*)c->utp1, (void*)c->smap, (void*)c->sptr, (void*)c->sphn,
                  (void*)c->sent, (void*)c->wl, (void*)c->bpt, (void*)c->bk,
                  (void*)c->pptr, (void*)c->pc, (void*)c->pv, (void*)c->pv1, (void*)c->ublk,
                  (void*)c->uzr, (void*)c->optr, (void*)c->ovoff, (void*)c->oc, (void*)c->ov, (void*)c->ov1,
                  (void*)c->nhs, (void*)c->rpart, (void*)c->otypes, (void*)c->otab, (void*)c->oint,
                  (void*)c->dh_idx, (void*)c->dh_g, (void*)c->dxc, (void*)c->oscr})
    dfree(c, p);
  c->prep_valid = false;
  c->coef_cap = 0;   // projection buffers depend on msp / chunk count
  // per-type table offsets
  std::vector<int> toff_q(c->ntype), toff_d0(c->ntype), toff_e(c->ntype);
  std::vector<double> q, d0, e;
  for (int t = 0; t < c->ntype; ++t) {
    const int ms = c->type_ms[t];
    const size_t m2 = (size_t)ms * ms;
    if (c->hq[t].empty()) {
      c->hq[t].assign((size_t)c->nl2 * m2, 0.0);
      c->hd0[t].assign(m2, 0.0);
      c->he[t].assign(m2 * m2, 0.0);
    }
    toff_q[t] = (int)q.size();
    toff_d0[t] = (int)d0.size();
    toff_e[t] = (int)e.size();
    q.insert(q.end(), c->hq[t].begin(), c->hq[t].end());
    d0.insert(d0.end(), c->hd0[t].begin(), c->hd0[t].end());
    e.insert(e.end(), c->he[t].begin(), c->he[t].end());
  }
  // libpaw on-site tables: one OnsiteType per type that has them, their
  // arrays concatenated; the D^hat q_ij^L of those types appended to qtab
  std::vector<int> type_ot(c->ntype, -1), toff_qd(c->ntype, -1);
  std::vector<OnsiteType> ots;
  std::vector<double> ot_d;
  std::vector<int> ot_i;
  c->onr_max = 0;
  for (int t = 0; t < c->ntype; ++t) {
    const auto& h = c->hon[t];
    if (h.nr == 0) continue;
    OnsiteType T{};
    T.nr = h.nr;
    T.lm = h.lm;
    T.nij = h.nij;
    T.usetcore = h.usetcore;
    auto put = [&](const std::vector<double>& v) {
      const long long off = (long long)ot_d.size();
      ot_d.insert(ot_d.end(), v.begin(), v.end());
      return off;
    };
    T.rad = put(h.rad);
    T.w = put(h.w);
    T.inv2 = put(h.inv2);
    T.pp = put(h.pp);
    T.tpp = put(h.tpp);
    T.core = put(h.core);
    T.tcore = put(h.tcore);
    T.gnt = put(h.gnt);
    T.kln = (int)ot_i.size();
    ot_i.insert(ot_i.end(), h.kln.begin(), h.kln.end());
    type_ot[t] = (int)ots.size();
    ots.push_back(T);
    toff_qd[t] = (int)q.size();
    q.insert(q.end(), h.qd.begin(), h.qd.end());
    c->onr_max = std::max(c->onr_max, h.nr);
  }
  c->any_onsite = !ots.empty();
  c->hatoms.assign(natom, AtomDesc{});
  std::vector<int> dhidx;
  std::vector<double> dhg;
  std::vector<double> P, P1, s, sm1, smh, sph, sp1;
  std::vector<int> sidx;
  std::vector<double> sg;
  long long tab = 0;
  int coff = 0, moff = 0, soff = 0;
  const int dims[3] = {c->nx, c->ny, c->nz};
  for (int a = 0; a < natom; ++a) {
    const int t = c->atom_type[a];
    AtomDesc& A = c->hatoms[a];
    A.ms = c->type_ms[t];
    A.itype = t;
    A.wrap = 0;
    for (int k = 0; k < 3; ++k) {
      A.start[k] = c->atom_start[3 * a + k];
      A.nr[k] = c->type_nr[3 * t + k];
      if (A.start[k] < 0 || A.start[k] + A.nr[k] > dims[k]) A.wrap = 1;
    }
    A.npt = A.nr[0] * A.nr[1] * A.nr[2];
    A.tab = tab;
    A.coff = coff;
    A.moff = moff;
    A.sph_off = soff;
    A.sph_n = (int)c->hsph_idx[a].size();
    A.toff_q = toff_q[t];
    A.toff_d0 = toff_d0[t];
    A.toff_e = toff_e[t];
    A.ot = type_ot[t];
    // D^hat: libpaw's sphere and q_ij^L for a type with on-site tables
    // (pawdijhat over pawfgrtab), else the compensation sphere and qijlm
    A.dh_off = (int)dhidx.size();
    A.dh_goff = (int)dhg.size();
    if (A.ot >= 0) {
      if (c->hdh_idx[a].empty() && c->hdh_g[a].empty())
        return fail(OPAW_ERR_STATE, "D^hat sphere of atom %d not set (opaw_set_dij_sphere)", a + 1);
      A.dh_n = (int)c->hdh_idx[a].size();
      A.dh_nl = c->hon[t].lm;
      A.toff_qh = toff_qd[t];
      dhidx.insert(dhidx.end(), c->hdh_idx[a].begin(), c->hdh_idx[a].end());
      dhg.insert(dhg.end(), c->hdh_g[a].begin(), c->hdh_g[a].end());
    } else {
      A.dh_n = (int)c->hsph_idx[a].size();
      A.dh_nl = c->nl2;
      A.toff_qh = toff_q[t];
      dhidx.insert(dhidx.end(), c->hsph_idx[a].begin(), c->hsph_idx[a].end());
      dhg.insert(dhg.end(), c->hsph_g[a].begin(), c->hsph_g[a].end());
    }
    P.insert(P.end(), c->hp[a].begin(), c->hp[a].end());
    P1.insert(P1.end(), c->hp1[a].begin(), c->hp1[a].end());
    for (int i = 0; i < A.ms; ++i) {
      const double si = c->hs[a][i];
      s.push_back(si);
      sm1.push_back(c->hsinv[a][i]);
      smh.push_back(std::pow(1.0 + si, -0.5) - 1.0);
      sph.push_back(std::pow(1.0 + si, 0.5) - 1.0);
      sp1.push_back(std::pow(1.0 + si, 1.0) - 1.0);
    }
    sidx.insert(sidx.end(), c->hsph_idx[a].begin(), c->hsph_idx[a].end());
    sg.insert(sg.end(), c->hsph_g[a].begin(), c->hsph_g[a].end());
    tab += (long long)A.ms * A.npt;
    coff += A.ms;
    moff += A.ms * A.ms;
    soff += A.sph_n;
  }
  c->ctot = coff;
  c->mtot = moff;
  c->sphtot = soff;
  RCHK(upload(c, &c->P, P));
  RCHK(upload(c, &c->P1, P1));
  RCHK(upload(c, &c->sn[0], sm1));
  RCHK(upload(c, &c->sn[1], smh));
  RCHK(upload(c, &c->sn[2], sph));
  RCHK(upload(c, &c->sn[3], sp1));
  RCHK(dalloc_t(c, &c->sn_custom, std::max(1, coff)));
  RCHK(upload(c, &c->sph_idx, sidx));
  RCHK(upload(c, &c->sph_g, sg));
  RCHK(upload(c, &c->qtab, q));
  RCHK(upload(c, &c->d0tab, d0));
  RCHK(upload(c, &c->etab, e));
  RCHK(upload(c, &c->dh_idx, dhidx));
  RCHK(upload(c, &c->dh_g, dhg));
  c->otypes = nullptr;
  c->otab = nullptr;
  c->oint = nullptr;
  c->dxc = nullptr;
  c->oscr = nullptr;
  if (c->any_onsite) {
    RCHK(upload(c, &c->otypes, ots));
    RCHK(upload(c, &c->otab, ot_d));
    RCHK(upload(c, &c->oint, ot_i));
    RCHK(dalloc_t(c, &c->dxc, (size_t)c->nsys * std::max(1, moff)));
    RCHK(dalloc_t(c, &c->oscr, (size_t)std::max(1, natom) * c->nsys * onsite_scratch_doubles(c->onr_max)));
    HCHK(hipMemset(c->dxc, 0, (size_t)c->nsys * std::max(1, moff) * sizeof(double)));
  }
  // projection chunks and coefficient map
  {
    std::vector<int2> chunks, achunk(std::max(1, natom)), coef;
    int maxms = 1;
    for (int a = 0; a < natom; ++a) {
      const AtomDesc& A = c->hatoms[a];
      maxms = std::max(maxms, A.ms);
      achunk[a] = make_int2((int)chunks.size(), 0);
      for (int p0 = 0; p0 < A.npt; p0 += kProjChunk) {
        chunks.push_back(make_int2(a, p0));
        achunk[a].y++;
      }
      for (int i = 0; i < A.ms; ++i) coef.push_back(make_int2(a, i));
    }
    c->msp = maxms <= 8 ? 8 : 16;
    c->nchunks = (int)chunks.size();
    RCHK(upload(c, &c->chunks, chunks));
    RCHK(upload(c, &c->achunk, achunk));
    RCHK(upload(c, &c->coef, coef));
  }
  // union of the projector boxes: grid point -> (atom, table values) entries
  {
    const long long N = c->N;
    std::vector<std::vector<std::pair<int, int>>> at;   // per union point: (atom, local point)
    std::vector<int> umap(N, -1), upt;
    for (int a = 0; a < natom; ++a) {
      const AtomDesc& A = c->hatoms[a];
      for (int p = 0; p < A.npt; ++p) {
        const int lx = p % A.nr[0], ly = (p / A.nr[0]) % A.nr[1], lz = p / (A.nr[0] * A.nr[1]);
        const int gx = ((A.start[0] + lx) % c->nx + c->nx) % c->nx;
        const int gy = ((A.start[1] + ly) % c->ny + c->ny) % c->ny;
        const int gz = ((A.start[2] + lz) % c->nz + c->nz) % c->nz;
        const long long pt = gx + (long long)c->nx * (gy + (long long)c->ny * gz);
        if (umap[pt] < 0) {
          umap[pt] = (int)upt.size();
          upt.push_back((int)pt);
          at.emplace_back();
        }
        at[umap[pt]].push_back({a, p});
      }
    }
    // renumber union points in grid order
    std::vector<int> order(upt.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return upt[x] < upt[y]; });
    std::vector<int> uptr(1, 0), upt2;
    std::vector<UEntry> ue;
    std::vector<double> tp, tp1;
    // box overlaps O_ab(i, j) = sum over common points of p_ai p1_bj (times
    // dv below), accumulated in grid order
    // and O11_ab(i, j) of p1_ai p1_bj on the same atom pairs
    std::vector<std::vector<double>> Oab((size_t)natom * natom), Oab1((size_t)natom * natom);
    for (int k : order) {
      const int pt = upt[k];
      for (auto& ea : at[k])
        for (auto& eb : at[k]) {
          const AtomDesc& Aa = c->hatoms[ea.first];
          const AtomDesc& Ab = c->hatoms[eb.first];
          std::vector<double>& O = Oab[(size_t)ea.first * natom + eb.first];
          std::vector<double>& O1 = Oab1[(size_t)ea.first * natom + eb.first];
          if (O.empty()) O.assign((size_t)Aa.ms * Ab.ms, 0.0);
          if (O1.empty()) O1.assign((size_t)Aa.ms * Ab.ms, 0.0);
          for (int i = 0; i < Aa.ms; ++i) {
            const double pa = c->hp[ea.first][(size_t)i * Aa.npt + ea.second];
            const double qa = c->hp1[ea.first][(size_t)i * Aa.npt + ea.second];
            for (int j = 0; j < Ab.ms; ++j) {
              const double qb = c->hp1[eb.first][(size_t)j * Ab.npt + eb.second];
              O[(size_t)i * Ab.ms + j] += pa * qb;
              O1[(size_t)i * Ab.ms + j] += qa * qb;
            }
          }
        }
      umap[pt] = (int)upt2.size();
      upt2.push_back(pt);
      for (auto& ap : at[k]) {   // atoms were appended in atom order
        const AtomDesc& A = c->hatoms[ap.first];
        ue.push_back(UEntry{A.coff, A.ms, (int)tp.size(), ap.first});
        for (int i = 0; i < A.ms; ++i) {
          tp.push_back(c->hp[ap.first][(size_t)i * A.npt + ap.second]);
          tp1.push_back(c->hp1[ap.first][(size_t)i * A.npt + ap.second]);
        }
      }
      uptr.push_back((int)ue.size());
    }
    c->U = (int)upt2.size();
    {
      const double dv = c->dx * c->dy * c->dz;
      std::vector<int> optr(1, 0), ovoff, oc;
      std::vector<double> ov, ov1;
      for (int a = 0; a < natom; ++a) {
        const AtomDesc& Aa = c->hatoms[a];
        std::vector<std::pair<int, int>> cols;   // (b, j)
        for (int b2 = 0; b2 < natom; ++b2)
          if (!Oab[(size_t)a * natom + b2].empty())
            for (int j = 0; j < c->hatoms[b2].ms; ++j) cols.push_back({b2, j});
        const int La = (int)cols.size();
        ovoff.push_back((int)ov.size());
        for (auto& cj : cols) oc.push_back(c->hatoms[cj.first].coff + cj.second);
        optr.push_back((int)oc.size());
        for (int i = 0; i < Aa.ms; ++i)
          for (int l = 0; l < La; ++l) {
            const int b2 = cols[l].first, j = cols[l].second;
            ov.push_back(dv * Oab[(size_t)a * natom + b2][(size_t)i * c->hatoms[b2].ms + j]);
            ov1.push_back(dv * Oab1[(size_t)a * natom + b2][(size_t)i * c->hatoms[b2].ms + j]);
          }
      }
      RCHK(upload(c, &c->optr, optr));
      RCHK(upload(c, &c->ovoff, ovoff));
      RCHK(upload(c, &c->oc, oc));
      RCHK(upload(c, &c->ov, ov));
      RCHK(upload(c, &c->ov1, ov1));
    }
    // the same entries flattened to (coefficient, value) pairs per union point
    {
      std::vector<int> pptr(1, 0), pc;
      std::vector<double> pv, pv1;
      for (size_t k = 0; k < upt2.size(); ++k) {
        for (int e = uptr[k]; e < uptr[k + 1]; ++e)
          for (int i = 0; i < ue[e].ms; ++i) {
            pc.push_back(ue[e].coff + i);
            pv.push_back(tp[ue[e].toff + i]);
            pv1.push_back(tp1[ue[e].toff + i]);
          }
        pptr.push_back((int)pc.size());
      }
      // blocks of consecutive points for the LDS kernel; none when one
      // point alone has more pairs than a block holds
      std::vector<int2> ublk;
      bool fits = true;
      for (size_t k = 0; k + 1 < pptr.size(); ++k)
        if (pptr[k + 1] - pptr[k] > kUnionBlockPairs) fits = false;
      if (fits) {
        int first = 0;
        const int U = (int)pptr.size() - 1;
        while (first < U) {
          int cnt = 0;
          while (first + cnt < U && cnt < kUnionBlockRows &&
                 pptr[first + cnt + 1] - pptr[first] <= kUnionBlockPairs)
            ++cnt;
          ublk.push_back(make_int2(first, cnt));
          first += cnt;
        }
      }
      c->nublk = (int)ublk.size();
      if (c->nublk > 0) RCHK(upload(c, &c->ublk, ublk));
      // per block: the coefficients its pairs reference (sorted), and each
      // pair's coefficient as an index into that list, so a block stages
      // only the weight rows it uses
      if (c->nublk > 0) {
        std::vector<int> ucptr(1, 0), ucol, pcl(pc.size(), 0);
        c->ucmax = 0;
        for (const int2& bk2 : ublk) {
          const int q0 = pptr[bk2.x], q1 = pptr[bk2.x + bk2.y];
          std::vector<int> cols(pc.begin() + q0, pc.begin() + q1);
          std::sort(cols.begin(), cols.end());
          cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
          for (int q = q0; q < q1; ++q)
            pcl[q] = (int)(std::lower_bound(cols.begin(), cols.end(), pc[q]) - cols.begin());
          ucol.insert(ucol.end(), cols.begin(), cols.end());
          ucptr.push_back((int)ucol.size());
          c->ucmax = std::max(c->ucmax, (int)cols.size());
        }
        RCHK(upload(c, &c->ucptr, ucptr));
        RCHK(upload(c, &c->ucol, ucol));
        RCHK(upload(c, &c->pcl, pcl));
      }
      RCHK(upload(c, &c->pptr, pptr));
      RCHK(upload(c, &c->pc, pc));
      RCHK(upload(c, &c->pv, pv));
      RCHK(upload(c, &c->pv1, pv1));
    }
    // per box point: grid point and union index (atoms concatenated)
    std::vector<int> bpt, bk;
    for (int a = 0; a < natom; ++a) {
      AtomDesc& A = c->hatoms[a];
      A.boff = (int)bpt.size();
      for (int p = 0; p < A.npt; ++p) {
        const int lx = p % A.nr[0], ly = (p / A.nr[0]) % A.nr[1], lz = p / (A.nr[0] * A.nr[1]);
        const int gx = ((A.start[0] + lx) % c->nx + c->nx) % c->nx;
        const int gy = ((A.start[1] + ly) % c->ny + c->ny) % c->ny;
        const int gz = ((A.start[2] + lz) % c->nz + c->nz) % c->nz;
        const int pt = gx + c->nx * (gy + c->ny * gz);
        bpt.push_back(pt);
        bk.push_back(umap[pt]);
      }
    }
    RCHK(upload(c, &c->bpt, bpt));
    RCHK(upload(c, &c->bk, bk));
    RCHK(upload(c, &c->umap, umap));
    RCHK(upload(c, &c->uptr, uptr));
    RCHK(upload(c, &c->upt, upt2));
    {
      const long long plane = (long long)c->nx * c->ny;
      std::vector<int> uzr(c->nz + 1);
      for (int z = 0; z <= c->nz; ++z)
        uzr[z] = (int)(std::lower_bound(upt2.begin(), upt2.end(), (long long)z * plane) - upt2.begin());
      RCHK(upload(c, &c->uzr, uzr));
    }
    RCHK(upload(c, &c->uent, ue));
    RCHK(upload(c, &c->utp, tp));
    RCHK(upload(c, &c->utp1, tp1));
  }
  // union of the compensation spheres
  {
    std::vector<std::vector<int2>> at;
    std::vector<int> smap(c->N, -1), spt;
    std::vector<int> sphn(std::max(1, natom), 0);
    for (int a = 0; a < natom; ++a) {
      const AtomDesc& A = c->hatoms[a];
      sphn[a] = A.sph_n;
      for (int p = 0; p < A.sph_n; ++p) {
        const int pt = c->hsph_idx[a][p];
        if (smap[pt] < 0) {
          smap[pt] = (int)spt.size();
          spt.push_back(pt);
          at.emplace_back();
        }
        at[smap[pt]].push_back(make_int2(a, A.sph_off * c->nl2 + p));
      }
    }
    std::vector<int> sptr(1, 0);
    std::vector<int2> se;
    for (size_t k = 0; k < spt.size(); ++k) {
      std::stable_sort(at[k].begin(), at[k].end(), [](int2 x, int2 y) { return x.x < y.x; });
      se.insert(se.end(), at[k].begin(), at[k].end());
      sptr.push_back((int)se.size());
    }
    c->S = (int)spt.size();
    RCHK(dalloc_t(c, &c->nhs, (size_t)c->nsys * std::max(1, c->S)));
    // rebuild partial sums: density (one per 256 points) and nhat (one per
    // 64 sphere points), per system
    RCHK(dalloc_t(c, &c->rpart, (size_t)c->nsys * ((c->N + 255) / 256 + (c->S + 63) / 64 + 1)));
    c->nhat_lazy.assign(c->nsys, 0);
    RCHK(upload(c, &c->smap, smap));
    RCHK(upload(c, &c->sptr, sptr));
    RCHK(upload(c, &c->sent, se));
    RCHK(upload(c, &c->sphn, sphn));
    RCHK(dalloc_t(c, &c->wl, (size_t)c->nsys * std::max(1, natom) * c->nl2));
  }
  RCHK(upload(c, &c->datoms, c->hatoms));
  RCHK(dalloc_t(c, &c->dij, (size_t)c->nsys * std::max(1, moff)));
  RCHK(dalloc_t(c, &c->rhoij, (size_t)c->nsys * std::max(1, moff)));
  HCHK(hipMemset(c->dij, 0, (size_t)c->nsys * std::max(1, moff) * sizeof(double)));
  HCHK(hipMemset(c->rhoij, 0, (size_t)c->nsys * std::max(1, moff) * sizeof(double)));
  c->tables_ready = true;
  return OPAW_OK;
}

int opaw_set_potentials(opaw_ctx* c, const double* vloc_tot, const double* ncoret, double rnel, int usexcnhat) {
  CTX_CHECK(c);
  if (vloc_tot)
    HCHK(hipMemcpy(c->vloc, vloc_tot, c->N * sizeof(double), hipMemcpyHostToDevice));
  else
    HCHK(hipMemset(c->vloc, 0, c->N * sizeof(double)));
  if (ncoret)
    HCHK(hipMemcpy(c->ncoret, ncoret, c->N * sizeof(double), hipMemcpyHostToDevice));
  else
    HCHK(hipMemset(c->ncoret, 0, c->N * sizeof(double)));
  c->rnel = rnel;
  c->usexcnhat = usexcnhat;
  c->pot_ready = true;
  return OPAW_OK;
}

int opaw_vk_prep(opaw_ctx* c) {
  CTX_CHECK(c);
  const GridGeom& B = c->gbig;
  const long long NB = (long long)B.nx * B.ny * B.nz;
  const double aa = 7.0 / std::min({B.nx * c->dx, B.ny * c->dy, B.nz * c->dz});
  hipLaunchKernelGGL(k_vk_real, dim3(blocks_for(NB)), dim3(256), 0, c->st, c->hwork, B.nx, B.ny, B.nz, c->dx, c->dy,
                     c->dz, aa, c->periodic);
  HCHK(hipGetLastError());
  if (!c->periodic) {
    // fft3d_general(sign=1): unnormalised e^{+ikr} transform
    HCHK(fft3d<double>(B, c->st, +1, c->hwork, c->hwork, 1, c->tw64[B.nx], c->tw64[B.ny], c->tw64[B.nz]));
  }
  hipLaunchKernelGGL(k_vk_finish, dim3(blocks_for(NB)), dim3(256), 0, c->st, c->hwork, c->vk, B.nx, B.ny, B.nz, c->dx,
                     c->dy, c->dz, aa, c->periodic);
  HCHK(hipGetLastError());
  HCHK(hipStreamSynchronize(c->st));
  c->vk_ready = true;
  return OPAW_OK;
}

static int check_sys(opaw_ctx* c, int sys) {
  if (sys < 0 || sys >= c->nsys) return fail(OPAW_ERR_ARG, "system %d outside 0..%d", sys, c->nsys - 1);
  return OPAW_OK;
}

// materialise the full nhat field of system sys from the rebuild's sparse
// sphere values (the rebuild itself never needs it)
static int nhat_full(opaw_ctx* c, int sys) {
  if (sys < (int)c->nhat_lazy.size() && c->nhat_lazy[sys]) {
    HCHK(nhat_expand(c->st, c->smap, c->nhs + (size_t)sys * std::max(1, c->S), c->fac + sys, c->N,
                     c->nhat + (size_t)sys * c->N));
    c->nhat_lazy[sys] = 0;
  }
  return OPAW_OK;
}

int opaw_write_vks(opaw_ctx* c, int sys, const double* vks) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!vks) return fail(OPAW_ERR_ARG, "null vks");
  HCHK(hipMemcpy(c->vks + sys * c->N, vks, c->N * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_copy_d2f, dim3(blocks_for(c->N)), dim3(256), 0, c->st, c->vks + sys * c->N,
                     c->vksf + sys * c->N, c->N);
  HCHK(hipGetLastError());
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

int opaw_read_vks(opaw_ctx* c, int sys, double* vks) { return opaw_read_field(c, OPAW_FIELD_VKS, sys, vks); }

int opaw_read_field(opaw_ctx* c, int which, int sys, double* out) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!out) return fail(OPAW_ERR_ARG, "null output");
  if (which == OPAW_FIELD_NHAT) RCHK(nhat_full(c, sys));
  const double* src = nullptr;
  switch (which) {
    case OPAW_FIELD_DENS: src = c->dens; break;
    case OPAW_FIELD_NHAT: src = c->nhat; break;
    case OPAW_FIELD_VXC: src = c->vxc; break;
    case OPAW_FIELD_VH: src = c->vh; break;
    case OPAW_FIELD_VKS: src = c->vks; break;
    default: return fail(OPAW_ERR_ARG, "unknown field %d", which);
  }
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(out, src + sys * c->N, c->N * sizeof(double), hipMemcpyDeviceToHost));
  return OPAW_OK;
}

static int check_atom(opaw_ctx* c, int ia) {
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  if (ia < 1 || ia > c->natom) return fail(OPAW_ERR_ARG, "atom index %d outside 1..%d", ia, c->natom);
  return OPAW_OK;
}

int opaw_write_dij(opaw_ctx* c, int sys, int ia, const double* dij) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(check_atom(c, ia));
  if (!dij) return fail(OPAW_ERR_ARG, "null dij");
  const AtomDesc& A = c->hatoms[ia - 1];
  HCHK(hipMemcpy(c->dij + (size_t)sys * c->mtot + A.moff, dij, (size_t)A.ms * A.ms * sizeof(double),
                 hipMemcpyHostToDevice));
  return OPAW_OK;
}

int opaw_read_dij(opaw_ctx* c, int sys, int ia, double* dij) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(check_atom(c, ia));
  if (!dij) return fail(OPAW_ERR_ARG, "null dij");
  const AtomDesc& A = c->hatoms[ia - 1];
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(dij, c->dij + (size_t)sys * c->mtot + A.moff, (size_t)A.ms * A.ms * sizeof(double),
                 hipMemcpyDeviceToHost));
  return OPAW_OK;
}

int opaw_read_rhoij(opaw_ctx* c, int sys, int ia, double* rho) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(check_atom(c, ia));
  if (!rho) return fail(OPAW_ERR_ARG, "null rhoij");
  const AtomDesc& A = c->hatoms[ia - 1];
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(rho, c->rhoij + (size_t)sys * c->mtot + A.moff, (size_t)A.ms * A.ms * sizeof(double),
                 hipMemcpyDeviceToHost));
  return OPAW_OK;
}

int opaw_write_orbitals(opaw_ctx* c, int sys, const void* host) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!host) return fail(OPAW_ERR_ARG, "null orbitals");
  const size_t bytes = (size_t)c->nb * c->N * c->esize();
  c->prep_valid = false;
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy((char*)c->phi + sys * bytes, host, bytes, hipMemcpyHostToDevice));
  return OPAW_OK;
}

int opaw_read_orbitals(opaw_ctx* c, int sys, void* host) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!host) return fail(OPAW_ERR_ARG, "null orbitals");
  const size_t bytes = (size_t)c->nb * c->N * c->esize();
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(host, (char*)c->phi + sys * bytes, bytes, hipMemcpyDeviceToHost));
  return OPAW_OK;
}

void* opaw_orbitals_device(opaw_ctx* c, int sys) {
  if (!c || sys < 0 || sys >= c->nsys) return nullptr;
  return (char*)c->phi + (size_t)sys * c->nb * c->N * c->esize();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// operator flows (templated on the orbital precision)
static BoxTables box_tables(opaw_ctx* c) {
  BoxTables b{};
  b.atoms = c->datoms;
  b.natom = c->natom;
  b.ctot = c->ctot;
  b.msp = c->msp;
  b.nchunks = c->nchunks;
  b.chunks = c->chunks;
  b.achunk = c->achunk;
  b.coef = c->coef;
  b.p = c->P;
  b.p1 = c->P1;
  b.bpt = c->bpt;
  b.bk = c->bk;
  b.pptr = c->pptr;
  b.pc = c->pc;
  b.pv = c->pv;
  b.pv1 = c->pv1;
  b.ublk = c->ublk;
  b.nublk = c->nublk;
  b.ucptr = c->ucptr;
  b.ucol = c->ucol;
  b.pcl = c->pcl;
  b.ucmax = c->ucmax;
  b.optr = c->optr;
  b.ovoff = c->ovoff;
  b.oc = c->oc;
  b.ov = c->ov;
  b.ov1 = c->ov1;
  b.U = UnionTabs{c->umap, c->uptr, c->uent, c->upt, c->utp, c->utp1, c->U};
  return b;
}

static SphereTabs sphere_tabs(opaw_ctx* c) { return SphereTabs{c->smap, c->sptr, c->sent, c->sphn}; }

static int need_tables(opaw_ctx* c) {
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "PAW tables not finalised (opaw_finalize_tables)");
  if (!c->ek_ready) return fail(OPAW_ERR_STATE, "kinetic table not set (opaw_ek_prep / opaw_set_ek)");
  return OPAW_OK;
}

static int sn_table(opaw_ctx* c, double n, const double** out) {
  if (n == -1.0) *out = c->sn[0];
  else if (n == -0.5) *out = c->sn[1];
  else if (n == 0.5) *out = c->sn[2];
  else if (n == 1.0) *out = c->sn[3];
  else {
    std::vector<double> v;
    for (int a = 0; a < c->natom; ++a)
      for (int i = 0; i < c->hatoms[a].ms; ++i) v.push_back(std::pow(1.0 + c->hs[a][i], n) - 1.0);
    HCHK(hipStreamSynchronize(c->st));
    if (!v.empty()) HCHK(hipMemcpy(c->sn_custom, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    *out = c->sn_custom;
  }
  return OPAW_OK;
}

static int ensure_coef(opaw_ctx* c, int nf) {
  if (nf <= c->coef_cap) return OPAW_OK;
  const size_t cb = (size_t)nf * std::max(1, c->ctot) * sizeof(c64);
  for (DevBuf* b : {&c->cA, &c->wA, &c->cB, &c->wB, &c->cC, &c->wC}) RCHK(ensure(c, *b, cb));
  RCHK(ensure(c, c->ppart, (size_t)3 * std::max(1, c->nchunks) * nf * c->msp * 2 * sizeof(double)));
  const size_t ub = (size_t)nf * std::max(1, c->U) * sizeof(c64);
  for (DevBuf* b : {&c->corrA, &c->corrNL, &c->corrC, &c->hU}) RCHK(ensure(c, *b, ub));
  c->coef_cap = nf;
  return OPAW_OK;
}

template <typename R>
struct Fields;
template <>
struct Fields<double> {
  static const double* vks(opaw_ctx* c) { return c->vks; }
  static const double* ek(opaw_ctx* c) { return c->ek; }
};
template <>
struct Fields<float> {
  static const float* vks(opaw_ctx* c) { return c->vksf; }
  static const float* ek(opaw_ctx* c) { return c->ekf; }
};

static double dvol(opaw_ctx* c) { return c->dx * c->dy * c->dz; }

// coefficients A (p1 of in, weights sn(-1/2), compact correction corrA) and
// B (p of in + corrA = p of S^-1/2 in); the B weights are formed later with
// the then-current D_ij
// st2 starts after everything enqueued on st so far
static int fork_side(opaw_ctx* c) {
  HCHK(hipEventRecord(c->ev_fork, c->st));
  HCHK(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
  return OPAW_OK;
}

// ev_join marks the end of the side work enqueued so far
static int mark_side(opaw_ctx* c) {
  HCHK(hipEventRecord(c->ev_join, c->st2));
  return OPAW_OK;
}

// st continues only after the side work
static int join_side(opaw_ctx* c) {
  RCHK(mark_side(c));
  HCHK(hipStreamWaitEvent(c->st, c->ev_join, 0));
  return OPAW_OK;
}

// after a failed time step that left a pipelined rebuild unconsumed: fold
// the side stream back into st
static int settle_side(opaw_ctx* c, int r) {
  if (c->pot_pending) {
    c->pot_pending = false;
    (void)join_side(c);
  }
  return r;
}

// Coefficients of `in` for shs and the rebuild: A = P1^T in (weights
// wA = sn(-1/2) A, compact correction corrA = S^-1/2 in - in) and
// B = P^T S^-1/2 in = P^T in + O wA.  With side != 0 the corrA pass runs on
// the side stream (the caller joins it).
template <typename R>
static int flow_prep(opaw_ctx* c, const cx<R>* in, int nf, bool side = false) {
  RCHK(ensure_coef(c, nf));
  const BoxTables bt = box_tables(c);
  HCHK(box_coeffs_dual<R>(c->g, bt, c->st, in, nf, dvol(c), (double*)c->ppart.p, (c64*)c->cA.p, c->sn[1],
                          (c64*)c->wA.p, (c64*)c->cB.p));
  if (side) {
    RCHK(fork_side(c));
    HCHK(union_corr(bt, c->st2, nf, 1, (const c64*)c->wA.p, (c64*)c->corrA.p));
  } else {
    HCHK(union_corr(bt, c->st, nf, 1, (const c64*)c->wA.p, (c64*)c->corrA.p));
  }
  HCHK(box_bfix(bt, c->st, nf, (const c64*)c->wA.p, (c64*)c->cB.p));
  return OPAW_OK;
}

// out = S^n in  (in may equal out)
template <typename R>
static int flow_sn_phi(opaw_ctx* c, const cx<R>* in, cx<R>* out, int nf, double n) {
  RCHK(ensure_coef(c, nf));
  c->prep_valid = false;
  const double* sn = nullptr;
  RCHK(sn_table(c, n, &sn));
  const BoxTables bt = box_tables(c);
  HCHK(box_coeffs<R>(c->g, bt, c->st, in, nf, 1, nullptr, dvol(c), (double*)c->ppart.p, (c64*)c->cA.p, 1, sn,
                     nullptr, nf, c->mtot, (c64*)c->wA.p));
  HCHK(union_corr(bt, c->st, nf, 1, (const c64*)c->wA.p, (c64*)c->corrA.p));
  if (in != out) HCHK(hipMemcpyAsync(out, in, (size_t)nf * c->N * sizeof(cx<R>), hipMemcpyDeviceToDevice, c->st));
  HCHK(union_apply<R>(c->g, bt, c->st, out, nf, (const c64*)c->corrA.p));
  return OPAW_OK;
}

// The kinetic pass of an shs stage, out = T u + V_KS u + corrNL with
// u = in + corrA, and corrNL the union add-back of the weights D(sys) cB.
//  - After a pipelined rebuild (pot_pending), st holds the rebuild's
//    potential chain, whose D_ij kernel has formed the weights: corrNL
//    follows it on st; the forward plane and z passes run on st2 as soon as
//    corrA exists (ev_corr), and the inverse pass runs on st after them
//    (ev_mid).  The critical chain stays on st, whose only wait (for the
//    forward passes) has normally been met by the time st reaches it.
//  - Otherwise the weights and corrNL run on st2 while st runs the forward
//    passes, and the inverse pass waits for them (ev_join).
template <typename R>
static int kinetic_stage(opaw_ctx* c, int sys0, const cx<R>* in, cx<R>* out, int nf, int nf_per_sys,
                         const PlaneRk4<R>* rk) {
  RCHK(ensure(c, c->wtmp, (size_t)nf * c->N * sizeof(cx<R>)));
  const BoxTables bt = box_tables(c);
  const R* vks = Fields<R>::vks(c) + (size_t)sys0 * c->N;
  const bool pending = c->pot_pending && sys0 == 0 && nf == c->nsys * c->nb && nf_per_sys == c->nb;
  c->pot_pending = false;
  if (pending) {
    HCHK(union_corr(bt, c->st, nf, 0, (const c64*)c->wB.p, (c64*)c->corrNL.p));
    HCHK(hipStreamWaitEvent(c->st2, c->ev_corr, 0));
    HCHK(apply_t_plus_v<R>(c->g, c->st2, in, out, (cx<R>*)c->wtmp.p, nf, nf_per_sys, Fields<R>::ek(c), vks,
                           TW<R>::get(c, c->nx), TW<R>::get(c, c->ny), TW<R>::get(c, c->nz), c->umap, c->upt,
                           c->uzr, (const c64*)c->corrA.p, (const c64*)c->corrNL.p, nf, nullptr, rk, c->st,
                           c->ev_mid));
    return OPAW_OK;
  }
  RCHK(fork_side(c));
  HCHK(box_weights(bt, c->st2, (const c64*)c->cB.p, nf, 2, nullptr, c->dij + (size_t)sys0 * c->mtot, nf_per_sys,
                   c->mtot, (c64*)c->wB.p));
  HCHK(union_corr(bt, c->st2, nf, 0, (const c64*)c->wB.p, (c64*)c->corrNL.p));
  RCHK(mark_side(c));
  HCHK(apply_t_plus_v<R>(c->g, c->st, in, out, (cx<R>*)c->wtmp.p, nf, nf_per_sys, Fields<R>::ek(c), vks,
                         TW<R>::get(c, c->nx), TW<R>::get(c, c->ny), TW<R>::get(c, c->nz), c->umap, c->upt, c->uzr,
                         (const c64*)c->corrA.p, (const c64*)c->corrNL.p, nf, c->ev_join, rk));
  return OPAW_OK;
}

// out = H(sys) in for nf fields; field f uses system sys0 + f / nf_per_sys
template <typename R>
static int flow_h_phi(opaw_ctx* c, int sys0, const cx<R>* in, cx<R>* out, int nf, int nf_per_sys) {
  RCHK(ensure_coef(c, nf));
  RCHK(ensure(c, c->wtmp, (size_t)nf * c->N * sizeof(cx<R>)));
  c->prep_valid = false;
  const BoxTables bt = box_tables(c);
  HCHK(box_coeffs<R>(c->g, bt, c->st, in, nf, 0, nullptr, dvol(c), (double*)c->ppart.p, (c64*)c->cB.p, 2, nullptr,
                     c->dij + (size_t)sys0 * c->mtot, nf_per_sys, c->mtot, (c64*)c->wB.p));
  HCHK(union_corr(bt, c->st, nf, 0, (const c64*)c->wB.p, (c64*)c->corrNL.p));
  HCHK(apply_t_plus_v<R>(c->g, c->st, in, out, (cx<R>*)c->wtmp.p, nf, nf_per_sys, Fields<R>::ek(c),
                         Fields<R>::vks(c) + (size_t)sys0 * c->N, TW<R>::get(c, c->nx), TW<R>::get(c, c->ny),
                         TW<R>::get(c, c->nz), c->umap, c->upt, c->uzr, nullptr, (const c64*)c->corrNL.p, nf,
                         nullptr));
  return OPAW_OK;
}

// h = S^-1/2 H S^-1/2 in from the A/B coefficients of `in` (weights B from
// D_ij now): H S^-1/2 in, then the C coefficients of that (p1, weights
// sn(-1/2)) and their add-back into h in place
template <typename R>
static int flow_hs(opaw_ctx* c, int sys0, const cx<R>* in, cx<R>* h, int nf, int nf_per_sys) {
  RCHK(kinetic_stage<R>(c, sys0, in, h, nf, nf_per_sys, nullptr));
  const BoxTables bt = box_tables(c);
  HCHK(box_coeffs<R>(c->g, bt, c->st, h, nf, 1, nullptr, dvol(c), (double*)c->ppart.p, (c64*)c->cC.p, 1, c->sn[1],
                     nullptr, nf, c->mtot, (c64*)c->wC.p));
  // the outer S^-1/2 of shs: h += corrC on the union of boxes
  HCHK(union_corr_apply<R>(c->g, bt, c->st, nf, 1, (const c64*)c->wC.p, (c64*)c->corrC.p, h));
  return OPAW_OK;
}

// One RK4 stage 1-3 with the update fused into the last kinetic pass (plane
// path): y = p + c k and acc = k (first) or acc += 2k, k = -i shs(yin), from
// the A/B coefficients of yin.  The kinetic pass writes h only on the union
// of boxes (hU); its C coefficients give the outer S^-1/2 add-back, which is
// linear in h and is applied to y and acc afterwards.
//
// The next rebuild's coefficients of y come from the same launch: with y0 the
// pass's output and y = y0 + c (-i) P1 wC the stage's,
//   P1^T y dv = P1^T y0 dv + c (-i) O11 wC        (box_afix)
//   P^T S^-1/2 y dv = P^T y0 dv + O (wA + c (-i) wC)   (box_bfix)
// so A, B (of y0) and C (of h) are one projection, and the add-back of wC
// and the compact correction corrA = P1 wA one union pass.  The rebuild
// that follows consumes them (coef_ready) and starts at the density.
template <typename R>
static int flow_hs_rk4(opaw_ctx* c, const cx<R>* yin, cx<R>* y, cx<R>* acc, const cx<R>* p, int nf, bool first,
                       double cst) {
  PlaneRk4<R> rk;
  rk.p = p;
  rk.acc = acc;
  rk.hU = (c64*)c->hU.p;
  rk.c = cst;
  rk.first = first ? 1 : 0;
  RCHK(kinetic_stage<R>(c, 0, yin, y, nf, c->nb, &rk));
  const BoxTables bt = box_tables(c);
  const double a2 = first ? 1.0 : 2.0;
  if (c->pot_ready && c->vk_ready) {
    const c64* wC = (const c64*)c->wC.p;
    HCHK(box_coeffs_triple<R>(c->g, bt, c->st, y, (const c64*)c->hU.p, nf, dvol(c), (double*)c->ppart.p,
                              (c64*)c->cA.p, (c64*)c->cB.p, (c64*)c->cC.p, c->sn[1], (c64*)c->wC.p));
    HCHK(box_afix(bt, c->st, nf, wC, cst, c->sn[1], (c64*)c->cA.p, (c64*)c->wA.p));
    RCHK(fork_side(c));
    HCHK(box_bfix(bt, c->st2, nf, (const c64*)c->wA.p, (c64*)c->cB.p, wC, cst));
    HCHK(union_rk4_both<R>(c->g, bt, c->st, nf, wC, (const c64*)c->wA.p, (c64*)c->corrC.p, (c64*)c->corrA.p, y,
                           acc, cst, a2));
    HCHK(hipEventRecord(c->ev_corr, c->st));
    c->coef_ready = y;
    return OPAW_OK;
  }
  HCHK(box_coeffs_compact(c->g, bt, c->st, (const c64*)c->hU.p, nf, dvol(c), (double*)c->ppart.p, (c64*)c->cC.p,
                          c->sn[1], (c64*)c->wC.p));
  HCHK(union_rk4_apply<R>(c->g, bt, c->st, nf, (const c64*)c->wC.p, (c64*)c->corrC.p, y, acc, cst, a2));
  return OPAW_OK;
}

// out = S^-1/2 H S^-1/2 in
template <typename R>
static int flow_shs(opaw_ctx* c, int sys0, const cx<R>* in, cx<R>* out, int nf, int nf_per_sys) {
  RCHK(flow_prep<R>(c, in, nf));
  c->prep_valid = false;
  RCHK(flow_hs<R>(c, sys0, in, out, nf, nf_per_sys));
  return OPAW_OK;
}

// The per-stage rebuild for systems [s0, s0+ns) from their orbitals phi:
// psi~ = S^-1/2 phi (never stored: its coefficients A and B and its density are
// gathered), rho_ij, density, nhat + rnel renormalisation, V_H + V_xc -> V_KS,
// D_ij (src/ncpaw_make_hamiltonian.f90:213-246).
//
// pipelined (the time loop, whose next call is an shs stage): st runs the
// critical chain (coefficients, corrA, density, renormalisation, Hartree,
// V_KS, D_ij with the next stage's weights) and st2 the B fix-up, rho_ij and
// nhat (handed back by ev_nhat); the stage consumes the state
// (kinetic_stage), running its forward kinetic passes on st2 (they need only
// corrA, marked by ev_corr) while st finishes the chain.
template <typename R>
static int flow_rebuild(opaw_ctx* c, int s0, int ns, const cx<R>* phi, bool pipelined = false) {
  if (!c->pot_ready) return fail(OPAW_ERR_STATE, "potentials not set (opaw_set_potentials)");
  if (!c->vk_ready) return fail(OPAW_ERR_STATE, "Hartree kernel not set (opaw_vk_prep)");
  if (c->pot_pending) return fail(OPAW_ERR_STATE, "a pipelined rebuild was not consumed");
  if (pipelined && (s0 != 0 || ns != c->nsys)) return fail(OPAW_ERR_ARG, "a pipelined rebuild covers every system");
  const int nf = ns * c->nb;
  const long long N = c->N;
  const double dv = dvol(c);
  hipStream_t sd, sr;   // streams of the density pass and of rho_ij + nhat
  const bool ready = pipelined && c->coef_ready == (const void*)phi;
  c->coef_ready = nullptr;
  if (ready) {
    // the RK4 stage before formed A, B (fixing on st2) and corrA (ev_corr)
    sd = c->st;
    sr = c->st2;
  } else if (pipelined) {
    RCHK(ensure_coef(c, nf));
    const BoxTables bt = box_tables(c);
    HCHK(box_coeffs_dual<R>(c->g, bt, c->st, phi, nf, dv, (double*)c->ppart.p, (c64*)c->cA.p, c->sn[1],
                            (c64*)c->wA.p, (c64*)c->cB.p));
    RCHK(fork_side(c));
    HCHK(box_bfix(bt, c->st2, nf, (const c64*)c->wA.p, (c64*)c->cB.p));
    HCHK(union_corr(bt, c->st, nf, 1, (const c64*)c->wA.p, (c64*)c->corrA.p));
    HCHK(hipEventRecord(c->ev_corr, c->st));
    sd = c->st;
    sr = c->st2;
  } else {
    // the coefficients with corrA and then the density pass (which needs
    // corrA) on the side stream, overlapping B, rho_ij and nhat on st
    RCHK(flow_prep<R>(c, phi, nf, true));
    sd = c->st2;
    sr = c->st;
  }
  const int dstride = (int)((N + 255) / 256);
  double* dpart = c->rpart;
  double* npart = c->rpart + (size_t)c->nsys * dstride;
  int nd = 0, nn = 0;
  prof_begin(sd, KC_DENS);
  HCHK(density_sum<R>(sd, phi, N, ns, c->nb, c->dens + s0 * N, c->umap, (const c64*)c->corrA.p, nf, dpart, &nd));
  prof_end(sd, KC_DENS);
  prof_begin(sr, KC_DENS);
  HCHK(rhoij_nhatw(sr, c->datoms, c->natom, (const c64*)c->cB.p, c->ctot, ns, c->nb, c->mtot,
                   c->rhoij + (size_t)s0 * c->mtot, c->qtab, c->nl2, c->wl));
  HCHK(nhat_sparse(sr, sphere_tabs(c), c->S, c->sph_g, c->wl, c->natom, c->nl2, ns,
                   c->nhs + (size_t)s0 * std::max(1, c->S), npart, &nn));
  prof_end(sr, KC_DENS);
  const hipStream_t sp = c->st;   // stream of the potential chain
  if (pipelined) {
    HCHK(hipEventRecord(c->ev_nhat, c->st2));
    HCHK(hipStreamWaitEvent(c->st, c->ev_nhat, 0));
  } else {
    RCHK(join_side(c));
  }
  // on-site Dij^xc from rho_ij: on the side stream beside the Hartree chain
  // when pipelined (joined before D_ij), else in order on st
  if (c->any_onsite) {
    const hipStream_t so = pipelined ? c->st2 : c->st;
    prof_begin(so, KC_POT);
    HCHK(onsite_xc(so, c->datoms, c->natom, c->otypes, c->otab, c->oint, c->rhoij + (size_t)s0 * c->mtot, c->mtot,
                   ns, c->oscr, c->onr_max, c->dxc + (size_t)s0 * c->mtot));
    prof_end(so, KC_POT);
    if (pipelined) HCHK(hipEventRecord(c->ev_onsite, c->st2));
  }
  // the Hartree solve stays FP64 in both precisions: the kick response is the
  // difference of two nearly equal systems, and an FP32 V_H moves the mixed
  // dipole trace by ~4e-4 (relative) over 2000 steps
  const int npair = (ns + 1) / 2;
  if (!c->periodic && hartree_fused_supported(c->g, c->scale)) {
    // renormalisation + pair packing fused into the first Hartree pass; the
    // renormalised density is stored by the V_KS kernel
    const int S1 = std::max(1, c->S);
    const HartreeRebuild H{c->dens + s0 * N, c->smap, c->nhs + (size_t)s0 * S1, S1, dpart, nd, npart, nn, dv,
                           c->rnel, 1, c->fac + s0};
    for (int q = 0; q < npair; ++q)
      HCHK(hartree_fused(c->g, sp, H, q, ns, c->hwork, c->pair + q * N, c->vk, c->tw64[c->gbig.nx],
                         c->tw64[c->gbig.ny], c->tw64[c->gbig.nz]));
    for (int s = s0; s < s0 + ns; ++s) c->nhat_lazy[s] = 1;
    prof_begin(sp, KC_POT);
    HCHK(vks_build(sp, c->dens + s0 * N, c->ncoret, c->pair, nullptr, c->vloc, N, ns, c->vxc + s0 * N,
                   c->vh + s0 * N, c->vks + s0 * N, c->vksf + s0 * N, c->fac + s0));
    prof_end(sp, KC_POT);
  } else {
    prof_begin(sp, KC_POT);
    HCHK(renorm_pack(sp, c->dens + s0 * N, c->smap, c->nhs + (size_t)s0 * std::max(1, c->S), std::max(1, c->S),
                     dpart, nd, npart, nn, dv, c->rnel, 1, N, ns, c->pair, c->fac + s0));
    prof_end(sp, KC_POT);
    for (int s = s0; s < s0 + ns; ++s) c->nhat_lazy[s] = 1;
    for (int q = 0; q < npair; ++q)
      HCHK(hartree_conv(c->g, c->scale, sp, c->pair + q * N, c->pair + q * N, c->hwork, c->vk,
                        c->tw64[c->gbig.nx], c->tw64[c->gbig.ny], c->tw64[c->gbig.nz]));
    if (c->periodic) HCHK(pair_means(sp, c->pair, N, ns, c->mpart, c->means));
    prof_begin(sp, KC_POT);
    HCHK(vks_build(sp, c->dens + s0 * N, c->ncoret, c->pair, c->periodic ? c->means : nullptr, c->vloc, N, ns,
                   c->vxc + s0 * N, c->vh + s0 * N, c->vks + s0 * N, c->vksf + s0 * N));
    prof_end(sp, KC_POT);
  }
  // pipelined: the D_ij kernel also forms the next stage's nonlocal weights
  // wB = D cB (cB of these orbitals, from the prep above)
  const DijWeights wt{(const c64*)c->cB.p, (c64*)c->wB.p, c->ctot, c->nsys * c->nb, c->nb, s0 * c->nb};
  if (pipelined && c->any_onsite) HCHK(hipStreamWaitEvent(sp, c->ev_onsite, 0));
  prof_begin(sp, KC_POT);
  HCHK(dij(sp, c->datoms, c->natom, c->rhoij + (size_t)s0 * c->mtot, c->mtot, c->qtab, c->d0tab, c->etab,
           c->dh_idx, c->dh_g, c->vks + s0 * N, c->vxc + s0 * N, c->usexcnhat == 0 ? 1 : 0, N, ns, dv,
           c->dxc ? c->dxc + (size_t)s0 * c->mtot : nullptr, c->dij + (size_t)s0 * c->mtot,
           pipelined ? &wt : nullptr));
  prof_end(sp, KC_POT);
  c->pot_pending = pipelined;
  c->prep_valid = (s0 == 0 && ns == c->nsys && (const void*)phi == c->phi);
  return OPAW_OK;
}

// one RK4 step of all systems (src/rk4.f90:1-103).  Stage s applies
// shs to y_s from the A/B coefficients of y_s (computed by the previous
// rebuild), updates acc and y with the outer S^-1/2 gathered into the update,
// and rebuilds H from the new y.
template <typename R>
static int flow_rk4(opaw_ctx* c, double dt, double* dip_out) {
  const int nf = c->nsys * c->nb;
  const long long n = (long long)nf * c->N;
  const double dv = dvol(c);
  RCHK(ensure(c, c->wy, n * sizeof(cx<R>)));
  RCHK(ensure(c, c->wacc, n * sizeof(cx<R>)));
  RCHK(ensure(c, c->wh, n * sizeof(cx<R>)));
  cx<R>* p = (cx<R>*)c->phi;
  cx<R>* y = (cx<R>*)c->wy.p;
  cx<R>* acc = (cx<R>*)c->wacc.p;
  cx<R>* h = (cx<R>*)c->wh.p;
  if (!c->prep_valid) RCHK(flow_prep<R>(c, p, nf));
  // after a pipelined rebuild st carries the critical chain: the norms go to
  // st2, ahead of the first stage's forward passes (everything after those
  // waits for them)
  const hipStream_t sn = c->pot_pending ? c->st2 : c->st;
  prof_begin(sn, KC_DENS);
  HCHK(field_norms<R>(sn, p, c->N, nf, dv, c->part, c->nrm0));
  prof_end(sn, KC_DENS);
  const double cst[4] = {dt / 2.0, dt / 2.0, dt, dt};
  const bool fused = plane_supported<R>(c->nx, c->ny);
  for (int stage = 1; stage <= 4; ++stage) {
    const cx<R>* yin = (stage == 1) ? p : y;
    if (fused && stage < 4) {
      // stages 1-3: the update rides in the last kinetic pass
      RCHK(flow_hs_rk4<R>(c, yin, y, acc, p, nf, stage == 1, cst[stage - 1]));
      RCHK(flow_rebuild<R>(c, 0, c->nsys, y, true));
      continue;
    }
    RCHK(flow_hs<R>(c, 0, yin, h, nf, c->nb));
    // stage 4 also forms the post-step norms (norm1) of every orbital
    prof_begin(c->st, KC_RK4);
    HCHK(rk4_update<R>(c->st, stage, c->N, nf, h, p, acc, y, cst[stage - 1], dt, c->part, dv, c->nrm1));
    prof_end(c->st, KC_RK4);
    if (stage < 4) RCHK(flow_rebuild<R>(c, 0, c->nsys, y, true));
  }
  // norm restore; inside the time loop fused with the dipole of the step
  prof_begin(c->st, KC_DENS);
  if (dip_out)
    HCHK(scale_dipole<R>(c->st, p, c->nx, c->ny, c->nz, c->nsys, c->nb, c->dx, c->dy, c->dz, c->nrm0, c->nrm1,
                         c->part, dip_out));
  else
    HCHK(scale_fields<R>(c->st, p, c->N, nf, c->nrm0, c->nrm1));
  prof_end(c->st, KC_DENS);
  c->prep_valid = false;
  return OPAW_OK;
}

// ---------------------------------------------------------------------------
// host Hermitian eigensolver, FP64 (Householder tridiagonalisation + implicit
// QL, opaw_linalg.h).  A is n x n row-major A[i*n+j] = A(i,j); returns
// ascending eigenvalues and column eigenvectors V[i*n+k] = V(i,k).
static void herm_eig(int n, std::vector<std::complex<double>> A, std::vector<double>& w,
                     std::vector<std::complex<double>>& V) {
  herm_eig_tridiag(n, std::move(A), w, V);
}

// hm = dv * <a_i | b_j> for nb fields; downloads the matrix (row-major)
template <typename R>
static int gram_host(opaw_ctx* c, const cx<R>* a, const cx<R>* b, std::vector<std::complex<double>>& hm) {
  const int nb = c->nb;
  HCHK(gram<R>(c->st, a, b, c->N, nb, c->dx * c->dy * c->dz, c->gpart, c->gmat));
  hm.resize((size_t)nb * nb);
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(hm.data(), c->gmat, (size_t)nb * nb * sizeof(c64), hipMemcpyDeviceToHost));
  return OPAW_OK;
}

// phi <- phi T with T(i,j) (column-major upload), through the wtmp buffer
template <typename R>
static int rotate_resident(opaw_ctx* c, int sys, const std::vector<std::complex<double>>& Tcol) {
  const int nb = c->nb;
  RCHK(ensure(c, c->wtmp, (size_t)nb * c->N * sizeof(cx<R>)));
  HCHK(hipMemcpy(c->tmat, Tcol.data(), (size_t)nb * nb * sizeof(c64), hipMemcpyHostToDevice));
  cx<R>* phi = (cx<R>*)c->phi + (size_t)sys * nb * c->N;
  HCHK(rotate<R>(c->st, phi, (cx<R>*)c->wtmp.p, c->N, nb, c->tmat));
  HCHK(hipMemcpyAsync(phi, c->wtmp.p, (size_t)nb * c->N * sizeof(cx<R>), hipMemcpyDeviceToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  c->prep_valid = false;
  return OPAW_OK;
}

template <typename R>
static int flow_orthog(opaw_ctx* c, int sys, int h_type, double* eig) {
  const int nb = c->nb;
  cx<R>* phi = (cx<R>*)c->phi + (size_t)sys * nb * c->N;
  const cx<R>* sp = phi;
  if (h_type == 0) {
    RCHK(ensure(c, c->wu, (size_t)nb * c->N * sizeof(cx<R>)));
    RCHK(flow_sn_phi<R>(c, phi, (cx<R>*)c->wu.p, nb, 1.0));
    sp = (const cx<R>*)c->wu.p;
  }
  std::vector<std::complex<double>> hm, V;
  RCHK(gram_host<R>(c, sp, phi, hm));
  std::vector<double> w;
  herm_eig(nb, hm, w, V);
  for (int k = 0; k < nb; ++k)
    if (!(w[k] > 0.0)) return fail(OPAW_ERR_ARG, "overlap eigenvalue %d is %g (linearly dependent orbitals)", k + 1, w[k]);
  std::vector<std::complex<double>> T((size_t)nb * nb);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) T[i + (size_t)nb * j] = V[(size_t)i * nb + j] / std::sqrt(w[j]);
  RCHK(rotate_resident<R>(c, sys, T));
  if (eig) std::copy(w.begin(), w.end(), eig);
  return OPAW_OK;
}

template <typename R>
static int flow_diagh(opaw_ctx* c, int sys, int h_type, double* eig) {
  const int nb = c->nb;
  RCHK(need_tables(c));
  cx<R>* phi = (cx<R>*)c->phi + (size_t)sys * nb * c->N;
  RCHK(ensure(c, c->wh, (size_t)nb * c->N * sizeof(cx<R>)));
  cx<R>* hp = (cx<R>*)c->wh.p;
  if (h_type == 1)
    RCHK(flow_shs<R>(c, sys, phi, hp, nb, nb));
  else
    RCHK(flow_h_phi<R>(c, sys, phi, hp, nb, nb));
  std::vector<std::complex<double>> hm, V;
  RCHK(gram_host<R>(c, hp, phi, hm));
  std::vector<double> w;
  herm_eig(nb, hm, w, V);
  std::vector<std::complex<double>> T((size_t)nb * nb);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) T[i + (size_t)nb * j] = V[(size_t)i * nb + j];
  RCHK(rotate_resident<R>(c, sys, T));
  if (eig) std::copy(w.begin(), w.end(), eig);
  return OPAW_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

int opaw_h_phi(opaw_ctx* c, int sys, const void* pin, void* hp, int nbatch) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_tables(c));
  if (!pin || !hp || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad h_phi arguments");
  if (pin == hp) return fail(OPAW_ERR_ARG, "h_phi needs distinct input and output (src/hphi.f90:1-10)");
  if (c->prec == OPAW_PREC_F64) return flow_h_phi<double>(c, sys, (const c64*)pin, (c64*)hp, nbatch, nbatch);
  return flow_h_phi<float>(c, sys, (const c32*)pin, (c32*)hp, nbatch, nbatch);
}

int opaw_sn_phi(opaw_ctx* c, const void* pin, void* sp, int nbatch, double n) {
  CTX_CHECK(c);
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  if (!pin || !sp || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad sn_phi arguments");
  if (c->prec == OPAW_PREC_F64) return flow_sn_phi<double>(c, (const c64*)pin, (c64*)sp, nbatch, n);
  return flow_sn_phi<float>(c, (const c32*)pin, (c32*)sp, nbatch, n);
}

int opaw_shs(opaw_ctx* c, int sys, const void* pin, void* pout, int nbatch) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_tables(c));
  if (!pin || !pout || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad shs arguments");
  if (pin == pout) return fail(OPAW_ERR_ARG, "shs needs distinct input and output (src/shs.f90:3-12)");
  if (c->prec == OPAW_PREC_F64) return flow_shs<double>(c, sys, (const c64*)pin, (c64*)pout, nbatch, nbatch);
  return flow_shs<float>(c, sys, (const c32*)pin, (c32*)pout, nbatch, nbatch);
}

int opaw_proj(opaw_ctx* c, const void* pin, int nbatch, int which_p1, void* ca) {
  CTX_CHECK(c);
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  if (!pin || !ca || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad proj arguments");
  RCHK(ensure_coef(c, nbatch));
  const BoxTables bt = box_tables(c);
  if (c->prec == OPAW_PREC_F64)
    HCHK(box_coeffs<double>(c->g, bt, c->st, (const c64*)pin, nbatch, which_p1 ? 1 : 0, nullptr, dvol(c),
                            (double*)c->ppart.p, (c64*)ca, 0, nullptr, nullptr, nbatch, c->mtot, nullptr));
  else
    HCHK(box_coeffs<float>(c->g, bt, c->st, (const c32*)pin, nbatch, which_p1 ? 1 : 0, nullptr, dvol(c),
                           (double*)c->ppart.p, (c64*)ca, 0, nullptr, nullptr, nbatch, c->mtot, nullptr));
  return OPAW_OK;
}

// ---------------------------------------------------------------------------
// host-array entry points: complex*16 fields in host memory, in the
// reference's layout (x fastest), converted to and from the context
// precision around the device operators
static int upload_c16(opaw_ctx* c, const double* host, void* dev, long long nelem) {
  HCHK(hipStreamSynchronize(c->st));
  if (c->prec == OPAW_PREC_F64) {
    HCHK(hipMemcpy(dev, host, (size_t)nelem * sizeof(c64), hipMemcpyHostToDevice));
    return OPAW_OK;
  }
  std::vector<c32> t((size_t)nelem);
  for (long long i = 0; i < nelem; ++i) t[i] = mk<float>((float)host[2 * i], (float)host[2 * i + 1]);
  HCHK(hipMemcpy(dev, t.data(), (size_t)nelem * sizeof(c32), hipMemcpyHostToDevice));
  return OPAW_OK;
}

static int download_c16(opaw_ctx* c, const void* dev, double* host, long long nelem) {
  HCHK(hipStreamSynchronize(c->st));
  if (c->prec == OPAW_PREC_F64) {
    HCHK(hipMemcpy(host, dev, (size_t)nelem * sizeof(c64), hipMemcpyDeviceToHost));
    return OPAW_OK;
  }
  std::vector<c32> t((size_t)nelem);
  HCHK(hipMemcpy(t.data(), dev, (size_t)nelem * sizeof(c32), hipMemcpyDeviceToHost));
  for (long long i = 0; i < nelem; ++i) {
    host[2 * i] = (double)t[i].x;
    host[2 * i + 1] = (double)t[i].y;
  }
  return OPAW_OK;
}

// staging for nbatch fields in and out
static int stage_fields(opaw_ctx* c, int nbatch) {
  const size_t bytes = (size_t)nbatch * c->N * c->esize();
  RCHK(ensure(c, c->hin, bytes));
  RCHK(ensure(c, c->hout, bytes));
  return OPAW_OK;
}

int opaw_set_dt(opaw_ctx* c, double dt) {
  CTX_CHECK(c);
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(OPAW_ERR_ARG, "dt must be positive and finite");
  c->dt = dt;
  return OPAW_OK;
}

int opaw_write_orbitals_c16(opaw_ctx* c, int sys, const double* host) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!host) return fail(OPAW_ERR_ARG, "null orbitals");
  c->prep_valid = false;
  return upload_c16(c, host, opaw_orbitals_device(c, sys), (long long)c->nb * c->N);
}

int opaw_read_orbitals_c16(opaw_ctx* c, int sys, double* host) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!host) return fail(OPAW_ERR_ARG, "null orbitals");
  return download_c16(c, opaw_orbitals_device(c, sys), host, (long long)c->nb * c->N);
}

int opaw_h_phi_host(opaw_ctx* c, int sys, const double* pin, double* hp, int nbatch) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!pin || !hp || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad h_phi arguments");
  RCHK(stage_fields(c, nbatch));
  RCHK(upload_c16(c, pin, c->hin.p, (long long)nbatch * c->N));
  RCHK(opaw_h_phi(c, sys, c->hin.p, c->hout.p, nbatch));
  return download_c16(c, c->hout.p, hp, (long long)nbatch * c->N);
}

int opaw_shs_host(opaw_ctx* c, int sys, const double* pin, double* pout, int nbatch) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!pin || !pout || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad shs arguments");
  RCHK(stage_fields(c, nbatch));
  RCHK(upload_c16(c, pin, c->hin.p, (long long)nbatch * c->N));
  RCHK(opaw_shs(c, sys, c->hin.p, c->hout.p, nbatch));
  return download_c16(c, c->hout.p, pout, (long long)nbatch * c->N);
}

int opaw_sn_phi_host(opaw_ctx* c, const double* pin, double* sp, int nbatch, double n) {
  CTX_CHECK(c);
  if (!pin || !sp || nbatch <= 0) return fail(OPAW_ERR_ARG, "bad sn_phi arguments");
  RCHK(stage_fields(c, nbatch));
  RCHK(upload_c16(c, pin, c->hin.p, (long long)nbatch * c->N));
  RCHK(opaw_sn_phi(c, c->hin.p, c->hout.p, nbatch, n));
  return download_c16(c, c->hout.p, sp, (long long)nbatch * c->N);
}

int opaw_proj_host(opaw_ctx* c, int ia, const double* pin, int which_p1, double* ca, int ms) {
  CTX_CHECK(c);
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  if (ia < 1 || ia > c->natom) return fail(OPAW_ERR_ARG, "atom %d outside 1..%d", ia, c->natom);
  const AtomDesc& A = c->hatoms[ia - 1];
  if (ms != A.ms) return fail(OPAW_ERR_ARG, "atom %d has %d projectors, not %d", ia, A.ms, ms);
  if (!pin || !ca) return fail(OPAW_ERR_ARG, "bad proj arguments");
  RCHK(stage_fields(c, 1));
  RCHK(ensure(c, c->hca, (size_t)std::max(1, c->ctot) * sizeof(c64)));
  RCHK(upload_c16(c, pin, c->hin.p, c->N));
  RCHK(opaw_proj(c, c->hin.p, 1, which_p1, c->hca.p));
  std::vector<c64> all((size_t)std::max(1, c->ctot));
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(all.data(), c->hca.p, all.size() * sizeof(c64), hipMemcpyDeviceToHost));
  for (int i = 0; i < ms; ++i) {
    ca[2 * i] = all[A.coff + i].x;
    ca[2 * i + 1] = all[A.coff + i].y;
  }
  return OPAW_OK;
}

int opaw_rk4_prop_shs_host(opaw_ctx* c, double* p, double* pp, long long n, int nb) {
  CTX_CHECK(c);
  if (!p || !pp) return fail(OPAW_ERR_ARG, "null p / pp");
  if (n != c->N || nb != c->nb) return fail(OPAW_ERR_ARG, "p(n, nb) is (%lld, %d), the context holds (%lld, %d)", n, nb, c->N, c->nb);
  if (c->nsys < 2) return fail(OPAW_ERR_STATE, "rk4_prop_shs propagates phi_bar and phi_bar_pert (nsys >= 2)");
  if (!(c->dt > 0.0)) return fail(OPAW_ERR_STATE, "dt not set (opaw_set_dt)");
  // p and pp must be the orbitals the current Hamiltonian was built from: the
  // host copies are compared with the resident ones, word by word
  const long long ne = (long long)nb * c->N;
  RCHK(stage_fields(c, nb));
  RCHK(ensure(c, c->hflag, sizeof(int)));
  const size_t bytes = (size_t)ne * c->esize();
  for (int s = 0; s < 2; ++s) {
    RCHK(upload_c16(c, s == 0 ? p : pp, c->hin.p, ne));
    HCHK(hipMemsetAsync(c->hflag.p, 0, sizeof(int), c->st));
    HCHK(words_differ(c->st, (const unsigned*)c->hin.p, (const unsigned*)opaw_orbitals_device(c, s),
                      (long long)(bytes / sizeof(unsigned)), (int*)c->hflag.p));
    int flag = 0;
    HCHK(hipStreamSynchronize(c->st));
    HCHK(hipMemcpy(&flag, c->hflag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag)
      return fail(OPAW_ERR_STATE,
                  "%s differs from the resident orbitals the Hamiltonian was built from: write it "
                  "(opaw_write_orbitals_c16) and rebuild (opaw_make_hamiltonian_tddft) first",
                  s == 0 ? "p (phi_bar)" : "pp (phi_bar_pert)");
  }
  RCHK(opaw_rk4_prop_shs(c, c->dt));
  RCHK(download_c16(c, opaw_orbitals_device(c, 0), p, ne));
  return download_c16(c, opaw_orbitals_device(c, 1), pp, ne);
}

int opaw_fft3d(opaw_ctx* c, int dir, const void* in, void* out, int nbatch) {
  CTX_CHECK(c);
  if (!in || !out || nbatch <= 0 || (dir != 1 && dir != -1)) return fail(OPAW_ERR_ARG, "bad fft3d arguments");
  if (c->prec == OPAW_PREC_F64)
    HCHK(fft3d<double>(c->g, c->st, dir, (const c64*)in, (c64*)out, nbatch, c->tw64[c->nx], c->tw64[c->ny],
                       c->tw64[c->nz]));
  else
    HCHK(fft3d<float>(c->g, c->st, dir, (const c32*)in, (c32*)out, nbatch, c->tw32[c->nx], c->tw32[c->ny],
                      c->tw32[c->nz]));
  return OPAW_OK;
}

int opaw_tddft_rhoij_2(opaw_ctx* c, int sys, const void* psi) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  if (!psi) return fail(OPAW_ERR_ARG, "null psi");
  const int nb = c->nocc;   // occupied orbitals only
  RCHK(ensure_coef(c, nb));
  c->prep_valid = false;
  RCHK(opaw_proj(c, psi, nb, 0, c->cB.p));
  HCHK(rhoij(c->st, c->datoms, c->natom, (const c64*)c->cB.p, c->ctot, 1, nb, c->mtot,
             c->rhoij + (size_t)sys * c->mtot));
  return OPAW_OK;
}

int opaw_tddft_dens_2(opaw_ctx* c, int sys, const void* psi) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!psi) return fail(OPAW_ERR_ARG, "null psi");
  if (c->prec == OPAW_PREC_F64)
    HCHK(density<double>(c->st, (const c64*)psi, c->N, 1, c->nocc, c->dens + sys * c->N, nullptr, nullptr, 0));
  else
    HCHK(density<float>(c->st, (const c32*)psi, c->N, 1, c->nocc, c->dens + sys * c->N, nullptr, nullptr, 0));
  return OPAW_OK;
}

int opaw_get_nhat(opaw_ctx* c, int sys) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!c->tables_ready || !c->pot_ready) return fail(OPAW_ERR_STATE, "tables / potentials not set");
  const long long N = c->N;
  HCHK(nhat(c->st, c->datoms, c->natom, c->rhoij + (size_t)sys * c->mtot, c->mtot, c->qtab, c->nl2, sphere_tabs(c),
            c->sph_g, N, 1, c->wl, c->nhat + sys * N));
  if (sys < (int)c->nhat_lazy.size()) c->nhat_lazy[sys] = 0;
  HCHK(renorm_factor(c->st, c->nhat + sys * N, c->dens + sys * N, N, 1, dvol(c), c->rnel, c->part, c->fac + sys));
  HCHK(scale_pack(c->st, c->dens + sys * N, c->nhat + sys * N, c->fac + sys, N, 1, c->pair));
  return OPAW_OK;
}

int opaw_get_pot_tddft_2(opaw_ctx* c, int sys) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!c->pot_ready || !c->vk_ready) return fail(OPAW_ERR_STATE, "potentials / Hartree kernel not set");
  const long long N = c->N;
  RCHK(nhat_full(c, sys));
  HCHK(scale_pack(c->st, c->dens + sys * N, c->nhat + sys * N, nullptr, N, 1, c->pair));
  HCHK(hartree_conv(c->g, c->scale, c->st, c->pair, c->pair, c->hwork, c->vk, c->tw64[c->gbig.nx],
                    c->tw64[c->gbig.ny], c->tw64[c->gbig.nz]));
  if (c->periodic) HCHK(pair_means(c->st, c->pair, N, 1, c->part, c->means));
  HCHK(vks_build(c->st, c->dens + sys * N, c->ncoret, c->pair, c->periodic ? c->means : nullptr, c->vloc, N, 1,
                 c->vxc + sys * N, c->vh + sys * N, c->vks + sys * N, c->vksf + sys * N));
  return OPAW_OK;
}

int opaw_get_dij(opaw_ctx* c, int sys) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (!c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  const long long N = c->N;
  if (c->any_onsite)
    HCHK(onsite_xc(c->st, c->datoms, c->natom, c->otypes, c->otab, c->oint, c->rhoij + (size_t)sys * c->mtot,
                   c->mtot, 1, c->oscr, c->onr_max, c->dxc + (size_t)sys * c->mtot));
  HCHK(dij(c->st, c->datoms, c->natom, c->rhoij + (size_t)sys * c->mtot, c->mtot, c->qtab, c->d0tab, c->etab,
           c->dh_idx, c->dh_g, c->vks + sys * N, c->vxc + sys * N, c->usexcnhat == 0 ? 1 : 0, N, 1,
           c->dx * c->dy * c->dz, c->dxc ? c->dxc + (size_t)sys * c->mtot : nullptr, c->dij + (size_t)sys * c->mtot));
  return OPAW_OK;
}

int opaw_make_hamiltonian_tddft(opaw_ctx* c, int sys) {
  CTX_CHECK(c);
  if (c->nocc != c->nb) return fail(OPAW_ERR_STATE, "the propagation runs on occupied orbitals only (nocc = nb)");
  RCHK(need_tables(c));
  int s0 = 0, ns = c->nsys;
  if (sys >= 0) {
    RCHK(check_sys(c, sys));
    s0 = sys;
    ns = 1;
  }
  if (c->prec == OPAW_PREC_F64)
    return flow_rebuild<double>(c, s0, ns, (const c64*)c->phi + (size_t)s0 * c->nb * c->N);
  return flow_rebuild<float>(c, s0, ns, (const c32*)c->phi + (size_t)s0 * c->nb * c->N);
}

int opaw_excite_psi_bar_pert(opaw_ctx* c, int from_sys, int to_sys, double sm, int ipol) {
  CTX_CHECK(c);
  RCHK(check_sys(c, from_sys));
  RCHK(check_sys(c, to_sys));
  if (ipol < 1 || ipol > 3) return fail(OPAW_ERR_ARG, "ipol %d outside 1..3", ipol);
  const double d = ipol == 1 ? c->dx : ipol == 2 ? c->dy : c->dz;
  const size_t fs = (size_t)c->nb * c->N;
  c->prep_valid = false;
  if (c->prec == OPAW_PREC_F64)
    HCHK(kick<double>(c->st, (const c64*)c->phi + from_sys * fs, (c64*)c->phi + to_sys * fs, c->nx, c->ny, c->nz, c->nb,
                      ipol - 1, d, sm));
  else
    HCHK(kick<float>(c->st, (const c32*)c->phi + from_sys * fs, (c32*)c->phi + to_sys * fs, c->nx, c->ny, c->nz, c->nb,
                     ipol - 1, d, sm));
  return OPAW_OK;
}

int opaw_rk4_prop_shs(opaw_ctx* c, double dt) {
  CTX_CHECK(c);
  if (c->nocc != c->nb) return fail(OPAW_ERR_STATE, "the propagation runs on occupied orbitals only (nocc = nb)");
  RCHK(need_tables(c));
  if (c->prec == OPAW_PREC_F64) return settle_side(c, flow_rk4<double>(c, dt, nullptr));
  return settle_side(c, flow_rk4<float>(c, dt, nullptr));
}

static int dipoles_to(opaw_ctx* c, double* dst) {
  prof_begin(c->st, KC_DENS);
  if (c->prec == OPAW_PREC_F64)
    HCHK(dipole<double>(c->st, (const c64*)c->phi, c->nx, c->ny, c->nz, c->nsys, c->nb, c->dx, c->dy, c->dz,
                        c->part, dst));
  else
    HCHK(dipole<float>(c->st, (const c32*)c->phi, c->nx, c->ny, c->nz, c->nsys, c->nb, c->dx, c->dy, c->dz, c->part,
                       dst));
  prof_end(c->st, KC_DENS);
  return OPAW_OK;
}

int opaw_plot_dip(opaw_ctx* c, double sm, double* d_dip, double* dip, double* dip_pert) {
  CTX_CHECK(c);
  if (c->dipcap < 1) {
    dfree(c, c->dipbuf);
    RCHK(dalloc_t(c, &c->dipbuf, (size_t)3 * c->nsys));
    c->dipcap = 1;
  }
  RCHK(dipoles_to(c, c->dipbuf));
  std::vector<double> h((size_t)3 * c->nsys);
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(h.data(), c->dipbuf, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < 3; ++k) {
    if (dip) dip[k] = h[k];
    if (dip_pert) dip_pert[k] = c->nsys > 1 ? h[3 + k] : 0.0;
    if (d_dip) d_dip[k] = c->nsys > 1 ? (h[3 + k] - h[k]) / sm : 0.0;
  }
  return OPAW_OK;
}

// one iteration of the time loop for every system of c: the rebuild, then
// rk4_prop_shs with the dipole of the step written to slot (3 nsys doubles,
// device).  Enqueues only; no host synchronisation.
static int tddft_step(opaw_ctx* c, double dt, double* slot) {
  int r;
  if (c->prec == OPAW_PREC_F64) {
    r = flow_rebuild<double>(c, 0, c->nsys, (const c64*)c->phi, true);
    if (r == OPAW_OK) r = flow_rk4<double>(c, dt, slot);
  } else {
    r = flow_rebuild<float>(c, 0, c->nsys, (const c32*)c->phi, true);
    if (r == OPAW_OK) r = flow_rk4<float>(c, dt, slot);
  }
  return settle_side(c, r);
}

static int tddft_ready(opaw_ctx* c, int nt) {
  if (c->nocc != c->nb) return fail(OPAW_ERR_STATE, "the propagation runs on occupied orbitals only (nocc = nb)");
  RCHK(need_tables(c));
  if (!c->pot_ready) return fail(OPAW_ERR_STATE, "potentials not set (opaw_set_potentials)");
  if (!c->vk_ready) return fail(OPAW_ERR_STATE, "Hartree kernel not set (opaw_vk_prep)");
  if (nt <= 0) return fail(OPAW_ERR_ARG, "nt must be positive");
  if (c->dipcap < nt) {
    dfree(c, c->dipbuf);
    c->dipbuf = nullptr;
    c->dipcap = 0;
    RCHK(dalloc_t(c, &c->dipbuf, (size_t)3 * c->nsys * nt));
    c->dipcap = nt;
  }
  return OPAW_OK;
}

// dipole series of the last nt steps: dip (nt, 3 nsys) and, from systems 0
// and 1, d_dip (nt, 3) = (dip_pert - dip) / sm; either may be NULL
static int dipoles_download(opaw_ctx* c, int nt, double sm, double* d_dip, double* dip) {
  if (!d_dip && !dip) return OPAW_OK;
  std::vector<double> h((size_t)3 * c->nsys * nt);
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(h.data(), c->dipbuf, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int it = 0; it < nt; ++it) {
    const double* r = h.data() + (size_t)3 * c->nsys * it;
    if (dip)
      for (int k = 0; k < 3 * c->nsys; ++k) dip[(size_t)3 * c->nsys * it + k] = r[k];
    if (d_dip)
      for (int k = 0; k < 3; ++k) d_dip[(size_t)3 * it + k] = c->nsys > 1 ? (r[3 + k] - r[k]) / sm : 0.0;
  }
  return OPAW_OK;
}

int opaw_tddft_run(opaw_ctx* c, int nt, double dt, double sm, double* d_dip, double* dip) {
  CTX_CHECK(c);
  RCHK(tddft_ready(c, nt));
  for (int it = 0; it < nt; ++it) RCHK(tddft_step(c, dt, c->dipbuf + (size_t)3 * c->nsys * it));
  return dipoles_download(c, nt, sm, d_dip, dip);
}

// one time step of c recorded (not run) from its stream into an executable
// graph; the dipole of a replay lands in c->dipstep
static int capture_step(opaw_ctx* c, double dt, hipGraphExec_t* exec) {
  *exec = nullptr;
  HCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  const int r = tddft_step(c, dt, c->dipstep);
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(c->st, &gr);
  if (r != OPAW_OK || e != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return r != OPAW_OK ? r : fail(OPAW_ERR_HIP, "stream capture of a time step: %s", hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(exec, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (ei != hipSuccess) {
    *exec = nullptr;
    return fail(OPAW_ERR_HIP, "graph instantiation of a time step: %s", hipGetErrorString(ei));
  }
  return OPAW_OK;
}

// the time loop of several contexts, each on its own stream: step 0 runs
// eagerly (it also sizes every work buffer), then either the remaining steps
// are enqueued context after context, or one step per context is captured
// and its graph replayed
static int batch_loop(opaw_ctx* const* cs, int nctx, int nt, double dt, int graph) {
  for (int k = 0; k < nctx; ++k) {
    ProfScope ps(cs[k]);
    RCHK(tddft_step(cs[k], dt, cs[k]->dipbuf));
  }
  if (nt == 1) return OPAW_OK;
  if (!graph) {
    for (int it = 1; it < nt; ++it)
      for (int k = 0; k < nctx; ++k) {
        ProfScope ps(cs[k]);
        RCHK(tddft_step(cs[k], dt, cs[k]->dipbuf + (size_t)3 * cs[k]->nsys * it));
      }
    return OPAW_OK;
  }
  std::vector<hipGraphExec_t> ex(nctx, nullptr);
  int r = OPAW_OK;
  for (int k = 0; k < nctx && r == OPAW_OK; ++k) r = capture_step(cs[k], dt, &ex[k]);
  for (int it = 1; it < nt && r == OPAW_OK; ++it)
    for (int k = 0; k < nctx && r == OPAW_OK; ++k) {
      opaw_ctx* c = cs[k];
      const size_t nd = (size_t)3 * c->nsys;
      hipError_t e = hipGraphLaunch(ex[k], c->st);
      if (e == hipSuccess)
        e = hipMemcpyAsync(c->dipbuf + nd * it, c->dipstep, nd * sizeof(double), hipMemcpyDeviceToDevice, c->st);
      if (e != hipSuccess) r = fail(OPAW_ERR_HIP, "graph replay of a time step: %s", hipGetErrorString(e));
    }
  // the replays must have ended before their graphs go
  for (int k = 0; k < nctx; ++k)
    if (ex[k]) {
      const hipError_t e = hipStreamSynchronize(cs[k]->st);
      if (e != hipSuccess && r == OPAW_OK) r = fail(OPAW_ERR_HIP, "graph replay: %s", hipGetErrorString(e));
      (void)hipGraphExecDestroy(ex[k]);
    }
  return r;
}

int opaw_tddft_run_batch(opaw_ctx* const* cs, int nctx, int nt, double dt, int graph, double* dip) {
  if (!cs || nctx <= 0) return fail(OPAW_ERR_ARG, "no contexts");
  for (int k = 0; k < nctx; ++k) {
    opaw_ctx* c = cs[k];
    if (!c) return fail(OPAW_ERR_ARG, "null context %d", k);
    for (int j = 0; j < k; ++j)
      if (cs[j] == c) return fail(OPAW_ERR_ARG, "context %d appears twice", k);
    ProfScope ps(c);
    RCHK(tddft_ready(c, nt));
    if (graph && c->prof.on) return fail(OPAW_ERR_STATE, "graph replay with kernel timing enabled (context %d)", k);
    if (!c->dipstep) RCHK(dalloc_t(c, &c->dipstep, (size_t)3 * c->nsys));
    if (!c->bst) HCHK(hipStreamCreateWithFlags(&c->bst, hipStreamNonBlocking));
  }
  // each context moves from its caller's stream to its own and back
  std::vector<hipStream_t> home(nctx);
  for (int k = 0; k < nctx; ++k) {
    opaw_ctx* c = cs[k];
    home[k] = c->st;
    HCHK(hipEventRecord(c->ev_fork, c->st));
    HCHK(hipStreamWaitEvent(c->bst, c->ev_fork, 0));
    c->st = c->bst;
  }
  int r = batch_loop(cs, nctx, nt, dt, graph);
  for (int k = 0; k < nctx; ++k) {
    opaw_ctx* c = cs[k];
    hipError_t e = hipEventRecord(c->ev_join, c->bst);
    if (e == hipSuccess) e = hipStreamWaitEvent(home[k], c->ev_join, 0);
    if (e != hipSuccess && r == OPAW_OK) r = fail(OPAW_ERR_HIP, "stream hand-back: %s", hipGetErrorString(e));
    c->st = home[k];
  }
  if (r != OPAW_OK || !dip) return r;
  size_t off = 0;
  for (int k = 0; k < nctx; ++k) {
    RCHK(dipoles_download(cs[k], nt, 0.0, nullptr, dip + off));
    off += (size_t)3 * cs[k]->nsys * nt;
  }
  return OPAW_OK;
}

int opaw_orthog_phi(opaw_ctx* c, int sys, int h_type, double* eig) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (h_type == 0 && !c->tables_ready) return fail(OPAW_ERR_STATE, "tables not finalised");
  if (c->prec == OPAW_PREC_F64) return flow_orthog<double>(c, sys, h_type, eig);
  return flow_orthog<float>(c, sys, h_type, eig);
}

int opaw_diagh(opaw_ctx* c, int sys, int h_type, double* eig) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  if (c->prec == OPAW_PREC_F64) return flow_diagh<double>(c, sys, h_type, eig);
  return flow_diagh<float>(c, sys, h_type, eig);
}

int opaw_profile_enable(opaw_ctx* c, int mask) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  c->prof.mask = (unsigned)mask;
  c->prof.on = mask != 0;
  return OPAW_OK;
}

int opaw_profile_reset(opaw_ctx* c) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  c->prof.harvest();
  for (int k = 0; k < OPAW_K_COUNT; ++k) {
    c->prof.total[k] = 0;
    c->prof.count[k] = 0;
  }
  return OPAW_OK;
}

int opaw_profile_read(opaw_ctx* c, int kclass, double* total_ms, long long* launches) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  if (kclass < 0 || kclass >= OPAW_K_COUNT) return fail(OPAW_ERR_ARG, "bad kernel class");
  c->prof.harvest();
  if (total_ms) *total_ms = c->prof.total[kclass];
  if (launches) *launches = c->prof.count[kclass];
  return OPAW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// services of the context to the file formats (opaw_io.hip): orbitals of one
// system as complex*16 on the host, whatever the storage precision
namespace opaw {

int io_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int ctx_shape(opaw_ctx* c, int* nsys, int* nb, long long* n) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  *nsys = c->nsys;
  *nb = c->nb;
  *n = c->N;
  return OPAW_OK;
}

int ctx_orbitals_to_host(opaw_ctx* c, int sys, std::vector<std::complex<double>>& out, int* nb, long long* n) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  RCHK(check_sys(c, sys));
  const size_t cnt = (size_t)c->nb * c->N;
  out.resize(cnt);
  if (c->prec == OPAW_PREC_F64) {
    RCHK(opaw_read_orbitals(c, sys, out.data()));
  } else {
    std::vector<std::complex<float>> f(cnt);
    RCHK(opaw_read_orbitals(c, sys, f.data()));
    for (size_t k = 0; k < cnt; ++k) out[k] = std::complex<double>(f[k].real(), f[k].imag());
  }
  *nb = c->nb;
  *n = c->N;
  return OPAW_OK;
}

int ctx_orbitals_from_host(opaw_ctx* c, int sys, const std::vector<std::complex<double>>& in) {
  if (!c) return fail(OPAW_ERR_ARG, "null context");
  RCHK(check_sys(c, sys));
  const size_t cnt = (size_t)c->nb * c->N;
  if (in.size() < cnt) return fail(OPAW_ERR_ARG, "orbital buffer too small");
  if (c->prec == OPAW_PREC_F64) return opaw_write_orbitals(c, sys, in.data());
  std::vector<std::complex<float>> f(cnt);
  for (size_t k = 0; k < cnt; ++k) f[k] = std::complex<float>((float)in[k].real(), (float)in[k].imag());
  return opaw_write_orbitals(c, sys, f.data());
}

}  // namespace opaw

// ---------------------------------------------------------------------------
// ground state (FP64 contexts): the SCF pieces of the reference's main loop
// (src/main.f90:66-95) and a block preconditioned solver whose k-space
// preconditioner carries a learned correction
static int need_f64(opaw_ctx* c) {
  if (c->prec != OPAW_PREC_F64) return fail(OPAW_ERR_ARG, "the ground-state solver runs in FP64 contexts");
  return OPAW_OK;
}

static double* gs_part(opaw_ctx* c, size_t count, int* rc) {
  *rc = ensure(c, c->gsPart, count * sizeof(double));
  return (double*)c->gsPart.p;
}

// dense solve B x = rhs (partial pivoting), B row-major n x n
static bool small_solve(int n, std::vector<double> B, std::vector<double>& x) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(B[(size_t)i * n + k]) > std::fabs(B[(size_t)p * n + k])) p = i;
    if (B[(size_t)p * n + k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(B[(size_t)k * n + j], B[(size_t)p * n + j]);
      std::swap(x[k], x[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = B[(size_t)i * n + k] / B[(size_t)k * n + k];
      for (int j = k; j < n; ++j) B[(size_t)i * n + j] -= f * B[(size_t)k * n + j];
      x[i] -= f * x[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = x[k];
    for (int j = k + 1; j < n; ++j) s -= B[(size_t)k * n + j] * x[j];
    x[k] = s / B[(size_t)k * n + k];
  }
  return true;
}

// diis_general (src/diis.f90:12-153) on V_KS of system sys
static int diis_step(opaw_ctx* c, int sys) {
  auto& D = c->diis[sys];
  const int L = c->diis_L;
  const long long N = c->N;
  double* vec = c->vks + (size_t)sys * N;
  const size_t fb = (size_t)N * sizeof(double);
  if (!D.bank) {
    RCHK(dalloc_t(c, &D.bank, (size_t)L * N));
    RCHK(dalloc_t(c, &D.err, (size_t)L * N));
  }
  if (D.stage == 0) {
    D.stage = 1;
    D.K = 0;
    D.Kb = 1;
    HCHK(hipMemcpyAsync(D.bank, vec, fb, hipMemcpyDeviceToDevice, c->st));
    return OPAW_OK;
  }
  // stackerr
  if (D.K == L)
    for (int k = 0; k < L - 1; ++k)
      HCHK(hipMemcpyAsync(D.err + (size_t)k * N, D.err + (size_t)(k + 1) * N, fb, hipMemcpyDeviceToDevice, c->st));
  D.K = std::min(D.K + 1, L);
  const int K = D.K;
  HCHK(sub_real(c->st, vec, D.bank + (size_t)(D.Kb - 1) * N, N, D.err + (size_t)(K - 1) * N));
  // minimize: [B 1; 1 0] C = e_M
  int rc;
  double* part = gs_part(c, (size_t)kGsChunks * K * K, &rc);
  RCHK(rc);
  HCHK(real_dots(c->st, D.err, N, K, part));
  std::vector<double> hp((size_t)kGsChunks * K * K);
  HCHK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  const int M = K + 1;
  std::vector<double> B((size_t)M * M, 0.0), C(M, 0.0);
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) {
      double s = 0.0;
      for (int ch = 0; ch < kGsChunks; ++ch) s += hp[((size_t)ch * K + i) * K + j];
      B[(size_t)i * M + j] = s * dvol(c);
    }
  for (int i = 0; i < K; ++i) B[(size_t)i * M + K] = B[(size_t)K * M + i] = 1.0;
  C[K] = 1.0;
  if (!small_solve(M, B, C)) return fail(OPAW_ERR_ARG, "DIIS matrix is singular");
  double sc = 0.0;
  for (int i = 0; i < K; ++i) sc += C[i];
  for (int i = 0; i < K; ++i) C[i] /= sc;
  RCHK(ensure(c, c->gsSmall, (size_t)M * sizeof(double)));
  HCHK(hipMemcpyAsync(c->gsSmall.p, C.data(), K * sizeof(double), hipMemcpyHostToDevice, c->st));
  // predict_output, then bankstack
  HCHK(diis_predict(c->st, D.bank, D.err, (const double*)c->gsSmall.p, K, c->diis_mix, N, vec));
  if (D.Kb == L)
    for (int k = 0; k < L - 1; ++k)
      HCHK(hipMemcpyAsync(D.bank + (size_t)k * N, D.bank + (size_t)(k + 1) * N, fb, hipMemcpyDeviceToDevice, c->st));
  if (D.Kb != K) return fail(OPAW_ERR_STATE, "DIIS bank and error stacks out of step (Kb %d, K %d)", D.Kb, K);
  D.Kb = std::min(D.Kb + 1, L);
  HCHK(hipMemcpyAsync(D.bank + (size_t)(D.Kb - 1) * N, vec, fb, hipMemcpyDeviceToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

// 1 x 1 Gram of two single fields: dv <a|b>
static int dot1(opaw_ctx* c, const c64* a, const c64* b, std::complex<double>* out) {
  HCHK(gram<double>(c->st, a, b, c->N, 1, dvol(c), c->gpart, c->gmat));
  HCHK(hipMemcpyAsync(out, c->gmat, sizeof(c64), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

// S^-1/2 H S^-1/2 (h_type 1) or H (h_type 0) of nf fields of system sys
static int apply_h(opaw_ctx* c, int sys, int h_type, const c64* in, c64* out, int nf) {
  if (h_type == 1) return flow_shs<double>(c, sys, in, out, nf, nf);
  return flow_h_phi<double>(c, sys, in, out, nf, nf);
}

// generalised eigenproblem H C = S C w on an m-dimensional basis (row-major
// host matrices): Cholesky when the unit-scaled overlap is safely positive
// definite, else directions with overlap below 1e-12 of the largest are
// dropped; returns the lowest nout eigenpairs, C column-major m x nout
static int subspace_eig(int m, const std::vector<std::complex<double>>& S0, const std::vector<std::complex<double>>& H0,
                        int nout, std::vector<std::complex<double>>& Ccol, std::vector<double>& w) {
  const int r = pencil_lowest(m, S0, H0, nout, 1e-12, Ccol, w);
  if (r < nout) return fail(OPAW_ERR_ARG, "subspace rank %d below the %d bands", r, nout);
  return OPAW_OK;
}

// gram of m fields (row-major m x m on the host), with the solver's buffers
static int gram_m(opaw_ctx* c, const c64* a, const c64* b, int m, std::vector<std::complex<double>>& out) {
  RCHK(ensure(c, c->gsG, gram_partials(m) * sizeof(c64)));
  RCHK(ensure(c, c->gsM, (size_t)m * m * sizeof(c64)));
  HCHK(gram<double>(c->st, a, b, c->N, m, dvol(c), (c64*)c->gsG.p, (c64*)c->gsM.p));
  out.resize((size_t)m * m);
  HCHK(hipMemcpyAsync(out.data(), c->gsM.p, (size_t)m * m * sizeof(c64), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

// per-band kinetic energy sum_k ek |f(k)|^2 / sum_k |f(k)|^2 of nf k-space fields
static int band_kinetic(opaw_ctx* c, const c64* fk, int nf, std::vector<double>& t) {
  int rc;
  double* part = gs_part(c, (size_t)kGsChunks * nf * 2, &rc);
  RCHK(rc);
  HCHK(kinetic_partials(c->st, fk, c->ek, c->N, nf, part));
  std::vector<double> h((size_t)kGsChunks * nf * 2);
  HCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  t.assign(nf, 0.0);
  for (int i = 0; i < nf; ++i) {
    double a = 0.0, b = 0.0;
    for (int ch = 0; ch < kGsChunks; ++ch) {
      a += h[((size_t)ch * nf + i) * 2];
      b += h[((size_t)ch * nf + i) * 2 + 1];
    }
    t[i] = (b > 0.0 ? a / b : 0.0) * (double)c->N;   // ek carries 1/N
    if (!(t[i] > 1e-6)) t[i] = 1e-6;
  }
  return OPAW_OK;
}

static int upload_small(opaw_ctx* c, size_t offset_doubles, const std::vector<double>& v) {
  RCHK(ensure(c, c->gsSmall, (offset_doubles + v.size()) * sizeof(double)));
  HCHK(hipMemcpyAsync((double*)c->gsSmall.p + offset_doubles, v.data(), v.size() * sizeof(double),
                      hipMemcpyHostToDevice, c->st));
  return OPAW_OK;
}

extern "C" {

int opaw_set_nocc(opaw_ctx* c, int nocc) {
  CTX_CHECK(c);
  if (nocc < 1 || nocc > c->nb) return fail(OPAW_ERR_ARG, "nocc must be in 1..nb");
  c->nocc = nocc;
  return OPAW_OK;
}

int opaw_set_diis(opaw_ctx* c, int ndi, double mix) {
  CTX_CHECK(c);
  if (ndi < 0 || ndi > 64) return fail(OPAW_ERR_ARG, "ndi must be in 0..64");
  for (auto& D : c->diis) {
    dfree(c, D.bank);
    dfree(c, D.err);
    D = opaw_ctx::Diis();
  }
  c->diis_L = ndi;
  c->diis_mix = mix;
  return OPAW_OK;
}

int opaw_set_subspace_device(opaw_ctx* c, int min_m) {
  CTX_CHECK(c);
  if (min_m < 0) return fail(OPAW_ERR_ARG, "min_m must be >= 0");
  c->rr_dev_min = min_m;
  return OPAW_OK;
}

int opaw_ncpaw_make_hamiltonian(opaw_ctx* c, int sys, int use_diis) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  RCHK(need_tables(c));
  const long long N = c->N;
  const int no = c->nocc;
  RCHK(ensure(c, c->wu, (size_t)no * N * sizeof(c64)));
  const c64* phi = (const c64*)c->phi + (size_t)sys * c->nb * N;
  RCHK(flow_sn_phi<double>(c, phi, (c64*)c->wu.p, no, -0.5));
  RCHK(opaw_tddft_rhoij_2(c, sys, c->wu.p));
  RCHK(opaw_tddft_dens_2(c, sys, c->wu.p));
  RCHK(opaw_get_nhat(c, sys));
  RCHK(opaw_get_pot_tddft_2(c, sys));
  if (use_diis && c->diis_L > 0) RCHK(diis_step(c, sys));
  if (c->vksf) HCHK(to_f32(c->st, c->vks + (size_t)sys * N, c->vksf + (size_t)sys * N, N));
  RCHK(opaw_get_dij(c, sys));
  c->prep_valid = false;
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

int opaw_ncpaw_make_hamiltonian_density(opaw_ctx* c, int sys, const double* dens, int use_diis) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  RCHK(need_tables(c));
  if (!dens) return fail(OPAW_ERR_ARG, "null density");
  const long long N = c->N;
  // iscf = 1: the starting density, rho_ij = 0 (so nhat = 0 and the on-site
  // terms see the core densities only)
  HCHK(hipMemcpyAsync(c->dens + sys * N, dens, N * sizeof(double), hipMemcpyHostToDevice, c->st));
  HCHK(fill(c->st, c->rhoij + (size_t)sys * c->mtot, c->mtot, 0.0));
  HCHK(fill(c->st, c->nhat + (size_t)sys * N, N, 0.0));
  if (sys < (int)c->nhat_lazy.size()) c->nhat_lazy[sys] = 0;
  RCHK(opaw_get_pot_tddft_2(c, sys));
  if (use_diis && c->diis_L > 0) RCHK(diis_step(c, sys));
  if (c->vksf) HCHK(to_f32(c->st, c->vks + (size_t)sys * N, c->vksf + (size_t)sys * N, N));
  RCHK(opaw_get_dij(c, sys));
  c->prep_valid = false;
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

int opaw_hminmax(opaw_ctx* c, int sys, int h_type, int niter, double fctr, unsigned long long seed, double* hmin,
                 double* hmax) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  RCHK(need_tables(c));
  if (niter < 1 || !hmin || !hmax) return fail(OPAW_ERR_ARG, "bad hminmax arguments");
  const long long N = c->N;
  RCHK(ensure(c, c->gsT, 2 * (size_t)N * sizeof(c64)));
  c64* z = (c64*)c->gsT.p;
  c64* hz = z + N;
  std::vector<std::complex<double>> host((size_t)N);
  double e[2] = {0.0, 0.0};
  for (int pass = 0; pass < 2; ++pass) {
    unsigned long long x = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(pass + 1);
    for (long long r = 0; r < N; ++r) {   // splitmix64, uniform in [0, 1)
      unsigned long long v = (x += 0x9E3779B97F4A7C15ull);
      v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
      v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
      v ^= v >> 31;
      host[(size_t)r] = std::complex<double>((double)(v >> 11) * 0x1.0p-53, 0.0);
    }
    HCHK(hipMemcpy(z, host.data(), (size_t)N * sizeof(c64), hipMemcpyHostToDevice));
    for (int it = 0; it < niter; ++it) {
      std::complex<double> nz;
      RCHK(dot1(c, z, z, &nz));
      HCHK(scale_complex(c->st, z, N, 1.0 / std::sqrt(nz.real())));
      RCHK(apply_h(c, sys, h_type, z, hz, 1));
      if (pass == 1) HCHK(axpy_complex(c->st, hz, z, e[0], N, hz));
      std::complex<double> q, n2;
      RCHK(dot1(c, z, hz, &q));
      RCHK(dot1(c, z, z, &n2));
      e[pass] = q.real() / n2.real();
      HCHK(hipMemcpyAsync(z, hz, (size_t)N * sizeof(c64), hipMemcpyDeviceToDevice, c->st));
    }
  }
  // get_hminmax (src/hminmax.f90:89-98)
  const double e2 = e[1] + e[0];
  double hx = std::max(e[0], e2), hn = std::min(e[0], e2);
  const double havg = 0.5 * (hn + hx), dh = (hx - hn) * fctr;
  *hmax = havg + dh / 2.0;
  *hmin = havg - dh / 2.0;
  return OPAW_OK;
}

int opaw_chebyshev_filter(opaw_ctx* c, int sys, int h_type, int degree, double hbot, double hmax) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  RCHK(need_tables(c));
  if (degree < 1 || !(hmax > hbot)) return fail(OPAW_ERR_ARG, "bad Chebyshev filter arguments");
  const long long N = c->N;
  const int nb = c->nb;
  const size_t fb = (size_t)nb * N * sizeof(c64);
  RCHK(ensure(c, c->gsY, 3 * fb));
  c64* phi = (c64*)c->phi + (size_t)sys * nb * N;
  c64* A = (c64*)c->gsY.p;
  c64* B = A + (size_t)nb * N;
  c64* Cc = B + (size_t)nb * N;
  HCHK(hipMemcpyAsync(A, phi, fb, hipMemcpyDeviceToDevice, c->st));
  HCHK(hipMemcpyAsync(B, phi, fb, hipMemcpyDeviceToDevice, c->st));
  const double havg = 0.5 * (hmax + hbot), dh = 0.5 * (hmax - hbot);
  for (int k = 1; k <= degree; ++k) {
    RCHK(apply_h(c, sys, h_type, B, Cc, nb));
    HCHK(cheb_combine<double>(c->st, Cc, B, A, (long long)nb * N, havg, dh, k == 1));
    c64* t = A;
    A = B;
    B = Cc;
    Cc = t;
  }
  HCHK(hipMemcpyAsync(phi, B, fb, hipMemcpyDeviceToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  c->prep_valid = false;
  return OPAW_OK;
}

int opaw_residuals(opaw_ctx* c, int sys, int h_type, const double* eig, double* res) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  RCHK(need_tables(c));
  if (!eig || !res) return fail(OPAW_ERR_ARG, "null argument");
  const long long N = c->N;
  const int nb = c->nb;
  RCHK(ensure(c, c->gsHY, (size_t)nb * N * sizeof(c64)));
  const c64* phi = (const c64*)c->phi + (size_t)sys * nb * N;
  c64* hp = (c64*)c->gsHY.p;
  RCHK(apply_h(c, sys, h_type, phi, hp, nb));
  RCHK(upload_small(c, 0, std::vector<double>(eig, eig + nb)));
  int rc;
  double* part = gs_part(c, (size_t)kGsChunks * nb, &rc);
  RCHK(rc);
  HCHK(residual_partials<double>(c->st, hp, phi, (const double*)c->gsSmall.p, N, nb, part));
  std::vector<double> h((size_t)kGsChunks * nb);
  HCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  for (int i = 0; i < nb; ++i) {
    double s = 0.0;
    for (int ch = 0; ch < kGsChunks; ++ch) s += h[(size_t)ch * nb + i];
    res[i] = std::sqrt(s * dvol(c));
  }
  return OPAW_OK;
}

// Block preconditioned eigensolver (LOBPCG) on the resident orbitals of
// system sys (h_type 1: shs; the orbitals must be orthonormal).  Each
// iteration: Rayleigh quotients e_i, residuals r_i = H phi_i - e_i phi_i,
// w_i = M_i r_i in k space (Teter-Payne-Allan ratio of x = |k|^2/2 / t_i, t_i
// the band's kinetic energy, times exp(g(log(1+x), e_i/t_i)) when an MLP is
// given), then Rayleigh-Ritz on [phi, w, p] with p the previous step's
// (w, p) part of the new vectors.  Stops when the residual norms of the
// first nconv bands are <= tol; hist[it] = their largest at iteration it.
static int ensure_shells(opaw_ctx* c);

// pinned host memory (hipHostMalloc) for small transfers, grown on demand
// and freed with the context
static int ensure_pinned(opaw_ctx* c, size_t bytes) {
  if (c->pin_bytes >= bytes) return OPAW_OK;
  if (c->pin) {
    HCHK(hipStreamSynchronize(c->st));
    HCHK(hipHostFree(c->pin));
    c->pin = nullptr;
    c->pin_bytes = 0;
  }
  HCHK(hipHostMalloc(&c->pin, bytes, hipHostMallocDefault));
  c->pin_bytes = bytes;
  return OPAW_OK;
}

int opaw_block_solve(opaw_ctx* c, int sys, int h_type, int maxit, double tol, int nconv, const double* mlp, int nh,
                     int record, double* eig, double* hist, int* iters) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  RCHK(need_tables(c));
  if (maxit < 1 || !eig) return fail(OPAW_ERR_ARG, "bad block_solve arguments");
  MlpShape shape;
  shape.nh = nh;
  if (mlp && (nh < 1 || nh > 16)) return fail(OPAW_ERR_ARG, "MLP hidden width must be 1..16");
  const long long N = c->N;
  const int nb = c->nb;
  if (nconv <= 0 || nconv > nb) nconv = nb;
  const size_t fb = (size_t)nb * N * sizeof(c64);
  // two [phi | w | p] blocks and their H images: the Rayleigh-Ritz rotation
  // writes the next iteration's phi and p into the other block
  RCHK(ensure(c, c->gsY, 6 * fb));
  RCHK(ensure(c, c->gsHY, 6 * fb));
  RCHK(ensure(c, c->gsK, fb));
  const size_t off = (size_t)nb * N;
  c64* phi = (c64*)c->phi + (size_t)sys * nb * N;
  c64* Y = (c64*)c->gsY.p;        // [phi | w | p]
  c64* HY = (c64*)c->gsHY.p;      // H of the same
  c64* Yn = Y + 3 * off;
  c64* HYn = HY + 3 * off;
  c64* Kb = (c64*)c->gsK.p;
  const c64 *twx = c->tw64[c->nx], *twy = c->tw64[c->ny], *twz = c->tw64[c->nz];
  const int npar = mlp_param_count(shape);
  double* dmlp = nullptr;
  RCHK(ensure(c, c->gsSmall, (size_t)(3 * nb + (mlp ? npar : 0)) * sizeof(double)));
  if (mlp) {
    dmlp = (double*)c->gsSmall.p + 3 * nb;
    HCHK(hipMemcpy(dmlp, mlp, npar * sizeof(double), hipMemcpyHostToDevice));
  }
  double* deig = (double*)c->gsSmall.p;
  double* dt = deig + nb;
  double* dy = dt + nb;
  RCHK(ensure_shells(c));
  RCHK(ensure(c, c->gsTab, (size_t)nb * c->nsh * sizeof(double)));
  // pinned staging, one area per transfer: the partial sums (down), e (down
  // or up) then t and y (up), the device pencil's flags (down), the host
  // pencil's Gram matrices S and H (down) and coefficients C and Cp (up).
  // Each area is rewritten only after a stream synchronisation that follows
  // its previous transfer.
  const int mmax = 3 * nb;
  const bool any_host = 2 * nb < c->rr_dev_min;
  const bool any_dev = mmax >= c->rr_dev_min;
  const size_t pin_h = (size_t)kGsChunks * nb * 3 * sizeof(double);
  const size_t pin_ety = (3 * (size_t)nb * sizeof(double) + 15) / 16 * 16;   // keeps what follows 16-byte aligned
  const size_t pin_fl = 16;
  const size_t pin_sh = any_host ? 2 * (size_t)mmax * mmax * sizeof(c64) : 0;
  const size_t pin_c = any_host ? (size_t)(mmax + (mmax - nb)) * nb * sizeof(c64) : 0;
  RCHK(ensure_pinned(c, pin_h + pin_ety + pin_fl + pin_sh + pin_c));
  double* ph = (double*)c->pin;
  double* pety = (double*)((char*)c->pin + pin_h);
  int* pflags = (int*)((char*)pety + pin_ety);
  c64* psh = (c64*)((char*)pflags + pin_fl);
  c64* pc = psh + (any_host ? 2 * (size_t)mmax * mmax : 0);
  // the device pencil's work and the coefficients C (m x nb) and Cp (its
  // rows nb..m, (m - nb) x nb), for the largest m = 3 nb
  if (any_dev) RCHK(ensure(c, c->gsRR, rr_work_bytes(mmax)));
  RCHK(ensure(c, c->gsSmallC, (size_t)(mmax + (mmax - nb)) * nb * sizeof(c64)));
  HCHK(hipMemcpyAsync(Y, phi, fb, hipMemcpyDeviceToDevice, c->st));
  RCHK(apply_h(c, sys, h_type, Y, HY, nb));
  std::vector<double> e(nb), t(nb), y(nb);
  {
    // Rayleigh quotients of the start; later iterations carry the Ritz values
    std::vector<std::complex<double>> G;
    RCHK(gram_m(c, Y, HY, nb, G));
    for (int i = 0; i < nb; ++i) e[i] = G[(size_t)i * nb + i].real();
    std::copy(e.begin(), e.end(), pety);
    HCHK(hipMemcpyAsync(deig, pety, nb * sizeof(double), hipMemcpyHostToDevice, c->st));
    HCHK(hipStreamSynchronize(c->st));
  }
  bool have_p = false;
  int m = 0;   // dimension of the last device Rayleigh-Ritz, 0 once its flags are checked
  int rc;
  // one partial-sum buffer for the residual norms and the kinetic sums
  double* part = gs_part(c, (size_t)kGsChunks * nb * 3, &rc);
  RCHK(rc);
  const double* h = ph;
  // the Ritz values and the flags of the last Rayleigh-Ritz, read after a
  // synchronisation
  auto check_rr = [&]() -> int {
    if (m == 0) return OPAW_OK;
    std::copy(pety, pety + nb, e.begin());
    const int rank = pflags[0], info_s = pflags[1], info_h = pflags[2];
    if (info_s != 0 || info_h != 0)
      return fail(OPAW_ERR_HIP, "subspace eigensolver did not converge (info %d, %d)", info_s, info_h);
    if (rank < nb) return fail(OPAW_ERR_ARG, "subspace rank %d below the %d bands", rank, nb);
    m = 0;
    return OPAW_OK;
  };
  auto fetch_rr = [&]() -> int {
    if (m == 0) return OPAW_OK;
    HCHK(hipMemcpyAsync(pety, deig, nb * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HCHK(hipMemcpyAsync(pflags, rr_flags(c->gsRR.p, m), 3 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    return OPAW_OK;
  };
  int it = 0;
  for (; it < maxit; ++it) {
    c64* W = Y + off;
    HCHK(residual_vectors(c->st, HY, Y, deig, N, nb, W));
    HCHK(residual_partials<double>(c->st, HY, Y, deig, N, nb, part));
    // band kinetic energies from the same download
    HCHK(fft3d<double>(c->g, c->st, -1, Y, Kb, nb, twx, twy, twz));
    HCHK(kinetic_partials(c->st, Kb, c->ek, N, nb, part + (size_t)kGsChunks * nb));
    HCHK(hipMemcpyAsync(ph, part, pin_h, hipMemcpyDeviceToHost, c->st));
    RCHK(fetch_rr());
    HCHK(hipStreamSynchronize(c->st));
    RCHK(check_rr());
    double worst = 0.0;
    for (int i = 0; i < nconv; ++i) {
      double sm = 0.0;
      for (int ch = 0; ch < kGsChunks; ++ch) sm += h[(size_t)ch * nb + i];
      worst = std::max(worst, std::sqrt(sm * dvol(c)));
    }
    if (hist) hist[it] = worst;
    if (worst <= tol) break;
    const double* kp = h + (size_t)kGsChunks * nb;
    for (int i = 0; i < nb; ++i) {
      double a = 0.0, b = 0.0;
      for (int ch = 0; ch < kGsChunks; ++ch) {
        a += kp[((size_t)ch * nb + i) * 2];
        b += kp[((size_t)ch * nb + i) * 2 + 1];
      }
      t[i] = (b > 0.0 ? a / b : 0.0) * (double)N;   // ek carries 1/N
      if (!(t[i] > 1e-6)) t[i] = 1e-6;
      y[i] = e[i] / t[i];
    }
    std::copy(t.begin(), t.end(), pety + nb);
    std::copy(y.begin(), y.end(), pety + 2 * nb);
    HCHK(hipMemcpyAsync(dt, pety + nb, 2 * nb * sizeof(double), hipMemcpyHostToDevice, c->st));   // dy = dt + nb
    if (record > 0 && it % record == 0) {
      opaw_ctx::Snap sp;
      sp.sys = sys;
      sp.nb = nb;
      sp.t = t;
      sp.y = y;
      RCHK(ensure(c, sp.phi, fb));
      RCHK(ensure(c, sp.r, fb));
      HCHK(hipMemcpyAsync(sp.phi.p, Y, fb, hipMemcpyDeviceToDevice, c->st));
      HCHK(hipMemcpyAsync(sp.r.p, W, fb, hipMemcpyDeviceToDevice, c->st));
      c->snaps.push_back(sp);
    }
    HCHK(fft3d<double>(c->g, c->st, -1, W, Kb, nb, twx, twy, twz));
    HCHK(precond_apply_shells(c->st, Kb, N, nb, (const int*)c->shOf.p, (const double*)c->shEk.p, c->nsh, dt, dy, dmlp,
                              shape, (double*)c->gsTab.p));
    HCHK(fft3d<double>(c->g, c->st, 1, Kb, W, nb, twx, twy, twz));
    RCHK(apply_h(c, sys, h_type, W, HY + off, nb));
    // Rayleigh-Ritz on [phi, w, p] (LOBPCG; no p in the first iteration):
    // the Gram matrices, the pencil's lowest nb eigenpairs (Ritz values into
    // deig for the next residuals) and the rotations
    const int mi = have_p ? 3 * nb : 2 * nb;
    RCHK(ensure(c, c->gsG, gram_partials(mi) * sizeof(c64)));
    RCHK(ensure(c, c->gsM, (size_t)2 * mi * mi * sizeof(c64)));
    c64* dS = (c64*)c->gsM.p;
    c64* dH = dS + (size_t)mi * mi;
    HCHK(gram<double>(c->st, Y, Y, N, mi, dvol(c), (c64*)c->gsG.p, dS));
    HCHK(gram<double>(c->st, Y, HY, N, mi, dvol(c), (c64*)c->gsG.p, dH));
    c64* dC = (c64*)c->gsSmallC.p;
    c64* dCp = dC + (size_t)mi * nb;
    if (mi >= c->rr_dev_min) {
      // on the device: its rank and info are read at the next synchronisation
      const int rr = rr_lowest(&c->blas, c->st, mi, nb, 1e-12, dS, dH, c->gsRR.p, dC, deig);
      if (rr != 0)
        return fail(OPAW_ERR_HIP, "device Rayleigh-Ritz (m %d): %s %d", mi, rr > 0 ? "rocBLAS status" : "HIP error",
                    rr > 0 ? rr : -rr);
      HCHK(hipMemcpy2DAsync(dCp, (size_t)(mi - nb) * sizeof(c64), dC + nb, (size_t)mi * sizeof(c64),
                            (size_t)(mi - nb) * sizeof(c64), nb, hipMemcpyDeviceToDevice, c->st));
      m = mi;
    } else {
      // on the host: both Gram matrices in one download, the coefficients
      // and the Ritz values in two uploads
      std::vector<std::complex<double>> S((size_t)mi * mi), H((size_t)mi * mi), Ccol;
      HCHK(hipMemcpyAsync(psh, dS, 2 * (size_t)mi * mi * sizeof(c64), hipMemcpyDeviceToHost, c->st));
      HCHK(hipStreamSynchronize(c->st));
      std::memcpy((void*)S.data(), psh, (size_t)mi * mi * sizeof(c64));
      std::memcpy((void*)H.data(), psh + (size_t)mi * mi, (size_t)mi * mi * sizeof(c64));
      for (int a = 0; a < mi; ++a)
        for (int b = a + 1; b < mi; ++b) {
          const std::complex<double> hs = 0.5 * (S[(size_t)a * mi + b] + std::conj(S[(size_t)b * mi + a]));
          S[(size_t)a * mi + b] = hs;
          S[(size_t)b * mi + a] = std::conj(hs);
        }
      std::vector<double> w;
      RCHK(subspace_eig(mi, S, H, nb, Ccol, w));
      std::vector<std::complex<double>> Cp((size_t)(mi - nb) * nb);   // rows nb..m of C
      for (int j = 0; j < nb; ++j)
        for (int i = nb; i < mi; ++i) Cp[(i - nb) + (size_t)(mi - nb) * j] = Ccol[i + (size_t)mi * j];
      std::memcpy((void*)pc, Ccol.data(), Ccol.size() * sizeof(c64));
      std::memcpy((void*)(pc + Ccol.size()), Cp.data(), Cp.size() * sizeof(c64));
      HCHK(hipMemcpyAsync(dC, pc, ((size_t)mi + (mi - nb)) * nb * sizeof(c64), hipMemcpyHostToDevice, c->st));   // dCp follows dC
      e = w;
      std::copy(e.begin(), e.end(), pety);
      HCHK(hipMemcpyAsync(deig, pety, nb * sizeof(double), hipMemcpyHostToDevice, c->st));
    }
    HCHK(rotate_rect(c->st, Y, Yn, N, mi, nb, dC));
    HCHK(rotate_rect(c->st, HY, HYn, N, mi, nb, dC));
    HCHK(rotate_rect(c->st, Y + off, Yn + 2 * off, N, mi - nb, nb, dCp));
    HCHK(rotate_rect(c->st, HY + off, HYn + 2 * off, N, mi - nb, nb, dCp));
    std::swap(Y, Yn);
    std::swap(HY, HYn);
    have_p = true;
  }
  // a last Rayleigh-Ritz not followed by a residual pass
  RCHK(fetch_rr());
  HCHK(hipStreamSynchronize(c->st));
  RCHK(check_rr());
  if (iters) *iters = it;
  HCHK(hipMemcpyAsync(phi, Y, fb, hipMemcpyDeviceToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  std::copy(e.begin(), e.end(), eig);
  c->prep_valid = false;
  return OPAW_OK;
}

// k shells: the k points sorted by ek, grouped by equal value
static int ensure_shells(opaw_ctx* c) {
  if (c->nsh > 0) return OPAW_OK;
  const long long N = c->N;
  std::vector<double> ek((size_t)N);
  HCHK(hipMemcpy(ek.data(), c->ek, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int> perm((size_t)N);
  for (long long k = 0; k < N; ++k) perm[(size_t)k] = (int)k;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return ek[a] < ek[b]; });
  std::vector<int> off;
  std::vector<double> eks;
  for (long long q = 0; q < N; ++q)
    if (q == 0 || ek[perm[(size_t)q]] != ek[perm[(size_t)q - 1]]) {
      off.push_back((int)q);
      eks.push_back(ek[perm[(size_t)q]]);
    }
  off.push_back((int)N);
  std::vector<int> of((size_t)N);
  for (size_t s = 0; s + 1 < off.size(); ++s)
    for (int q = off[s]; q < off[s + 1]; ++q) of[(size_t)perm[(size_t)q]] = (int)s;
  RCHK(ensure(c, c->shOf, (size_t)N * sizeof(int)));
  HCHK(hipMemcpy(c->shOf.p, of.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice));
  RCHK(ensure(c, c->shPerm, (size_t)N * sizeof(int)));
  RCHK(ensure(c, c->shOff, off.size() * sizeof(int)));
  RCHK(ensure(c, c->shEk, eks.size() * sizeof(double)));
  HCHK(hipMemcpy(c->shPerm.p, perm.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice));
  HCHK(hipMemcpy(c->shOff.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  HCHK(hipMemcpy(c->shEk.p, eks.data(), eks.size() * sizeof(double), hipMemcpyHostToDevice));
  c->nsh = (int)eks.size();
  return OPAW_OK;
}

// Turns the recorded (phi_t, r_t) of system sys into training pairs with the
// resident (converged, orthonormal) orbitals phi*: e_t = phi* (phi*^H phi_t dv)
// - phi_t, the part of phi_t the step should remove; r_t and e_t go to k space.
int opaw_precond_prepare(opaw_ctx* c, int sys) {
  CTX_CHECK(c);
  RCHK(check_sys(c, sys));
  RCHK(need_f64(c));
  const long long N = c->N;
  const int nb = c->nb;
  const size_t fb = (size_t)nb * N * sizeof(c64);
  const c64* ref = (const c64*)c->phi + (size_t)sys * nb * N;
  RCHK(ensure(c, c->gsT, 2 * fb));
  c64* T0 = (c64*)c->gsT.p;
  const c64 *twx = c->tw64[c->nx], *twy = c->tw64[c->ny], *twz = c->tw64[c->nz];
  for (auto& sp : c->snaps) {
    if (sp.ready || sp.sys != sys) continue;
    std::vector<std::complex<double>> G;   // G[a][j] = dv <ref_a | phi_j>
    RCHK(ensure(c, c->gsG, gram_partials(nb) * sizeof(c64)));
    RCHK(ensure(c, c->gsM, (size_t)nb * nb * sizeof(c64)));
    HCHK(gram<double>(c->st, ref, (const c64*)sp.phi.p, N, nb, dvol(c), (c64*)c->gsG.p, (c64*)c->gsM.p));
    // rotate_rect wants T[a + nb j] = G[a][j]: the row-major gram is that transposed
    G.resize((size_t)nb * nb);
    HCHK(hipMemcpyAsync(G.data(), c->gsM.p, G.size() * sizeof(c64), hipMemcpyDeviceToHost, c->st));
    HCHK(hipStreamSynchronize(c->st));
    std::vector<std::complex<double>> Tc((size_t)nb * nb);
    for (int a = 0; a < nb; ++a)
      for (int j = 0; j < nb; ++j) Tc[a + (size_t)nb * j] = G[(size_t)a * nb + j];
    HCHK(hipMemcpyAsync(c->gsM.p, Tc.data(), Tc.size() * sizeof(c64), hipMemcpyHostToDevice, c->st));
    HCHK(rotate_rect(c->st, ref, T0, N, nb, nb, (const c64*)c->gsM.p));
    HCHK(axpy_complex(c->st, T0, (const c64*)sp.phi.p, 1.0, (long long)nb * N, T0));
    // k space: phi slot <- e_t(k), r slot <- r_t(k)
    HCHK(fft3d<double>(c->g, c->st, -1, T0, (c64*)sp.phi.p, nb, twx, twy, twz));
    HCHK(fft3d<double>(c->g, c->st, -1, (const c64*)sp.r.p, T0, nb, twx, twy, twz));
    HCHK(hipMemcpyAsync(sp.r.p, T0, fb, hipMemcpyDeviceToDevice, c->st));
    RCHK(ensure(c, sp.ty, 2 * (size_t)nb * sizeof(double)));
    HCHK(hipMemcpyAsync(sp.ty.p, sp.t.data(), nb * sizeof(double), hipMemcpyHostToDevice, c->st));
    HCHK(hipMemcpyAsync((double*)sp.ty.p + nb, sp.y.data(), nb * sizeof(double), hipMemcpyHostToDevice, c->st));
    // shell sums, and |e_i|^2 from the full-sample sums (independent of M)
    RCHK(ensure_shells(c));
    const int nsh = c->nsh;
    RCHK(ensure(c, sp.agg, (size_t)nb * nsh * (sizeof(c64) + sizeof(double))));
    c64* z = (c64*)sp.agg.p;
    double* w2 = (double*)(z + (size_t)nb * nsh);
    HCHK(shell_aggregate(c->st, (const c64*)sp.r.p, (const c64*)sp.phi.p, N, nb, (const int*)c->shPerm.p,
                         (const int*)c->shOff.p, nsh, z, w2));
    {
      int rc;
      double* part = gs_part(c, (size_t)kGsChunks * nb * 4, &rc);
      RCHK(rc);
      const double* dt = (const double*)sp.ty.p;
      MlpShape shape;
      HCHK(precond_loss_sums(c->st, (const c64*)sp.r.p, (const c64*)sp.phi.p, N, nb, c->ek, dt, dt + nb, nullptr,
                             shape, part));
      std::vector<double> h((size_t)kGsChunks * nb * 4);
      HCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
      HCHK(hipStreamSynchronize(c->st));
      sp.ee.assign(nb, 0.0);
      for (int i = 0; i < nb; ++i)
        for (int ch = 0; ch < kGsChunks; ++ch) sp.ee[i] += h[((size_t)ch * nb + i) * 4 + 3];
    }
    HCHK(hipStreamSynchronize(c->st));
    sp.ready = true;
    if (!c->snap_keep) {
      dfree(c, sp.phi.p);
      dfree(c, sp.r.p);
      sp.phi = DevBuf{};
      sp.r = DevBuf{};
    }
  }
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

int opaw_precond_clear(opaw_ctx* c) {
  CTX_CHECK(c);
  for (auto& sp : c->snaps) {
    dfree(c, sp.phi.p);
    dfree(c, sp.r.p);
    dfree(c, sp.ty.p);
    dfree(c, sp.agg.p);
  }
  c->snaps.clear();
  return OPAW_OK;
}

int opaw_precond_count(opaw_ctx* c, int* nsnap, int* nready) {
  CTX_CHECK(c);
  int r = 0;
  for (auto& sp : c->snaps) r += sp.ready ? 1 : 0;
  if (nsnap) *nsnap = (int)c->snaps.size();
  if (nready) *nready = r;
  return OPAW_OK;
}

// Loss of the preconditioner over the prepared pairs and its gradient:
// per pair and band, loss = 1 - |<e, M r>|^2 / (|e|^2 |M r|^2) (k space; a
// scale-free misalignment, since Rayleigh-Ritz only sees the direction of
// M r); loss = mean over pairs and bands with |e| above 1e-12 of the largest.
// grad (mlp_param_count) may be NULL.
int opaw_precond_loss(opaw_ctx* c, const double* mlp, int nh, double* loss, double* grad) {
  CTX_CHECK(c);
  RCHK(need_f64(c));
  if (!mlp || !loss || nh < 1 || nh > 16) return fail(OPAW_ERR_ARG, "bad precond_loss arguments");
  MlpShape shape;
  shape.nh = nh;
  const int npar = mlp_param_count(shape);
  const long long N = c->N;
  std::vector<opaw_ctx::Snap*> ready;
  int nbmax = 0;
  for (auto& sp : c->snaps)
    if (sp.ready) {
      ready.push_back(&sp);
      nbmax = std::max(nbmax, sp.nb);
    }
  const int np = (int)ready.size();
  if (np == 0) return fail(OPAW_ERR_STATE, "no prepared training pairs");
  const int nsh = c->nsh;
  constexpr int kGradChunks = 512;
  // device scratch: parameters, gradient accumulator, per-pair coefficients
  // (3 per band), per-pair per-band loss sums (3)
  RCHK(ensure(c, c->gsSmall, (size_t)(2 * npar + 6 * (size_t)np * nbmax) * sizeof(double)));
  double* dP = (double*)c->gsSmall.p;
  double* dacc = dP + npar;
  double* dcoef = dacc + npar;
  double* dsum = dcoef + 3 * (size_t)np * nbmax;
  HCHK(hipMemcpyAsync(dP, mlp, npar * sizeof(double), hipMemcpyHostToDevice, c->st));
  for (int k = 0; k < np; ++k) {
    const auto* sp = ready[k];
    const double* dt = (const double*)sp->ty.p;
    const c64* z = (const c64*)sp->agg.p;
    const double* w2 = (const double*)(z + (size_t)sp->nb * nsh);
    HCHK(loss_shells(c->st, z, w2, nsh, sp->nb, (const double*)c->shEk.p, N, dt, dt + sp->nb, dP, shape,
                     dsum + (size_t)k * nbmax * 3));
  }
  std::vector<double> h((size_t)np * nbmax * 3);
  HCHK(hipMemcpyAsync(h.data(), dsum, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  double eemax = 0.0;
  for (auto* sp : ready)
    for (double e : sp->ee) eemax = std::max(eemax, e);
  int count = 0;
  for (int k = 0; k < np; ++k)
    for (int i = 0; i < ready[k]->nb; ++i)
      if (ready[k]->ee[i] > 1e-12 * eemax && h[((size_t)k * nbmax + i) * 3 + 2] > 0.0) ++count;
  if (count == 0) return fail(OPAW_ERR_STATE, "every training pair has a zero error vector");
  double L = 0.0;
  std::vector<double> coef((size_t)3 * np * nbmax, 0.0);
  for (int k = 0; k < np; ++k) {
    const int nb = ready[k]->nb;
    double* cA = coef.data() + (size_t)3 * k * nbmax;   // 2 per band, then cB
    double* cB = cA + 2 * nbmax;
    for (int i = 0; i < nb; ++i) {
      const double* q = &h[((size_t)k * nbmax + i) * 3];
      const double ar = q[0], ai = q[1], bb = q[2], ee = ready[k]->ee[i];
      if (!(ee > 1e-12 * eemax && bb > 0.0)) continue;
      const double a2 = ar * ar + ai * ai;
      L += 1.0 - a2 / (ee * bb);
      // dL/dg = [-2 Re(conj(A) z) + 2 |A|^2 m w2 / B] m / (Ee B), over count
      const double sc = 1.0 / (ee * bb * count);
      cA[2 * i] = -2.0 * ar * sc;       // conj(A) = ar - i ai: Re(conj(A) z) = ar Re z + ai Im z
      cA[2 * i + 1] = 2.0 * ai * sc;    // kernel forms cA0 Re z - cA1 Im z
      cB[i] = 2.0 * a2 / bb * sc;
    }
  }
  *loss = L / count;
  if (!grad) return OPAW_OK;
  int rc;
  double* gpart = gs_part(c, (size_t)kGradChunks * npar, &rc);
  RCHK(rc);
  HCHK(hipMemcpyAsync(dcoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
  HCHK(hipMemsetAsync(dacc, 0, npar * sizeof(double), c->st));
  for (int k = 0; k < np; ++k) {
    const auto* sp = ready[k];
    const double* dt = (const double*)sp->ty.p;
    const c64* z = (const c64*)sp->agg.p;
    const double* w2 = (const double*)(z + (size_t)sp->nb * nsh);
    const double* cA = dcoef + (size_t)3 * k * nbmax;
    const int nchunk = (int)std::min<long long>(kGradChunks, ((long long)sp->nb * nsh + 63) / 64);
    HCHK(grad_shells(c->st, z, w2, nsh, sp->nb, (const double*)c->shEk.p, N, dt, dt + sp->nb, dP, shape, cA,
                     cA + 2 * nbmax, gpart, nchunk));
    HCHK(colsum_accumulate(c->st, gpart, nchunk, npar, dacc));
  }
  HCHK(hipMemcpyAsync(grad, dacc, npar * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  return OPAW_OK;
}

// prepared pair idx (0-based among the prepared ones): r(k) and e(k) of its
// nb bands (complex*16, nb x N, k space, unnormalised forward FFT), the band
// kinetic energies t and features y; any output may be NULL
int opaw_precond_pair(opaw_ctx* c, int idx, void* rk, void* ek, double* t, double* y, int* nb) {
  CTX_CHECK(c);
  int k = -1;
  for (auto& sp : c->snaps) {
    if (!sp.ready) continue;
    if (++k != idx) continue;
    const size_t fb = (size_t)sp.nb * c->N * sizeof(c64);
    if ((rk || ek) && !sp.r.p) return fail(OPAW_ERR_STATE, "pair %d kept only its shell sums", idx);
    HCHK(hipStreamSynchronize(c->st));
    if (rk) HCHK(hipMemcpy(rk, sp.r.p, fb, hipMemcpyDeviceToHost));
    if (ek) HCHK(hipMemcpy(ek, sp.phi.p, fb, hipMemcpyDeviceToHost));
    if (t) std::copy(sp.t.begin(), sp.t.end(), t);
    if (y) std::copy(sp.y.begin(), sp.y.end(), y);
    if (nb) *nb = sp.nb;
    return OPAW_OK;
  }
  return fail(OPAW_ERR_ARG, "no prepared pair %d", idx);
}

int opaw_precond_keep_fields(opaw_ctx* c, int keep) {
  CTX_CHECK(c);
  c->snap_keep = keep != 0;
  return OPAW_OK;
}

int opaw_mlp_param_count(int nh) {
  MlpShape s;
  s.nh = nh;
  return mlp_param_count(s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// local potential of the nonperiodic (Martyna-Tuckerman) path:
// vloc_tuma_prep -> vloc_prep (src/vloc/vloc_prep.f90:1-79) with make_cvkb
// (src/vloc/make_cvkb.f90) and add_formfactor (src/vloc/formfactor1.f90)
extern "C" int opaw_vloc_tuma(opaw_ctx* c, int ntype, const int* nr_of_type, const double* rr_cat,
                              const double* vr_cat, int natom, const int* atom_type, const double* coords,
                              double* vloc_out) {
  CTX_CHECK(c);
  if (c->periodic) return fail(OPAW_ERR_ARG, "vloc_tuma_prep is the nonperiodic path");
  if (ntype < 1 || natom < 1 || !nr_of_type || !rr_cat || !vr_cat || !atom_type || !coords || !vloc_out)
    return fail(OPAW_ERR_ARG, "bad vloc arguments");
  for (int a = 0; a < natom; ++a)
    if (atom_type[a] < 1 || atom_type[a] > ntype) return fail(OPAW_ERR_ARG, "atom %d has type %d", a + 1, atom_type[a]);
  const GridGeom& B = c->gbig;
  const long long Nb = (long long)B.nx * B.ny * B.nz;
  const long long N = c->N;
  std::vector<void*> tmp;
  auto cleanup = [&]() {
    for (void* p : tmp) dfree(c, p);
  };
  auto alloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (dalloc(c, &p, bytes) != OPAW_OK) return nullptr;
    tmp.push_back(p);
    return p;
  };
  c64* A = (c64*)alloc((size_t)Nb * sizeof(c64));
  c64* cvn = (c64*)alloc((size_t)Nb * sizeof(c64));
  double* dv_out = (double*)alloc((size_t)N * sizeof(double));
  if (!A || !cvn || !dv_out) {
    cleanup();
    return OPAW_ERR_ALLOC;
  }
  int rc = OPAW_OK;
  auto H = [&](hipError_t e) {
    if (e != hipSuccess && rc == OPAW_OK) rc = fail(OPAW_ERR_HIP, "%s", hipGetErrorString(e));
    return rc == OPAW_OK;
  };
  const c64 *twx = c->tw64[B.nx], *twy = c->tw64[B.ny], *twz = c->tw64[B.nz];
  H(hipMemsetAsync(cvn, 0, (size_t)Nb * sizeof(c64), c->st));
  size_t off = 0;
  for (int t = 0; t < ntype && rc == OPAW_OK; ++t) {
    const int nr = nr_of_type[t];
    if (nr < 3) {
      rc = fail(OPAW_ERR_ARG, "type %d: radial grid too short", t + 1);
      break;
    }
    const double* rr = rr_cat + off;
    const double* vr = vr_cat + off;
    off += nr;
    // make_cvkb: spline of vr on max(rr, 0.1 dr1) (SPLINE_dn, natural)
    const double dr1 = rr[1] - rr[0];
    std::vector<double> xs(nr), y2(nr, 0.0), u(nr, 0.0);
    for (int i = 0; i < nr; ++i) xs[i] = std::max(rr[i], 0.1 * dr1);
    for (int i = 1; i < nr - 1; ++i) {
      const double sig = (xs[i] - xs[i - 1]) / (xs[i + 1] - xs[i - 1]);
      const double p = sig * y2[i - 1] + 2.0;
      y2[i] = (sig - 1.0) / p;
      u[i] = (6.0 * ((vr[i + 1] - vr[i]) / (xs[i + 1] - xs[i]) - (vr[i] - vr[i - 1]) / (xs[i] - xs[i - 1])) /
                  (xs[i + 1] - xs[i - 1]) -
              sig * u[i - 1]) /
             p;
    }
    y2[nr - 1] = 0.0;
    for (int k = nr - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
    const double charge = vr[nr - 1] * rr[nr - 1];
    double* d = (double*)alloc((size_t)4 * nr * sizeof(double));
    std::vector<double> at;
    for (int a = 0; a < natom; ++a)
      if (atom_type[a] == t + 1) at.insert(at.end(), coords + 3 * a, coords + 3 * a + 3);
    double* dc = at.empty() ? nullptr : (double*)alloc(at.size() * sizeof(double));
    if (!d || (!at.empty() && !dc)) {
      rc = OPAW_ERR_ALLOC;
      break;
    }
    if (!H(hipMemcpy(d, rr, nr * sizeof(double), hipMemcpyHostToDevice)) ||
        !H(hipMemcpy(d + nr, xs.data(), nr * sizeof(double), hipMemcpyHostToDevice)) ||
        !H(hipMemcpy(d + 2 * nr, vr, nr * sizeof(double), hipMemcpyHostToDevice)) ||
        !H(hipMemcpy(d + 3 * nr, y2.data(), nr * sizeof(double), hipMemcpyHostToDevice)))
      break;
    if (at.empty()) continue;
    if (!H(hipMemcpy(dc, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice))) break;
    if (!H(vloc_sample(c->st, d, d + nr, d + 2 * nr, d + 3 * nr, nr, charge, dr1, B.nx, B.ny, B.nz, c->dx, c->dy,
                       c->dz, A)))
      break;
    if (!H(fft3d<double>(B, c->st, -1, A, c->hwork, 1, twx, twy, twz))) break;
    if (!H(vloc_accumulate(c->st, c->hwork, cvn, B.nx, B.ny, B.nz, c->dx, c->dy, c->dz, dc, (int)at.size() / 3)))
      break;
  }
  if (rc == OPAW_OK && H(fft3d<double>(B, c->st, 1, cvn, A, 1, twx, twy, twz)) &&
      H(vloc_crop(c->st, A, B.nx, B.ny, B.nz, c->nx, c->ny, c->nz, 1.0 / ((double)Nb * dvol(c)), dv_out)) &&
      H(hipStreamSynchronize(c->st)))
    H(hipMemcpy(vloc_out, dv_out, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
  cleanup();
  return rc;
}
