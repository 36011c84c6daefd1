/*
 * opaw_hip.h -- C ABI of the MI355X (gfx950) OPAW-TDDFT propagation hot path.
 *
 * The reference (Fortran 90 + MPI) keeps the hot path behind external
 * subroutines that act on module state (main_mod, atom_mod, tddft_mod).  This
 * ABI keeps that shape: an opaque context owns the device-resident copy of the
 * module state (grid, kinetic table, projector boxes, potentials, D_ij, the
 * propagated orbitals) and every entry point names the reference subroutine
 * it replaces.  Plain C types only; arrays are passed in the reference's
 * Fortran layouts (column-major, x fastest, complex interleaved).
 *
 * Conventions
 *   - every function returns 0 (OPAW_OK) or a negative status; the message of
 *     the last failure on the calling thread is in opaw_last_error().  The
 *     reference aborts with `stop` instead (e.g. src/hphi.f90:38).
 *   - atom indices `ia` and type indices `it` are 1-based, like the reference.
 *   - "device" buffers are HIP device pointers holding nbatch complex fields of
 *     nx*ny*nz points each: complex*16 in an OPAW_PREC_F64 context,
 *     complex*8 in an OPAW_PREC_MIXED context.
 *   - all work is enqueued on the context stream (opaw_set_stream); results in
 *     caller buffers are ready after opaw_synchronize().
 *   - host arrays passed to set/read functions are copied synchronously.
 */
#ifndef OPAW_HIP_H
#define OPAW_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct opaw_ctx opaw_ctx;

enum {
  OPAW_OK = 0,
  OPAW_ERR_ARG = -1,         /* bad argument / shape (reference: stop) */
  OPAW_ERR_HIP = -2,         /* HIP runtime error */
  OPAW_ERR_ALLOC = -3,       /* device allocation failed */
  OPAW_ERR_STATE = -4,       /* tables / potentials not set yet */
  OPAW_ERR_UNSUPPORTED = -5  /* grid size without a compiled FFT length */
};

enum { OPAW_PREC_F64 = 0, OPAW_PREC_MIXED = 1 };

/* fields readable with opaw_read_field */
enum { OPAW_FIELD_DENS = 0, OPAW_FIELD_NHAT = 1, OPAW_FIELD_VXC = 2, OPAW_FIELD_VH = 3, OPAW_FIELD_VKS = 4 };

/* kernel families timed by the profiler (opaw_profile_read) */
enum {
  OPAW_K_FFT_X = 0,      /* kinetic: xy-plane (or x-line) forward FFT        */
  OPAW_K_FFT_Y = 1,      /* kinetic: y line FFTs (non-plane grids only)     */
  OPAW_K_FFT_Z = 2,      /* kinetic: z line FFT, * ek, inverse (fused)      */
  OPAW_K_FFT_XH = 3,     /* kinetic: xy inverse + vks*psi + nonlocal epilogue */
  OPAW_K_HARTREE = 4,    /* Martyna-Tuckerman Hartree FFT passes            */
  OPAW_K_BOX_PROJ = 5,   /* projector contractions P^T psi                 */
  OPAW_K_BOX_ADD = 6,    /* projection finish, weights, union corrections  */
  OPAW_K_RK4 = 7,        /* RK4 stage update kernel                         */
  OPAW_K_DENS = 8,       /* density, rho_ij, nhat, norms + rescale, dipoles */
  OPAW_K_POT = 9,        /* LDA, V_KS, D_ij, packing                        */
  OPAW_K_OTHER = 10,     /* plain fft3d and everything else                 */
  OPAW_K_COUNT = 11
};

const char* opaw_last_error(void);
int opaw_version(void);
int opaw_fft_length_supported(int n);
/* 1 when the kinetic operator of nx x ny planes runs as the one-visit xy-plane
 * kernel at this precision (OPAW_PREC_F64 / OPAW_PREC_MIXED), 0 when it runs
 * as line passes.  A property of this library's kernels; the reference's
 * FFTW calls (src/lib/library_fftw_v3.f90:56-83) have no such split. */
int opaw_plane_path(int nx, int ny, int precision);

/* ---- context: grid and orbital arrays ------------------------------------
 * Replaces the grid part of main_mod (src/4_main.f90:5-50) and the orbital
 * arrays phi_bar / phi_bar_pert of tddft_mod (src/5_tddft.f90, allocated in
 * src/tddft_alloc.f90:15-16).  nsys systems (the reference propagates 2:
 * unperturbed and kicked) of nb occupied orbitals each.  periodic != 0 is the
 * reference `periodic` flag (Hartree scale_vh = 1, else 2: src/main.f90:351). */
int opaw_create(opaw_ctx** out, int nx, int ny, int nz, double dx, double dy, double dz, int periodic,
                int precision, int nsys, int nb);
int opaw_destroy(opaw_ctx* ctx);
int opaw_set_stream(opaw_ctx* ctx, void* hip_stream);
int opaw_synchronize(opaw_ctx* ctx);
long long opaw_device_bytes(opaw_ctx* ctx);

/* ---- kinetic table -------------------------------------------------------
 * ek_prep (src/ek.f90:1-51): ek = min(|k|^2/2, ekcut)/N with k folded into
 * (-pi/d, pi/d]; computed on the device.  opaw_set_ek uploads a caller table
 * ek(nx,ny,nz) already divided by N. */
int opaw_ek_prep(opaw_ctx* ctx, double ekcut);
int opaw_set_ek(opaw_ctx* ctx, const double* ek);

/* ---- atoms and PAW tables (atom_mod / paw_mod, src/3_atom.f90, src/2_paw.f90)
 * atom_map(natom) 1-based types; ir_start(3,natom) 1-based box origins as
 * from get_irs (src/get_irs.f90:22-41; a box outside the grid is an error
 * unless periodic, where it wraps); nrough(3,ntype); mstates(ntype);
 * nl2 = (2*lmax+1)^2 compensation channels. */
int opaw_set_atoms(opaw_ctx* ctx, int natom, const int* atom_map, const int* ir_start, int ntype,
                   const int* nrough, const int* mstates, int nl2);
/* local_p3d(nr1,nr2,nr3,ms) and local_p3d1(...) (real at the gamma point:
 * src/p3d.f90:93-103 stores them as complex with zero phase), s(ms), sinv(ms)
 * from transform_p3d (src/transform_p3d.f90:46-56). */
int opaw_set_projectors(opaw_ctx* ctx, int ia, const double* local_p3d, const double* local_p3d1,
                        const double* s, const double* sinv);
/* compensation sphere: local_grid(ngrid,3) 1-based, local_g3d(ngrid,nl2)
 * (src/3_atom.f90:24-26, src/g3d.f90). */
int opaw_set_sphere(opaw_ctx* ctx, int ia, int ngrid, const int* local_grid, const double* local_g3d);
/* per-type qijlm(nl2,ms,ms) of the compensation charge (src/2_paw.f90), and
 * the constant and Hartree parts of the on-site D_ij, dij0(ms,ms) +
 * eijkl(ms,ms,ms,ms).rho_sel (libpaw's pawtab%dij0 and pawdijhartree,
 * src/libpaw/4_m_pawdij.F90:1228-1313, unpacked).  A type without
 * opaw_set_onsite tables keeps just these (with D^hat over the compensation
 * sphere): the model of the synthetic workloads. */
int opaw_set_type(opaw_ctx* ctx, int it, const double* qijlm, const double* dij0, const double* eijkl);
/* libpaw's on-site tables of type it, which replace the model's on-site part
 * by what get_dij (src/get_dij.f90:14-30) computes through pawdenpot
 * (src/libpaw/m_paw_denpot.F90:136) and pawdij
 * (src/libpaw/4_m_pawdij.F90:168) for LDA with pawxcdev = 1: with them
 * D_ij = dij0 + dijhartree + dijxc + dijhat, where dij0 and eijkl
 * (dijhartree) come from opaw_set_type (libpaw's pawtab%dij0 and eijkl,
 * unpacked to (ms,ms) and (ms,ms,ms,ms)), dijxc is computed on the device at
 * every rebuild from rho_ij, and dijhat uses the atom's D^hat sphere
 * (opaw_set_dij_sphere) with qijl instead of the compensation sphere.
 * Radial tables on nr points of the dataset's mesh (pawtab%mesh_size):
 * rad(nr) with rad(1) = 0, w(nr) the simp_gen weights (zero beyond
 * int_meshsz), phiphj / tphitphj (nr, nij) the products phi_i phi_j of the
 * radial pairs, coredens / tcoredens (nr) the core densities (tcoredens
 * read only when usetcore), gnt (lm_size, ms, ms) the real Gaunt
 * coefficient of (lm, lm_i, lm_j), kln (ms, ms) the 0-based radial pair of
 * (i, j), qijl (lm_size, ms, ms) pawtab%qijl.  lm_size <= 9 and nij <= 10
 * (partial waves with l <= 1, up to four of them), else
 * OPAW_ERR_UNSUPPORTED. */
int opaw_set_onsite(opaw_ctx* ctx, int it, int nr, int lm_size, int nij, int usetcore, const double* rad,
                    const double* w, const double* phiphj, const double* tphitphj, const double* coredens,
                    const double* tcoredens, const double* gnt, const int* kln, const double* qijl);
/* pawfgrtab of atom ia for pawdijhat (nhatgrid with optcut = 0,
 * src/libpaw/finegrid.f90; pawgylm, src/libpaw/2_m_paw_finegrid.F90:92):
 * the nfgd grid points within rshp, ifftsph(nfgd) 1-based flat indices
 * (x fastest), and gylm(nfgd, lm_size) = g_l(r) S_lm.  The atom's type must
 * have on-site tables. */
int opaw_set_dij_sphere(opaw_ctx* ctx, int ia, int nfgd, const int* ifftsph, const double* gylm);
/* builds the plane lists and uploads everything; call after the set_* above */
int opaw_finalize_tables(opaw_ctx* ctx);

/* ---- potentials -----------------------------------------------------------
 * vloc_tot and ncoret (src/vloc.f90), rnel (src/rnel.f90); usexcnhat selects
 * the D^hat potential (0: vks - vxc, as libpaw pawdij with usexcnhat=0). */
int opaw_set_potentials(opaw_ctx* ctx, const double* vloc_tot, const double* ncoret, double rnel,
                        int usexcnhat);
/* Martyna-Tuckerman (or periodic 4pi/k^2) Hartree kernel on the
 * scale_vh-times box: vk_prep (src/vk_scale.f90:1-48, src/vk_TuMa.f90:2-117),
 * computed on the device. */
int opaw_vk_prep(opaw_ctx* ctx);

/* ---- per-system state (vks / vks_pert, dij / dij_pert, rhoij, ...) ------- */
int opaw_write_vks(opaw_ctx* ctx, int sys, const double* vks);
int opaw_read_vks(opaw_ctx* ctx, int sys, double* vks);
int opaw_write_dij(opaw_ctx* ctx, int sys, int ia, const double* dij);
int opaw_read_dij(opaw_ctx* ctx, int sys, int ia, double* dij);
int opaw_read_rhoij(opaw_ctx* ctx, int sys, int ia, double* rhoij);
int opaw_read_field(opaw_ctx* ctx, int which, int sys, double* out);

/* ---- resident orbitals: phi_bar(:,:,:,1:nb) of system sys ---------------- */
int opaw_write_orbitals(opaw_ctx* ctx, int sys, const void* host);
int opaw_read_orbitals(opaw_ctx* ctx, int sys, void* host);
void* opaw_orbitals_device(opaw_ctx* ctx, int sys);

/* ---- operators on caller device buffers ----------------------------------
 * h_phi / h_phi_pert (src/hphi.f90:1-66, src/hphi_pert.f90): hp = T pin +
 * vks(sys) pin + sum_a p_a D_a(sys) <p_a|pin> dv, for nbatch fields. */
int opaw_h_phi(opaw_ctx* ctx, int sys, const void* pin, void* hp, int nbatch);
/* sn_phi (src/snphi.f90:1-49): sp = pin + sum_a p1_a ((1+s)^n - 1) <p1_a|pin> dv */
int opaw_sn_phi(opaw_ctx* ctx, const void* pin, void* sp, int nbatch, double n);
/* shs / shs_pert (src/shs.f90:3-70): pout = S^-1/2 H(sys) S^-1/2 pin */
int opaw_shs(opaw_ctx* ctx, int sys, const void* pin, void* pout, int nbatch);
/* proj / proj1 (src/proj.f90, src/proj1.f90) for all atoms: ca is a device
 * complex*16 array (nbatch, sum_a ms_a), atom a at offset sum_{b<a} ms_b. */
int opaw_proj(opaw_ctx* ctx, const void* pin, int nbatch, int which_p1, void* ca);
/* fft3d_forward (dir=-1) / fft3d_backward (dir=+1), unnormalised
 * (src/lib/library_fftw_v3.f90:56-83) */
int opaw_fft3d(opaw_ctx* ctx, int dir, const void* in, void* out, int nbatch);

/* ---- per-stage Hamiltonian rebuild (src/ncpaw_make_hamiltonian.f90) ------
 * psi: device fields of system sys (nb of them), already S^-1/2 transformed
 * where the reference transforms them.  Each call updates the resident
 * per-system state exactly like the module-state routine it names. */
int opaw_tddft_rhoij_2(opaw_ctx* ctx, int sys, const void* psi);  /* src/tddft_rhoij.f90:75-115 */
int opaw_tddft_dens_2(opaw_ctx* ctx, int sys, const void* psi);   /* src/tddft_dens.f90:44-85 */
int opaw_get_nhat(opaw_ctx* ctx, int sys);                        /* src/nhat.f90:1-43 */
int opaw_get_pot_tddft_2(opaw_ctx* ctx, int sys);                 /* src/tddft_pot.f90:110-181 */
int opaw_get_dij(opaw_ctx* ctx, int sys);                         /* src/get_dij.f90:1-156 */
/* ncpaw_make_hamiltonian_tddft (all systems, :118-167) when sys < 0, else
 * the _unpert / _pert variant (:213-281) for that system; acts on the
 * resident orbitals. */
int opaw_make_hamiltonian_tddft(opaw_ctx* ctx, int sys);

/* ---- host-array operators (drop-ins for a Fortran host) ---------------------
 * The reference's operators take complex*16 arrays in host memory; these
 * forms take the same (pin(nx,ny,nz), x fastest, nbatch fields back to back),
 * convert to the context precision, run the device operator above and
 * return complex*16.  The Fortran drop-ins h_phi / h_phi_pert, shs /
 * shs_pert, sn_phi, proj / proj1 and rk4_prop_shs call these.
 * h_phi(pin,hp,ik) / h_phi_pert (src/hphi.f90:1-10, src/hphi_pert.f90:1):
 *   sys 0 / sys 1 */
int opaw_h_phi_host(opaw_ctx* ctx, int sys, const double* pin, double* hp, int nbatch);
/* shs(pin,pout,ik) / shs_pert (src/shs.f90:3-12, src/shs.f90:50): sys 0 / 1 */
int opaw_shs_host(opaw_ctx* ctx, int sys, const double* pin, double* pout, int nbatch);
/* sn_phi(pin,sp,ik,n) (src/snphi.f90:1-11) */
int opaw_sn_phi_host(opaw_ctx* ctx, const double* pin, double* sp, int nbatch, double n);
/* proj(ia,pin,ca,ms,ik) / proj1 (src/proj.f90:1-9, src/proj1.f90:1-9): the
 * ms coefficients of atom ia (1-based) of one field, ca(is) = dv sum pin
 * conj(p_is) over the atom's box; ms must equal the atom's projector count */
int opaw_proj_host(opaw_ctx* ctx, int ia, const double* pin, int which_p1, double* ca, int ms);
/* complex*16 forms of opaw_write_orbitals / opaw_read_orbitals, whatever the
 * context precision (scatter_phi_bar / gather_phi_bar, src/tddft_scatter.f90) */
int opaw_write_orbitals_c16(opaw_ctx* ctx, int sys, const double* host);
int opaw_read_orbitals_c16(opaw_ctx* ctx, int sys, double* host);
/* tddft_mod's dt (src/tddft_mod.f90), read by opaw_rk4_prop_shs_host */
int opaw_set_dt(opaw_ctx* ctx, double dt);
/* rk4_prop_shs(p,pp,n,nb) (src/rk4.f90:1-12) on host arrays p = phi_bar,
 * pp = phi_bar_pert (n = nx ny nz points, nb orbitals): p and pp must equal
 * the resident orbitals of systems 0 and 1 that the current Hamiltonian was
 * built from (else OPAW_ERR_STATE, nothing changed: write them with
 * opaw_write_orbitals_c16 and rebuild first); one step with the context's
 * dt, the results returned in p and pp. */
int opaw_rk4_prop_shs_host(opaw_ctx* ctx, double* p, double* pp, long long n, int nb);

/* ---- time propagation ------------------------------------------------------ */
/* excite_psi_bar_pert (src/tddft_excite.f90:1-16): orbitals(to) =
 * orbitals(from) * exp(-i sm r_ipol), ipol 1..3 */
int opaw_excite_psi_bar_pert(opaw_ctx* ctx, int from_sys, int to_sys, double sm, int ipol);
/* rk4_prop_shs (src/rk4.f90:1-103): one RK4 step of every system with the
 * per-stage self-consistent rebuild and the per-orbital norm restore. */
int opaw_rk4_prop_shs(opaw_ctx* ctx, double dt);
/* plot_dip (src/tddft_plot.f90:1-155): dipoles of systems 0 and 1 and
 * d_dip = (dip_pert - dip)/sm; any output pointer may be NULL. */
int opaw_plot_dip(opaw_ctx* ctx, double sm, double* d_dip, double* dip, double* dip_pert);
/* the time loop of opaw_tddft (src/tddft.f90:67-91, dyn=3): per step the
 * rebuild of all systems, rk4_prop_shs, and the dipole; d_dip (nt,3) and
 * dip (nt, 3*nsys) may be NULL.  No host synchronisation inside the loop. */
int opaw_tddft_run(opaw_ctx* ctx, int nt, double dt, double sm, double* d_dip, double* dip);
/* The same time loop for several independent contexts at once (molecules of
 * their own grids, orbital counts and kick systems), replacing one
 * opaw_tddft (src/tddft.f90:67-91) run per molecule.  Each context runs on a
 * stream of its own, handed over from and back to its caller's stream, so the
 * kernels of different molecules overlap on the device.  graph != 0: after
 * the first step, one time step per context is captured as a HIP graph and
 * replayed for the remaining nt-1 steps.  dip receives, context after
 * context, each one's (nt, 3*nsys) series (may be NULL).  No context may be
 * used by another thread during the call. */
int opaw_tddft_run_batch(opaw_ctx* const* ctxs, int nctx, int nt, double dt, int graph, double* dip);

/* ---- orthogonalisation / subspace diagonalisation ------------------------- */
/* orthog_phi (src/orthog_phi.f90:1-194): canonical phi U lambda^-1/2 of the
 * resident orbitals of system sys; h_type 1: S = 1, h_type 0: S via sn_phi(1).
 * eig (nb) receives the overlap eigenvalues (ascending). */
int opaw_orthog_phi(opaw_ctx* ctx, int sys, int h_type, double* eig);
/* diagh (src/diag.f90:1-180): Rayleigh-Ritz on shs (h_type 1) or h_phi
 * (h_type 0); orbitals rotated in place, eig (nb) ascending. */
int opaw_diagh(opaw_ctx* ctx, int sys, int h_type, double* eig);

/* ---- on-disk formats (host code; no device work except the restart copies)
 * Binary files are Fortran unformatted sequential files (4-byte record
 * markers, 4-byte integers, 8-byte reals); text files are read with Fortran
 * list-directed input.  Orbital arrays here are host complex*16, nstates
 * fields of nx*ny*nz points, x fastest. */
/* writewf (src/write_wf.f90:1-120; short_header = 0: wf_bar.bin / .txt with
 * orbmn_wf and ihomo) or write_wf_for_sgw_debug (src/write_wf.f90:122-219;
 * short_header = 1: the 8 header records that read_wf_tddft reads).  The
 * stored orbital is sign(Re phi) |phi| (src/write_wf.f90:55-63). */
int opaw_write_wf(const char* path, int binary, int short_header, int nx, int ny, int nz, double dx, double dy,
                  double dz, int nstates, int ihomo, const double* eig, const void* phi);
/* read_wf_tddft (src/read_wf.f90:130-228): checks the grid, the spacings and
 * nsp = 1 against the caller's (src/read_wf.f90:181-187), accepts both header
 * layouts above; returns nstates (<= max_states), ihomo (0 when the file has
 * none), eig (nstates) and phi (nstates fields, imaginary part 0); eig, phi
 * and ihomo may be NULL. */
int opaw_read_wf(const char* path, int binary, int nx, int ny, int nz, double dx, double dy, double dz,
                 int max_states, int* nstates, int* ihomo, double* eig, void* phi);
/* eig / eig_bar: `write(1,*) eig` (src/diag.f90:109-115), read back with
 * `read(1,*) eig` (src/write_wf.f90:33) */
int opaw_write_eig(const char* path, int n, const double* eig);
int opaw_read_eig(const char* path, int n, double* eig);
/* plot_dip's files (src/tddft_plot.f90:100-105) for nt steps starting at
 * it = it0+1: d_dip.dat lines `(' ',4e18.8)` of (it-1)*dt, d_dip(3) and
 * dip.dat lines (it-1)*dt, dip(3*nsys) (list-directed); d_dip (nt,3) and
 * dip (nt, 3*nsys) as returned by opaw_tddft_run.  append = 1 continues
 * existing files; either path/array pair may be NULL. */
int opaw_write_dipoles(const char* d_dip_path, const char* dip_path, int append, int it0, int nt, double dt,
                       const double* d_dip, const double* dip, int nsys);
/* write_to_restart (src/tddft_write_to_restart.f90:1-67): dir/info holds
 * `nb nds it` (nds = 1: one process), dir/phi_bar.bin system 0 and
 * dir/phi_bar_pert.bin system 1 (phi_bar_pert<s>.bin for further systems),
 * one complex*16 record per orbital.  Creates dir when missing. */
int opaw_write_restart(opaw_ctx* ctx, const char* dir, int it);
/* read_phi_bar (src/read_phi_bar.f90:1-51): loads every system's orbitals
 * from dir and returns it (the time loop restarts at it+1,
 * src/tddft.f90:26-28); nb and nds of dir/info must match the context. */
int opaw_read_restart(opaw_ctx* ctx, const char* dir, int* it);

/* ---- absorption spectrum (post-processing of the d_dip series; the
 * reference leaves it to external scripts, README.md:7-11).  For nsig host
 * series d (series s at d + s*nt, sample n at t = n*dt, the rows of d_dip.dat):
 *   alpha_s(w) = -dt * sum_n c_n f(t_n) d_s(t_n) exp(i w t_n),
 *   c_n = 1 - [n = 0]/2 - [n = nt-1]/2 (trapezoid; 0 when nt = 1);
 *   f = 1 (damp_kind 0), exp(-t/tau) (1),
 *   exp(-t^2/(2 tau^2)) (2);  w_k = w0 + k*dw (Hartree), k < nw;
 *   strength_s(w) = (2 w / pi) Im alpha_s(w).
 * With the reference's kick phi exp(-i sm r) (src/tddft_excite.f90:11-12)
 * and d = d_dip, alpha is the polarizability (the electrons' dipole is
 * minus the density dipole of src/tddft_plot.f90:40-57) and strength > 0
 * at absorption peaks.
 * alpha (nsig, nw) complex interleaved and strength (nsig, nw) are host
 * arrays; either may be NULL.  Runs on the current device, synchronously. */
int opaw_spectrum(const double* d, int nsig, int nt, double dt, int damp_kind, double tau, double w0, double dw,
                  int nw, double* alpha, double* strength);
/* opaw_spectrum on caller DEVICE buffers with no allocation: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops).  d (nsig * nt), alpha
 * (2 nsig nw) and strength (nsig nw) are device memory of the current
 * device (checked; OPAW_ERR_ARG otherwise, alpha and strength may not be
 * NULL); the weighted series lives in the library's static per-device
 * scratch array of 2^21 doubles (16 MiB), so nsig * nt <= 2^21 (else
 * OPAW_ERR_ARG).  Enqueued on `stream` (a hipStream_t; NULL: the null
 * stream) with no host synchronisation; results are the same bits as
 * opaw_spectrum's.  Calls that overlap on one device share the scratch
 * array and race. */
int opaw_spectrum_noalloc(const double* d, int nsig, int nt, double dt, int damp_kind, double tau, double w0,
                          double dw, int nw, double* alpha, double* strength, void* stream);

/* ---- PAW set-up, grid-sized part (prep_paw, src/prepare_paw.f90:1-49).
 * Stateless; host arrays in and out; runs on the current device.
 * get_local_p3d (src/p3d.f90:2-114) for one atom: for each of ms m-states,
 * rad(nr, ms) / rad1(nr, ms) (Fortran layout: the radial ptilde / ptilde1 of
 * the state's radial function) linearly interpolated in r and times the real
 * Y_lm of lm(2, ms) = (l, m), sampled at the fine points
 * corner + (i-1) dxf (nfine(3) of them, minimum image of the atom at coord
 * in the box of half lengths half(3)), then contracted onto the rough box
 * with the spline matrices bx(nrough(1), nfine(1)), by, bz of get_b
 * (src/read_paw.f90:277-345).  Outputs local_p3d(nrough, ms) and
 * local_p3d1 (before transform_p3d). */
int opaw_paw_box(int nr, const double* rr, int ms, const double* rad, const double* rad1, const int* lm,
                 const int* nrough, const int* nfine, const double* dxf, const double* corner, const double* coord,
                 const double* half, const double* bx, const double* by, const double* bz, double* local_p3d,
                 double* local_p3d1);
/* map_localgrid (src/map_localgrid.f90:4-58) and get_localg3d
 * (src/g3d.f90:1-119) for one atom: the grid points with (minimum-image)
 * distance < rcut in grid order, local_grid(ngrid, 3) 1-based, and
 * local_g3d(ngrid, nl2) = g_l(r) Y_L for r <= rcomp (else 0) from the shape
 * functions gl(nrc, 2 nl - 1).  With max_points = 0 only ngrid is returned
 * (size query); otherwise ngrid <= max_points is required. */
/* vloc_tuma_prep -> vloc_prep (src/vloc/vloc_prep.f90:1-79, make_cvkb
 * src/vloc/make_cvkb.f90, add_formfactor src/vloc/formfactor1.f90): the local
 * potential of a nonperiodic context.  Per type t, the radial potential
 * vr (the dataset's vloc, already times Y_00) on rr (nr_of_type[t] points;
 * rr_cat / vr_cat concatenate the types) is splined onto the doubled box,
 * Fourier transformed, shifted and summed over the atoms of the type with
 * structure factors; the inverse transform cropped to the grid is vloc_out
 * (nx*ny*nz).  atom_type is 1-based, coords (3, natom) in bohr. */
int opaw_vloc_tuma(opaw_ctx* ctx, int ntype, const int* nr_of_type, const double* rr_cat, const double* vr_cat,
                   int natom, const int* atom_type, const double* coords, double* vloc_out);
int opaw_paw_sphere(int nx, int ny, int nz, double dx, double dy, double dz, const double* coord, double rcut,
                    double rcomp, int nr, const double* rr, int nl2, int nrc, const double* gl, int max_points,
                    int* ngrid, int* local_grid, double* local_g3d);

/* ---- ground state (FP64 contexts) -----------------------------------------
 * The SCF pieces of the reference's main loop (src/main.f90:66-95):
 * ncpaw_make_hamiltonian, set_hminmax, chebyr_m, orthog_phi, diagh.
 * h_type 1: the operator is shs (orthogonal orbitals), 0: h_phi. */
/* orbitals 1..nocc of each system make its density and rho_ij
 * (acc_dens1 / acc_coeff1, src/update_dens.f90:60-79); default nb.  The
 * propagation entry points require nocc = nb. */
int opaw_set_nocc(opaw_ctx* ctx, int nocc);
/* DIIS history length ndi (src/0_parameter.f90:8, 4) and mixing mix_diis of
 * the V_KS extrapolation; resets the history.  ndi = 0 disables it. */
int opaw_set_diis(opaw_ctx* ctx, int ndi, double mix);
/* ncpaw_make_hamiltonian (src/ncpaw_make_hamiltonian.f90:1-34, h_type 1):
 * psi~ = S^-1/2 phi of the occupied orbitals, rho_ij, density, nhat with the
 * rnel renormalisation, V_KS (get_pot, src/pot.f90:1-99) followed by
 * diis_general (src/diis.f90:12-153) when use_diis, then D_ij. */
int opaw_ncpaw_make_hamiltonian(opaw_ctx* ctx, int sys, int use_diis);
/* The first SCF iteration (src/main.f90:46-69): iscf = 1 skips update_dens,
 * so the Hamiltonian comes from the starting density dens (nx*ny*nz host
 * array; the reference's is atm2fft's superposition of atomic pseudo
 * valence densities, src/vloc.f90:24-29) with rho_ij = 0 and nhat = 0
 * (src/alloc_atom.f90:24, src/main.f90:46-52): get_pot (src/pot.f90) with
 * DIIS when use_diis, then get_dij. */
int opaw_ncpaw_make_hamiltonian_density(opaw_ctx* ctx, int sys, const double* dens, int use_diis);
/* get_hminmax (src/hminmax.f90:18-108): two power iterations of niter steps
 * from random real starts (splitmix64 from seed), the interval widened by
 * fctr (1.2 in the reference). */
int opaw_hminmax(opaw_ctx* ctx, int sys, int h_type, int niter, double fctr, unsigned long long seed, double* hmin,
                 double* hmax);
/* chebyr_m (src/chebyr_m.f90:1-48): the degree-nch Chebyshev filter on
 * [hbot, hmax] applied to all nb resident orbitals of system sys. */
int opaw_chebyshev_filter(opaw_ctx* ctx, int sys, int h_type, int degree, double hbot, double hmax);
/* res_i = |H phi_i - eig_i phi_i| (with dv) of the resident orbitals. */
int opaw_residuals(opaw_ctx* ctx, int sys, int h_type, const double* eig, double* res);
/* Block preconditioned eigensolver on the resident (orthonormal) orbitals;
 * no counterpart in the reference, which has no preconditioner.  Per
 * iteration: Rayleigh quotients, residuals r, w = M r in k space with
 * M = TPA(x) exp(g(log(1+x), e/t)), x = |k|^2/2 / t, t the band kinetic
 * energy, g the MLP of mlp (NULL: plain Teter-Payne-Allan), then
 * Rayleigh-Ritz on [phi, w, p] (LOBPCG).  Stops when the residuals of the
 * first nconv bands (0: all) are <= tol, or after maxit.
 * eig (nb) out; hist (maxit) the largest residual per iteration (may be
 * NULL); iters the iterations done.  record > 0 keeps (phi, r) every
 * record-th iteration as training material.  The context keeps a device
 * work space of 13 nb N complex FP64 values for it (two [phi, w, p] blocks,
 * their H images, one k-space block). */
int opaw_block_solve(opaw_ctx* ctx, int sys, int h_type, int maxit, double tol, int nconv, const double* mlp, int nh,
                     int record, double* eig, double* hist, int* iters);
/* Where opaw_block_solve solves its Rayleigh-Ritz pencils (dimension m = 2 nb
 * or 3 nb): on the device (rocSOLVER's zheevd, rocBLAS) for m >= min_m, on
 * the host below; no counterpart in the reference.  The device route keeps
 * the Gram matrices in HBM and always drops overlap directions below 1e-12
 * of the largest; the host route takes Cholesky when the scaled overlap is
 * safely positive definite.  Default 144, where the two cost the same on
 * an MI355X host; 0: always the device. */
int opaw_set_subspace_device(opaw_ctx* ctx, int min_m);
/* MLP of the preconditioner: 2 -> nh -> nh -> 1 with tanh hidden layers;
 * parameters W1 (nh x 2), b1, W2 (nh x nh), b2, w3 (nh), b3. */
int opaw_mlp_param_count(int nh);
/* Training pairs: prepare turns the recorded (phi_t, r_t) of system sys into
 * (r_t(k), e_t(k)) with e_t = phi* (phi*^H phi_t) dv - phi_t, phi* the
 * resident converged orbitals; loss = mean over pairs and bands of
 * 1 - |<e, M r>|^2 / (|e|^2 |M r|^2) and its gradient w.r.t. the MLP
 * parameters (grad may be NULL); pair downloads one prepared pair. */
int opaw_precond_prepare(opaw_ctx* ctx, int sys);
int opaw_precond_clear(opaw_ctx* ctx);
int opaw_precond_count(opaw_ctx* ctx, int* nsnap, int* nready);
int opaw_precond_loss(opaw_ctx* ctx, const double* mlp, int nh, double* loss, double* grad);
int opaw_precond_pair(opaw_ctx* ctx, int idx, void* rk, void* ek, double* t, double* y, int* nb);
/* Whether opaw_precond_prepare keeps each pair's k-space fields (default 1,
 * for opaw_precond_pair) or frees them after reducing them to the shell sums
 * the loss and gradient use (0: two band sets of device memory less per pair,
 * for training sets of many large systems). */
int opaw_precond_keep_fields(opaw_ctx* ctx, int keep);

/* ---- profiling: HIP events around the launches of selected kernel families
 * mask: bit k enables family k (OPAW_K_*); 0 disables, -1 enables all. */
int opaw_profile_enable(opaw_ctx* ctx, int mask);
int opaw_profile_reset(opaw_ctx* ctx);
int opaw_profile_read(opaw_ctx* ctx, int kclass, double* total_ms, long long* launches);

#ifdef __cplusplus
}
#endif
#endif /* OPAW_HIP_H */
