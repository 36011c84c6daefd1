"""Host mirror of the reference's hot-path interface over the HIP C ABI.

``OpawGPU`` owns one device context (grid, tables, potentials, resident
orbitals of ``nsys`` systems) and exposes the reference's subroutine names:

=========================  ============================================
reference (src/...)        here
=========================  ============================================
h_phi / h_phi_pert         h_phi(pin, sys=0|1)
sn_phi                     sn_phi(pin, n)
shs / shs_pert             shs(pin, sys=0|1)
proj / proj1               proj(pin, which_p1=False)  (all atoms)
tddft_rhoij_2(pert)        tddft_rhoij_2(psi, sys)
tddft_dens_2(pert)         tddft_dens_2(psi, sys)
get_nhat / get_nhat_pert   get_nhat(sys)
get_pot_tddft_2(pert)      get_pot_tddft_2(sys)
get_dij / get_dij_pert     get_dij(sys)
ncpaw_make_hamiltonian_*   ncpaw_make_hamiltonian_tddft(sys=-1|0|1)
rk4_prop_shs               rk4_prop_shs(dt)
plot_dip                   plot_dip(sm)
excite_psi_bar_pert        excite_psi_bar_pert(sm, ipol)
opaw_tddft (time loop)     opaw_tddft(nt, dt, sm)
orthog_phi / diagh         orthog_phi(sys, h_type) / diagh(sys, h_type)
=========================  ============================================

Device buffers are torch CUDA tensors (plumbing only): complex128 in the
FP64 context, complex64 in the mixed context, shape (nbatch, nx*ny*nz) with
the grid flattened in Fortran order.  Every call goes to the HIP library;
there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import warnings
from typing import Optional

import numpy as np

from . import lib
from .synth import SystemSpec


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OpawGPU:
    def __init__(self, spec: SystemSpec, precision: str = "f64", nsys: int = 2, nb: Optional[int] = None,
                 stream=None, usexcnhat: int = 0):
        import torch
        self.torch = torch
        self.spec = spec
        standins = tuple((getattr(spec, "extra", None) or {}).get("standins", ()))
        if standins:
            # a dataset-built system whose tables include stand-ins for the
            # reference's (v_loc on periodic grids): runs are not the reference's
            warnings.warn(f"{getattr(spec, 'name', 'system')}: stand-in tables {standins} in place of the "
                          "reference's", stacklevel=2)
        self.nsys = nsys
        self.nb = nb if nb is not None else spec.nocc
        self.N = spec.n
        self.prec = lib.OPAW_PREC_F64 if precision == "f64" else lib.OPAW_PREC_MIXED
        self.cdtype = torch.complex128 if self.prec == lib.OPAW_PREC_F64 else torch.complex64
        self.np_cdtype = np.complex128 if self.prec == lib.OPAW_PREC_F64 else np.complex64
        self.L = lib.load()
        ctx = C.c_void_p()
        lib.call("opaw_create", C.byref(ctx), spec.nx, spec.ny, spec.nz, spec.dx, spec.dy, spec.dz,
                 1 if spec.periodic else 0, self.prec, nsys, self.nb)
        self.ctx = ctx
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream().cuda_stream
        if stream is not None:
            lib.call("opaw_set_stream", self.ctx, C.c_void_p(stream))
        self._upload_tables(usexcnhat)

    # ------------------------------------------------------------------ setup
    def _upload_tables(self, usexcnhat):
        s = self.spec
        lib.call("opaw_ek_prep", self.ctx, s.ekcut)
        natom = len(s.atoms)
        types = s.types
        atom_map = _i32([a.itype + 1 for a in s.atoms])
        ir_start = _i32([[int(a.start[0]) + 1, int(a.start[1]) + 1, int(a.start[2]) + 1] for a in s.atoms]).reshape(-1)
        nr_type = []
        for t in range(len(types)):
            nrs = [a.nr for a in s.atoms if a.itype == t]
            nr_type.append([int(x) for x in nrs[0]] if nrs else [1, 1, 1])
        nrough = _i32(nr_type).reshape(-1)
        mstates = _i32([t.ms for t in types])
        lib.call("opaw_set_atoms", self.ctx, natom, _ip(atom_map), _ip(ir_start), len(types), _ip(nrough),
                 _ip(mstates), s.nl2)
        self._keep = []
        for ia, a in enumerate(s.atoms, start=1):
            p, p1, ss, si = _f64(a.p), _f64(a.p1), _f64(a.s), _f64(a.sinv)
            lib.call("opaw_set_projectors", self.ctx, ia, _dp(p), _dp(p1), _dp(ss), _dp(si))
            idx = np.asarray(a.sph_idx, dtype=np.int64)
            ix = idx % s.nx
            iy = (idx // s.nx) % s.ny
            iz = idx // (s.nx * s.ny)
            grid = _i32(np.stack([ix + 1, iy + 1, iz + 1]))           # local_grid(ngrid,3)
            g3d = _f64(np.asarray(a.sph_g).T)                         # local_g3d(ngrid,nl2)
            lib.call("opaw_set_sphere", self.ctx, ia, int(idx.size), _ip(grid), _dp(g3d))
        for it, t in enumerate(types, start=1):
            q = _f64(np.transpose(t.qijlm, (2, 1, 0)))                # qijlm(nl2,ms,ms)
            d0 = _f64(t.dij0.T)
            e = _f64(np.transpose(t.eijkl, (3, 2, 1, 0)))             # eijkl(ms,ms,ms,ms)
            lib.call("opaw_set_type", self.ctx, it, _dp(q), _dp(d0), _dp(e))
        for it, t in enumerate(types, start=1):
            o = getattr(t, "onsite", None)
            if o is None:
                continue
            arrs = dict(rad=_f64(o["rad"]), w=_f64(o["w"]), pp=_f64(np.asarray(o["phiphj"]).T),
                        tpp=_f64(np.asarray(o["tphitphj"]).T), core=_f64(o["coredens"]),
                        tcore=_f64(o["tcoredens"]), gnt=_f64(np.transpose(o["gnt"], (0, 2, 1))),
                        kln=_i32(np.asarray(o["kln"]).T), qd=_f64(np.transpose(o["qd"], (2, 1, 0))))
            self._keep.append(arrs)
            lib.call("opaw_set_onsite", self.ctx, it, int(o["nr"]), int(o["lm"]), int(o["nij"]), int(o["usetcore"]),
                     _dp(arrs["rad"]), _dp(arrs["w"]), _dp(arrs["pp"]), _dp(arrs["tpp"]), _dp(arrs["core"]),
                     _dp(arrs["tcore"]), _dp(arrs["gnt"]), _ip(arrs["kln"]), _dp(arrs["qd"]))
        for ia, a in enumerate(s.atoms, start=1):
            if getattr(a, "dsph_idx", None) is None:
                continue
            idx = _i32(np.asarray(a.dsph_idx) + 1)
            g = _f64(np.asarray(a.dsph_g).T)                          # gylm(nfgd, lm)
            lib.call("opaw_set_dij_sphere", self.ctx, ia, int(idx.size), _ip(idx), _dp(g))
        lib.call("opaw_finalize_tables", self.ctx)
        lib.call("opaw_set_potentials", self.ctx, _dp(_f64(s.vloc)), _dp(_f64(s.ncoret)), float(s.rnel),
                 int(usexcnhat))
        lib.call("opaw_vk_prep", self.ctx)
        self.ms = [a.ms for a in s.atoms]
        self.coff = np.concatenate([[0], np.cumsum(self.ms)])[:-1].astype(int)
        self.ctot = int(sum(self.ms))
        if s.vks0 is not None:
            for sys in range(self.nsys):
                self.write_vks(sys, s.vks0)
        if s.dij0_atoms is not None:
            for sys in range(self.nsys):
                for ia, d in enumerate(s.dij0_atoms, start=1):
                    self.write_dij(sys, ia, d)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.opaw_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- buffers
    def empty(self, nbatch: int):
        return self.torch.empty((nbatch, self.N), dtype=self.cdtype, device="cuda")

    def to_device(self, a: np.ndarray):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=self.np_cdtype)).to("cuda")

    def synchronize(self):
        lib.call("opaw_synchronize", self.ctx)

    @staticmethod
    def _ptr(t):
        return C.c_void_p(t.data_ptr())

    # ------------------------------------------------------------- state io
    def write_orbitals(self, sys: int, phi: np.ndarray):
        a = np.ascontiguousarray(phi, dtype=self.np_cdtype)
        assert a.shape == (self.nb, self.N)
        lib.call("opaw_write_orbitals", self.ctx, sys, a.ctypes.data_as(C.c_void_p))

    def read_orbitals(self, sys: int) -> np.ndarray:
        out = np.empty((self.nb, self.N), dtype=self.np_cdtype)
        lib.call("opaw_read_orbitals", self.ctx, sys, out.ctypes.data_as(C.c_void_p))
        return out

    def write_vks(self, sys: int, vks: np.ndarray):
        lib.call("opaw_write_vks", self.ctx, sys, _dp(_f64(vks)))

    def write_dij(self, sys: int, ia: int, dij: np.ndarray):
        lib.call("opaw_write_dij", self.ctx, sys, ia, _dp(_f64(np.asarray(dij).T)))

    def read_dij(self, sys: int, ia: int) -> np.ndarray:
        ms = self.ms[ia - 1]
        out = np.empty(ms * ms)
        lib.call("opaw_read_dij", self.ctx, sys, ia, _dp(out))
        return out.reshape(ms, ms).T.copy()

    def read_rhoij(self, sys: int, ia: int) -> np.ndarray:
        ms = self.ms[ia - 1]
        out = np.empty(ms * ms)
        lib.call("opaw_read_rhoij", self.ctx, sys, ia, _dp(out))
        return out.reshape(ms, ms).T.copy()

    def read_field(self, name: str, sys: int) -> np.ndarray:
        out = np.empty(self.N)
        lib.call("opaw_read_field", self.ctx, lib.FIELD[name], sys, _dp(out))
        return out

    # ---------------------------------------------------------- operators
    def h_phi(self, pin, sys: int = 0, out=None):
        out = self.empty(pin.shape[0]) if out is None else out
        lib.call("opaw_h_phi", self.ctx, sys, self._ptr(pin), self._ptr(out), pin.shape[0])
        return out

    def h_phi_pert(self, pin, out=None):
        return self.h_phi(pin, 1, out)

    def sn_phi(self, pin, n: float, out=None):
        out = self.empty(pin.shape[0]) if out is None else out
        lib.call("opaw_sn_phi", self.ctx, self._ptr(pin), self._ptr(out), pin.shape[0], float(n))
        return out

    def shs(self, pin, sys: int = 0, out=None):
        out = self.empty(pin.shape[0]) if out is None else out
        lib.call("opaw_shs", self.ctx, sys, self._ptr(pin), self._ptr(out), pin.shape[0])
        return out

    def shs_pert(self, pin, out=None):
        return self.shs(pin, 1, out)

    def proj(self, pin, which_p1: bool = False):
        ca = self.torch.empty((pin.shape[0], self.ctot), dtype=self.torch.complex128, device="cuda")
        lib.call("opaw_proj", self.ctx, self._ptr(pin), pin.shape[0], 1 if which_p1 else 0, self._ptr(ca))
        return ca

    def proj1(self, pin):
        return self.proj(pin, True)

    # ---------------------------------- host-array operators (Fortran drop-ins)
    def _host_fields(self, pin) -> np.ndarray:
        return np.ascontiguousarray(pin, dtype=np.complex128).reshape(-1, self.spec.n)

    def h_phi_host(self, pin, sys: int = 0) -> np.ndarray:
        """h_phi / h_phi_pert on complex*16 host arrays (nbatch, N)."""
        x = self._host_fields(pin)
        out = np.empty_like(x)
        lib.call("opaw_h_phi_host", self.ctx, sys, _dp(x.view(np.float64)), _dp(out.view(np.float64)), x.shape[0])
        return out

    def shs_host(self, pin, sys: int = 0) -> np.ndarray:
        x = self._host_fields(pin)
        out = np.empty_like(x)
        lib.call("opaw_shs_host", self.ctx, sys, _dp(x.view(np.float64)), _dp(out.view(np.float64)), x.shape[0])
        return out

    def sn_phi_host(self, pin, n: float) -> np.ndarray:
        x = self._host_fields(pin)
        out = np.empty_like(x)
        lib.call("opaw_sn_phi_host", self.ctx, _dp(x.view(np.float64)), _dp(out.view(np.float64)), x.shape[0],
                 float(n))
        return out

    def proj_host(self, ia: int, pin, ms: int, which_p1: bool = False) -> np.ndarray:
        """proj(ia, pin, ca, ms, ik) with 1-based ia, one field."""
        x = self._host_fields(pin)
        ca = np.zeros(ms, dtype=np.complex128)
        lib.call("opaw_proj_host", self.ctx, int(ia), _dp(x.view(np.float64)), 1 if which_p1 else 0,
                 _dp(ca.view(np.float64)), int(ms))
        return ca

    def write_orbitals_c16(self, sys: int, phi):
        x = self._host_fields(phi)
        lib.call("opaw_write_orbitals_c16", self.ctx, sys, _dp(x.view(np.float64)))

    def read_orbitals_c16(self, sys: int) -> np.ndarray:
        out = np.empty((self.nb, self.spec.n), dtype=np.complex128)
        lib.call("opaw_read_orbitals_c16", self.ctx, sys, _dp(out.view(np.float64)))
        return out

    def set_dt(self, dt: float):
        lib.call("opaw_set_dt", self.ctx, float(dt))

    def rk4_prop_shs_host(self, p: np.ndarray, pp: np.ndarray):
        """rk4_prop_shs(p, pp, n, nb): p, pp complex128 (nb, N), updated in place."""
        for a in (p, pp):
            if a.dtype != np.complex128 or not a.flags.c_contiguous:
                raise ValueError("p / pp must be C-contiguous complex128 (nb, N)")
        lib.call("opaw_rk4_prop_shs_host", self.ctx, _dp(p.view(np.float64)), _dp(pp.view(np.float64)),
                 int(self.spec.n), int(p.shape[0]))

    def fft3d(self, pin, direction: int, out=None):
        out = self.empty(pin.shape[0]) if out is None else out
        lib.call("opaw_fft3d", self.ctx, int(direction), self._ptr(pin), self._ptr(out), pin.shape[0])
        return out

    # ------------------------------------------------------ rebuild pieces
    def tddft_rhoij_2(self, psi, sys: int):
        lib.call("opaw_tddft_rhoij_2", self.ctx, sys, self._ptr(psi))

    def tddft_dens_2(self, psi, sys: int):
        lib.call("opaw_tddft_dens_2", self.ctx, sys, self._ptr(psi))

    def get_nhat(self, sys: int):
        lib.call("opaw_get_nhat", self.ctx, sys)

    def get_pot_tddft_2(self, sys: int):
        lib.call("opaw_get_pot_tddft_2", self.ctx, sys)

    def get_dij(self, sys: int):
        lib.call("opaw_get_dij", self.ctx, sys)

    def ncpaw_make_hamiltonian_tddft(self, sys: int = -1):
        lib.call("opaw_make_hamiltonian_tddft", self.ctx, sys)

    # ----------------------------------------------------------- propagation
    def excite_psi_bar_pert(self, sm: float, ipol: int, from_sys: int = 0, to_sys: int = 1):
        lib.call("opaw_excite_psi_bar_pert", self.ctx, from_sys, to_sys, float(sm), int(ipol))

    def rk4_prop_shs(self, dt: float):
        lib.call("opaw_rk4_prop_shs", self.ctx, float(dt))

    def plot_dip(self, sm: float):
        d = np.zeros(3)
        a = np.zeros(3)
        b = np.zeros(3)
        lib.call("opaw_plot_dip", self.ctx, float(sm), _dp(d), _dp(a), _dp(b))
        return d, a, b

    def opaw_tddft(self, nt: int, dt: float, sm: float):
        d = np.zeros((nt, 3))
        dip = np.zeros((nt, 3 * self.nsys))
        lib.call("opaw_tddft_run", self.ctx, int(nt), float(dt), float(sm), _dp(d), _dp(dip))
        return d, dip

    def orthog_phi(self, sys: int = 0, h_type: int = 1):
        eig = np.zeros(self.nb)
        lib.call("opaw_orthog_phi", self.ctx, sys, h_type, _dp(eig))
        return eig

    def diagh(self, sys: int = 0, h_type: int = 1):
        eig = np.zeros(self.nb)
        lib.call("opaw_diagh", self.ctx, sys, h_type, _dp(eig))
        return eig

    # ---------------------------------------------------------- ground state
    def set_nocc(self, nocc: int):
        lib.call("opaw_set_nocc", self.ctx, int(nocc))

    def set_diis(self, ndi: int = 4, mix: float = 0.15):
        lib.call("opaw_set_diis", self.ctx, int(ndi), float(mix))

    def set_subspace_device(self, min_m: int = 144):
        """block_solve's Rayleigh-Ritz pencils of dimension >= min_m run on
        the device (rocSOLVER), smaller ones on the host; 0: always the
        device."""
        lib.call("opaw_set_subspace_device", self.ctx, int(min_m))

    def ncpaw_make_hamiltonian(self, sys: int = 0, diis: bool = True):
        lib.call("opaw_ncpaw_make_hamiltonian", self.ctx, sys, int(diis))

    def ncpaw_make_hamiltonian_density(self, dens: np.ndarray, sys: int = 0, diis: bool = True):
        """SCF iteration 1: the Hamiltonian of a starting density, rho_ij = 0."""
        lib.call("opaw_ncpaw_make_hamiltonian_density", self.ctx, sys, _dp(_f64(dens)), int(diis))

    def hminmax(self, sys: int = 0, h_type: int = 1, niter: int = 25, fctr: float = 1.2, seed: int = 12345):
        lo, hi = C.c_double(), C.c_double()
        lib.call("opaw_hminmax", self.ctx, sys, h_type, int(niter), float(fctr), C.c_ulonglong(seed),
                 C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def chebyshev_filter(self, hbot: float, hmax: float, degree: int = 10, sys: int = 0, h_type: int = 1):
        lib.call("opaw_chebyshev_filter", self.ctx, sys, h_type, int(degree), float(hbot), float(hmax))

    def residuals(self, eig, sys: int = 0, h_type: int = 1):
        e = np.ascontiguousarray(eig, dtype=np.float64)
        r = np.zeros(self.nb)
        lib.call("opaw_residuals", self.ctx, sys, h_type, _dp(e), _dp(r))
        return r

    def block_solve(self, maxit: int, tol: float, mlp=None, nh: int = 16, record: int = 0, sys: int = 0,
                    h_type: int = 1, nconv: int = 0):
        eig = np.zeros(self.nb)
        hist = np.zeros(maxit)
        its = C.c_int()
        p = None if mlp is None else np.ascontiguousarray(mlp, dtype=np.float64)
        lib.call("opaw_block_solve", self.ctx, sys, h_type, int(maxit), float(tol), int(nconv),
                 None if p is None else _dp(p), int(nh), int(record), _dp(eig), _dp(hist), C.byref(its))
        return eig, hist[:max(its.value + (its.value < maxit), 1)], its.value

    def precond_prepare(self, sys: int = 0):
        lib.call("opaw_precond_prepare", self.ctx, sys)

    def precond_keep_fields(self, keep: bool):
        lib.call("opaw_precond_keep_fields", self.ctx, 1 if keep else 0)

    def precond_clear(self):
        lib.call("opaw_precond_clear", self.ctx)

    def precond_count(self):
        a, b = C.c_int(), C.c_int()
        lib.call("opaw_precond_count", self.ctx, C.byref(a), C.byref(b))
        return a.value, b.value

    def precond_loss(self, mlp, nh: int = 16, grad: bool = True):
        p = np.ascontiguousarray(mlp, dtype=np.float64)
        loss = C.c_double()
        g = np.zeros_like(p) if grad else None
        lib.call("opaw_precond_loss", self.ctx, _dp(p), int(nh), C.byref(loss), None if g is None else _dp(g))
        return loss.value, g

    def precond_pair(self, idx: int):
        nb = C.c_int()
        lib.call("opaw_precond_pair", self.ctx, int(idx), None, None, None, None, C.byref(nb))
        rk = np.zeros((nb.value, self.N), np.complex128)
        ek = np.zeros_like(rk)
        t = np.zeros(nb.value)
        y = np.zeros(nb.value)
        lib.call("opaw_precond_pair", self.ctx, int(idx), rk.ctypes.data_as(C.c_void_p),
                 ek.ctypes.data_as(C.c_void_p), _dp(t), _dp(y), C.byref(nb))
        return rk, ek, t, y

    # -------------------------------------------------------------- restart
    def write_to_restart(self, directory: str, it: int):
        """write_to_restart (src/tddft_write_to_restart.f90:1-67)."""
        lib.call("opaw_write_restart", self.ctx, str(directory).encode(), int(it))

    def read_phi_bar(self, directory: str) -> int:
        """read_phi_bar (src/read_phi_bar.f90:1-51); returns it_restart."""
        it = C.c_int()
        lib.call("opaw_read_restart", self.ctx, str(directory).encode(), C.byref(it))
        return it.value

    # -------------------------------------------------------------- profile
    def profile(self, classes=None):
        """classes: None / True = every kernel family, False / [] = off, or an
        iterable of family names (lib.KCLASS keys)."""
        if classes is None or classes is True:
            mask = -1
        elif classes is False:
            mask = 0
        else:
            mask = 0
            for name in classes:
                mask |= 1 << lib.KCLASS[name]
        lib.call("opaw_profile_enable", self.ctx, int(mask))

    def profile_reset(self):
        lib.call("opaw_profile_reset", self.ctx)

    def profile_read(self, name: str):
        ms = C.c_double()
        n = C.c_longlong()
        lib.call("opaw_profile_read", self.ctx, lib.KCLASS[name], C.byref(ms), C.byref(n))
        return ms.value, n.value
