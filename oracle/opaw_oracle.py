"""FP64 CPU restatement of the OPAW-TDDFT real-time propagation hot path.

TEST INFRASTRUCTURE: this module is the oracle the HIP path is checked
against.  Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import it.  It is never on the product path.

Every routine restates one reference subroutine (file:line below, paths
relative to the reference repository root).  Arrays are flat in Fortran order
(x fastest): a grid field has length N = nx*ny*nz with flat index
ix + nx*(iy + ny*iz); a batch of orbitals is (nb, N) complex128.  Box tables
are (ms, nr3, nr2, nr1) = Fortran local_p3d(nr1,nr2,nr3,ms).

Pinning (what independently fixes each routine; the tests under tests/ hold
the checks):
  ek_prep           closed form + brute-force DFT of the Laplacian (tiny grid)
  fft3d_*           numpy.fft (pocketfft, an independent library) + brute DFT
  h_phi             dense-matrix brute force on a 6x5x4 grid; hermiticity
  proj / proj1      brute force loops; periodic wrap vs an unwrapped embedding
  sn_phi            S^-1/2 S^-1/2 = S^-1 and S^1/2 S^-1/2 = 1 on isolated atoms
  shs               hermiticity of the dense shs matrix; eigenvalues vs eigh
  rk4_prop_shs      frozen-H step == 4th-order Taylor of exp(-i dt H) + norm
  tddft_rhoij_2     brute-force loops
  tddft_dens_2      brute-force loops
  get_nhat          brute-force loops + sum(n + nhat) dv == rnel
  vh_sub (MT)       Gaussian charge vs Q erf(r/(sqrt2 s))/r (isolated system)
  vh_sub (periodic) mean zero + Poisson equation in reciprocal space
  lda_xc            Slater exchange closed form; PW92 correlation closed form
                    (libxc ids 1/12 density threshold: parity unpinned)
  get_dij           brute-force loops; clip and symmetric fill
  plot_dip          closed form on a separable density
  excite            closed form
  orthog_phi        Gram == identity afterwards; subspace preserved
  diagh             eigenvalues vs numpy eigh of the dense shs matrix
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# ---------------------------------------------------------------------------
# constants of the reference (src/0_parameter.f90)
PARAM_DIJ_MAX = 500.0          # src/0_parameter.f90 param_dij_max
RHOIJ_SELECT = 1e-10           # src/get_dij.f90:82 (|rhoij| > 1d-10 selected)
XC_DENS_THRESHOLD = 1e-15      # libxc per-point density cutoff (parity unpinned)


# ---------------------------------------------------------------------------
# data model (mirrors main_mod / atom_mod / tddft_mod state)
@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    periodic: bool

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def dv(self) -> float:
        return self.dx * self.dy * self.dz

    @property
    def scale_vh(self) -> int:
        # src/main.f90:351-352
        return 1 if self.periodic else 2


@dataclass
class AtomBox:
    """One atom: its projector box (src/3_atom.f90:17-60)."""
    itype: int
    start: np.ndarray            # 0-based ir_start (3,), may be <0 when periodic
    nr: np.ndarray               # nrough (3,)
    p: np.ndarray                # (ms, nr3, nr2, nr1) local_p3d (real at gamma)
    p1: np.ndarray               # (ms, nr3, nr2, nr1) local_p3d1
    s: np.ndarray                # (ms,) overlap eigenvalues
    sinv: np.ndarray             # (ms,) 1/(1+s)-1
    # compensation-charge sphere (src/3_atom.f90:24-26): flat grid indices
    sph_idx: Optional[np.ndarray] = None   # (ngrid,) int64
    sph_g: Optional[np.ndarray] = None     # (ngrid, nl2) local_g3d
    # libpaw's D^hat sphere (pawfgrtab: points within rshp, g_l S_lm)
    dsph_idx: Optional[np.ndarray] = None
    dsph_g: Optional[np.ndarray] = None

    @property
    def ms(self) -> int:
        return self.p.shape[0]


@dataclass
class AtomType:
    """Per-type tables (src/2_paw.f90:8-69)."""
    ms: int
    qijlm: np.ndarray            # (nl2, ms, ms)
    dij0: np.ndarray             # (ms, ms) on-site constant part
    eijkl: np.ndarray            # (ms, ms, ms, ms) on-site Hartree kernel
    onsite: Optional[dict] = None  # libpaw on-site tables; None: the model below


@dataclass
class SysState:
    """One propagated system (unperturbed or kicked)."""
    vks: np.ndarray                       # (N,)
    dij: List[np.ndarray]                 # per atom (ms, ms)
    rhoij: List[np.ndarray] = field(default_factory=list)
    dens: Optional[np.ndarray] = None
    nhat: Optional[np.ndarray] = None
    vxc: Optional[np.ndarray] = None
    vh: Optional[np.ndarray] = None


@dataclass
class OpawState:
    grid: Grid
    ek: np.ndarray                        # (N,) already divided by N
    atoms: List[AtomBox]
    types: List[AtomType]
    vloc: np.ndarray                      # (N,) vloc_tot
    ncoret: np.ndarray                    # (N,) pseudo core density
    vk: Optional[np.ndarray]              # doubled-box Hartree kernel (flat)
    rnel: float
    systems: List[SysState]
    dt: float = 0.05
    usexcnhat: int = 0                    # 0: D^hat uses vks - vxc


# ---------------------------------------------------------------------------
# grid helpers
def fidx(g: Grid, ix, iy, iz):
    return ix + g.nx * (iy + g.ny * iz)


def to3(g: Grid, f):
    """flat Fortran-order field -> (nz, ny, nx) view."""
    return f.reshape(f.shape[:-1] + (g.nz, g.ny, g.nx))


def box_flat_index(g: Grid, a: AtomBox) -> np.ndarray:
    """Flat grid index of every box point, (nr3, nr2, nr1) order, with the
    periodic wrap mod(ix+nx-1,nx)+1 of src/proj.f90:31-33 (identical to the
    contiguous slice when the box lies inside the grid)."""
    lx = (a.start[0] + np.arange(a.nr[0])) % g.nx
    ly = (a.start[1] + np.arange(a.nr[1])) % g.ny
    lz = (a.start[2] + np.arange(a.nr[2])) % g.nz
    return (lx[None, None, :] + g.nx * (ly[None, :, None] + g.ny * lz[:, None, None])).reshape(-1)


# ---------------------------------------------------------------------------
# kinetic table: src/ek.f90:20-50
def ek_prep(g: Grid, ekcut: float) -> np.ndarray:
    pi = math.acos(-1.0)
    dkx = 2.0 * pi / (g.nx * g.dx)
    dky = 2.0 * pi / (g.ny * g.dy)
    dkz = 2.0 * pi / (g.nz * g.dz)

    def fold(n, dk, d):
        k = np.arange(n) * dk
        return np.where(k > pi / d, k - 2.0 * pi / d, k)

    kx = fold(g.nx, dkx, g.dx)
    ky = fold(g.ny, dky, g.dy)
    kz = fold(g.nz, dkz, g.dz)
    ek = (kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2) / 2.0
    ek = np.where(ek > ekcut, ekcut, ek)
    return (ek / (float(g.nx) * float(g.ny) * float(g.nz))).reshape(-1)


def ekcut_default(g: Grid) -> float:
    """Largest kinetic energy inside the inscribed sphere of the FFT box,
    0.5*(pi/dmax)^2*(1-1e-10): the ecut that src/libpaw/getcut.f90 returns
    for boxcut=1 on an orthorhombic grid."""
    dmax = max(g.dx, g.dy, g.dz)
    return 0.5 * (math.pi / dmax) ** 2 * (1.0 - 1e-10)


# ---------------------------------------------------------------------------
# Threads of the 3D transforms.  1 (the default, used by every test): numpy's
# pocketfft.  More: scipy.fft's threaded pocketfft, which only the timed CPU
# baseline asks for (the reference runs one FFTW transform per MPI rank).
FFT_WORKERS = 1


def _fftn(a, axes=None):
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.fftn(a, axes=axes, workers=FFT_WORKERS)
    return np.fft.fftn(a, axes=axes)


def _ifftn(a, axes=None):
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.ifftn(a, axes=axes, workers=FFT_WORKERS)
    return np.fft.ifftn(a, axes=axes)


# FFTW semantics: src/lib/library_fftw_v3.f90:56-83 (unnormalised, forward e^-i)
def fft3d_forward(g: Grid, f):
    return _fftn(to3(g, f), axes=(-3, -2, -1)).reshape(f.shape)


def fft3d_backward(g: Grid, f):
    return (_ifftn(to3(g, f), axes=(-3, -2, -1)) * g.n).reshape(f.shape)


# ---------------------------------------------------------------------------
# projections: src/proj.f90:1-44 and src/proj1.f90:1-46
def proj(st: OpawState, ia: int, pin: np.ndarray, which: str = "p") -> np.ndarray:
    """ca(is) = sum_box pin * conj(p_is) * dv.  pin may be (N,) or (nb, N)."""
    g = st.grid
    a = st.atoms[ia]
    tab = a.p if which == "p" else a.p1
    idx = box_flat_index(g, a)
    loc = pin[..., idx]
    return loc @ tab.reshape(a.ms, -1).T.conj() * g.dv


def proj1(st, ia, pin):
    return proj(st, ia, pin, "p1")


# ---------------------------------------------------------------------------
# H.psi: src/hphi.f90:1-66 (h_phi + addvnl); sys selects vks/dij
# (sys=1 is h_phi_pert, src/hphi_pert.f90:1-62)
def h_phi(st: OpawState, pin: np.ndarray, sys: int = 0) -> np.ndarray:
    g = st.grid
    S = st.systems[sys]
    cout = fft3d_forward(g, pin) * st.ek
    hp = fft3d_backward(g, cout) + pin * S.vks
    for ia, a in enumerate(st.atoms):
        ca = proj(st, ia, pin)                       # (..., ms)
        arr = ca @ S.dij[ia].T                       # sum_j dij(is,j) ca(j)
        idx = box_flat_index(g, a)
        add = arr @ a.p.reshape(a.ms, -1)            # (..., box)
        np.add.at(hp, (Ellipsis, idx), add) if hp.ndim == 1 else _add_rows(hp, idx, add)
    return hp


def _add_rows(hp, idx, add):
    for b in range(hp.shape[0]):
        np.add.at(hp[b], idx, add[b])


def _box_add(hp, idx, add):
    if hp.ndim == 1:
        np.add.at(hp, idx, add)
    else:
        _add_rows(hp, idx, add)


# ---------------------------------------------------------------------------
# S^n psi: src/snphi.f90:1-49 (per-atom sum over the SAME input)
def sn_phi(st: OpawState, pin: np.ndarray, n: float) -> np.ndarray:
    g = st.grid
    sp = pin.copy()
    for ia, a in enumerate(st.atoms):
        ca = proj1(st, ia, pin)
        if n == -1:
            sn = a.sinv
        else:
            sn = (1.0 + a.s) ** n - 1.0
        idx = box_flat_index(g, a)
        add = (ca * sn) @ a.p1.reshape(a.ms, -1)
        _box_add(sp, idx, add)
    return sp


# src/shs.f90:3-23 (shs) and :50-70 (shs_pert)
def shs(st: OpawState, pin: np.ndarray, sys: int = 0) -> np.ndarray:
    tmp = sn_phi(st, pin, -0.5)
    pout = h_phi(st, tmp, sys)
    return sn_phi(st, pout, -0.5)


# ---------------------------------------------------------------------------
# density / rhoij: src/tddft_dens.f90:44-85, src/tddft_rhoij.f90:75-115
def tddft_dens_2(st: OpawState, phi: np.ndarray) -> np.ndarray:
    return 2.0 * np.sum(np.abs(phi) ** 2, axis=0)


def tddft_rhoij_2(st: OpawState, phi: np.ndarray) -> List[np.ndarray]:
    out = []
    for ia in range(len(st.atoms)):
        ca = proj(st, ia, phi)                        # (nb, ms)
        out.append(2.0 * np.real(ca.conj().T @ ca))   # sum_b conj(c_i) c_j * 2
    return out


# src/nhat.f90:1-43 (get_nhat) / src/nhat_pert.f90:1-43
def get_nhat(st: OpawState, rhoij: List[np.ndarray], dens: np.ndarray, renorm: bool = True):
    g = st.grid
    nhat = np.zeros(g.n)
    for ia, a in enumerate(st.atoms):
        t = st.types[a.itype]
        # sum_ij rhoij(i,j) * sum_L qijlm(L,i,j) g3d(pt,L)
        wL = np.einsum("ij,lij->l", rhoij[ia], t.qijlm)
        np.add.at(nhat, a.sph_idx, a.sph_g @ wL)
    if renorm:
        s = np.sum(nhat) * g.dv + np.sum(dens) * g.dv
        fac = st.rnel / s
        nhat = nhat * fac
        dens = dens * fac
    return nhat, dens


# ---------------------------------------------------------------------------
# Martyna-Tuckerman kernel: src/vk_TuMa.f90:2-117 + src/vk_scale.f90:23-48
def prep_vk(nxb, nyb, nzb, dx, dy, dz, periodic) -> np.ndarray:
    from scipy.special import erf
    pi = math.acos(-1.0)
    aa = 7.0 / min(nxb * dx, nyb * dy, nzb * dz)
    r0 = np.array([-nxb / 2.0 * dx, -nyb / 2.0 * dy, -nzb / 2.0 * dz])
    dv = dx * dy * dz
    dkx = 2.0 * pi / (nxb * dx)
    dky = 2.0 * pi / (nyb * dy)
    dkz = 2.0 * pi / (nzb * dz)
    shape = (nzb, nyb, nxb)
    if not periodic:
        x = r0[0] + np.arange(nxb) * dx
        y = r0[1] + np.arange(nyb) * dy
        z = r0[2] + np.arange(nzb) * dz
        r2 = x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2
        rp = np.sqrt(np.maximum(r2, 1e-26))
        cvk = erf(aa * rp) / rp * dv
        # fft3d_general(sign=1) = conj(FFT(conj(.))): unnormalised e^{+ikr}
        cvk = np.fft.ifftn(cvk.astype(np.complex128)) * (nxb * nyb * nzb)
    else:
        cvk = np.zeros(shape, np.complex128)

    def fold(n, dk, d):
        k = np.arange(n) * dk
        return np.where(k > pi / d, k - 2.0 * pi / d, k)

    kx = fold(nxb, dkx, dx)[None, None, :]
    ky = fold(nyb, dky, dy)[None, :, None]
    kz = fold(nzb, dkz, dz)[:, None, None]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    if not periodic:
        cvk = cvk * np.exp(1j * (r0[0] * kx + r0[1] * ky + r0[2] * kz))
        k2s = k2 / (4.0 * aa ** 2)
        with np.errstate(divide="ignore", invalid="ignore"):
            big = 4.0 * pi / k2 * (1.0 - np.exp(-k2s))
        small = 4.0 * pi / (4.0 * aa ** 2) * (1.0 - 0.5 * k2s + k2s ** 2 / 6.0)
        cterm = np.where(k2s > 1e-8, big, small)
    else:
        with np.errstate(divide="ignore", invalid="ignore"):
            cterm = np.where(k2 > 1e-8, 4.0 * pi / k2, 0.0)
    cvk = cvk + cterm
    vk = np.real(cvk) / (float(nxb) * float(nyb) * float(nzb))
    return vk.reshape(-1)


def vk_prep(g: Grid) -> np.ndarray:
    s = g.scale_vh
    return prep_vk(s * g.nx, s * g.ny, s * g.nz, g.dx, g.dy, g.dz, g.periodic)


# src/vh_big.f90:1-75 + src/lib/library_fft3d_extra.f90:8-31 (rconv)
def vh_sub(st: OpawState, rho: np.ndarray) -> np.ndarray:
    g = st.grid
    s = g.scale_vh
    nxb, nyb, nzb = s * g.nx, s * g.ny, s * g.nz
    ox, oy, oz = nxb // 2 - g.nx // 2, nyb // 2 - g.ny // 2, nzb // 2 - g.nz // 2
    densb = np.zeros((nzb, nyb, nxb))
    densb[oz:oz + g.nz, oy:oy + g.ny, ox:ox + g.nx] = to3(g, rho)
    vkb = st.vk.reshape(nzb, nyb, nxb)
    # c2r(r2c(densb) * vk) unnormalised both ways
    vhb = np.real(_ifftn(_fftn(densb) * vkb)) * (nxb * nyb * nzb)
    return vhb[oz:oz + g.nz, oy:oy + g.ny, ox:ox + g.nx].reshape(-1).copy()


# ---------------------------------------------------------------------------
# LDA exchange-correlation (libxc LDA_X id 1 + LDA_C_PW id 12, unpolarised)
# as called from src/tddft_pot.f90:163-177 -> src/XCI/libxc1.f90:27-64
PW92 = dict(a=0.031091, alpha1=0.21370, beta1=7.5957, beta2=3.5876,
            beta3=1.6382, beta4=0.49294)


def lda_xc(n: np.ndarray):
    """returns (vxc, exc) per point; exc is energy per particle."""
    n = np.asarray(n, dtype=np.float64)
    ok = n > XC_DENS_THRESHOLD
    nn = np.where(ok, n, 1.0)
    pi = math.pi
    cx = (3.0 / pi) ** (1.0 / 3.0)
    n13 = np.cbrt(nn)
    ex = -0.75 * cx * n13
    vx = -cx * n13
    rs = (3.0 / (4.0 * pi * nn)) ** (1.0 / 3.0)
    p = PW92
    srs = np.sqrt(rs)
    q0 = -2.0 * p["a"] * (1.0 + p["alpha1"] * rs)
    q1 = 2.0 * p["a"] * (p["beta1"] * srs + p["beta2"] * rs + p["beta3"] * rs * srs + p["beta4"] * rs * rs)
    dq1 = p["a"] * (p["beta1"] / srs + 2.0 * p["beta2"] + 3.0 * p["beta3"] * srs + 4.0 * p["beta4"] * rs)
    lg = np.log1p(1.0 / q1)
    ec = q0 * lg
    decdrs = -2.0 * p["a"] * p["alpha1"] * lg - q0 * dq1 / (q1 * q1 + q1)
    vc = ec - rs / 3.0 * decdrs
    vxc = np.where(ok, vx + vc, 0.0)
    exc = np.where(ok, ex + ec, 0.0)
    return vxc, exc


# src/tddft_pot.f90:110-181 (get_pot_tddft_2)
def get_pot_tddft_2(st: OpawState, dens: np.ndarray, nhat: np.ndarray):
    vxc, _ = lda_xc(dens + st.ncoret)
    vh = vh_sub(st, dens + nhat)
    if st.grid.periodic:
        vh = vh - np.sum(vh) / vh.size
    vks = vh + vxc + st.vloc
    return vks, vxc, vh


# ---------------------------------------------------------------------------
# D_ij: src/get_dij.f90:1-156.  A type with libpaw tables runs pawdenpot +
# pawdij (libpaw_oracle.onsite_dij: dij0 + dijhartree + dijxc + dijhat over
# libpaw's own D^hat sphere).  A synthetic type keeps the model
#   D = dij0 + sum_kl eijkl(i,j,k,l) rhoij_sel(k,l) + D^hat
# with D^hat over the compensation sphere.
def get_dij(st: OpawState, rhoij: List[np.ndarray], vks: np.ndarray, vxc: np.ndarray):
    g = st.grid
    pot = vks - vxc if st.usexcnhat == 0 else vks
    out = []
    for ia, a in enumerate(st.atoms):
        t = st.types[a.itype]
        if t.onsite is not None:
            from . import libpaw_oracle as LO
            d = LO.onsite_dij(t.onsite, rhoij[ia], pot, a.dsph_idx, a.dsph_g, g.dv)
        else:
            rho = np.where(np.abs(rhoij[ia]) > RHOIJ_SELECT, rhoij[ia], 0.0)
            d = t.dij0 + np.einsum("ijkl,kl->ij", t.eijkl, rho)
            prod = (pot[a.sph_idx] @ a.sph_g) * g.dv          # (nl2,)
            d = d + np.einsum("lij,l->ij", t.qijlm, prod)
        # transfer_dij: packed (row<=col) element, clip, symmetric fill
        up = np.triu(d)
        full = up + np.triu(d, 1).T
        full = np.where(np.abs(full) > PARAM_DIJ_MAX, np.sign(full) * PARAM_DIJ_MAX, full)
        out.append(full)
    return out


# ---------------------------------------------------------------------------
# per-stage rebuild: src/ncpaw_make_hamiltonian.f90:213-246 (unpert) and
# :248-281 (pert); updates st.systems[sys] in place.
def rebuild(st: OpawState, phi: np.ndarray, sys: int):
    S = st.systems[sys]
    psit = sn_phi(st, phi, -0.5)                 # psi~ = S^-1/2 phi_bar
    rhoij = tddft_rhoij_2(st, psit)
    dens = tddft_dens_2(st, psit)
    nhat, dens = get_nhat(st, rhoij, dens)
    vks, vxc, vh = get_pot_tddft_2(st, dens, nhat)
    S.dij = get_dij(st, rhoij, vks, vxc)
    S.vks, S.vxc, S.vh = vks, vxc, vh
    S.rhoij, S.dens, S.nhat = rhoij, dens, nhat


# ---------------------------------------------------------------------------
# RK4 step: src/rk4.f90:1-103 (rk4_prop_shs); one system at a time.
def rk4_prop_shs(st: OpawState, p: np.ndarray, sys: int, selfconsistent: bool = True) -> np.ndarray:
    g = st.grid
    dt = st.dt
    p2 = p.copy()
    nrm0 = np.sqrt(g.dv * np.sum(np.abs(p2) ** 2, axis=1))

    def stage(pin, h, last):
        k = -1j * shs(st, pin, sys)
        y = None
        if not last:
            y = p2 + h * k
            if selfconsistent:
                rebuild(st, y, sys)
        return k, y

    k1, y = stage(p2, dt / 2.0, False)
    k2, y = stage(y, dt / 2.0, False)
    k3, y = stage(y, dt, False)
    k4, _ = stage(y, dt, True)
    pn = p2 + dt / 6.0 * (k1 + 2.0 * (k2 + k3) + k4)
    nrm1 = np.sqrt(g.dv * np.sum(np.abs(pn) ** 2, axis=1))
    return pn * (nrm0 / nrm1)[:, None]


# ---------------------------------------------------------------------------
# dipole: src/tddft_plot.f90:1-155 (h_type=1: rho = 2 sum |phi_bar|^2)
def position_axes(g: Grid):
    rx = (np.arange(g.nx) - g.nx // 2) * g.dx
    ry = (np.arange(g.ny) - g.ny // 2) * g.dy
    rz = (np.arange(g.nz) - g.nz // 2) * g.dz
    return rx, ry, rz


def dipole(st: OpawState, phi: np.ndarray) -> np.ndarray:
    g = st.grid
    rho = to3(g, 2.0 * np.sum(np.abs(phi) ** 2, axis=0))
    rx, ry, rz = position_axes(g)
    return np.array([
        g.dv * np.sum(rho * rx[None, None, :]),
        g.dv * np.sum(rho * ry[None, :, None]),
        g.dv * np.sum(rho * rz[:, None, None]),
    ])


def plot_dip(st: OpawState, phi: np.ndarray, phi_pert: np.ndarray, sm: float):
    dip = dipole(st, phi)
    dip_pert = dipole(st, phi_pert)
    return (dip_pert - dip) / sm, dip, dip_pert


# src/tddft_excite.f90:1-16
def excite_psi_bar_pert(st: OpawState, phi: np.ndarray, sm: float, ipol: int) -> np.ndarray:
    g = st.grid
    rr = position_axes(g)[ipol]
    shape = [1, 1, 1]
    shape[2 - ipol] = -1
    ph = np.exp(-1j * sm * rr).reshape(shape)
    return (to3(g, phi) * ph).reshape(phi.shape)


# ---------------------------------------------------------------------------
# orthogonalisation: src/orthog_phi.f90:1-194 (canonical: phi U lambda^-1/2)
def orthog_phi(st: OpawState, phi: np.ndarray, h_type: int = 1):
    g = st.grid
    sp = phi if h_type == 1 else sn_phi(st, phi, 1.0)
    hm = (sp.conj() @ phi.T) * g.dv                # hm(i,j) = sum conj(sp_i) phi_j dv
    eig, vec = np.linalg.eigh(hm)
    out = (vec.T @ phi) / np.sqrt(eig)[:, None]
    return out, eig, hm


# src/diag.f90:1-180 (diagh, Rayleigh-Ritz on shs)
def diagh(st: OpawState, phi: np.ndarray, h_type: int = 1, sys: int = 0):
    g = st.grid
    hp = shs(st, phi, sys) if h_type == 1 else h_phi(st, phi, sys)
    hm = (hp.conj() @ phi.T) * g.dv                # hm(i,j) = sum conj(hp_i) phi_j dv
    eig, vec = np.linalg.eigh(hm)
    out = vec.T @ phi
    return out, eig, hm


# ---------------------------------------------------------------------------
# one full time step of the loop at src/tddft.f90:67-91 (dyn=3), for the
# unperturbed (0) and kicked (1) systems.
def tddft_step(st: OpawState, phis: List[np.ndarray], sm: float):
    for s, phi in enumerate(phis):
        rebuild(st, phi, s)                         # ncpaw_make_hamiltonian_tddft
    new = [rk4_prop_shs(st, phi, s) for s, phi in enumerate(phis)]
    d_dip, dip, dip_pert = plot_dip(st, new[0], new[1], sm)
    return new, d_dip, dip, dip_pert
