"""Seeded synthetic OPAW systems for parity tests and benchmarks.

A ``SystemSpec`` carries every table the propagation hot path reads, in the
reference's memory layouts (Fortran order, x fastest):

* grid, kinetic cutoff, time step, kick;
* per atom: projector box origin (0-based ir_start) and size (nrough), the
  projector table p (local_p3d), the grid-orthonormal projectors p1
  (local_p3d1) with their overlap eigenvalues s / sinv (src/transform_p3d.f90),
  the compensation-charge sphere points (local_grid) and shape table
  (local_g3d);
* per type: qijlm, and the on-site D_ij model (dij0 + eijkl . rhoij);
* vloc_tot, ncoret, rnel;
* the H.psi microbenchmark also carries a frozen V_KS and frozen D_ij.

The recipes follow the synthetic workload definitions: the naphthalene-shaped
64^3 Martyna-Tuckerman system and the 96^3 periodic H.psi system (40 atoms,
8 projectors each, r^l exp(-r^2/0.8) Y_lm radial form cut at 3 bohr).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class AtomSpec:
    itype: int
    coord: np.ndarray
    start: np.ndarray          # (3,) int, 0-based
    nr: np.ndarray             # (3,) int
    p: np.ndarray              # (ms, nr3, nr2, nr1)
    p1: np.ndarray             # (ms, nr3, nr2, nr1)
    s: np.ndarray              # (ms,)
    sinv: np.ndarray           # (ms,)
    sph_idx: np.ndarray        # (ngrid,) int64 flat grid index
    sph_g: np.ndarray          # (ngrid, nl2)
    # libpaw's D^hat sphere (pawfgrtab): flat indices and g_l S_lm (nfgd, lm)
    dsph_idx: Optional[np.ndarray] = None
    dsph_g: Optional[np.ndarray] = None

    @property
    def ms(self):
        return self.p.shape[0]


@dataclass
class TypeSpec:
    ms: int
    qijlm: np.ndarray          # (nl2, ms, ms)
    dij0: np.ndarray           # (ms, ms)
    eijkl: np.ndarray          # (ms, ms, ms, ms)
    # libpaw on-site tables (opaw_amd.libpaw.onsite_spec); None: model on-site D_ij
    onsite: Optional[dict] = None


@dataclass
class SystemSpec:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    periodic: bool
    ekcut: float
    atoms: List[AtomSpec]
    types: List[TypeSpec]
    vloc: np.ndarray
    ncoret: np.ndarray
    rnel: float
    nocc: int
    dt: float = 0.05
    sm: float = 1e-3
    ipol: int = 0
    vks0: Optional[np.ndarray] = None           # frozen V_KS (H.psi benchmark)
    dij0_atoms: Optional[List[np.ndarray]] = None  # frozen D_ij per atom
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    @property
    def dv(self):
        return self.dx * self.dy * self.dz

    @property
    def nl2(self):
        return self.types[0].qijlm.shape[0]


# ---------------------------------------------------------------------------
def ekcut_for(dx, dy, dz):
    dmax = max(dx, dy, dz)
    return 0.5 * (math.pi / dmax) ** 2 * (1.0 - 1e-10)


def real_ylm(lmax, x, y, z):
    """Real spherical harmonics up to lmax (<= 2), ABINIT order (m=-l..l with
    the (y, z, x) convention for l=1).  Returns (npts, (lmax+1)^2)."""
    r = np.sqrt(x * x + y * y + z * z)
    rs = np.where(r > 1e-12, r, 1.0)
    ux, uy, uz = x / rs, y / rs, z / rs
    ux = np.where(r > 1e-12, ux, 0.0)
    uy = np.where(r > 1e-12, uy, 0.0)
    uz = np.where(r > 1e-12, uz, 0.0)
    out = [np.full_like(x, 0.5 / math.sqrt(math.pi))]
    if lmax >= 1:
        c1 = math.sqrt(3.0 / (4.0 * math.pi))
        out += [c1 * uy, c1 * uz, c1 * ux]
    if lmax >= 2:
        c2 = math.sqrt(15.0 / (4.0 * math.pi))
        c20 = math.sqrt(5.0 / (16.0 * math.pi))
        out += [c2 * ux * uy, c2 * uy * uz, c20 * (3 * uz * uz - 1.0),
                c2 * ux * uz, 0.5 * c2 * (ux * ux - uy * uy)]
    return np.stack(out, axis=-1)


def _box_start(coord, nr, n, d):
    # src/get_irs.f90:22-27: ceiling((coord+xmax)/dx) - nrough/2 (0-based)
    xmax = n * d / 2.0
    return int(math.ceil((coord + xmax) / d)) - nr // 2


def _grid_coord(i, n, d):
    # position of 0-based grid index i: (i - n/2) d  (src/tddft_plot.f90:45)
    return (i - n // 2) * d


def _make_atom(rng, spec_grid, itype, coord, nr, radial, ms_l, rcut, sij_scale,
               sph_radius, nl2, gwidth):
    nx, ny, nz, dx, dy, dz, periodic = spec_grid
    nr = np.array(nr, dtype=np.int64)
    start = np.array([_box_start(coord[0], nr[0], nx, dx),
                      _box_start(coord[1], nr[1], ny, dy),
                      _box_start(coord[2], nr[2], nz, dz)], dtype=np.int64)
    if not periodic:
        if start.min() < 0 or start[0] + nr[0] > nx or start[1] + nr[1] > ny or start[2] + nr[2] > nz:
            raise ValueError("non periodic: projector box outside the grid")
    gx = _grid_coord(start[0] + np.arange(nr[0]), nx, dx) - coord[0]
    gy = _grid_coord(start[1] + np.arange(nr[1]), ny, dy) - coord[1]
    gz = _grid_coord(start[2] + np.arange(nr[2]), nz, dz) - coord[2]
    X = np.broadcast_to(gx[None, None, :], (nr[2], nr[1], nr[0]))
    Y = np.broadcast_to(gy[None, :, None], (nr[2], nr[1], nr[0]))
    Z = np.broadcast_to(gz[:, None, None], (nr[2], nr[1], nr[0]))
    R = np.sqrt(X * X + Y * Y + Z * Z)
    ylm = real_ylm(1, X, Y, Z)
    tabs = []
    for (l, nrad) in ms_l:
        for m in range(2 * l + 1):
            f = radial(R, l, nrad) * ylm[..., l * l + m]
            if rcut is not None:
                f = np.where(R <= rcut, f, 0.0)
            tabs.append(f)
    p = np.stack(tabs).astype(np.float64)                 # (ms, nr3, nr2, nr1)
    ms = p.shape[0]
    dv = dx * dy * dz
    # src/transform_p3d.f90: L = <p|p>dv ; p1 = p L^-1/2 V ; s = eig(L^1/2 sij L^1/2)
    P = p.reshape(ms, -1)
    L = P @ P.T * dv
    w, U = np.linalg.eigh(L)
    Lmh = (U / np.sqrt(w)) @ U.T
    Lph = (U * np.sqrt(w)) @ U.T
    a = rng.uniform(-1.0, 1.0, (ms, ms)) * sij_scale
    sij = 0.5 * (a + a.T)
    M = Lph @ sij @ Lph
    # keep 1+s comfortably positive (src/read_paw.f90:755 clamps at -1+0.01)
    sev, V = np.linalg.eigh(M)
    sev = np.maximum(sev, -0.6)
    P1 = (V.T @ Lmh) @ P
    p1 = P1.reshape(p.shape)
    s = sev
    sinv = 1.0 / (1.0 + s) - 1.0
    # compensation sphere: every grid point within sph_radius (wrapped)
    rcs = sph_radius
    ir = [int(math.ceil(rcs / d)) + 1 for d in (dx, dy, dz)]
    cidx = [int(round(coord[0] / dx)) + nx // 2, int(round(coord[1] / dy)) + ny // 2,
            int(round(coord[2] / dz)) + nz // 2]
    ox = np.arange(cidx[0] - ir[0], cidx[0] + ir[0] + 1)
    oy = np.arange(cidx[1] - ir[1], cidx[1] + ir[1] + 1)
    oz = np.arange(cidx[2] - ir[2], cidx[2] + ir[2] + 1)
    OX, OY, OZ = np.meshgrid(ox, oy, oz, indexing="ij")
    rx = _grid_coord(OX, nx, dx) - coord[0]
    ry = _grid_coord(OY, ny, dy) - coord[1]
    rz = _grid_coord(OZ, nz, dz) - coord[2]
    rr = np.sqrt(rx * rx + ry * ry + rz * rz)
    sel = rr < rcs
    OX, OY, OZ, rx, ry, rz, rr = (v[sel] for v in (OX, OY, OZ, rx, ry, rz, rr))
    if not periodic and (OX.min() < 0 or OY.min() < 0 or OZ.min() < 0 or
                         OX.max() >= nx or OY.max() >= ny or OZ.max() >= nz):
        raise ValueError("non periodic: compensation sphere outside the grid")
    flat = (OX % nx) + nx * ((OY % ny) + ny * (OZ % nz))
    order = np.argsort(flat, kind="stable")
    flat, rx, ry, rz, rr = flat[order], rx[order], ry[order], rz[order], rr[order]
    lmax = int(round(math.sqrt(nl2))) - 1
    yl = real_ylm(lmax, rx, ry, rz)
    gl = np.zeros((flat.size, nl2))
    for L in range(lmax + 1):
        shape = rr ** L * np.exp(-(rr / gwidth) ** 2)
        norm = np.sum(shape * rr ** L) * dv / (4.0 * math.pi / (2 * L + 1))
        for m in range(2 * L + 1):
            gl[:, L * L + m] = shape / norm * yl[:, L * L + m]
    return AtomSpec(itype=itype, coord=np.asarray(coord, float), start=start, nr=nr,
                    p=p, p1=p1, s=s, sinv=sinv, sph_idx=flat.astype(np.int64), sph_g=gl)


def _make_type(rng, ms, nl2, q00, onsite_scale):
    q = rng.uniform(-1.0, 1.0, (nl2, ms, ms)) * 0.02
    q = 0.5 * (q + np.transpose(q, (0, 2, 1)))
    q[0] += q00 * np.eye(ms)
    a = rng.uniform(-0.5, 0.5, (ms, ms))
    dij0 = 0.5 * (a + a.T)
    e = rng.uniform(0.0, 1.0, (ms, ms, ms, ms)) * onsite_scale
    e = e + e.transpose(1, 0, 2, 3)
    e = e + e.transpose(0, 1, 3, 2)
    e = e + e.transpose(2, 3, 0, 1)
    return TypeSpec(ms=ms, qijlm=q, dij0=dij0, eijkl=e / 8.0)


# ---------------------------------------------------------------------------
def polyacene_geometry(nring, cc=2.65, ch=2.05):
    """Planar linear polyacene C(4n+2)H(2n+4) in the xy plane, long axis x,
    regular hexagons of side cc, C-H bonds radial from the ring centre."""
    a = cc
    h = a * math.sqrt(3.0) / 2.0
    centers = [((2 * i - (nring - 1)) * h, 0.0) for i in range(nring)]
    pts = {}
    ring_of = {}
    for ci, (cx, cy) in enumerate(centers):
        for k in range(6):
            th = math.radians(30 + 60 * k)
            key = (round(cx + a * math.cos(th), 6), round(cy + a * math.sin(th), 6))
            pts.setdefault(key, []).append(ci)
            ring_of[key] = ci
    carbons = sorted(pts.keys())
    hyd = []
    for key in carbons:
        if len(pts[key]) == 1:
            cx, cy = centers[pts[key][0]]
            vx, vy = key[0] - cx, key[1] - cy
            nrm = math.hypot(vx, vy)
            hyd.append((key[0] + ch * vx / nrm, key[1] + ch * vy / nrm))
    C = np.array([[x, y, 0.0] for x, y in carbons])
    H = np.array([[x, y, 0.0] for x, y in hyd])
    return C, H


def coronene_geometry(cc=2.68, ch=2.05):
    """Planar coronene C24H12 (D6h) in the xy plane: a central hexagon of side
    cc and the six hexagons sharing its edges; C-H bonds radial from the
    centre of the one ring an outer carbon belongs to."""
    a = cc
    centers = [(0.0, 0.0)] + [(math.sqrt(3.0) * a * math.cos(math.radians(60 * k)),
                               math.sqrt(3.0) * a * math.sin(math.radians(60 * k))) for k in range(6)]
    pts = {}
    for ci, (cx, cy) in enumerate(centers):
        for k in range(6):
            th = math.radians(30 + 60 * k)
            key = (round(cx + a * math.cos(th), 6), round(cy + a * math.sin(th), 6))
            pts.setdefault(key, []).append(ci)
    carbons = sorted(pts.keys())
    hyd = []
    for key in carbons:
        if len(pts[key]) == 1:
            cx, cy = centers[pts[key][0]]
            vx, vy = key[0] - cx, key[1] - cy
            nrm = math.hypot(vx, vy)
            hyd.append((key[0] + ch * vx / nrm, key[1] + ch * vy / nrm))
    C = np.array([[x, y, 0.0] for x, y in carbons])
    H = np.array([[x, y, 0.0] for x, y in hyd])
    return C, H


def molecule_system(name="c10h8", nring=2, n=(64, 64, 64), d=0.5, seed=1234,
                    dt=0.05, sm=1e-3, ipol=0, geometry="polyacene", cc=None) -> SystemSpec:
    """Nonperiodic MT system with synthetic C (14^3 box, 8 projectors) and H
    (10^3 box, 5 projectors) tables: the polyacene with nring rings
    (naphthalene for nring=2; C-C 2.65 bohr) or coronene C24H12 (C-C 2.68)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    dx = dy = dz = float(d)
    if geometry == "polyacene":
        Cxyz, Hxyz = polyacene_geometry(nring, cc=cc or 2.65)
    elif geometry == "coronene":
        Cxyz, Hxyz = coronene_geometry(cc=cc or 2.68)
    else:
        raise ValueError(f"unknown geometry {geometry}")
    grid = (nx, ny, nz, dx, dy, dz, False)
    nl2 = 9

    def radial_c(R, l, nrad):
        return R ** (l + 2 * nrad) * np.exp(-R * R / 0.8)

    def radial_h(R, l, nrad):
        return R ** (l + 2 * nrad) * np.exp(-R * R / 0.6)

    types = [_make_type(rng, 8, nl2, 0.05, 0.02), _make_type(rng, 5, nl2, 0.03, 0.02)]
    atoms = []
    for c in Cxyz:
        atoms.append(_make_atom(rng, grid, 0, c, (14, 14, 14), radial_c,
                                [(0, 0), (0, 1), (1, 0), (1, 1)], None, 0.3, 1.6, nl2, 0.7))
    for hpos in Hxyz:
        atoms.append(_make_atom(rng, grid, 1, hpos, (10, 10, 10), radial_h,
                                [(0, 0), (0, 1), (1, 0)], None, 0.3, 1.2, nl2, 0.5))
    N = nx * ny * nz
    rx = (np.arange(nx) - nx // 2) * dx
    ry = (np.arange(ny) - ny // 2) * dy
    rz = (np.arange(nz) - nz // 2) * dz
    X = rx[None, None, :]
    Y = ry[:, None][None]
    Z = rz[:, None, None]
    vloc = np.zeros((nz, ny, nx))
    ncore = np.zeros((nz, ny, nx))
    from math import erf  # noqa: F401  (scalar erf not needed; vectorised below)
    for a in atoms:
        r = np.sqrt((X - a.coord[0]) ** 2 + (Y - a.coord[1]) ** 2 + (Z - a.coord[2]) ** 2)
        zv, rc = (4.0, 0.9) if a.itype == 0 else (1.0, 0.6)
        vloc += -zv * _erf_over_r(r, rc)
        if a.itype == 0:
            ncore += 0.3 * np.exp(-(r / 0.5) ** 2)
    nocc = (len(Cxyz) * 4 + len(Hxyz)) // 2
    return SystemSpec(name=name, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, periodic=False,
                      ekcut=ekcut_for(dx, dy, dz), atoms=atoms, types=types,
                      vloc=vloc.reshape(-1), ncoret=ncore.reshape(-1), rnel=2.0 * nocc,
                      nocc=nocc, dt=dt, sm=sm, ipol=ipol,
                      extra=dict(seed=seed, natom_c=len(Cxyz), natom_h=len(Hxyz), N=N))


def _erf_over_r(r, rc):
    from scipy.special import erf
    rs = np.where(r > 1e-12, r, 1.0)
    return np.where(r > 1e-12, erf(r / rc) / rs, 2.0 / (math.sqrt(math.pi) * rc))


def hpsi_system(n=(96, 96, 96), d=0.4, natom=40, seed=2024, nl2=9, nrad=2) -> SystemSpec:
    """Periodic synthetic H.psi system: V_KS = seeded sum of 40 Gaussians
    (depth -0.5..-2 Ha, width 0.5..1.5 bohr); natom atoms uniformly placed
    with >= 2 bohr separation, 4 * nrad projectors r^l exp(-r^2/0.8) Y_lm
    (nrad radial orders per l <= 1; 8 for nrad = 2), cut at 3 bohr; random
    symmetric D_ij in [-1, 1]."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    dx = dy = dz = float(d)
    L = np.array([nx * dx, ny * dy, nz * dz])
    grid = (nx, ny, nz, dx, dy, dz, True)
    coords = []
    while len(coords) < natom:
        c = (rng.uniform(0.0, 1.0, 3) - 0.5) * L
        ok = True
        for q in coords:
            dd = c - q
            dd -= L * np.round(dd / L)
            if np.sqrt(np.sum(dd * dd)) < 2.0:
                ok = False
                break
        if ok:
            coords.append(c)
    rcut = 3.0
    nrough = [2 * (int(math.ceil((rcut + 1.0) / dd)) + 1) for dd in (dx, dy, dz)]

    def radial(R, l, nrad):
        return R ** (l + 2 * nrad) * np.exp(-R * R / 0.8)

    ms_l = [(l, r) for l in (0, 1) for r in range(nrad)]
    types = [_make_type(rng, 4 * nrad, nl2, 0.05, 0.02)]
    atoms = [_make_atom(rng, grid, 0, c, nrough, radial, ms_l, rcut, 0.3, 1.6, nl2, 0.7) for c in coords]
    rx = (np.arange(nx) - nx // 2) * dx
    ry = (np.arange(ny) - ny // 2) * dy
    rz = (np.arange(nz) - nz // 2) * dz
    vks = np.zeros((nz, ny, nx))
    for _ in range(40):
        c = (rng.uniform(0.0, 1.0, 3) - 0.5) * L
        A = rng.uniform(-2.0, -0.5)
        w = rng.uniform(0.5, 1.5)
        ddx = rx - c[0]
        ddx -= L[0] * np.round(ddx / L[0])
        ddy = ry - c[1]
        ddy -= L[1] * np.round(ddy / L[1])
        ddz = rz - c[2]
        ddz -= L[2] * np.round(ddz / L[2])
        r2 = ddx[None, None, :] ** 2 + ddy[None, :, None] ** 2 + ddz[:, None, None] ** 2
        vks += A * np.exp(-r2 / (w * w))
    dij = []
    for a in atoms:
        m = rng.uniform(-1.0, 1.0, (a.ms, a.ms))
        dij.append(0.5 * (m + m.T))
    N = nx * ny * nz
    return SystemSpec(name="hpsi", nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, periodic=True,
                      ekcut=ekcut_for(dx, dy, dz), atoms=atoms, types=types,
                      vloc=vks.reshape(-1).copy(), ncoret=np.zeros(N), rnel=2.0 * 64,
                      nocc=64, vks0=vks.reshape(-1), dij0_atoms=dij,
                      extra=dict(seed=seed, N=N))


def random_orbitals(spec: SystemSpec, nb: int, seed: int = 7, sigma: float = 1.0,
                    envelope: Optional[float] = None, dtype=np.complex128):
    """nb complex orbitals ~ N(0,1), Gaussian-filtered (sigma bohr) in
    reciprocal space, optionally multiplied by a Gaussian envelope of width
    ``envelope`` around the origin; NOT orthonormalised (orthog_phi does that)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = spec.nx, spec.ny, spec.nz
    kx = 2 * np.pi * np.fft.fftfreq(nx, spec.dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, spec.dy)
    kz = 2 * np.pi * np.fft.fftfreq(nz, spec.dz)
    filt = np.exp(-0.5 * sigma ** 2 * (kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2))
    env = None
    if envelope is not None:
        rx = (np.arange(nx) - nx // 2) * spec.dx
        ry = (np.arange(ny) - ny // 2) * spec.dy
        rz = (np.arange(nz) - nz // 2) * spec.dz
        env = np.exp(-(rx[None, None, :] ** 2 + ry[None, :, None] ** 2 + rz[:, None, None] ** 2) / (2 * envelope ** 2))
    out = np.empty((nb, spec.n), dtype=dtype)
    for b in range(nb):
        z = rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))
        z = np.fft.ifftn(np.fft.fftn(z) * filt)
        if env is not None:
            z = z * env
        out[b] = z.reshape(-1)
    return out


def lowdin_orthonormalise(phi: np.ndarray, dv: float) -> np.ndarray:
    """phi (nb, N): returns phi G^-1/2 with G = phi^H phi dv (FP64)."""
    G = (phi.conj() @ phi.T) * dv
    w, U = np.linalg.eigh(G)
    T = (U / np.sqrt(w)) @ U.conj().T
    return T.T @ phi
