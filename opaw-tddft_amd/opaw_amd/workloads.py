"""The molecules of the benchmark configurations, built from the JTH
datasets, and the ground state that starts their propagation.

acene_geometry: polyacenes C(4n+2)H(2n+4) with regular hexagons of C-C bond
a and C-H bond h along the outward ring radius, rings fused along x, in the
z = 0 plane, centred at the origin (bohr).  n = 2 is naphthalene in its D2h
form (a = 2.65, h = 2.05 bohr for the C10H8 workload).  coronene_system:
C24H12 in its D6h form (C-C 2.68, C-H 2.05 bohr).

ground_state: the reference's main program up to writewf
(src/main.f90:46-112): nb = max(ceil(1.3 nocc) + 5, 20) bands
(src/rnel.f90:17-21, src/0_parameter.f90:20-21), seeded random start
orbitals (the reference draws them from its KISS generator,
src/phi0.f90), iteration 1 from the atomic superposition density with
rho_ij = 0, then Chebyshev filter + orthog_phi + diagh + DIIS until the
occupied residuals reach tol.  The occupied orbitals go to the TDDFT
context through the wf_bar format (writewf / read_wf_tddft), as the
reference hands them over.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import gs, paw, synth

CC_BOND = 2.65
CH_BOND = 2.05


def acene_geometry(nring: int, a: float = CC_BOND, h: float = CH_BOND) -> Tuple[List[str], np.ndarray]:
    """Symbols and coordinates (bohr) of the polyacene with nring rings."""
    if nring < 1:
        raise ValueError("nring >= 1")
    s3 = math.sqrt(3.0)
    centres = [((k - 0.5 * (nring - 1)) * a * s3, 0.0) for k in range(nring)]
    carbons: List[Tuple[float, float]] = []
    owner: List[List[int]] = []
    for k, (cx, cy) in enumerate(centres):
        for deg in (30, 90, 150, 210, 270, 330):
            t = math.radians(deg)
            p = (cx + a * math.cos(t), cy + a * math.sin(t))
            for i, q in enumerate(carbons):
                if abs(q[0] - p[0]) < 1e-9 and abs(q[1] - p[1]) < 1e-9:
                    owner[i].append(k)
                    break
            else:
                carbons.append(p)
                owner.append([k])
    syms = ["C"] * len(carbons)
    xyz = [[x, y, 0.0] for (x, y) in carbons]
    for (x, y), own in zip(carbons, owner):
        if len(own) == 1:
            cx, cy = centres[own[0]]
            u = np.array([x - cx, y - cy])
            u /= np.linalg.norm(u)
            syms.append("H")
            xyz.append([x + h * u[0], y + h * u[1], 0.0])
    return syms, np.array(xyz)


def jth_datasets(geom: paw.Geometry, directory: str) -> dict:
    return {s: paw.read_dataset(os.path.join(directory, f"{s}.LDA_PW-JTH.xml"), half_box=geom.half)
            for s in ("H", "C")}


def acene_system(nring: int, n: Tuple[int, int, int], d: float, directory: str, name: Optional[str] = None,
                 distort: Optional[Tuple[float, int]] = None):
    """A polyacene on an n grid of spacing d, nonperiodic (Martyna-Tuckerman),
    nocc = (valence electrons) / 2.  distort = (sigma, seed): every atom
    moved by a seeded Gaussian of standard deviation sigma (bohr) per axis."""
    syms, xyz = acene_geometry(nring)
    if distort is not None:
        sigma, seed = distort
        xyz = xyz + np.random.default_rng(seed).normal(0.0, sigma, xyz.shape)
    geom = paw.Geometry(n[0], n[1], n[2], d, d, d, False)
    ds = jth_datasets(geom, directory)
    nval = int(round(sum(ds[s].val for s in syms)))
    return paw.system_from_datasets(geom, syms, xyz, ds, nocc=nval // 2, name=name or f"acene{nring}-jth")


def coronene_system(n: Tuple[int, int, int], d: float, directory: str, name: Optional[str] = None,
                    cc: float = 2.68, ch: float = 2.05):
    """Coronene C24H12 in its D6h form (C-C cc, C-H ch bohr) from the JTH
    datasets on an n grid of spacing d, nonperiodic (Martyna-Tuckerman);
    nocc = 54."""
    C, H = synth.coronene_geometry(cc=cc, ch=ch)
    syms = ["C"] * len(C) + ["H"] * len(H)
    xyz = np.concatenate([C, H])
    geom = paw.Geometry(n[0], n[1], n[2], d, d, d, False)
    ds = jth_datasets(geom, directory)
    nval = int(round(sum(ds[s].val for s in syms)))
    return paw.system_from_datasets(geom, syms, xyz, ds, nocc=nval // 2, name=name or "coronene-jth")


def nbands(nocc: int) -> int:
    """get_rnel (src/rnel.f90:17-21) with rnel_1 = 1.3, rnel_2 = 5, rnel_3 = 20."""
    return max(math.ceil(nocc * 1.3) + 5, 20)


def real_phase(phi: np.ndarray) -> np.ndarray:
    """Each orbital times the global phase that makes its largest component
    real and positive.  At the gamma point H is real, so a non-degenerate
    eigenvector is then real up to rounding, and the wf_bar format
    (sign(Re phi) |phi|) stores it without loss."""
    out = np.array(phi, dtype=np.complex128, copy=True)
    for k in range(out.shape[0]):
        j = int(np.argmax(np.abs(out[k])))
        out[k] *= np.conj(out[k, j]) / abs(out[k, j])
    return out


@dataclass
class GroundState:
    phi: np.ndarray            # (nocc, N) occupied orbitals, complex128
    eig: np.ndarray            # (nb,) eigenvalues of the last diagh
    scf: gs.ScfResult


def ground_state(spec, tol: float = 1e-8, nscf: int = 60, seed: int = 2024, nb: Optional[int] = None,
                 wf_path: Optional[str] = None, degree: int = 10) -> GroundState:
    """The ground-state SCF on the GPU (FP64, h_type 1), Chebyshev filter of
    the given degree (the reference's nch; 10 by default, 16 for coronene at
    0.45 bohr, whose SCF stalls near 1e-7 with 10).  With wf_path the
    occupied orbitals are written there in the wf_bar.bin layout and read
    back, the hand-over the reference makes (src/write_wf.f90:122-219,
    src/read_wf.f90:130-228)."""
    from . import formats
    from .api import OpawGPU
    nocc = spec.nocc
    nb = nb or nbands(nocc)
    rnel_keep = spec.rnel
    spec.rnel = 2.0 * nocc
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=seed, envelope=4.0), spec.dv)
    g = OpawGPU(spec, "f64", nsys=1, nb=nb)
    try:
        g.set_nocc(nocc)
        g.set_diis(4, 0.15)
        g.write_orbitals(0, phi)
        res = gs.scf_chebyshev(g, nocc, nscf=nscf, tol=tol, degree=degree, dens0=spec.extra.get("dens0"))
        out = real_phase(g.read_orbitals(0)[:nocc])
    finally:
        g.close()
        spec.rnel = rnel_keep
    if wf_path:
        grid, spacing = (spec.nx, spec.ny, spec.nz), (spec.dx, spec.dy, spec.dz)
        formats.write_wf(wf_path, out, res.eig[:nocc], grid, spacing, binary=True, short_header=True)
        out = formats.read_wf(wf_path, grid, spacing, nocc, binary=True)[0]
    return GroundState(phi=out, eig=res.eig, scf=res)
