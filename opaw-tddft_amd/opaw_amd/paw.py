"""Host side of the PAW set-up (prep_paw, src/prepare_paw.f90:1-49).

Reads the atom input and the JTH XML datasets and does the radial work of
readpaw (src/read_paw.f90), transform_pphi (src/transform_pphi.f90), the
Gaunt coefficients of cg3 (src/cg.f90:71-148), the real spherical harmonics
of src/ylm.f90, the spline B matrices of get_b (src/read_paw.f90:277-345)
and the box origins of get_irs (src/get_irs.f90).  The grid-sized work --
projectors sampled on the fine grid and contracted onto the box
(get_local_p3d, src/p3d.f90) and the compensation spheres (map_localgrid,
get_localg3d: src/map_localgrid.f90, src/g3d.f90) -- runs on the GPU
through opaw_paw_box / opaw_paw_sphere; the orthogonalisation of the box
projectors (transform_p3d, src/transform_p3d.f90) is a small host step.

Literal constants keep the precision the reference gives them: a Fortran
default-real literal is a single-precision number even where it is assigned
to real*8 (pi in src/ylm.f90:13, 4 pi in src/cg.f90:88, the Y_00 factor of
src/read_paw.f90:199, tollsij in src/0_parameter.f90:18).
"""
from __future__ import annotations

import ctypes as C
import math
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import lib

F32 = lambda v: float(np.float32(v))          # a Fortran default-real literal

PI_YLM = F32(3.14159265359)                   # src/ylm.f90:13
FOURPI_CG = F32(12.5663706144)                # src/cg.f90:88
VLOC_Y00 = F32(0.282094792)                   # src/read_paw.f90:199
TOLLSIJ = F32(0.01)                           # src/0_parameter.f90:18
SQ4PI = math.sqrt(12.5663706144)              # src/read_paw.f90:188,699 (double literals)
PI_GL = 3.14159265359                         # src/read_paw.f90:118 (double literal)
RPAD = 1.0                                    # src/0_parameter.f90:14
RPAD_R = 2.0                                  # src/0_parameter.f90:15


# ---------------------------------------------------------------------------
# small numerical pieces of the reference

def interpolate(arr: np.ndarray, targ: float) -> Tuple[int, int, float, float]:
    """src/interpolate.f90: bisection for arr[i] <= targ < arr[j], j = i+1,
    and the linear weights; indices 0-based here."""
    i, j = 0, len(arr) - 1
    while j - i > 1:
        jm = (i + j + 2) // 2 - 1            # Fortran (i+j)/2 on 1-based indices
        if targ < arr[jm]:
            j = jm
        else:
            i = jm
    wi = (arr[j] - targ) / (arr[j] - arr[i])
    wj = (targ - arr[i]) / (arr[j] - arr[i])
    return i, j, wi, wj


def ylm(l: int, m: int, xx: float, yy: float, zz: float) -> float:
    """Real spherical harmonics of src/ylm.f90:4-107 (l <= 4), with the
    single-precision pi of the reference."""
    pi = PI_YLM
    p1 = 1.0 / (2.0 * math.sqrt(pi))
    p2 = p1 * math.sqrt(3.0)
    p3 = math.sqrt(15.0) / (2.0 * math.sqrt(pi))
    if l == 0:
        return p1
    if l == 1:
        return p2 * {-1: yy, 0: zz, 1: xx}[m]
    if l == 2:
        return {-2: p3 * xx * yy, -1: p3 * yy * zz,
                0: p3 * (2.0 * zz ** 2 - xx ** 2 - yy ** 2) / (2.0 * math.sqrt(3.0)),
                1: p3 * xx * zz, 2: p3 * (xx ** 2 - yy ** 2) / 2.0}[m]
    if l == 3:
        return {-3: 1 / 4 * math.sqrt(35 / 2 / pi) * yy * (3 * xx * xx - yy * yy),
                -2: 1 / 2 * math.sqrt(105 / pi) * xx * yy * zz,
                -1: 1 / 4 * math.sqrt(21 / 2 / pi) * yy * (4 * zz * zz - xx * xx - yy * yy),
                0: 1 / 4 * math.sqrt(7 / pi) * zz * (2 * zz * zz - 3 * xx * xx - 3 * yy * yy),
                1: 1 / 4 * math.sqrt(21 / 2 / pi) * xx * (4 * zz * zz - xx * xx - yy * yy),
                2: 1 / 4 * math.sqrt(105 / pi) * (xx * xx - yy * yy) * zz,
                3: 1 / 4 * math.sqrt(35 / 2 / pi) * xx * (xx * xx - 3 * yy * yy)}[m]
    if l == 4:
        return {-4: 3 / 4 * math.sqrt(35 / pi) * xx * yy * (xx * xx - yy * yy),
                -3: 3 / 4 * math.sqrt(35 / 2 / pi) * (3 * xx * xx - yy * yy) * yy * zz,
                -2: 3 / 4 * math.sqrt(5 / pi) * xx * yy * (7 * zz * zz - 1),
                -1: 3 / 4 * math.sqrt(5 / 2 / pi) * yy * zz * (7 * zz * zz - 3),
                0: 3 / 16 / math.sqrt(pi) * (35 * zz ** 4 - 30 * zz * zz + 3),
                1: 3 / 4 * math.sqrt(5 / 2 / pi) * xx * zz * (7 * zz * zz - 3),
                2: 3 / 8 * math.sqrt(5 / pi) * (xx * xx - yy * yy) * (7 * zz * zz - 1),
                3: 3 / 4 * math.sqrt(35 / 2 / pi) * (xx * xx - 3 * yy * yy) * xx * zz,
                4: 3 / 16 * math.sqrt(35 / pi) * (xx * xx * (xx * xx - 3 * yy * yy) -
                                                 yy * yy * (3 * xx * xx - yy * yy))}[m]
    raise ValueError("l>4 not implemented yet")


def _fact(n: int) -> float:
    """src/cg.f90:177-185: the product 1..n (1 for n <= 0)."""
    f = 1.0
    for i in range(1, n + 1):
        f *= float(i)
    return f


def cg_coeff(l, m, l1, m1, l2, m2) -> float:
    """Clebsch-Gordan <l1 m1 l2 m2 | l m> by the Racah sum of
    src/cg.f90:1-69."""
    if m < -l or m > l or m1 < -l1 or m1 > l1 or m2 < -l2 or m2 > l2:
        raise ValueError(f"error in cg: m,l,m1,l1,m2,l2 {m} {l} {m1} {l1} {m2} {l2}")
    if abs(l1 - l2) > l or l1 + l2 < l or m != m1 + m2:
        return 0.0
    num = (2 * l + 1) * _fact(l + l1 - l2) * _fact(l - l1 + l2) * _fact(l1 + l2 - l) * _fact(l + m) * _fact(l - m)
    den = _fact(l1 + l2 + l + 1) * _fact(l1 - m1) * _fact(l1 + m1) * _fact(l2 - m2) * _fact(l2 + m2)
    s = 0.0
    for i in range(0, l + m + 1):
        if l - l1 + l2 - i < 0 or l + m - i < 0 or i + l1 - l2 - m < 0:
            continue
        s += ((-1.0) ** (i + l2 + m2) * _fact(l2 + l + m1 - i) * _fact(l1 - m1 + i)) / (
            _fact(i) * _fact(l - l1 + l2 - i) * _fact(l + m - i) * _fact(i + l1 - l2 - m))
    return math.sqrt(num / den) * s


def _aa(m: int) -> Tuple[complex, complex]:
    """get_aa (src/cg.f90:122-137): real harmonic (l,m) as a(1) Y_lm + a(2) Y_l,-m."""
    if m < 0:
        return 1j / math.sqrt(2.0), 1j / math.sqrt(2.0) * (-1) ** (m + 1)
    if m == 0:
        return 1.0 + 0j, 0j
    return (-1) ** m / math.sqrt(2.0) + 0j, 1 / math.sqrt(2.0) + 0j


def cg3(l1, m1, l2, m2, l3, m3) -> float:
    """Integral of three real spherical harmonics, src/cg.f90:71-120."""
    k = (m1 < 0) + (m2 < 0) + (m3 < 0)
    if k in (1, 3):
        return 0.0
    pf = math.sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) / (2.0 * l3 + 1.0) / FOURPI_CG)
    cg0 = cg_coeff(l3, 0, l1, 0, l2, 0)
    cgg = np.zeros((2, 2, 2))
    for a in range(2):
        for b in range(2):
            for c in range(2):
                cgg[a, b, c] = cg_coeff(l3, m3 * (1 - 2 * c), l1, m1 * (1 - 2 * a), l2, m2 * (1 - 2 * b))
    A = np.conj(np.array(_aa(m1)))
    B = np.conj(np.array(_aa(m2)))
    Cc = np.array(_aa(m3))
    tmp = 0j
    for a in range(2):
        for b in range(2):
            for c in range(2):
                tmp += cgg[a, b, c] * A[a] * B[b] * Cc[c]
    tmp *= cg0
    return pf * tmp.real


def natural_spline_y2(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Second derivatives of the natural cubic spline through (x, y):
    SPLINE_dn (src/lib/library_NRF_spline.f:11-55), the Numerical Recipes
    routine with yp1 = ypn = 1e32, i.e. natural end conditions."""
    n = len(x)
    y2 = np.zeros(n)
    u = np.zeros(n)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[i - 1] + 2.0
        y2[i] = (sig - 1.0) / p
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]))
                / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
    y2[n - 1] = 0.0
    for k in range(n - 2, -1, -1):
        y2[k] = y2[k] * y2[k + 1] + u[k]
    return y2


def natural_spline_eval(x: np.ndarray, y: np.ndarray, y2: np.ndarray, t: float) -> float:
    """SPLINT_dn (src/lib/library_NRF_splint.f:6-33): bisection for the
    interval, then the cubic in the interval."""
    klo, khi = 0, len(x) - 1
    while khi - klo > 1:
        k = (khi + klo) // 2
        if x[k] > t:
            khi = k
        else:
            klo = k
    h = x[khi] - x[klo]
    a = (x[khi] - t) / h
    b = (t - x[klo]) / h
    return a * y[klo] + b * y[khi] + ((a ** 3 - a) * y2[klo] + (b ** 3 - b) * y2[khi]) * h * h / 6.0


def b_matrix(nrough: int, nfine: int, d: float) -> Tuple[np.ndarray, float]:
    """get_b (src/read_paw.f90:277-345) along one axis: B[i, j] = value at
    fine point j of the spline through the rough-grid delta at i, times
    dxf/d.  Returns (B (nrough, nfine), dxf)."""
    dxf = d * float(nrough - 1) / float(nfine - 1)
    xx = np.arange(nrough) * d
    B = np.zeros((nrough, nfine))
    for i in range(nrough):
        yy = np.zeros(nrough)
        yy[i] = 1.0
        y2 = natural_spline_y2(xx, yy)
        for j in range(nfine):
            B[i, j] = natural_spline_eval(xx, yy, y2, j * dxf) * dxf / d
    return B, dxf


# ---------------------------------------------------------------------------
# the JTH dataset (readpaw, src/read_paw.f90:1-785)

class _Lines:
    """Line cursor with the reference's searchstring / rewind semantics."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def rewind(self):
        self.pos = 0

    def search(self, key: str) -> str:
        while self.pos < len(self.lines):
            line = self.lines[self.pos]
            self.pos += 1
            if key in line:
                return line
        raise ValueError(f"error : string not found: {key}")

    def next_line(self) -> str:
        if self.pos >= len(self.lines):
            raise ValueError("unexpected end of dataset")
        line = self.lines[self.pos]
        self.pos += 1
        return line

    def values(self, n: int) -> np.ndarray:
        """List-directed read of n reals starting at the next record."""
        out: List[float] = []
        while len(out) < n:
            toks = self.next_line().replace(",", " ").split()
            for t in toks:
                if len(out) == n:
                    break
                out.append(float(t.replace("D", "E").replace("d", "e")))
        return np.array(out)


def _attr(line: str, key: str) -> str:
    m = re.search(r"\b" + re.escape(key) + r'\s*=\s*"?\s*([^"\s/>]+)', line)
    if not m:
        raise ValueError(f"attribute {key} missing in: {line.strip()}")
    return m.group(1)


@dataclass
class PawDataset:
    filename: str
    symbol: str
    zat: float
    core: float
    val: float
    rcut: float
    lstate: np.ndarray             # (ns,) l of each radial state
    occ: np.ndarray                # (ns,) occupations, divided by 2l+1 for l > 0
    rr: np.ndarray                 # (nr,) radial grid, clamped to >= 1e-8
    dr: np.ndarray                 # (nr,) dr/di
    phi: np.ndarray                # (ns, nr) all-electron partial waves
    phitilde: np.ndarray           # (ns, nr) pseudo partial waves
    ptilde: np.ndarray             # (ns, nr) projectors
    vloc: np.ndarray               # (nr,) blochl local ionic potential * Y_00
    ncoretilde: np.ndarray         # (nr,) what readpaw stores under this name
    comp_type: str
    rcomp: float
    nrcomp: int                    # number of radial points of the shape functions
    nrcut: int                     # 1-based upper bracket of rcut
    nr_int: int                    # 1-based upper bracket of rcut + rpad_r
    nl: int
    lstate_diff: np.ndarray
    ms_ls: np.ndarray              # (ms,) 0-based radial state of each m-state
    mstate: np.ndarray             # (ms,) m of each m-state
    mat_t: np.ndarray = None
    mat_sp: np.ndarray = None
    ptilde1: np.ndarray = None
    phi1: np.ndarray = None
    phitilde1: np.ndarray = None
    gl: np.ndarray = None          # (nrcomp, 2 nl - 1)
    qijlm: np.ndarray = None       # (nl2, ms, ms)
    sij: np.ndarray = None         # (ms, ms)
    extra: dict = field(default_factory=dict)

    @property
    def nr(self) -> int:
        return len(self.rr)

    @property
    def ns(self) -> int:
        return len(self.lstate)

    @property
    def ms(self) -> int:
        return len(self.ms_ls)

    @property
    def nl2(self) -> int:
        return (2 * self.nl - 1) ** 2


def read_dataset(path: str, spacing: Sequence[float] = None, half_box: Sequence[float] = None) -> PawDataset:
    """readpaw: everything it reads and derives from one JTH XML file.
    half_box = (xmax, ymax, zmax) enables the rcut checks of get_rcut."""
    text = Path(path).read_text()
    f = _Lines(text)
    line = f.search("atom")
    symbol = _attr(line, "symbol")
    zat, core, val = (float(_attr(line, k)) for k in ("Z", "core", "valence"))
    line = f.search("<paw_radius")
    rcut = float(_attr(line, "rc"))
    if half_box is not None and any(rcut > h for h in half_box):
        raise ValueError("error:rcut>xmax")
    f.rewind()
    f.search("<valence_states>")
    states = []
    while True:
        line = f.next_line()
        if "</valence_states>" in line:
            break
        states.append(line)
    ns = len(states)
    line = f.search("radial_grid")
    nr = int(_attr(line, "iend")) - int(_attr(line, "istart")) + 1
    lstate = np.array([int(_attr(s, "l")) for s in states])
    occ = np.array([float(_attr(s, "f")) if re.search(r"\bf\s*=", s) else 0.0 for s in states])
    # get_nl: distinct l in increasing order
    lstate_diff = np.array(sorted(set(lstate.tolist())))
    nl = len(lstate_diff)
    # get_nm: m-states (m = -l..l) state by state
    ms_ls, mstate = [], []
    for i, l in enumerate(lstate):
        if l > 2:
            raise ValueError("l>2 not implemented yet")
        for m in range(-l, l + 1):
            ms_ls.append(i)
            mstate.append(m)
        if l > 0:
            occ[i] /= (2 * l + 1)
    f.rewind()
    f.search("<valence_states>")
    f.search("radial_grid")
    if "<values>" not in f.next_line():
        raise ValueError("error : wrong line before reading rr")
    rr = np.maximum(f.values(nr), 1e-8)
    if "</values>" not in f.next_line():
        raise ValueError("error : wrong line after reading rr")
    if "<derivatives>" not in f.next_line():
        raise ValueError("error : wrong line before reading dr")
    dr = f.values(nr)
    if "</derivatives>" not in f.next_line():
        raise ValueError("error : wrong line after reading dr")
    _, nrcut, _, _ = interpolate(rr, rcut)
    _, nr_int, _, _ = interpolate(rr, rcut + RPAD_R)
    phi = np.zeros((ns, nr))
    phitilde = np.zeros((ns, nr))
    ptilde = np.zeros((ns, nr))
    for i in range(ns):
        f.search("<ae_partial_wave")
        phi[i] = f.values(nr)
        f.search("<pseudo_partial_wave")
        phitilde[i] = f.values(nr)
        f.search("<projector_function")
        ptilde[i] = f.values(nr)
    f.rewind()
    f.search("<blochl_local_io")
    vloc = f.values(nr) * VLOC_Y00
    f.rewind()
    line = f.search("<shape_function").replace('"', " ")
    comp_type = line[line.index("type") + 5:].split()[0]
    if comp_type not in ("sinc", "gauss", "exp", "bessel"):
        raise ValueError(f"unknown shape function: {comp_type}")
    rcomp = float(line[line.index("rc") + 3:].split()[0])
    _, nrcomp, _, _ = interpolate(rr, rcomp)
    f.rewind()
    f.search("<pseudo_valence_dens")
    ncoretilde = f.values(nr) / SQ4PI
    ds = PawDataset(filename=str(path), symbol=symbol, zat=zat, core=core, val=val, rcut=rcut,
                    lstate=lstate, occ=occ, rr=rr, dr=dr, phi=phi, phitilde=phitilde, ptilde=ptilde,
                    vloc=vloc, ncoretilde=ncoretilde, comp_type=comp_type, rcomp=rcomp,
                    nrcomp=nrcomp + 1, nrcut=nrcut + 1, nr_int=nr_int + 1, nl=nl, lstate_diff=lstate_diff,
                    ms_ls=np.array(ms_ls), mstate=np.array(mstate))
    radial_matrices(ds)
    transform_pphi(ds)
    ds.gl = shape_functions(ds)
    ds.qijlm, ds.sij = calc_qijlm(ds)
    return ds


def radial_matrices(ds: PawDataset):
    """mat_sp and mat_t of read_p_phi (src/read_paw.f90:509-526) over the
    first nr_int radial points."""
    n = ds.nr_int
    w = ds.rr[:n] * ds.rr[:n] * ds.dr[:n]
    ds.mat_sp = (ds.ptilde[:, :n] * w) @ ds.ptilde[:, :n].T
    ds.mat_t = (ds.phi[:, :n] * w) @ ds.phi[:, :n].T - (ds.phitilde[:, :n] * w) @ ds.phitilde[:, :n].T


def _sym_eig(a: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Eigenvalues ascending and orthonormal eigenvectors of a symmetric
    matrix (mat_diag, src/lib/library_1_matrix_all.f90:213-288); each
    eigenvector's sign is fixed by making its largest component positive."""
    ev, vc = np.linalg.eigh(0.5 * (a + a.T))
    for k in range(vc.shape[1]):
        j = int(np.argmax(np.abs(vc[:, k])))
        if vc[j, k] < 0:
            vc[:, k] = -vc[:, k]
    return ev, vc


def transform_pphi(ds: PawDataset):
    """src/transform_pphi.f90: per angular momentum, projectors made
    orthonormal (p1 = O^-1/2 p, phi1 = O^1/2 phi) and then rotated to make
    the augmentation overlap mat_t diagonal (the equations after (A.2)-(A.4)
    of the OPAW paper cited there)."""
    ds.ptilde1 = np.zeros_like(ds.ptilde)
    ds.phi1 = np.zeros_like(ds.phi)
    ds.phitilde1 = np.zeros_like(ds.phitilde)
    mat_t = ds.mat_t.copy()
    for l in ds.lstate_diff:
        idx = np.nonzero(ds.lstate == l)[0]
        mat = ds.mat_sp[np.ix_(idx, idx)]
        eig, vec = _sym_eig(mat)
        sq = np.sign(eig) * np.sqrt(np.abs(eig))
        m1 = (vec / sq) @ vec.T
        m2 = (vec * sq) @ vec.T
        mat3 = mat_t[np.ix_(idx, idx)]
        ds.ptilde1[idx] = m1 @ ds.ptilde[idx]
        ds.phi1[idx] = m2.T @ ds.phi[idx]
        ds.phitilde1[idx] = m2.T @ ds.phitilde[idx]
        mat = m2.T @ mat3 @ m2
        eig, vec = _sym_eig(mat)
        ds.ptilde1[idx] = vec.T @ ds.ptilde1[idx]
        ds.phi1[idx] = vec.T @ ds.phi1[idx]
        ds.phitilde1[idx] = vec.T @ ds.phitilde1[idx]
        mat_t[np.ix_(idx, idx)] = vec.T @ mat @ vec
    ds.extra["mat_t_transformed"] = mat_t


def shape_functions(ds: PawDataset) -> np.ndarray:
    """get_gl (src/read_paw.f90:111-182) for the sinc shape function:
    g_l(r) = (sin(pi r/rc)/(pi r/rc))^2 r^l on the first nrcomp points,
    normalised to sum r^(l+2) dr g_l = 1."""
    if ds.comp_type != "sinc":
        raise ValueError(f"shape function {ds.comp_type}: only sinc is supported "
                         "(the reference stops for gauss and exp; bessel needs libpaw's atompaw_shapebes)")
    nrc = ds.nrcomp
    r = ds.rr[:nrc]
    gl = np.zeros((nrc, 2 * ds.nl - 1))
    for i in range(1, 2 * ds.nl):
        arg = PI_GL * r[1:] / ds.rcomp
        if i == 1:
            gl[0, 0] = 1.0
            gl[1:, 0] = (np.sin(arg) / arg) ** 2
        else:
            gl[0, i - 1] = 0.0
            gl[1:, i - 1] = (np.sin(arg) / arg) ** 2 * r[1:] ** (i - 1)
        norm = np.sum(r ** (i + 1) * ds.dr[:nrc] * gl[:, i - 1])
        gl[:, i - 1] /= norm
    return gl


def calc_qijlm(ds: PawDataset) -> Tuple[np.ndarray, np.ndarray]:
    """calc_qijlm (src/read_paw.f90:695-783): s_ij from the L = 0 moment of
    the transformed partial waves (must be diagonal; clipped at -1+tollsij),
    then q_ij^L of the original partial waves times the Gaunt coefficient."""
    ms = ds.ms
    nl2 = ds.nl2

    def moments(a, b, ll):
        integrand = np.empty((ds.ns, ds.ns, ds.nr))
        w = ds.rr ** (ll + 2) * ds.dr
        for i in range(ds.ns):
            for j in range(ds.ns):
                integrand[i, j] = (a[i] * a[j] - b[i] * b[j]) * w
        integrand[:, :, 0] = 0.0
        return integrand.sum(axis=2)

    def fill(nij, L):
        ll = int(math.floor(math.sqrt(L)))
        mm = L - ll * ll - ll
        q = np.zeros((ms, ms))
        for is_ in range(ms):
            for js in range(ms):
                il, jl = ds.ms_ls[is_], ds.ms_ls[js]
                q[is_, js] = nij[il, jl] * cg3(ll, mm, int(ds.lstate[il]), int(ds.mstate[is_]),
                                               int(ds.lstate[jl]), int(ds.mstate[js]))
        return q

    q0 = fill(moments(ds.phi1, ds.phitilde1, 0), 0)
    off = q0 - np.diag(np.diag(q0))
    if np.any(np.abs(off) > 1e-6):
        raise ValueError("need to use paw files with diagonalized sij!!")
    sij = q0 * SQ4PI
    for i in range(ms):
        if sij[i, i] < -1.0 + TOLLSIJ:
            sij[i, i] = -1.0 + TOLLSIJ
    qijlm = np.zeros((nl2, ms, ms))
    cache = {}
    for L in range(nl2):
        ll = int(math.floor(math.sqrt(L)))
        if ll not in cache:
            cache[ll] = moments(ds.phi, ds.phitilde, ll)
        qijlm[L] = fill(cache[ll], L)
    return qijlm, sij


# ---------------------------------------------------------------------------
# the atom input and the box geometry

@dataclass
class Geometry:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    periodic: bool
    finegrid: float = 0.15

    @property
    def half(self) -> Tuple[float, float, float]:
        """xmax, ymax, zmax = n d / 2 (src/main.f90:353-355)."""
        return (self.nx * self.dx / 2.0, self.ny * self.dy / 2.0, self.nz * self.dz / 2.0)

    @property
    def nfovnr(self) -> int:
        """src/main.f90:342-344."""
        n = max(math.ceil(self.dx / self.finegrid), 1)
        n = max(n, math.ceil(self.dy / self.finegrid))
        return max(n, math.ceil(self.dz / self.finegrid))


Z_OF = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10,
        "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18}


def read_cnt(path: str, geom: Geometry) -> Tuple[List[str], np.ndarray]:
    """cnt.ini (src/read_atominfo.f90:16-37): symbol x y z per line (bohr);
    atoms must lie within [-L/2, L/2]."""
    syms, xyz = [], []
    for line in Path(path).read_text().splitlines():
        t = line.split()
        if not t:
            continue
        syms.append(t[0])
        xyz.append([float(v.replace("D", "E").replace("d", "e")) for v in t[1:4]])
    xyz = np.array(xyz)
    h = geom.half
    for i, c in enumerate(xyz):
        if abs(c[0]) > h[0] or abs(c[1]) > h[1] or abs(c[2]) > h[2]:
            raise ValueError(f"for now put atoms between -L/2 to L/2 (atom {i + 1})")
    return syms, xyz


def type_map(symbols: Sequence[str]) -> Tuple[List[str], np.ndarray]:
    """map_atominfo (src/map_atominfo.f90): types are the distinct atomic
    numbers in increasing order; returns (type symbols, 0-based type of each
    atom)."""
    zs = sorted(set(Z_OF[s] for s in symbols))
    sym_of = {Z_OF[s]: s for s in symbols}
    tsyms = [sym_of[z] for z in zs]
    return tsyms, np.array([zs.index(Z_OF[s]) for s in symbols])


def box_shape(ds: PawDataset, geom: Geometry) -> Tuple[np.ndarray, np.ndarray]:
    """nrough, nfine of get_nrcut (src/read_paw.f90:247-265)."""
    d = (geom.dx, geom.dy, geom.dz)
    n = (geom.nx, geom.ny, geom.nz)
    nrough = np.array([(math.ceil((ds.rcut + RPAD) / d[k]) + 1) * 2 for k in range(3)])
    for k in range(3):
        if nrough[k] > n[k]:
            raise ValueError(f"nrough({k + 1})>n{'xyz'[k]}")
    return nrough, nrough * geom.nfovnr


def ir_start(coord: np.ndarray, nrough: np.ndarray, geom: Geometry) -> Tuple[np.ndarray, bool]:
    """get_irs (src/get_irs.f90:22-41): 1-based box origin; a box that
    leaves the grid is an error unless periodic (then it wraps)."""
    h = geom.half
    d = (geom.dx, geom.dy, geom.dz)
    n = (geom.nx, geom.ny, geom.nz)
    irs = np.array([math.ceil((coord[k] + h[k]) / d[k]) - nrough[k] // 2 + 1 for k in range(3)])
    ire = irs + nrough - 1
    edge = bool(irs.min() < 1 or ire[0] > n[0] or ire[1] > n[1] or ire[2] > n[2])
    if edge and not geom.periodic:
        raise ValueError("non periodic, local grid cannot got outside box")
    return irs, edge


# ---------------------------------------------------------------------------
# grid-sized work on the GPU

def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def box_projectors(ds: PawDataset, geom: Geometry, coord: np.ndarray, irs: np.ndarray
                   ) -> Tuple[np.ndarray, np.ndarray]:
    """get_local_p3d (src/p3d.f90:2-114) for one atom on the GPU: ptilde and
    ptilde1 times Y_lm sampled on the fine grid, contracted onto the rough
    box with the B matrices.  Returns (p3d, p3d1), each (ms, nr3, nr2, nr1)."""
    nrough, nfine = box_shape(ds, geom)
    d = (geom.dx, geom.dy, geom.dz)
    bs, dxf = [], []
    for k in range(3):
        b, f = b_matrix(int(nrough[k]), int(nfine[k]), d[k])
        bs.append(np.ascontiguousarray(b.T))          # (nfine, nrough): Fortran b(nrough, nfine)
        dxf.append(f)
    h = geom.half
    corner = np.array([-h[k] + (irs[k] - 1) * d[k] for k in range(3)])
    rad = np.ascontiguousarray(ds.ptilde[ds.ms_ls])
    rad1 = np.ascontiguousarray(ds.ptilde1[ds.ms_ls])
    lm = np.ascontiguousarray(np.stack([ds.lstate[ds.ms_ls], ds.mstate], axis=1).astype(np.int32))
    ms = ds.ms
    p3d = np.zeros((ms, nrough[2], nrough[1], nrough[0]))
    p3d1 = np.zeros_like(p3d)
    lib.call("opaw_paw_box", ds.nr, _dp(np.ascontiguousarray(ds.rr)), ms, _dp(rad), _dp(rad1), _ip(lm),
             _ip(np.ascontiguousarray(nrough, dtype=np.int32)), _ip(np.ascontiguousarray(nfine, dtype=np.int32)),
             _dp(np.array(dxf)), _dp(corner), _dp(np.ascontiguousarray(coord, dtype=np.float64)),
             _dp(np.array(h)), _dp(bs[0]), _dp(bs[1]), _dp(bs[2]), _dp(p3d), _dp(p3d1))
    return p3d, p3d1


def sphere(ds: PawDataset, geom: Geometry, coord: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """map_localgrid + get_localg3d (src/map_localgrid.f90:4-58,
    src/g3d.f90:1-119) for one atom on the GPU: the grid points with
    r < rcut in grid order and g_L(r) Y_L on them.  Returns (flat 0-based
    grid indices (ngrid,), g3d (ngrid, nl2))."""
    cap = C.c_int()
    gl = np.ascontiguousarray(ds.gl.T)                 # (2nl-1, nrcomp)
    args = (geom.nx, geom.ny, geom.nz, geom.dx, geom.dy, geom.dz, _dp(np.ascontiguousarray(coord, dtype=np.float64)),
            ds.rcut, ds.rcomp, ds.nr, _dp(np.ascontiguousarray(ds.rr)), ds.nl2, ds.nrcomp, _dp(gl))
    lib.call("opaw_paw_sphere", *args, 0, C.byref(cap), None, None)
    ng = cap.value
    got = C.c_int()
    grid = np.zeros((3, max(ng, 1)), np.int32)          # local_grid(ngrid, 3)
    g3d = np.zeros((ds.nl2, max(ng, 1)))                # local_g3d(ngrid, nl2)
    lib.call("opaw_paw_sphere", *args, ng, C.byref(got), _ip(grid), _dp(g3d))
    assert got.value == ng
    ix, iy, iz = (grid[k, :ng].astype(np.int64) - 1 for k in range(3))
    return ix + geom.nx * (iy + geom.ny * iz), np.ascontiguousarray(g3d[:, :ng].T)


def transform_p3d(p3d1: np.ndarray, sij: np.ndarray, dv: float
                  ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """transform_p3d (src/transform_p3d.f90): the box projectors p1 made
    orthonormal on the grid (L^-1/2, L = <p1_i|p1_j> dv), then rotated to
    diagonalise L^1/2 s L^1/2; returns (p3d1, s, sinv) with s its
    eigenvalues and sinv = 1/(1+s) - 1."""
    ms = p3d1.shape[0]
    flat = p3d1.reshape(ms, -1)
    L = (flat @ flat.T) * dv
    eig, vec = _sym_eig(L)
    sq = np.sign(eig) * np.sqrt(np.abs(eig))
    m1 = (vec / sq) @ vec.T
    m2 = (vec * sq) @ vec.T
    flat = m1 @ flat
    mat = m2.T @ sij @ m2
    eig, vec = _sym_eig(mat)
    flat = vec.T @ flat
    s = eig.copy()
    sinv = 1.0 / (1.0 + eig) - 1.0
    return flat.reshape(p3d1.shape), s, sinv


# ---------------------------------------------------------------------------
# a whole system from the datasets

def read_radial_block(path: str, tag: str, nr: int) -> np.ndarray:
    """The nr values after the first line containing tag (as searchstring
    followed by a list-directed read)."""
    f = _Lines(Path(path).read_text())
    f.search(tag)
    return f.values(nr)


def kinetic_differences(ds: PawDataset) -> np.ndarray:
    """<kinetic_energy_differences> (ns x ns) of the dataset."""
    return read_radial_block(ds.filename, "<kinetic_energy_differences>", ds.ns * ds.ns).reshape(ds.ns, ds.ns)


def vloc_tuma(geom: Geometry, symbols: Sequence[str], coords: np.ndarray, datasets: Dict[str, PawDataset]
              ) -> np.ndarray:
    """vloc_tuma_prep (src/vloc/vloc_prep.f90:1-79) on the GPU for a
    nonperiodic grid: the datasets' local potentials (vloc, already times
    Y_00) placed at the atoms through the doubled-box transform of
    make_cvkb / add_formfactor.  Returns vloc_tot (nx*ny*nz)."""
    if geom.periodic:
        raise ValueError("vloc_tuma_prep is the nonperiodic path (periodic runs use libpaw's atm2fft)")
    tsyms, amap = type_map(symbols)
    nr = np.array([datasets[t].nr for t in tsyms], np.int32)
    rr = np.ascontiguousarray(np.concatenate([datasets[t].rr for t in tsyms]))
    vr = np.ascontiguousarray(np.concatenate([datasets[t].vloc for t in tsyms]))
    types = np.ascontiguousarray(amap + 1, dtype=np.int32)
    xyz = np.ascontiguousarray(np.asarray(coords, dtype=np.float64))
    out = np.zeros(geom.nx * geom.ny * geom.nz)
    L = lib.load()
    ctx = C.c_void_p()
    lib.call("opaw_create", C.byref(ctx), geom.nx, geom.ny, geom.nz, geom.dx, geom.dy, geom.dz, 0, 0, 1, 1)
    try:
        lib.call("opaw_vloc_tuma", ctx, len(tsyms), _ip(nr), _dp(rr), _dp(vr), len(symbols), _ip(types), _dp(xyz),
                 _dp(out))
    finally:
        L.opaw_destroy(ctx)
    return out


def system_from_datasets(geom: Geometry, symbols: Sequence[str], coords: np.ndarray,
                         datasets: Dict[str, PawDataset], nocc: int, name: str = "paw",
                         dt: float = 0.05, sm: float = 1e-3, ipol: int = 0):
    """The tables of prep_paw and prep_libpaw for every atom and type.

    OPAW's own tables (prep_paw, src/prepare_paw.f90): box projectors and the
    compensation spheres on the GPU, transform_p3d on the host, and for a
    nonperiodic grid the local potential of vloc_tuma_prep on the GPU
    (vloc_tuma).  libpaw's (prep_libpaw, src/prepare_libpaw.f90, restated in
    opaw_amd.libpaw): dij0, eijkl and the radial tables of the on-site
    D_ij, the D^hat spheres, and from atm2fft the pseudo-core density ncoret
    and the starting density (extra["dens0"], the superposition of the
    neutralised atomic pseudo valence densities, src/vloc.f90:24-33).  A
    periodic grid keeps one stand-in, listed in extra["standins"]: v_loc
    sampled from the radial tables, where the reference uses atm2fft's
    local-potential branch."""
    from . import libpaw, synth
    tsyms, amap = type_map(symbols)
    dv = geom.dx * geom.dy * geom.dz
    nl2 = max(datasets[s].nl2 for s in tsyms)
    lp = {s: libpaw.load_type(datasets[s].filename) for s in tsyms}
    types = []
    for s in tsyms:
        ds, t = datasets[s], lp[s]
        if ds.nl2 != nl2:
            raise ValueError("every type needs the same number of compensation channels")
        if t.lmn_size != ds.ms or any(t.lstate[ds.ms_ls[i]] != ds.lstate[ds.ms_ls[i]] for i in range(ds.ms)):
            raise ValueError(f"{s}: libpaw's lmn order differs from readpaw's m-states")
        types.append(synth.TypeSpec(ms=ds.ms, qijlm=ds.qijlm, dij0=t.unpack(t.dij0), eijkl=t.eijkl_full(),
                                    onsite=libpaw.onsite_spec(t)))
    atoms = []
    for ia, (sym, c) in enumerate(zip(symbols, coords)):
        ds = datasets[sym]
        nrough, _ = box_shape(ds, geom)
        irs, _ = ir_start(c, nrough, geom)
        p3d, p3d1 = box_projectors(ds, geom, c, irs)
        p1, s, sinv = transform_p3d(p3d1, ds.sij, dv)
        idx, g3d = sphere(ds, geom, c)
        didx, dg = libpaw.dij_sphere(geom.nx, geom.ny, geom.nz, geom.dx, geom.dy, geom.dz, c, lp[sym])
        atoms.append(synth.AtomSpec(itype=int(amap[ia]), coord=np.asarray(c, dtype=np.float64), start=irs - 1,
                                    nr=nrough, p=p3d, p1=p1, s=s, sinv=sinv, sph_idx=idx, sph_g=g3d,
                                    dsph_idx=didx, dsph_g=dg))
    ncore, dens0 = libpaw.grid_densities(geom.nx, geom.ny, geom.nz, geom.dx, geom.dy, geom.dz,
                                         np.asarray(coords, dtype=np.float64), amap, [lp[t] for t in tsyms])
    standins: Tuple[str, ...] = ()
    if geom.periodic:
        h = geom.half
        xs = [-h[k] + np.arange(n) * d for k, (n, d) in enumerate(((geom.nx, geom.dx), (geom.ny, geom.dy),
                                                                  (geom.nz, geom.dz)))]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        vloc = np.zeros(Z.shape)
        for sym, c in zip(symbols, coords):
            ds = datasets[sym]
            d2 = 0.0
            for v, ck, hk in ((X, c[0], h[0]), (Y, c[1], h[1]), (Z, c[2], h[2])):
                e = np.abs(v - ck)
                e = np.minimum(e, np.abs(2.0 * hk - e))
                d2 = d2 + e * e
            vloc += np.interp(np.sqrt(d2), ds.rr, ds.vloc)
        vloc = vloc.reshape(-1)
        standins = ("vloc",)
    else:
        vloc = vloc_tuma(geom, symbols, coords, datasets)
    from .synth import ekcut_for
    return synth.SystemSpec(name=name, nx=geom.nx, ny=geom.ny, nz=geom.nz, dx=geom.dx, dy=geom.dy, dz=geom.dz,
                            periodic=geom.periodic, ekcut=ekcut_for(geom.dx, geom.dy, geom.dz), atoms=atoms,
                            types=types, vloc=vloc, ncoret=ncore, rnel=2.0 * nocc,
                            nocc=nocc, dt=dt, sm=sm, ipol=ipol,
                            extra=dict(standins=standins, type_symbols=tsyms, dens0=dens0,
                                       natom_c=sum(1 for s in symbols if s == "C"),
                                       natom_h=sum(1 for s in symbols if s == "H"), N=geom.nx * geom.ny * geom.nz))
