"""Host side of the libpaw set-up that the on-site D_ij needs.

The reference computes D_ij with ABINIT's libpaw (src/get_dij.f90:14-30:
pawdenpot, then pawdij), after a set-up in prep_libpaw
(src/prepare_libpaw.f90:1-225): pawpsp_17in reads each JTH XML dataset, and
pawinit builds the per-type tables.  This module restates that set-up for
the case the reference runs (LDA ixc=7, pawxcdev=1, usexcnhat=0, nspden=1,
no spin-orbit, no PAW+U, no Fock, usepotzero=0):

  radial meshes       pawrad_init, simp_gen, pawrad_deducer0, pawrad_ifromr,
                      poisson, bound_deriv (src/libpaw/1_m_pawrad.F90)
  dataset blocks      pawpsp_17in, XML branch (src/libpaw/3_m_pawpsp.F90:2967-3801)
  dij0                atompaw_dij0, atompaw_vhnzc, atompaw_shpfun
                      (src/libpaw/2_m_paw_atom.F90:83-530)
  per-type tables     pawinit (src/libpaw/m_paw_init.F90:165-652): indlmn,
                      indklmn, shape functions, phiphj, q_ij^L, eijkl
  real Gaunt          realgaunt (src/libpaw/1_m_pawang.F90:690-832)
  q-space densities   pawpsp_cg, paw_spline (src/libpaw/3_m_pawpsp.F90:492-682,
                      src/libpaw/1_m_paw_numeric.F90:82-318), the valence
                      compensation of pawpsp_calc (3_m_pawpsp.F90:2000-2050)
  grid densities      atm2fft (src/libpaw/m_atm2fft.F90) with getph, zerosym,
                      getcut/bound (gsqcut), called from get_vloc_ncoret
                      (src/vloc.f90:24-33): the pseudo-core density ncoret and
                      the starting density (atomic superposition)
  D^hat spheres       nhatgrid/pawrfgd_fft + pawgylm (src/libpaw/finegrid.f90,
                      src/libpaw/2_m_paw_finegrid.F90:92-1368), optcut=0

All of this is set-up work that runs once per run on the host (the reference
runs it once on rank 0); the per-rebuild on-site terms run on the GPU from
these tables.  Numbers are kept in the reference's precision: the double
literals of pawinit (6.28318530718d0, sqrt(12.5663706144d0)) and getph
(6.28318530717959d0) are used as written, and invcb's single-precision
exponent is in the GPU and oracle restatements of xcpbe.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

PI = math.pi                       # m_libpaw_defs: pi to 40 digits
FOUR_PI = 4.0 * PI
SQFPI = math.sqrt(FOUR_PI)
TWO_PI_INIT = 6.28318530718        # src/libpaw/m_paw_init.F90:592 (eijkl)
SQ4PI_INIT = math.sqrt(12.5663706144)   # src/libpaw/m_paw_init.F90:465 (sij)
TWO_PI_PH = 6.28318530717959       # src/libpaw/getph.f90 (structure factors)
TOL8, TOL10, TOL12, TOL14, TOL16 = 1e-8, 1e-10, 1e-12, 1e-14, 1e-16
XC_DENPOS = 1e-14                  # src/get_dij.f90:20 (xc_denpos passed to pawdenpot)
LCUTDENS = 10                      # src/1_libpaw_mod.f90:50
MQGRID_VL = 3001                   # src/1_libpaw_mod.f90:48


# ---------------------------------------------------------------------------
# radial mesh r(i) = a (exp(d (i-1)) - 1), libpaw mesh_type 2

class RadMesh:
    """pawrad_init (src/libpaw/1_m_pawrad.F90:198-315) for mesh_type 2, the
    JTH grid "r=a*(exp(d*i)-1)": rad, radfact = dr/di / d, the Simpson
    weights simfact up to int_meshsz (the point nearest r_for_intg)."""

    def __init__(self, a: float, d: float, mesh_size: int, r_for_intg: Optional[float] = None):
        self.rstep = float(a)
        self.lstep = float(d)
        self.stepint = self.lstep
        self.mesh_size = int(mesh_size)
        i = np.arange(self.mesh_size, dtype=np.float64)
        self.rad = self.rstep * (np.exp(self.lstep * i) - 1.0)
        self.rad[0] = 0.0
        self.radfact = self.rad + self.rstep
        self.radfact[0] = self.rstep
        self.int_meshsz = self.mesh_size
        if r_for_intg is not None and r_for_intg > 0.0:
            self.int_meshsz = self._nearest(r_for_intg)
        self.simfact = self._simfact(self.int_meshsz)
        self.rmax = self.rad[-1]

    def copy_to(self, mesh_size: int, r_for_intg: Optional[float] = None) -> "RadMesh":
        return RadMesh(self.rstep, self.lstep, mesh_size, r_for_intg)

    def ifromr(self, rr: float) -> int:
        """pawrad_ifromr, 1-based."""
        return int(TOL8 + math.log(1.0 + rr / self.rstep) / self.lstep) + 1

    def _nearest(self, r: float) -> int:
        ir = min(self.ifromr(r), self.mesh_size)
        rad = self.rad
        if ir < self.mesh_size and abs(rad[ir] - r) < abs(rad[ir - 1] - r):
            ir += 1
        if ir > 1 and abs(rad[ir - 2] - r) < abs(rad[ir - 1] - r):
            ir -= 1
        return ir

    def _simfact(self, nint: int) -> np.ndarray:
        hh = self.stepint / 3.0
        sf = np.zeros(self.mesh_size)
        sf[nint - 1] = hh * self.radfact[nint - 1]
        isim = 3
        ir_last = 1
        for ir in range(nint, isim - 1, -2):              # Fortran do ir=nint,isim,-2
            sf[ir - 2] = 4.0 * hh * self.radfact[ir - 2]
            sf[ir - 3] = 2.0 * hh * self.radfact[ir - 3]
            ir_last = ir - 2
        sf[ir_last - 1] = 0.5 * sf[ir_last - 1]
        sf[nint:] = 0.0
        return sf

    def simp(self, func: np.ndarray, r_for_intg: Optional[float] = None) -> float:
        """simp_gen (src/libpaw/1_m_pawrad.F90:947-1041)."""
        func = np.asarray(func, dtype=np.float64)
        if r_for_intg is not None:
            nn = self._nearest(r_for_intg) if r_for_intg > 0.0 else self.mesh_size
            sf = self._simfact(nn)
        else:
            nn = self.int_meshsz
            sf = self.simfact
        if nn > len(func):
            raise ValueError("simp_gen: int_meshsz > size(func)")
        simp = 0.0
        for ii in range(nn):                   # the reference's sequential sum
            simp += func[ii] * sf[ii]
        return simp + self._resid(func, nn)

    def _resid(self, func: np.ndarray, nn: int) -> float:
        if nn % 2 == 0:
            rf = self.radfact
            return self.stepint / 3.0 * (1.25 * func[0] * rf[0] + 2.0 * func[1] * rf[1] - 0.25 * func[2] * rf[2])
        return 0.0

    def weights(self) -> np.ndarray:
        """simp_gen over int_meshsz as one weight per point (the end
        correction of an even point count folded into the first three)."""
        w = self.simfact.copy()
        if self.int_meshsz % 2 == 0:
            rf = self.radfact
            h = self.stepint / 3.0
            w[0] += h * 1.25 * rf[0]
            w[1] += h * 2.0 * rf[1]
            w[2] += -h * 0.25 * rf[2]
        return w

    @staticmethod
    def deducer0(func: np.ndarray) -> np.ndarray:
        """pawrad_deducer0 for mesh_type 2: f(1) = f(4) + 3 (f(2) - f(3))."""
        f = np.array(func, dtype=np.float64, copy=True)
        f[0] = f[3] + 3.0 * (f[1] - f[2])
        return f

    def poisson(self, den: np.ndarray, ll: int) -> np.ndarray:
        """poisson (src/libpaw/1_m_pawrad.F90:1356-1577), non-Numerov branch
        (mesh_type 2), no screening: r * V_l(r) of the radial density den
        (den = 4 pi r^2 n, or a product of u functions)."""
        den = np.asarray(den, dtype=np.float64)
        msz = len(den)
        if msz > self.mesh_size:
            raise ValueError("poisson: wrong sizes")
        rad, rf = self.rad, self.radfact
        nn = msz
        hh = self.stepint / 3.0
        while abs(den[nn - 1]) < TOL16 and nn > self.int_meshsz:
            nn -= 1
        mm = nn
        aa = np.zeros(nn + 1)       # 1-based work arrays, index 0 unused
        bb = np.zeros(nn + 1)
        cc = np.zeros(nn + 1)
        dd = np.zeros(nn + 1)
        radl = np.zeros(nn + 1)
        radl1 = np.zeros(nn + 1)
        for jr in range(nn, 1, -1):
            ir = nn - jr + 1
            radl[jr] = rad[jr - 1] ** ll
            radl1[jr] = rad[jr - 1] * radl[jr]
            aa[ir] = den[jr - 1] * rf[jr - 1] * radl[jr]
            bb[ir] = den[jr - 1] * rf[jr - 1] / radl1[jr]
        radl[1] = 0.0
        radl1[1] = 0.0
        ee = np.array([0.0, aa[nn - 1], aa[nn - 2], aa[nn - 3]])
        aa[nn] = ee[3] + 3.0 * (ee[1] - ee[2])
        ee = np.array([0.0, bb[nn - 1], bb[nn - 2], bb[nn - 3]])
        bb[nn] = ee[3] + 3.0 * (ee[1] - ee[2])
        cc[1] = 0.0
        dd[1] = 0.0
        for ir in range(3, mm + 1, 2):
            cc[ir] = cc[ir - 2] + hh * (aa[ir - 2] + 4.0 * aa[ir - 1] + aa[ir])
            cc[ir - 1] = cc[ir - 2] + hh * (1.25 * aa[ir - 2] + 2.0 * aa[ir - 1] - 0.25 * aa[ir])
            dd[ir] = dd[ir - 2] + hh * (bb[ir - 2] + 4.0 * bb[ir - 1] + bb[ir])
            dd[ir - 1] = dd[ir - 2] + hh * (1.25 * bb[ir - 2] + 2.0 * bb[ir - 1] - 0.25 * bb[ir])
        if mm % 2 == 0:
            cc[mm] = cc[mm - 1] + hh * (1.25 * aa[mm - 2] + 2.0 * aa[mm - 1] - 0.25 * aa[mm])
            dd[mm] = dd[mm - 1] + hh * (1.25 * bb[mm - 2] + 2.0 * bb[mm - 1] - 0.25 * bb[mm])
        rv = np.zeros(msz)
        for ir in range(2, nn + 1):
            jr = nn - ir + 1
            rv[ir - 1] = (dd[jr] * radl1[ir] + (cc[nn] - cc[jr]) / radl[ir]) / (2.0 * ll + 1.0)
        if nn < msz:
            rv[nn:] = rv[nn - 1]
        return rv

    def bound_deriv(self, func: np.ndarray, nn: int) -> Tuple[float, float]:
        """bound_deriv (src/libpaw/1_m_pawrad.F90:1291-1320)."""
        f = func
        rf = self.radfact
        yp1 = 1.0 / 12.0 / self.stepint / rf[0] * (-25.0 * f[0] + 48.0 * f[1] - 36.0 * f[2] + 16.0 * f[3]
                                                    - 3.0 * f[4])
        ypn = 1.0 / 12.0 / self.stepint * (3.0 * f[nn - 5] - 16.0 * f[nn - 4] + 36.0 * f[nn - 3] - 48.0 * f[nn - 2]
                                           + 25.0 * f[nn - 1]) / rf[nn - 1]
        return yp1, ypn


def paw_spline(t: np.ndarray, y: np.ndarray, ybcbeg: float, ybcend: float) -> np.ndarray:
    """paw_spline (src/libpaw/1_m_paw_numeric.F90:82-223): second
    derivatives of the cubic spline with end slopes ybcbeg/ybcend (natural
    when > 1e30)."""
    n = len(t)
    ypp = np.zeros(n)
    tmp = np.zeros(n)
    tiny = np.finfo(np.float64).tiny
    if ybcbeg > 1.0e30:
        ypp[0] = 0.0
        tmp[0] = 0.0
    else:
        ypp[0] = -0.5
        tmp[0] = (3.0 / (t[1] - t[0])) * ((y[1] - y[0]) / (t[1] - t[0]) - ybcbeg)
    if ybcend > 1.0e30:
        ypp[n - 1] = 0.0
        tmp[n - 1] = 0.0
    else:
        ypp[n - 1] = 0.5
        tmp[n - 1] = (3.0 / (t[n - 1] - t[n - 2])) * (ybcend - (y[n - 1] - y[n - 2]) / (t[n - 1] - t[n - 2]))
    for i in range(1, n - 1):
        ratio = (t[i] - t[i - 1]) / (t[i + 1] - t[i - 1])
        pinv = 1.0 / (ratio * ypp[i - 1] + 2.0)
        ypp[i] = (ratio - 1.0) * pinv
        tmp[i] = (6.0 * ((y[i + 1] - y[i]) / (t[i + 1] - t[i]) - (y[i] - y[i - 1]) / (t[i] - t[i - 1]))
                  / (t[i + 1] - t[i - 1]) - ratio * tmp[i - 1]) * pinv
        if abs(tmp[i]) < 1.0e5 * tiny:
            tmp[i] = 0.0
    ypp[n - 1] = (tmp[n - 1] - ypp[n - 1] * tmp[n - 2]) / (ypp[n - 1] * ypp[n - 2] + 1.0)
    for k in range(n - 2, -1, -1):
        ypp[k] = ypp[k] * ypp[k + 1] + tmp[k]
    return ypp


# ---------------------------------------------------------------------------
# real Gaunt coefficients

def real_ylm_values(lmax: int, ux: np.ndarray, uy: np.ndarray, uz: np.ndarray) -> np.ndarray:
    """Real spherical harmonics S_lm of a unit vector, l <= lmax (<= 4), in
    libpaw's order (lm = l^2 + l + m + 1) and sign convention (S_l,m>0 =
    sqrt2 (-1)^m Re Y_lm, S_l,m<0 = sqrt2 (-1)^m Im Y_l|m|: S_1,-1..1 are
    y, z, x times sqrt(3/4pi)), as initylmr makes them.  Built from the
    associated Legendre functions by recurrence."""
    ux, uy, uz = (np.asarray(v, dtype=np.float64) for v in (ux, uy, uz))
    ct = uz
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    phi = np.arctan2(uy, ux)
    out = np.zeros(ux.shape + ((lmax + 1) ** 2,))
    # P_l^m(cos t) without the Condon-Shortley phase
    P = {}
    P[(0, 0)] = np.ones_like(ct)
    for m in range(1, lmax + 1):
        P[(m, m)] = P[(m - 1, m - 1)] * (2 * m - 1) * st
    for m in range(0, lmax):
        P[(m + 1, m)] = ct * (2 * m + 1) * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * ct * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            nrm = math.sqrt((2 * l + 1) / FOUR_PI * math.factorial(l - am) / math.factorial(l + am))
            if m == 0:
                v = nrm * P[(l, 0)]
            elif m > 0:
                v = math.sqrt(2.0) * nrm * P[(l, am)] * np.cos(am * phi)
            else:
                v = math.sqrt(2.0) * nrm * P[(l, am)] * np.sin(am * phi)
            out[..., l * l + l + m] = v
    return out


def real_gaunt(lmax_out: int, lmax_in: int) -> np.ndarray:
    """G[lm3, lm1, lm2] = integral of S_lm1 S_lm2 S_lm3 over the sphere for
    l1, l2 <= lmax_in and l3 <= lmax_out: the values realgaunt stores in
    realgnt (src/libpaw/1_m_pawang.F90:690-832), entries below 1e-12 set to
    zero as its selection does.  Gauss-Legendre in cos(theta) times a
    uniform phi rule, exact for these polynomial degrees."""
    deg = lmax_out + 2 * lmax_in
    nth = deg // 2 + 2
    nph = deg + 2
    x, w = np.polynomial.legendre.leggauss(nth)
    ph = (np.arange(nph) + 0.5) * (2.0 * PI / nph)
    ct = np.repeat(x, nph)
    wt = np.repeat(w, nph) * (2.0 * PI / nph)
    st = np.sqrt(1.0 - ct * ct)
    pp = np.tile(ph, nth)
    lm = max(lmax_out, lmax_in)
    S = real_ylm_values(lm, st * np.cos(pp), st * np.sin(pp), ct)
    n3, n1 = (lmax_out + 1) ** 2, (lmax_in + 1) ** 2
    G = np.einsum("pa,pb,pc,p->cab", S[:, :n1], S[:, :n1], S[:, :n3], wt)
    G[np.abs(G) < TOL12] = 0.0
    return G


# ---------------------------------------------------------------------------
# dataset blocks (pawpsp_17in, XML branch)

_DATA_TAGS = ("ae_partial_wave", "pseudo_partial_wave", "projector_function", "ae_core_density",
              "pseudo_core_density", "pseudo_valence_density", "blochl_local_ionic_potential",
              "kinetic_energy_differences")


def _blocks(path: str) -> Tuple[str, Dict[str, List[Tuple[str, np.ndarray]]]]:
    """The numeric blocks of a PAW XML dataset, in file order per tag."""
    text = Path(path).read_text()
    out: Dict[str, List[Tuple[str, np.ndarray]]] = {}
    for tag in _DATA_TAGS:
        for m in re.finditer(r"<" + tag + r"\b([^>]*)>(.*?)</" + tag + ">", text, re.S):
            vals = m.group(2).replace("D", "E").replace("d", "e").split()
            out.setdefault(tag, []).append((m.group(1), np.array([float(v) for v in vals])))
    return text, out


def _attr(s: str, key: str) -> Optional[str]:
    m = re.search(r"\b" + re.escape(key) + r'\s*=\s*"\s*([^"]*?)\s*"', s)
    return m.group(1) if m else None


@dataclass
class LibpawType:
    """One dataset after pawpsp_17in and pawinit: the tables the on-site
    D_ij reads.  Radial arrays are on the partial-wave mesh; 0-based."""
    symbol: str
    znucl: float
    zion: float
    mesh: RadMesh                  # pawrad (partial waves), int_meshsz at rpaw
    core_mesh: RadMesh             # full grid, int_meshsz = mesh_size
    mesh_size: int                 # pawtab%mesh_size = ifromr(rpaw) + 5
    rpaw: float
    rshp: float
    lstate: np.ndarray             # (basis,) l of each basis function
    indlmn: np.ndarray             # (lmn, 3): l, m, ln (0-based ln)
    l_size: int
    lm_size: int                   # lcut_size^2
    phi: np.ndarray                # (basis, mmax) u = r R
    tphi: np.ndarray
    phiphj: np.ndarray             # (mesh_size, ij_size)
    tphitphj: np.ndarray
    coredens: np.ndarray           # (mesh_size,)
    tcoredens: np.ndarray          # (mesh_size,)
    usetcore: int
    ncore: np.ndarray              # (core mesh,) full ae core density / sqrt(4pi)
    tncore: np.ndarray
    tnvale: np.ndarray             # pseudo valence + compensation, / sqrt(4pi)
    vlocr: np.ndarray              # blochl local ionic potential / sqrt(4pi)
    shapefunc: np.ndarray          # (mesh_size, l_size)
    gnorm: np.ndarray              # (l_size,)
    kij: np.ndarray                # (lmn2,) packed
    dij0: np.ndarray               # (lmn2,) packed
    qijl: np.ndarray               # (l_size^2, lmn2)
    sij: np.ndarray                # (lmn2,)
    eijkl: np.ndarray              # (lmn2, lmn2) upper triangle (klmn <= klmn1) as pawinit fills it
    gaunt: np.ndarray              # G[lm, ilm, jlm] for lm < l_size^2
    indklmn: np.ndarray            # (lmn2, 4): klm (0-based pair), kln (0-based), lmin, lmax
    extra: dict = field(default_factory=dict)

    @property
    def lmn_size(self) -> int:
        return len(self.indlmn)

    @property
    def lmn2_size(self) -> int:
        n = self.lmn_size
        return n * (n + 1) // 2

    @property
    def basis_size(self) -> int:
        return len(self.lstate)

    @property
    def ij_size(self) -> int:
        b = self.basis_size
        return b * (b + 1) // 2

    # packed <-> matrix
    def pack_index(self) -> np.ndarray:
        """klmn (0-based) of every (ilmn, jlmn), symmetric."""
        n = self.lmn_size
        k = np.zeros((n, n), np.int64)
        for j in range(n):
            for i in range(j + 1):
                k[i, j] = k[j, i] = j * (j + 1) // 2 + i
        return k

    def unpack(self, packed: np.ndarray) -> np.ndarray:
        return np.asarray(packed)[self.pack_index()]

    def eijkl_full(self) -> np.ndarray:
        """E[i, j, k, l] with D^H_ij = sum_kl E[i,j,k,l] rho_kl over both
        triangles of rho (each off-diagonal counted twice = dltij), the
        symmetric use pawdijhartree makes of the stored upper triangle
        (src/libpaw/4_m_pawdij.F90:1270-1282)."""
        e = np.triu(self.eijkl) + np.triu(self.eijkl, 1).T
        K = self.pack_index()
        return e[K[:, :, None, None], K[None, None, :, :]]

    def gaunt_ij(self) -> np.ndarray:
        """(lm_size, lmn, lmn): realgnt of (lm, lm_i, lm_j) per m-state pair."""
        lmi = np.array([l * l + l + m for (l, m, _) in self.indlmn])
        return self.gaunt[:self.lm_size][:, lmi[:, None], lmi[None, :]]

    def kln_ij(self) -> np.ndarray:
        """(lmn, lmn) 0-based radial pair index of every m-state pair."""
        ln = np.array([x[2] for x in self.indlmn])
        a = np.minimum(ln[:, None], ln[None, :])
        b = np.maximum(ln[:, None], ln[None, :])
        return b * (b + 1) // 2 + a


def _shpfun(ll: int, mesh: RadMesh, size: int, rshp: float) -> Tuple[np.ndarray, float]:
    """atompaw_shpfun (src/libpaw/2_m_paw_atom.F90:83-180) for shape_type 2
    (sinc^2): g_l = [sin(pi r/rshp)/(pi r/rshp)]^2 r^l up to ishp, normalised
    so that simp_gen(g_l r^(2+l), r_for_intg=rshp) = 1.  Returns (g, norm)."""
    ishp = mesh.ifromr(rshp) - 1
    g = np.zeros(size)
    rad = mesh.rad
    g[0] = 1.0 if ll == 0 else 0.0
    for ir in range(2, ishp + 1):
        arg = PI * rad[ir - 1] / rshp
        g[ir - 1] = (math.sin(arg) / arg) ** 2 * (rad[ir - 1] ** ll if ll > 0 else 1.0)
    r2k = np.zeros(size)
    r2k[1:ishp] = g[1:ishp] * rad[1:ishp] ** (2 + ll)
    norm = 1.0 / mesh.simp(r2k, r_for_intg=rshp)
    g[:ishp] = g[:ishp] * norm
    return g, norm


def _vhnzc(ncore: np.ndarray, core_mesh: RadMesh, znucl: float) -> np.ndarray:
    """atompaw_vhnzc: Hartree potential of the core density minus Z/r."""
    nwk = ncore * FOUR_PI * core_mesh.rad ** 2
    v = core_mesh.poisson(nwk, 0)
    v[1:] = (v[1:] - znucl) / core_mesh.rad[1:len(v)]
    return RadMesh.deducer0(v)


def load_type(path: str) -> LibpawType:
    """pawpsp_17in (XML) + atompaw_dij0 + pawinit for one JTH dataset."""
    text, blk = _blocks(path)
    atom = re.search(r"<atom\b([^>]*)>", text).group(1)
    symbol = _attr(atom, "symbol")
    znucl = float(_attr(atom, "Z"))
    zion = float(_attr(atom, "valence"))
    rpaw = float(_attr(re.search(r"<paw_radius\b([^>]*)>", text).group(1), "rc"))
    shp = re.search(r"<shape_function\b([^>]*)>", text).group(1)
    if _attr(shp, "type") != "sinc":
        raise ValueError("only the sinc shape function (shape_type 2) of the JTH datasets is restated")
    rshp = float(_attr(shp, "rc"))
    grid = re.search(r"<radial_grid\b([^>]*)>", text).group(1)
    if _attr(grid, "eq") != "r=a*(exp(d*i)-1)":
        raise ValueError("only the JTH radial grid r=a*(exp(d*i)-1) is restated")
    a, d = float(_attr(grid, "a")), float(_attr(grid, "d"))
    istart, iend = int(_attr(grid, "istart")), int(_attr(grid, "iend"))
    mmax = iend - istart + 1 + (1 if istart == 1 else 0)
    states = re.findall(r"<state\b([^>]*)/>", text)
    lstate = np.array([int(_attr(s, "l")) for s in states])
    occ = np.array([float(_attr(s, "f")) if _attr(s, "f") is not None else 0.0 for s in states])
    nval = len(states)
    # pawpsp_17in: pawrad with int_meshsz at rpaw; mesh_size = ifromr(rpaw)+5
    mesh = RadMesh(a, d, mmax, r_for_intg=rpaw)
    mesh_size = min(mesh.ifromr(rpaw) + 5, mesh.mesh_size)
    core_mesh = RadMesh(a, d, mmax)
    rad = mesh.rad

    def arrays(tag):
        return [v for (_, v) in blk[tag]]

    phi = np.array([v[:mmax] * rad for v in arrays("ae_partial_wave")])
    tphi = np.array([v[:mmax] * rad for v in arrays("pseudo_partial_wave")])
    ncore = blk["ae_core_density"][0][1][:mmax] / SQFPI
    tncore = blk["pseudo_core_density"][0][1][:mmax] / SQFPI
    tnvale = blk["pseudo_valence_density"][0][1][:mmax] / SQFPI
    vlocr = blk["blochl_local_ionic_potential"][0][1][:mmax] / SQFPI
    kdiff = blk["kinetic_energy_differences"][0][1].reshape(nval, nval)
    usetcore = 0 if np.max(np.abs(tncore)) < 1e-6 else 1

    # indlmn: basis functions in file order, m = -l..l
    indlmn = []
    for ib, l in enumerate(lstate):
        for m in range(-l, l + 1):
            indlmn.append((int(l), m, ib))
    lmax = int(lstate.max())
    l_size = 2 * lmax + 1
    lcut_size = min(l_size, LCUTDENS + 1)
    lm_size = lcut_size * lcut_size
    lmn = len(indlmn)
    lmn2 = lmn * (lmn + 1) // 2
    G = real_gaunt(l_size - 1, lmax)                   # (l_size^2, (lmax+1)^2, (lmax+1)^2)
    lmidx = [l * l + l + m for (l, m, _) in indlmn]

    # kij and rhoij0 (pawpsp_17in)
    kij = np.zeros(lmn2)
    rhoij0 = np.zeros(lmn2)
    for j in range(lmn):
        for i in range(j + 1):
            k = j * (j + 1) // 2 + i
            if lmidx[i] == lmidx[j]:
                kij[k] = kdiff[indlmn[i][2], indlmn[j][2]]
    i0 = 0
    for ib, l in enumerate(lstate):
        nm = 2 * l + 1
        o = max(occ[ib], 0.0)
        for il in range(i0, i0 + nm):
            rhoij0[il * (il + 1) // 2 + il] = o / nm
        i0 += nm

    # atompaw_dij0 (vlocopt = 2: blochl local potential, opt_init = 0)
    meshsz = min(mesh.mesh_size, core_mesh.mesh_size, len(vlocr))
    dij0 = kij.copy()
    vhnzc = _vhnzc(ncore, core_mesh, znucl)
    vt = vlocr[:meshsz]
    for j in range(lmn):
        for i in range(j + 1):
            if lmidx[i] != lmidx[j]:
                continue
            k = j * (j + 1) // 2 + i
            il, jl = indlmn[i][2], indlmn[j][2]
            dij0[k] += mesh.simp(phi[il, :meshsz] * phi[jl, :meshsz] * vhnzc[:meshsz])
    for j in range(lmn):
        for i in range(j + 1):
            if lmidx[i] != lmidx[j]:
                continue
            k = j * (j + 1) // 2 + i
            il, jl = indlmn[i][2], indlmn[j][2]
            dij0[k] -= mesh.simp(tphi[il, :meshsz] * tphi[jl, :meshsz] * vt)
    shpf0, _ = _shpfun(0, mesh, mesh.mesh_size, rshp)
    intvh = mesh.simp(vt * shpf0[:meshsz] * rad[:meshsz] ** 2)
    for j in range(lmn):
        for i in range(j + 1):
            if lmidx[i] != lmidx[j]:
                continue
            k = j * (j + 1) // 2 + i
            il, jl = indlmn[i][2], indlmn[j][2]
            ff = phi[il, :meshsz] * phi[jl, :meshsz] - tphi[il, :meshsz] * tphi[jl, :meshsz]
            dij0[k] -= intvh * mesh.simp(ff)

    # pawpsp_calc: the valence density made neutral with the l=0 shape
    # function (the testval branch: tnvale negligible at the mesh end)
    vale = core_mesh
    if abs(tnvale[-1]) >= 1e-9:
        raise ValueError("valence density not negligible at the end of its mesh (rhoij0 branch not restated)")
    qq = zion / FOUR_PI - vale.simp(tnvale * vale.rad ** 2)
    nh, _ = _shpfun(0, vale, vale.mesh_size, rshp)
    tnvale_c = tnvale + qq * nh

    # pawinit: shape functions, phiphj, qijl, sij, eijkl
    shapefunc = np.zeros((mesh_size, l_size))
    gnorm = np.zeros(l_size)
    for il in range(l_size):
        g, nrm = _shpfun(il, mesh, mesh_size, rshp)
        shapefunc[:, il] = g
        gnorm[il] = nrm
    basis = len(lstate)
    ij_size = basis * (basis + 1) // 2
    phiphj = np.zeros((mesh_size, ij_size))
    tphitphj = np.zeros((mesh_size, ij_size))
    for jln in range(basis):
        for iln in range(jln + 1):
            kln = jln * (jln + 1) // 2 + iln
            phiphj[:, kln] = phi[iln, :mesh_size] * phi[jln, :mesh_size]
            tphitphj[:, kln] = tphi[iln, :mesh_size] * tphi[jln, :mesh_size]
    indklmn = np.zeros((lmn2, 4), np.int64)
    klm_diag = np.zeros(lmn2, bool)
    for j in range(lmn):
        jl, jln = indlmn[j][0], indlmn[j][2]
        jlm = lmidx[j]
        for i in range(j + 1):
            il, iln = indlmn[i][0], indlmn[i][2]
            ilm = lmidx[i]
            k = j * (j + 1) // 2 + i
            lo, hi = min(ilm, jlm), max(ilm, jlm)
            indklmn[k, 0] = hi * (hi + 1) // 2 + lo
            lo, hi = min(iln, jln), max(iln, jln)
            indklmn[k, 1] = hi * (hi + 1) // 2 + lo
            indklmn[k, 2] = min(abs(il - jl), LCUTDENS)
            indklmn[k, 3] = min(il + jl, LCUTDENS)
            klm_diag[k] = ilm == jlm
    # (i, j) m-states of each klmn, for Gaunt lookups
    pair_of = np.zeros((lmn2, 2), np.int64)
    for j in range(lmn):
        for i in range(j + 1):
            pair_of[j * (j + 1) // 2 + i] = (i, j)

    def gnt(lm3, k):
        i, j = pair_of[k]
        return G[lm3, lmidx[i], lmidx[j]]

    qijl = np.zeros((l_size * l_size, lmn2))
    sij = np.zeros(lmn2)
    for k in range(lmn2):
        kln = indklmn[k, 1]
        for ll in range(indklmn[k, 2], indklmn[k, 3] + 1, 2):
            ff = np.zeros(mesh_size)
            ff[1:] = (phiphj[1:, kln] - tphitphj[1:, kln]) * rad[1:mesh_size] ** ll
            intg = mesh.simp(ff)
            for mm in range(-ll, ll + 1):
                g = gnt(ll * ll + ll + mm, k)
                if g != 0.0:
                    qijl[ll * ll + ll + mm, k] = intg * g
        if klm_diag[k]:
            sij[k] = qijl[0, k] * SQ4PI_INIT
    # eijkl (Hartree), pawinit section 6
    vhatijl = np.zeros((lmn2, l_size))
    intvhatl = np.zeros(l_size)
    for il in range(l_size):
        ff = np.zeros(mesh_size)
        ff[1:] = shapefunc[1:, il] * rad[1:mesh_size] ** 2
        vhatl = mesh.poisson(ff, il)
        vhatl[0] = 0.0
        vhatl[1:] = 2.0 * vhatl[1:] / rad[1:mesh_size]
        intvhatl[il] = mesh.simp(vhatl * ff)
        for k in range(lmn2):
            vhatijl[k, il] = mesh.simp(vhatl * tphitphj[:, indklmn[k, 1]])
    eijkl = np.zeros((lmn2, lmn2))
    meshsz = mesh.int_meshsz
    ff = np.zeros(mesh_size)
    for k in range(lmn2):
        kln = indklmn[k, 1]
        for ll in range(indklmn[k, 2], indklmn[k, 3] + 1, 2):
            lm0 = ll * ll + ll
            ff[:meshsz] = phiphj[:meshsz, kln]
            ff[meshsz:] = 0.0
            gg = mesh.poisson(ff, ll)
            ff[:meshsz] = tphitphj[:meshsz, kln]
            hh = mesh.poisson(ff, ll)
            for k1 in range(k, lmn2):
                kln1 = indklmn[k1, 1]
                vh1 = 0.0
                if indklmn[k1, 2] <= ll <= indklmn[k1, 3]:
                    ff[0] = 0.0
                    ff[1:meshsz] = (phiphj[1:meshsz, kln1] * gg[1:meshsz] - tphitphj[1:meshsz, kln1] * hh[1:meshsz]
                                    ) * 2.0 / rad[1:meshsz]
                    vh1 = mesh.simp(ff)
                for mm in range(-ll, ll + 1):
                    rg, rg1 = gnt(lm0 + mm, k), gnt(lm0 + mm, k1)
                    if rg != 0.0 and rg1 != 0.0:
                        ql, ql1 = qijl[lm0 + mm, k], qijl[lm0 + mm, k1]
                        eijkl[k, k1] += (vh1 * rg * rg1 - vhatijl[k, ll] * rg * ql1 - vhatijl[k1, ll] * rg1 * ql
                                         - intvhatl[ll] * ql * ql1) * TWO_PI_INIT
    coredens = ncore[:mesh_size].copy()
    tcoredens = tncore[:mesh_size].copy() if usetcore else np.zeros(mesh_size)
    return LibpawType(symbol=symbol, znucl=znucl, zion=zion, mesh=mesh, core_mesh=core_mesh, mesh_size=mesh_size,
                      rpaw=rpaw, rshp=rshp, lstate=lstate, indlmn=np.array(indlmn, np.int64), l_size=l_size,
                      lm_size=lm_size, phi=phi, tphi=tphi, phiphj=phiphj, tphitphj=tphitphj, coredens=coredens,
                      tcoredens=tcoredens, usetcore=usetcore, ncore=ncore, tncore=tncore, tnvale=tnvale_c,
                      vlocr=vlocr, shapefunc=shapefunc, gnorm=gnorm, kij=kij, dij0=dij0, qijl=qijl, sij=sij,
                      eijkl=eijkl, gaunt=G, indklmn=indklmn,
                      extra=dict(rhoij0=rhoij0, vhnzc=vhnzc, intvh=intvh))


# ---------------------------------------------------------------------------
# grid-side set-up: gsqcut, q-space densities, atm2fft, D^hat spheres

def gsqcut_of(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, ek_factor: float = 1.0) -> float:
    """getcut with iboxcut = 0 (src/libpaw/getcut.f90) after bound
    (src/libpaw/bound.f90): boxsq is the smallest |G|^2 on the faces of the
    FFT box, ecut = 1/2 (2pi)^2 boxsq (ek_factor - 1e-10)."""
    L = (nx * dx, ny * dy, nz * dz)
    n = (nx, ny, nz)
    boxsq = min((n[k] // 2) ** 2 / L[k] ** 2 for k in range(3))
    two_pi = 2.0 * PI
    effcut = 0.5 * two_pi ** 2 * boxsq
    ecut = effcut * (ek_factor - 1e-10)
    sphsq = 2.0 * ecut / two_pi ** 2
    boxcut = math.sqrt(boxsq / sphsq)
    cutrad = min(2.0, boxcut)
    return cutrad ** 2 * (2.0 * ecut) / two_pi ** 2


def qgrid_vl(gsqcut: float) -> np.ndarray:
    """qgrid_vl of set_param_alloc (src/prepare_libpaw.f90:211-215)."""
    qmax = 1.2 * math.sqrt(gsqcut)
    dq = qmax / float(MQGRID_VL - 1)
    return np.arange(MQGRID_VL) * dq


def pawpsp_cg(qgrid: np.ndarray, mesh: RadMesh, nr: np.ndarray) -> Tuple[np.ndarray, float, float, float]:
    """pawpsp_cg (src/libpaw/3_m_pawpsp.F90:492-682): n(q) = 4pi int r^2
    n(r) j0(2 pi q r) dr with an exponential tail beyond the mesh; returns
    (nq, dnqdq0, yp1, ypn)."""
    msz = min(len(nr), mesh.mesh_size)
    rad = mesh.rad[:msz]
    nr = np.asarray(nr[:msz], dtype=np.float64)
    rnr = rad * nr
    rm = rad[-1]
    rf = mesh.radfact
    dn = 1.0 / (12.0 * mesh.stepint * rf[msz - 1]) * (3.0 * nr[msz - 5] - 16.0 * nr[msz - 4] + 36.0 * nr[msz - 3]
                                                       - 48.0 * nr[msz - 2] + 25.0 * nr[msz - 1])
    if dn < 0.0 and abs(rad[-1] * nr[-1]) > 1e-20:
        bexp = -dn / nr[-1]
        if bexp * rm > 50.0:
            bexp, aexp = 0.001, 0.0
        else:
            aexp = nr[-1] * math.exp(bexp * rm)
            if abs(aexp) < 1e-20:
                bexp, aexp = 0.001, 0.0
    else:
        bexp, aexp = 0.001, 0.0
    keep = np.abs(rnr) > 1e-20
    two_pi = 2.0 * PI
    nq = np.zeros(len(qgrid))
    ff = np.where(keep, rnr * rad, 0.0)
    r1torm = mesh.simp(np.concatenate([ff, np.zeros(mesh.mesh_size - msz)]) if msz < mesh.mesh_size else ff)
    rmtoin = aexp * math.exp(-bexp * rm) / bexp ** 3 * (2.0 + 2.0 * bexp * rm + bexp * bexp * rm * rm)
    nq[0] = FOUR_PI * (r1torm + rmtoin)
    w = mesh.weights()[:msz]
    for iq in range(1, len(qgrid)):
        arg = two_pi * qgrid[iq]
        ff = np.where(keep, np.sin(arg * rad) * rnr, 0.0)
        r1torm = float(np.dot(ff, w))
        rmtoin = aexp * math.exp(-bexp * rm) / (arg ** 2 + bexp ** 2) ** 2 * (
            arg * (2.0 * bexp + arg ** 2 * rm + bexp ** 2 * rm) * math.cos(arg * rm)
            + (arg ** 2 * (bexp * rm - 1.0) + bexp ** 2 * (bexp * rm + 1.0)) * math.sin(arg * rm))
        nq[iq] = 2.0 / qgrid[iq] * (r1torm + rmtoin)
    q = qgrid[-1]
    arg = two_pi * q
    ff = np.where(keep, (two_pi * rad * np.cos(arg * rad) - np.sin(arg * rad) / q) * rnr, 0.0)
    r1torm = float(np.dot(ff, w))
    rmtoin = -1.0 / (q * (arg ** 2 + bexp ** 2) ** 3) * aexp * math.exp(-bexp * rm) * (
        (arg ** 5 * rm - two_pi * arg ** 4 * q * rm * (bexp * rm - 2.0) + 2.0 * arg ** 3 * bexp * (bexp * rm + 1.0)
         + arg * bexp ** 3 * (bexp * rm + 2.0) - FOUR_PI * arg ** 2 * bexp * q * (bexp ** 2 * rm ** 2 - 3.0)
         - two_pi * bexp ** 3 * q * (bexp ** 2 * rm ** 2 + 2.0 * bexp * rm + 2.0)) * math.cos(arg * rm)
        + (2.0 * arg ** 2 * bexp ** 3 * rm + two_pi * arg ** 5 * q * rm ** 2 + arg ** 4 * (bexp * rm - 1.0)
           + bexp ** 4 * (bexp * rm + 1.0) + FOUR_PI * arg ** 3 * q * (bexp ** 2 * rm ** 2 + 2.0 * bexp * rm - 1.0)
           + two_pi * arg * bexp ** 2 * q * (bexp ** 2 * rm ** 2 + 4.0 * bexp * rm + 6.0)) * math.sin(arg * rm))
    ypn = 2.0 / q * (r1torm + rmtoin)
    ff = np.where(keep, rnr * rad ** 3, 0.0)
    r1torm = float(np.dot(ff, w))
    rmtoin = aexp * math.exp(-bexp * rm) / bexp ** 5 * (24.0 + 24.0 * bexp * rm + 12.0 * bexp ** 2 * rm ** 2
                                                        + 4.0 * bexp ** 3 * rm ** 3 + bexp ** 4 * rm ** 4)
    dnqdq0 = -(2.0 / 3.0) * two_pi ** 3 * (r1torm + rmtoin)
    return nq, dnqdq0, 0.0, ypn


def density_spline(t: LibpawType, which: str, qgrid: np.ndarray) -> np.ndarray:
    """tcorespl / tvalespl (src/libpaw/3_m_pawpsp.F90:2461-2497): (mq, 2)
    values and spline second derivatives; zero for a type without a
    pseudo core."""
    if which == "core":
        if not t.usetcore:
            return np.zeros((len(qgrid), 2))
        nr = t.tncore
    else:
        nr = t.tnvale
    nq, _, yp1, ypn = pawpsp_cg(qgrid, t.core_mesh, nr)
    return np.stack([nq, paw_spline(qgrid, nq, yp1, ypn)], axis=1)


def atm2fft(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, coords: np.ndarray,
            types: Sequence[int], splines: Sequence[np.ndarray], qgrid: np.ndarray, gsqcut: float) -> np.ndarray:
    """atm2fft with optv = 0 (src/libpaw/m_atm2fft.F90): the sum over atoms
    of the radial density's q-space spline times the structure factor, on
    the G vectors with |G|^2 <= gsqcut (1 + 1e-7), Nyquist planes zeroed
    (zerosym), inverse FFT (FFTW backward, unnormalised) divided by the cell
    volume.  coords in bohr on the grid whose point i sits at i d - L/2;
    returns (nx*ny*nz,) with x fastest."""
    n = (nx, ny, nz)
    L = (nx * dx, ny * dy, nz * dz)
    ucvol = L[0] * L[1] * L[2]
    xred = np.asarray(coords, dtype=np.float64) / np.array(L) + 0.5
    ig = [np.array([i - (i // (nk // 2 + 2)) * nk - 1 for i in range(1, nk + 1)]) for nk in n]
    g2 = (ig[2][:, None, None] ** 2 / L[2] ** 2 + ig[1][None, :, None] ** 2 / L[1] ** 2
          + ig[0][None, None, :] ** 2 / L[0] ** 2)
    cutoff = gsqcut * 1.0000001
    inside = g2 <= cutoff
    gmag = np.sqrt(g2)
    dq = (qgrid[-1] - qgrid[0]) / float(len(qgrid) - 1)
    dqm1 = 1.0 / dq
    dq2div6 = dq * dq / 6.0
    jj = np.minimum((1 + (gmag * dqm1).astype(np.int64)) - 1, len(qgrid) - 2)
    diff = gmag - qgrid[jj]
    bb = diff * dqm1
    aa = 1.0 - bb
    cc = aa * (aa ** 2 - 1.0) * dq2div6
    ddc = bb * (bb ** 2 - 1.0) * dq2div6
    work = np.zeros((nz, ny, nx), np.complex128)
    types = np.asarray(types)
    for it in sorted(set(types.tolist())):
        spl = splines[it]
        n_at = aa * spl[jj, 0] + bb * spl[jj + 1, 0] + cc * spl[jj, 1] + ddc * spl[jj + 1, 1]
        sf = np.zeros((nz, ny, nx), np.complex128)
        for ia in np.nonzero(types == it)[0]:
            p = [np.exp(1j * TWO_PI_PH * ig[k] * xred[ia, k]) for k in range(3)]
            ph = p[2][:, None, None] * p[1][None, :, None] * p[0][None, None, :]
            sf += np.conj(ph)
        work += np.where(inside, sf * n_at, 0.0)
    for k, nk in enumerate(n):
        if nk % 2 == 0:
            sl = [slice(None)] * 3
            sl[2 - k] = nk // 2
            work[tuple(sl)] = 0.0
    out = np.fft.ifftn(work) * (nx * ny * nz)
    return (out.real / ucvol).reshape(-1)


def dij_sphere(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, coord: np.ndarray, t: LibpawType
               ) -> Tuple[np.ndarray, np.ndarray]:
    """nhatgrid with optcut = 0 (src/libpaw/finegrid.f90:200-240 ->
    pawrfgd_fft, src/libpaw/2_m_paw_finegrid.F90:1238-1368) and pawgylm
    (:92-1081, shape_type 2, optgr0 only): the grid points within rshp of
    the atom (periodic images, libpaw's loop order: z, y, x) and g_l(r) S_lm
    on them.  Returns (flat 0-based grid index (nfgd,), gylm (nfgd, lm_size))."""
    n = (nx, ny, nz)
    L = (nx * dx, ny * dy, nz * dz)
    xred = np.asarray(coord, dtype=np.float64) / np.array(L) + 0.5
    rcut = t.rshp
    r2cut = 1.0000001 * rcut * rcut
    ishift, delta = 15, 0.99                  # pawrfgd_fft parameters
    lo, hi = [], []
    for k in range(3):
        rr = math.sqrt(r2cut / L[k] ** 2)
        lo.append(int((xred[k] - rr + ishift) * n[k] + delta) - ishift * n[k])
        hi.append(int((xred[k] + rr + ishift) * n[k]) - ishift * n[k])
    idx, rv = [], []
    for i3 in range(lo[2], hi[2] + 1):
        iz = (i3 + ishift * nz) % nz
        difz = i3 / nz - xred[2]
        for i2 in range(lo[1], hi[1] + 1):
            iy = (i2 + ishift * ny) % ny
            dify = i2 / ny - xred[1]
            for i1 in range(lo[0], hi[0] + 1):
                ix = (i1 + ishift * nx) % nx
                difx = i1 / nx - xred[0]
                r = (difx * L[0], dify * L[1], difz * L[2])
                if r[0] ** 2 + r[1] ** 2 + r[2] ** 2 <= r2cut:
                    idx.append(ix + nx * (iy + ny * iz))
                    rv.append(r)
    rv = np.array(rv, dtype=np.float64).reshape(-1, 3)
    rn = np.sqrt(np.sum(rv * rv, axis=1))
    lcut = int(round(math.sqrt(t.lm_size)))
    pi_over_rshp = PI / t.rshp
    rc = TOL12 + t.rshp
    g0 = np.zeros(len(rn))
    for ic, arg in enumerate(rn):
        if arg < 1e-3:                        # toldev: shapefunc2_0, the series of sinc^2
            x = pi_over_rshp * arg
            g0[ic] = 1.0 - x ** 2 / 3.0 + 2.0 * x ** 4 / 45.0
        elif arg <= rc:
            x = pi_over_rshp * arg
            g0[ic] = (math.sin(x) / x) ** 2
    safe = np.where(rn > TOL10, rn, 1.0)
    u = rv / safe[:, None]
    u[rn <= TOL10] = 0.0
    S = real_ylm_values(lcut - 1, u[:, 0], u[:, 1], u[:, 2])
    S[rn <= TOL10, 1:] = 0.0
    gylm = np.zeros((len(rn), t.lm_size))
    for l in range(lcut):
        gl = t.gnorm[l] * g0 * (rn ** l if l > 0 else 1.0)
        for m in range(-l, l + 1):
            gylm[:, l * l + l + m] = gl * S[:, l * l + l + m]
    return np.array(idx, np.int64), gylm


# ---------------------------------------------------------------------------
# hand-over to the device tables

def onsite_spec(t: LibpawType) -> dict:
    """The per-type tables opaw_set_onsite takes (and the oracle reads), in
    m-state matrix form: gnt (lm, ms, ms), kln (ms, ms), qd = q_ij^L
    (lm, ms, ms); plus the packed eijkl/dij0 and the mesh parameters."""
    ms = t.lmn_size
    K = t.pack_index()
    qd = np.zeros((t.lm_size, ms, ms))
    for lm in range(t.lm_size):
        qd[lm] = t.qijl[lm][K]
    rad = t.mesh.rad[:t.mesh_size].copy()
    return dict(nr=t.mesh_size, lm=t.lm_size, nij=t.ij_size, usetcore=t.usetcore,
                a=t.mesh.rstep, d=t.mesh.lstep, nint=t.mesh.int_meshsz,
                rad=rad, w=t.mesh.weights()[:t.mesh_size], phiphj=t.phiphj, tphitphj=t.tphitphj,
                coredens=t.coredens, tcoredens=t.tcoredens, gnt=t.gaunt_ij(), kln=t.kln_ij(), qd=qd,
                eijkl_packed=t.eijkl.copy(), dij0_packed=t.dij0.copy())


def grid_densities(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, coords: np.ndarray,
                   types: Sequence[int], tabs: Sequence[LibpawType]) -> Tuple[np.ndarray, np.ndarray]:
    """get_vloc_ncoret's two atm2fft calls (src/vloc.f90:24-33): the pseudo
    core density ncoret (optn2 = 1) and the starting density, the
    superposition of the neutralised pseudo valence densities (optn2 = 2).
    Returns (ncoret, dens0), each (nx*ny*nz,)."""
    gsq = gsqcut_of(nx, ny, nz, dx, dy, dz)
    qg = qgrid_vl(gsq)
    core = [density_spline(t, "core", qg) for t in tabs]
    vale = [density_spline(t, "vale", qg) for t in tabs]
    ncoret = atm2fft(nx, ny, nz, dx, dy, dz, coords, types, core, qg, gsq)
    dens0 = atm2fft(nx, ny, nz, dx, dy, dz, coords, types, vale, qg, gsq)
    return ncoret, dens0
