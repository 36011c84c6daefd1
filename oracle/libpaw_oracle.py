"""FP64 CPU restatement of libpaw's per-rebuild on-site D_ij (LDA).

TEST INFRASTRUCTURE: the oracle the HIP on-site kernel is checked against.
Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import it.

What the reference runs at every Hamiltonian rebuild (src/get_dij.f90:14-30,
rank 0): set_rhoij packs the upper triangle of rho_ij and selects
|rho| > 1e-10; pawdenpot (src/libpaw/m_paw_denpot.F90:136) builds the
on-site densities and their XC potentials; pawdij
(src/libpaw/4_m_pawdij.F90:168) sums dij0 + dijhartree + dijxc + dijhat;
transfer_dij clips to +-500 and fills the symmetric matrix.  The call sets
ixc = 7 (ABINIT's PW92 LDA), pawxcdev = 1 (first-order moment expansion,
pawxcm), usexcnhat = 0, nspden = 1, xc_denpos = 1e-14, nzlmopt = 0 (every
lm moment selected).

Restated here, each from the routine named:
  pawdensities     rho1, trho1 lm moments (m_paw_denpot.F90:1216-1574), opt_dens = 1
  invcb, xcpbe     rho^-1/3 by Newton iteration and the option-1 LSD
                   exchange + PW92 correlation (xc_invcb.F90, 0_xc_m_xcpbe.F90:881-945)
  pawxcm           vxc moments, pawxcdev = 1 (2_m_pawxc.F90:3763-4513); its
                   differences of v at rho (1 +- 1e-4) from a Taylor jet of v
                   (pawxcm_differences)
  pawdijxcm        (4_m_pawdij.F90:1959-2178)
  pawdijhartree    (4_m_pawdij.F90:1228-1313)
  pawdijhat        (4_m_pawdij.F90:2227-2514) with libpaw's own g_l S_lm and q_ij^L
  simp_gen         (1_m_pawrad.F90:947-1041), rebuilt here from the mesh
                   parameters (a, d, int_meshsz) rather than taken as weights

Tables (phi phj products, core densities, Gaunt coefficients, eijkl, dij0,
q_ij^L, the D^hat spheres) come from the set-up (libpaw's pawpsp_17in /
pawinit, restated on the host); their pins are in tests/test_libpaw_cpu.py.
"""
from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np

PI = math.pi
FOUR_PI = 4.0 * PI
SQFPI = math.sqrt(FOUR_PI)
DELTA = 1.0e-4                 # pawxcm finite-difference step (2_m_pawxc.F90: delta)
RHO_MIN = 1.0e-14              # pawxcm rho_min (tol14)
XC_DENPOS = 1.0e-14
RHOIJ_SELECT = 1e-10
PARAM_DIJ_MAX = 500.0
M1THRD = float(np.float32(-1.0 / 3.0))   # invcb: m1thrd = -1.0/3.0, a default-real literal

# xcpbe constants (0_xc_m_xcpbe.F90:200-260), option 1 (non-RPA)
_PIINV = 1.0 / PI
THREEFOURTH_DIVPI = 0.75 * _PIINV
SIXPI2_1_3 = (6.0 * PI ** 2) ** (1.0 / 3.0)
RSFAC = 0.6203504908994000
SQ_RSFAC = math.sqrt(RSFAC)
SQ_RSFAC_INV = 1.0 / SQ_RSFAC
TWOM1_3 = 2.0 ** (-1.0 / 3.0)
EC0_AA, EC0_A1, EC0_B1, EC0_B2, EC0_B3, EC0_B4 = 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294


# ---------------------------------------------------------------------------
# radial integration

def simp_weights(a: float, d: float, nr: int, nint: int) -> Tuple[np.ndarray, np.ndarray]:
    """Mesh r(i) = a (exp(d (i-1)) - 1): simfact up to nint (pawrad_init,
    1_m_pawrad.F90:271-284) and radfact."""
    i = np.arange(nr, dtype=np.float64)
    rad = a * (np.exp(d * i) - 1.0)
    rad[0] = 0.0
    radfact = rad + a
    radfact[0] = a
    hh = d / 3.0
    sf = np.zeros(nr)
    sf[nint - 1] = hh * radfact[nint - 1]
    ir_last = 1
    ir = nint
    while ir >= 3:
        sf[ir - 2] = 4.0 * hh * radfact[ir - 2]
        sf[ir - 3] = 2.0 * hh * radfact[ir - 3]
        ir_last = ir - 2
        ir -= 2
    sf[ir_last - 1] *= 0.5
    return sf, radfact


def simp_gen(func: np.ndarray, tab: Dict) -> float:
    """simp_gen over int_meshsz with the end correction of an even count."""
    sf, rf = tab["_sf"], tab["_rf"]
    nn = tab["nint"]
    simp = float(np.dot(func[:nn], sf[:nn]))
    resid = 0.0
    if nn % 2 == 0:
        resid = tab["d"] / 3.0 * (1.25 * func[0] * rf[0] + 2.0 * func[1] * rf[1] - 0.25 * func[2] * rf[2])
    return simp + resid


def _prep(tab: Dict) -> Dict:
    if "_sf" not in tab:
        sf, rf = simp_weights(tab["a"], tab["d"], tab["nr"], tab["nint"])
        tab["_sf"], tab["_rf"] = sf, rf
        i = np.arange(tab["nr"], dtype=np.float64)
        rad = tab["a"] * (np.exp(tab["d"] * i) - 1.0)
        rad[0] = 0.0
        tab["_rad"] = rad
    return tab


# ---------------------------------------------------------------------------
# ABINIT's LDA (ixc = 7 -> xcpbe option 1)

def invcb(rho: np.ndarray) -> np.ndarray:
    """invcb (src/libpaw/xc_invcb.F90): rho^(-1/3) point by point, each
    started from the previous point and refined by a rational iteration
    until (rho r^3 - 1)^2 < 1e-24; the first point and large jumps use the
    power with the single-precision exponent -1/3."""
    n = len(rho)
    out = np.zeros(n)
    r0 = rho[0]
    rhomtrd = math.copysign(abs(r0) ** M1THRD, r0)
    rhom1 = 1.0 / r0
    out[0] = rhomtrd
    c2_27, c5_9, c8_9 = 2.0 / 27.0, 5.0 / 9.0, 8.0 / 9.0
    for i in range(1, n):
        r = rho[i]
        prod = r * rhom1
        if prod < 0.01 or prod > 10.0:
            rhomtrd = math.copysign(abs(r) ** M1THRD, r)
            rhom1 = 1.0 / r
        else:
            dl = prod - 1.0
            test = False
            for _ in range(5):
                rhomtrd = ((1.0 + c5_9 * dl) / (1.0 + dl * (c8_9 + c2_27 * dl))) * rhomtrd
                rhom1 = rhomtrd * rhomtrd * rhomtrd
                dl = r * rhom1 - 1.0
                test = dl * dl < 1.0e-24
                if test:
                    break
            if not test:
                rhomtrd = math.copysign(abs(r) ** M1THRD, r)
        out[i] = rhomtrd
    return out


def xcpbe_lda(rho_up: np.ndarray, exact_invcb: bool = False) -> np.ndarray:
    """xcpbe option 1, nspden = 1, order 1: the potential of the LSD
    exchange + PW92 correlation for the spin-up density rho_up (total
    density 2 rho_up).  exact_invcb: the cube roots by invcb's sequential
    iteration (what the reference runs); otherwise 1/cbrt, which agrees
    with it to ~1e-13 relative at every point that invcb iterates
    (tests/test_libpaw_cpu.py pins the two against each other).  A
    long-double input is evaluated in long double (with cbrt)."""
    rho_up = np.asarray(rho_up)
    if rho_up.dtype != np.longdouble:
        rho_up = rho_up.astype(np.float64)
    rum13 = invcb(rho_up) if exact_invcb else 1.0 / np.cbrt(rho_up)    # rho_updnm1_3
    rhom1_3 = TWOM1_3 * rum13                # (2 rho_up)^-1/3
    rhotot = 2.0 * rho_up
    rhotmot = rhom1_3
    rhotmo6 = np.sqrt(rhotmot)
    rhoto6 = rhotot * rhotmot * rhotmot * rhotmo6
    ex_lsd = -THREEFOURTH_DIVPI * SIXPI2_1_3 * rum13 * rum13 * rho_up
    v = (4.0 / 3.0) * ex_lsd
    rs = RSFAC * rhotmot
    sqr_rs = SQ_RSFAC * rhotmo6
    rsm1_2 = SQ_RSFAC_INV * rhoto6
    q0 = -2.0 * EC0_AA * (1.0 + EC0_A1 * rs)
    q1 = 2.0 * EC0_AA * (EC0_B1 * sqr_rs + EC0_B2 * rs + EC0_B3 * rs * sqr_rs + EC0_B4 * rs * rs)
    q1p = EC0_AA * (EC0_B1 * rsm1_2 + 2.0 * EC0_B2 + 3.0 * EC0_B3 * sqr_rs + 4.0 * EC0_B4 * rs)
    den = 1.0 / (q1 * q1 + q1)
    lg = -np.log(q1 * q1 * den)
    ecrs = q0 * lg
    decrs_drs = -2.0 * EC0_AA * EC0_A1 * lg - q0 * q1p * den
    return v + ecrs - rs * (1.0 / 3.0) * decrs_drs


# ---------------------------------------------------------------------------
# pawxcm's differences of the potential, evaluated from a Taylor jet
#
# pawxcm takes v' and v'' of the spin-up potential v(rho) by differences at
# rho (1 +- delta), delta = 1e-4 (2_m_pawxc.F90:3945-3979):
#   d1 = (v(rho(1+d)) - v(rho(1-d))) / (4 d rho),
#   d2 = (v(rho(1+d)) + v(rho(1-d)) - 2 v(rho)) / (4 d^2 rho^2).
# With v(rho (1+s)) = sum_k c_k s^k these are exactly
#   d1 = (c1 + c3 d^2 + c5 d^4 + ...) / (2 rho),
#   d2 = (c2 + c4 d^2 + c6 d^4 + ...) / (2 rho^2),
# so the jet below (c_0..c_4; the first terms left out, c5 d^4 and c6 d^4,
# are below 1e-15 of the sums) gives the reference's differences without
# their cancellation, which multiplies a one-ulp difference in v by
# 1/d^2 = 1e8.  Two implementations whose cube roots or logarithms differ in
# the last bit then agree to rounding in d1 and d2.  Jets are arrays
# (K, npoints) of Taylor coefficients in s.

JET_K = 5


def _jmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    r = np.zeros_like(a)
    for k in range(a.shape[0]):
        for i in range(k + 1):
            r[k] += a[i] * b[k - i]
    return r


def _jdiv(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    r = np.zeros_like(a)
    for k in range(a.shape[0]):
        acc = a[k].copy()
        for i in range(1, k + 1):
            acc -= b[i] * r[k - i]
        r[k] = acc / b[0]
    return r


def _jlog(a: np.ndarray) -> np.ndarray:
    r = np.zeros_like(a)
    r[0] = np.log(a[0])
    for k in range(1, a.shape[0]):
        acc = a[k].copy()
        for i in range(1, k):
            acc -= (i / k) * r[i] * a[k - i]
        r[k] = acc / a[0]
    return r


def _jpow1p(p: float, base: np.ndarray, K: int) -> np.ndarray:
    """base (1 + s)^p: binomial coefficients times base"""
    r = np.zeros((K,) + np.shape(base))
    c = 1.0
    for k in range(K):
        r[k] = c * base
        c = c * (p - k) / (k + 1)
    return r


def _jconst(x, K: int, n: int) -> np.ndarray:
    r = np.zeros((K, n))
    r[0] = x
    return r


def xcpbe_lda_jet(rho_up: np.ndarray, exact_invcb: bool = False) -> np.ndarray:
    """Taylor coefficients c_k (JET_K, n) of xcpbe_lda(rho_up (1 + s)) in s,
    the same operations as xcpbe_lda carried through the jet (the cube root
    of rho_up once, its powers of (1 + s) by the binomial series)."""
    rho_up = np.asarray(rho_up, dtype=np.float64)
    K, n = JET_K, rho_up.shape[0]
    rum13_0 = invcb(rho_up) if exact_invcb else 1.0 / np.cbrt(rho_up)
    rum13 = _jpow1p(-1.0 / 3.0, rum13_0, K)                 # rho_up^-1/3
    rhotmot = TWOM1_3 * rum13                               # (2 rho_up)^-1/3
    rho = _jpow1p(1.0, rho_up, K)                           # rho_up (1 + s)
    rhotot = 2.0 * rho
    rhotmo6 = _jpow1p(-1.0 / 6.0, np.sqrt(TWOM1_3 * rum13_0), K)
    rhoto6 = _jmul(_jmul(_jmul(rhotot, rhotmot), rhotmot), rhotmo6)
    ex_lsd = -THREEFOURTH_DIVPI * SIXPI2_1_3 * _jmul(_jmul(rum13, rum13), rho)
    v = (4.0 / 3.0) * ex_lsd
    rs = RSFAC * rhotmot
    sqr_rs = SQ_RSFAC * rhotmo6
    rsm1_2 = SQ_RSFAC_INV * rhoto6
    one = _jconst(1.0, K, n)
    q0 = -2.0 * EC0_AA * (one + EC0_A1 * rs)
    rs_sq = _jmul(rs, sqr_rs)
    q1 = 2.0 * EC0_AA * (EC0_B1 * sqr_rs + EC0_B2 * rs + EC0_B3 * rs_sq + EC0_B4 * _jmul(rs, rs))
    q1p = EC0_AA * (EC0_B1 * rsm1_2 + _jconst(2.0 * EC0_B2, K, n) + 3.0 * EC0_B3 * sqr_rs + 4.0 * EC0_B4 * rs)
    q1q1 = _jmul(q1, q1)
    den = _jdiv(one, q1q1 + q1)
    lg = -_jlog(_jmul(q1q1, den))
    ecrs = _jmul(q0, lg)
    decrs_drs = -2.0 * EC0_AA * EC0_A1 * lg - _jmul(_jmul(q0, q1p), den)
    return v + ecrs - (1.0 / 3.0) * _jmul(rs, decrs_drs)


def pawxcm_differences(rhosph: np.ndarray, exact_invcb: bool = False) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(v, d1, d2) of pawxcm at the spin-up spherical densities rhosph (after
    mkdenpos): v(rho) and the two differences, zero where rho <= rho_min."""
    c = xcpbe_lda_jet(rhosph, exact_invcb)
    d2_ = DELTA * DELTA
    on = rhosph > RHO_MIN
    safe = np.where(on, rhosph, 1.0)
    d1 = np.where(on, (c[1] + c[3] * d2_) / (2.0 * safe), 0.0)
    d2 = np.where(on, (c[2] + c[4] * d2_) / (2.0 * safe * safe), 0.0)
    return c[0], d1, d2


# ---------------------------------------------------------------------------
# pawdensities, pawxcm, pawdijxcm, pawdijhartree, pawdijhat

def selected_packed(rho: np.ndarray) -> Tuple[list, list]:
    """set_rhoij (src/get_dij.f90:53-91): packed upper triangle, j outer,
    entries with |rho| > 1e-10.  Returns ([(i, j)], [value])."""
    ms = rho.shape[0]
    pairs, vals = [], []
    for j in range(ms):
        for i in range(j + 1):
            v = rho[i, j]
            if abs(v) > RHOIJ_SELECT:
                pairs.append((i, j))
                vals.append(v)
    return pairs, vals


def pawdensities(tab: Dict, rho: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """rho1(r, lm), trho1(r, lm) of the selected rho_ij (opt_dens = 1:
    no nhat), r = 0 by pawrad_deducer0."""
    tab = _prep(tab)
    nr, lmsz = tab["nr"], tab["lm"]
    rad = tab["_rad"]
    inv2 = np.zeros(nr)
    inv2[1:] = 1.0 / rad[1:] ** 2
    gnt, kln = tab["gnt"], tab["kln"]
    rho1 = np.zeros((nr, lmsz))
    trho1 = np.zeros((nr, lmsz))
    for (i, j), v in zip(*selected_packed(rho)):
        ro = v * (1.0 if i == j else 2.0)            # dltij
        k = kln[i, j]
        for lm in range(lmsz):
            g = gnt[lm, i, j]
            if g == 0.0:
                continue
            rg = ro * g
            rho1[1:, lm] += rg * tab["phiphj"][1:, k] * inv2[1:]
            trho1[1:, lm] += rg * tab["tphitphj"][1:, k] * inv2[1:]
    for lm in range(lmsz):
        for a in (rho1, trho1):
            a[0, lm] = a[3, lm] + 3.0 * (a[1, lm] - a[2, lm])
    return rho1, trho1


def pawxcm_vxc(rhor: np.ndarray, corexc: np.ndarray, usecore: int, tab: Dict, exact_invcb: bool = False
               ) -> np.ndarray:
    """pawxcm (pawxcdev = 1, nspden = 1, option 0) potential moments."""
    tab = _prep(tab)
    nr, lmsz = rhor.shape
    rho_updn = rhor.copy()
    if usecore:
        rho_updn[:, 0] += SQFPI * corexc
    invsqfpi2 = 0.5 / SQFPI
    rhosph = rho_updn[:, 0] * invsqfpi2
    rhosph = np.where(rhosph < XC_DENPOS, XC_DENPOS, rhosph)          # mkdenpos
    vxci, d1vxc, d2vxc = pawxcm_differences(rhosph, exact_invcb)
    v1sum = np.zeros(nr)
    for lm in range(1, lmsz):
        v1sum += rho_updn[:, lm] ** 2
    vxc = np.zeros((nr, lmsz))
    vxc[:, 0] = vxci * SQFPI + v1sum * d2vxc * invsqfpi2
    for lm in range(1, lmsz):
        vxc[:, lm] = d1vxc * rho_updn[:, lm]
    # linear interpolation across low-density gaps (lmselect(1))
    rad = tab["_rad"]
    tol_rho = XC_DENPOS * (1.0 + 1e-6)
    ir1 = ir2 = 0
    for ir in range(1, nr + 1):
        if rho_updn[ir - 1, 0] < tol_rho:
            if ir1 == 0:
                ir1 = ir - 1
            ir2 = ir + 1
        elif ir1 > 0:
            if ir1 > 1 or ir2 < nr:
                f = (vxc[ir2 - 1, 0] - vxc[ir1 - 1, 0]) / (rad[ir2 - 1] - rad[ir1 - 1])
                for jr in range(ir1 + 1, ir2):
                    vxc[jr - 1, 0] = vxc[ir1 - 1, 0] + f * (rad[jr - 1] - rad[ir1 - 1])
            ir1 = ir2 = 0
    return vxc


def pawdijxcm(tab: Dict, vxc1: np.ndarray, vxct1: np.ndarray) -> np.ndarray:
    """Dij^xc (ms, ms) from the potential moments (usexcnhat = 0)."""
    tab = _prep(tab)
    gnt, kln = tab["gnt"], tab["kln"]
    ms = kln.shape[0]
    nij = tab["phiphj"].shape[1]
    out = np.zeros((ms, ms))
    for lm in range(tab["lm"]):
        vij = np.array([simp_gen(vxc1[:, lm] * tab["phiphj"][:, k] - vxct1[:, lm] * tab["tphitphj"][:, k], tab)
                        for k in range(nij)])
        out += gnt[lm] * vij[kln]
    return out


def onsite_xc_dij(tab: Dict, rho: np.ndarray, exact_invcb: bool = False) -> np.ndarray:
    """pawdenpot's XC moments followed by pawdijxcm for one atom."""
    rho1, trho1 = pawdensities(tab, rho)
    vxc1 = pawxcm_vxc(rho1, tab["coredens"], 1, tab, exact_invcb)
    vxct1 = pawxcm_vxc(trho1, tab["tcoredens"], int(tab["usetcore"]), tab, exact_invcb)
    return pawdijxcm(tab, vxc1, vxct1)


def pawdijhartree(eijkl_packed: np.ndarray, rho: np.ndarray) -> np.ndarray:
    """dijhartree (packed, lmn2) from the stored upper triangle of eijkl."""
    ms = rho.shape[0]
    n2 = ms * (ms + 1) // 2
    out = np.zeros(n2)
    for (i, j), v in zip(*selected_packed(rho)):
        k = j * (j + 1) // 2 + i
        ro = v * (1.0 if i == j else 2.0)
        out[k] += ro * eijkl_packed[k, k]
        for k1 in range(k):
            out[k1] += ro * eijkl_packed[k1, k]
        for k1 in range(k + 1, n2):
            out[k1] += ro * eijkl_packed[k, k1]
    return out


def pawdijhat(pot: np.ndarray, idx: np.ndarray, gylm: np.ndarray, qd: np.ndarray, dv: float) -> np.ndarray:
    """Dij^hat (ms, ms): prod(lm) = dv sum pot g_l S_lm over the sphere,
    then sum_lm prod(lm) q_ij^lm (qd (lm, ms, ms))."""
    prod = (pot[np.asarray(idx)] @ gylm) * dv
    return np.einsum("lij,l->ij", qd[:gylm.shape[1]], prod)


def onsite_dij(tab: Dict, rho: np.ndarray, pot: np.ndarray, dh_idx: np.ndarray, dh_g: np.ndarray, dv: float
               ) -> np.ndarray:
    """pawdij's sum dij0 + dijhartree + dijxc + dijhat for one atom, as the
    (ms, ms) matrix of the packed result (before transfer_dij)."""
    ms = rho.shape[0]
    d = unpack(tab["dij0_packed"], ms)
    d = d + unpack(pawdijhartree(tab["eijkl_packed"], rho), ms)
    d = d + onsite_xc_dij(tab, rho)
    d = d + pawdijhat(pot, dh_idx, dh_g, tab["qd"], dv)
    return d


def unpack(packed: np.ndarray, ms: int) -> np.ndarray:
    out = np.zeros((ms, ms))
    for j in range(ms):
        for i in range(j + 1):
            out[i, j] = out[j, i] = packed[j * (j + 1) // 2 + i]
    return out
