"""CPU pins of the libpaw restatement behind the on-site D_ij.

Set-up (opaw_amd.libpaw, host product code) and the per-rebuild oracle
(oracle/libpaw_oracle.py), each checked against something that does not
restate it:
  - Simpson weights on the JTH log mesh: closed-form radial integrals;
  - the radial Poisson solver: the potential of a Gaussian charge,
    Q erf(r / (sqrt2 s)) / r, and scipy quadrature of the multipole formula;
  - real Gaunt coefficients: scipy's spherical harmonics on a separate rule;
  - xcpbe (ABINIT's ixc = 7): the libxc-form PW92 LDA of the grid oracle,
    written independently, and v = d(n eps)/dn;
  - invcb: numpy's cbrt;
  - pawxcm's differences of v at rho (1 +- 1e-4) from a Taylor jet: the
    same three-point formulas in long double;
  - pawxcm's first-order moment expansion: the exact angular projection of
    v_xc, error second order in the anisotropy;
  - Dij^xc: central differences of the exact (angular-quadrature) on-site
    XC energy, agreeing to second order in the anisotropy;
  - eijkl: the on-site Hartree energy by a brute-force double integral of
    the multipole Coulomb kernel;
  - shape functions: normalisation; valence + compensation charge = zion;
  - atm2fft: the Fourier-space sum equals the direct lattice sum of the
    band-limited radial density; integrates to the charge.
The datasets are the reference's naphthalene inputs (tests/golden/jth)."""
import math
import os

import numpy as np
import pytest
from scipy import integrate
from scipy.special import erf, sph_harm_y

import tests.helpers  # noqa: F401
from oracle import libpaw_oracle as LO
from oracle import opaw_oracle as O
from opaw_amd import libpaw as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jth")


@pytest.fixture(scope="module")
def tabs():
    return {s: L.load_type(os.path.join(GOLD, f"{s}.LDA_PW-JTH.xml")) for s in ("C", "H")}


def _mesh():
    return L.RadMesh(9.4548737315239002e-04, 5.6729242389143399e-03, 2001, r_for_intg=1.5073670273)


def test_simpson_weights_closed_form():
    m = _mesh()
    r = m.rad
    R = r[m.int_meshsz - 1]
    for k, f, exact in ((2, np.exp(-r) * r ** 2, 2.0 - math.exp(-R) * (R * R + 2 * R + 2)),
                        (0, np.sin(r) * r, math.sin(R) - R * math.cos(R))):
        got = m.simp(f)
        assert abs(got - exact) < 1e-9 * max(1.0, abs(exact)), (k, got, exact)
        # the weight form used on the device is the same integral
        assert abs(float(np.dot(m.weights(), f)) - got) < 1e-13 * max(1.0, abs(exact))
    # an even point count exercises the end correction
    m2 = L.RadMesh(m.rstep, m.lstep, 2001, r_for_intg=r[999])
    assert m2.int_meshsz % 2 == 0
    R2 = r[m2.int_meshsz - 1]
    f = np.exp(-r) * r ** 2
    assert abs(m2.simp(f) - (2.0 - math.exp(-R2) * (R2 * R2 + 2 * R2 + 2))) < 1e-9


def test_poisson_gaussian_closed_form():
    m = L.RadMesh(9.4548737315239002e-04, 5.6729242389143399e-03, 2001)
    r = m.rad
    s = 0.7
    n = np.exp(-r ** 2 / (2 * s * s)) / (2 * math.pi * s * s) ** 1.5
    rv = m.poisson(4 * math.pi * r ** 2 * n, 0)
    want = erf(r / (math.sqrt(2) * s))
    assert np.max(np.abs(rv - want)) < 2e-9


@pytest.mark.parametrize("ll", [1, 2])
def test_poisson_multipole_quadrature(ll):
    m = L.RadMesh(9.4548737315239002e-04, 5.6729242389143399e-03, 2001)
    r = m.rad
    f = lambda x: x ** (ll + 2) * math.exp(-x * x / 0.9)          # den(r), a product of u functions
    rv = m.poisson(np.array([f(x) for x in r]), ll)
    for i in (300, 700, 1000, 1200, 1500):
        x = r[i]
        inner = integrate.quad(lambda t: f(t) * t ** ll, 0, x, epsabs=0, epsrel=1e-13)[0]
        outer = integrate.quad(lambda t: f(t) / t ** (ll + 1), x, np.inf, epsabs=0, epsrel=1e-13)[0]
        want = (x ** (ll + 1) * outer + inner / x ** ll) / (2 * ll + 1)
        assert abs(rv[i] - want) < 1e-8 * max(1.0, abs(want)), (i, rv[i], want)


def _real_ylm_scipy(l, m, th, ph):
    if m == 0:
        return sph_harm_y(l, 0, th, ph).real
    y = sph_harm_y(l, abs(m), th, ph)
    return math.sqrt(2.0) * (-1) ** m * (y.real if m > 0 else y.imag)


def test_real_gaunt_vs_scipy():
    G = L.real_gaunt(2, 1)
    x, w = np.polynomial.legendre.leggauss(12)
    th = np.arccos(x)
    ph = np.linspace(0, 2 * math.pi, 25)[:-1]
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    W = w[:, None] * (2 * math.pi / 24)
    S = {(l, m): _real_ylm_scipy(l, m, TH, PH) for l in range(3) for m in range(-l, l + 1)}
    for l3 in range(3):
        for m3 in range(-l3, l3 + 1):
            for l1 in range(2):
                for m1 in range(-l1, l1 + 1):
                    for l2 in range(2):
                        for m2 in range(-l2, l2 + 1):
                            want = float(np.sum(W * S[(l1, m1)] * S[(l2, m2)] * S[(l3, m3)]))
                            got = G[l3 * l3 + l3 + m3, l1 * l1 + l1 + m1, l2 * l2 + l2 + m2]
                            assert abs(got - (want if abs(want) >= 1e-12 else 0.0)) < 1e-13
    # the l = 0 row is the identity / sqrt(4 pi)
    assert abs(G[0, 1, 1] - 1 / math.sqrt(4 * math.pi)) < 1e-15


def test_xcpbe_equals_independent_pw92():
    n = np.logspace(-9, 2, 400)
    v_abinit = LO.xcpbe_lda(n / 2)
    v_libxc, _ = O.lda_xc(n)
    assert np.max(np.abs(v_abinit - v_libxc) / np.abs(v_libxc)) < 1e-12
    # and v = d(n eps)/dn of the energy written out independently
    _, eps = O.lda_xc(n)
    h = 1e-6 * n
    _, ep = O.lda_xc(n + h)
    _, em = O.lda_xc(n - h)
    dnde = ((n + h) * ep - (n - h) * em) / (2 * h)
    assert np.max(np.abs(dnde - v_abinit) / np.abs(v_abinit)) < 1e-7


def test_invcb_matches_cbrt():
    r = np.exp(-np.linspace(0, 30, 1500)) * 3.0
    got = LO.invcb(r)
    want = 1.0 / np.cbrt(r)
    assert np.max(np.abs(got[1:] - want[1:]) / want[1:]) < 1e-12
    # the first point uses the single-precision exponent
    assert abs(got[0] - r[0] ** np.float32(-1.0 / 3.0)) < 1e-15 * abs(got[0])
    # D_ij^xc with the exact iteration equals the cbrt form: pawxcm's
    # differences come from the Taylor jet, so the cube roots' last-bit
    # differences are not amplified by 1/delta^2
    for sym in ("C", "H"):
        t = L.load_type(os.path.join(GOLD, f"{sym}.LDA_PW-JTH.xml"))
        tab = L.onsite_spec(t)
        rho = _rho_psd(t.lmn_size)
        d_exact = LO.onsite_xc_dij(dict(tab), rho, exact_invcb=True)
        d_fast = LO.onsite_xc_dij(dict(tab), rho)
        assert np.max(np.abs(d_exact - d_fast)) < 1e-11 * max(1.0, np.max(np.abs(d_fast)))


def test_pawxcm_jet_equals_the_reference_differences():
    """pawxcm's v, v' and v'' differences at rho (1 +- 1e-4) from the
    Taylor jet against the same three-point formulas evaluated in long
    double (80-bit; its own rounding, amplified by 1/delta^2, stays near
    1e-10).  Leaving out the c3 d^2 or c4 d^2 term of the jet moves v' or v''
    by 2e-9 / 4e-9; in double the three-point v'' is off by up to 2e-2."""
    rho = np.geomspace(1e-5, 50.0, 400)
    v, d1, d2 = LO.pawxcm_differences(rho)
    r = rho.astype(np.longdouble)
    dl = np.longdouble(LO.DELTA)
    vi = LO.xcpbe_lda(r)
    v1 = LO.xcpbe_lda((1 + dl) * r)
    v2 = LO.xcpbe_lda((1 - dl) * r)
    rinv = np.longdouble(0.5) / dl / r
    d1_ref = ((v1 - v2) * np.longdouble(0.5) * rinv).astype(np.float64)
    d2_ref = ((v1 + v2 - 2 * vi) * rinv * rinv).astype(np.float64)
    assert np.max(np.abs(v / vi.astype(np.float64) - 1)) < 1e-14
    assert np.max(np.abs(d1 / d1_ref - 1)) < 1e-12
    assert np.max(np.abs(d2 / d2_ref - 1)) < 1.5e-9
    # zero at and below rho_min, as pawxcm's rhoinv
    _, z1, z2 = LO.pawxcm_differences(np.array([LO.RHO_MIN, 0.5 * LO.RHO_MIN]))
    assert np.all(z1 == 0.0) and np.all(z2 == 0.0)


def _rho_psd(ms, seed=0):
    """rho_ij = 2 sum_o c_io c_jo of four orbitals' projections: positive
    semi-definite, so the on-site densities are non-negative."""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(ms, 4)) * 0.4
    c[1] *= 0.05
    c[5:] *= 0.05
    return 2.0 * c @ c.T


def _rho_c(eps=1.0):
    """A carbon rho_ij: the reference occupations (2s^2 2p^2 spread over m)
    plus an anisotropic part of size eps."""
    rng = np.random.default_rng(7)
    rho = np.diag([2.0, 0.05, 2 / 3, 2 / 3, 2 / 3, 0.02, 0.02, 0.02])
    a = rng.normal(size=(8, 8)) * 0.05
    a = a + a.T
    return rho + eps * a


def _angular_rule(nth=14, nph=28):
    x, w = np.polynomial.legendre.leggauss(nth)
    ph = (np.arange(nph) + 0.5) * 2 * math.pi / nph
    ct = np.repeat(x, nph)
    st = np.sqrt(1 - ct * ct)
    pp = np.tile(ph, nth)
    S = L.real_ylm_values(2, st * np.cos(pp), st * np.sin(pp), ct)
    return S, np.repeat(w, nph) * 2 * math.pi / nph


def _exact_vxc_moments(rho1, core, S, wq):
    n = rho1 @ S.T + core[:, None]                       # (nr, nq)
    n = np.maximum(n, 1e-14)
    v, _ = O.lda_xc(n)
    return (v * wq) @ S                                  # (nr, lm)


def test_pawxcm_first_order_vs_angular_projection(tabs):
    t = tabs["C"]
    tab = L.onsite_spec(t)
    S, wq = _angular_rule()
    errs = []
    for eps in (0.2, 0.1, 0.05):
        rho1, _ = LO.pawdensities(dict(tab), _rho_c(eps))
        v = LO.pawxcm_vxc(rho1, t.coredens, 1, dict(tab))
        ve = _exact_vxc_moments(rho1, t.coredens, S, wq)
        sl = slice(10, t.mesh.int_meshsz)               # the sphere, away from r = 0
        errs.append(np.max(np.abs(v[sl] - ve[sl])))
    # spherical density: exact
    rho_s = np.diag(np.diag(_rho_c(0.0)))
    rho_s[2:5, 2:5] = np.eye(3) * (2 / 3)
    rho1, _ = LO.pawdensities(dict(tab), rho_s)
    v = LO.pawxcm_vxc(rho1, t.coredens, 1, dict(tab))
    assert np.max(np.abs(v[10:, 1:])) < 1e-13
    ve = _exact_vxc_moments(rho1, t.coredens, S, wq)
    assert np.max(np.abs(v[10:t.mesh.int_meshsz, 0] - ve[10:t.mesh.int_meshsz, 0])) < 1e-7 * np.max(np.abs(ve[10:, 0]))
    # anisotropic: error falls ~4x per halving (second order)
    assert errs[1] < errs[0] / 3.0 and errs[2] < errs[1] / 3.0, errs


def _exact_exc_energy(t, tab, rho, S, wq):
    """E_xc of the on-site densities n1 + core minus that of tn1 + tcore,
    angular quadrature of the exact LDA, radial simp_gen."""
    rho1, trho1 = LO.pawdensities(dict(tab), rho)
    out = 0.0
    for dens, core in ((rho1, t.coredens), (trho1, t.tcoredens)):
        n = np.maximum(dens @ S.T + core[:, None], 1e-14)
        _, eps = O.lda_xc(n)
        f = ((n * eps) * wq).sum(axis=1) * t.mesh.rad[:t.mesh_size] ** 2
        out += t.mesh.simp(f) * (1 if core is t.coredens else -1)
    return out


def test_dijxc_is_derivative_of_exact_energy(tabs):
    """D^xc_ij = dE_xc/drho_ij (each off-diagonal pair moves together): the
    moment expansion agrees with central differences of the exact on-site
    XC energy to second order in the anisotropy."""
    t = tabs["C"]
    tab = L.onsite_spec(t)
    S, wq = _angular_rule()
    errs = []
    for eps in (0.2, 0.1):
        rho = _rho_c(eps)
        d = LO.onsite_xc_dij(dict(tab), rho)
        err = 0.0
        for (i, j) in ((0, 0), (0, 1), (2, 2), (2, 3), (0, 5), (3, 6), (4, 4)):
            h = 1e-5
            e = np.zeros((8, 8))
            e[i, j] = e[j, i] = h
            fd = (_exact_exc_energy(t, tab, rho + e, S, wq) - _exact_exc_energy(t, tab, rho - e, S, wq)) / (2 * h)
            want = fd if i == j else fd / 2.0
            err = max(err, abs(d[i, j] - want))
        errs.append(err)
    assert errs[0] < 2e-3, errs
    assert errs[1] < errs[0] / 2.5, errs


def _coulomb_energy(rho_lm, r, jac, l):
    """1/2 (rho|rho) of one multipole: the potential from the two cumulative
    radial integrals (scipy's cumulative Simpson in the mesh index, dr =
    jac di), then 1/2 int rho V r^2 dr."""
    from scipy.integrate import cumulative_simpson
    f_in = rho_lm * r ** (l + 2) * jac
    rs = np.where(r > 0, r, 1.0)
    f_out = np.where(r > 0, rho_lm * rs ** (1 - l), 0.0) * jac
    inner = cumulative_simpson(f_in, dx=1.0, initial=0.0)
    outer_all = cumulative_simpson(f_out, dx=1.0, initial=0.0)
    outer = outer_all[-1] - outer_all
    V = 4 * math.pi / (2 * l + 1) * (inner / rs ** (l + 1) + rs ** l * outer)
    V[0] = 0.0
    from scipy.integrate import simpson
    return 0.5 * simpson(rho_lm * V * r ** 2 * jac, dx=1.0)


def test_eijkl_is_onsite_hartree_energy(tabs):
    """1/2 sum E_ijkl rho_ij rho_kl = 1/2 [(n1|n1) - (tn1 + nhat|tn1 + nhat)]
    per multipole (Coulomb kernel 4pi/(2l+1) r<^l / r>^(l+1)), the right side
    by scipy's cumulative Simpson rule on the mesh (not libpaw's poisson)."""
    for sym in ("C", "H"):
        t = tabs[sym]
        tab = L.onsite_spec(t)
        ms = t.lmn_size
        rho = _rho_psd(ms, seed=5)
        E = t.eijkl_full()
        e_tab = 0.5 * np.einsum("ij,ijkl,kl->", rho, E, rho)
        rho1, trho1 = LO.pawdensities(dict(tab), rho)
        nint = t.mesh.int_meshsz
        r = t.mesh.rad[:nint]
        jac = t.mesh.radfact[:nint] * t.mesh.lstep
        qd = tab["qd"]
        e_dir = 0.0
        for lm in range(t.lm_size):
            l = int(math.isqrt(lm))
            nh = np.einsum("ij,ij->", rho, qd[lm]) * t.shapefunc[:nint, l]
            e_dir += _coulomb_energy(rho1[:nint, lm], r, jac, l)
            e_dir -= _coulomb_energy(trho1[:nint, lm] + nh, r, jac, l)
        assert abs(e_tab - e_dir) < 1e-6 * max(1.0, abs(e_dir)), (sym, e_tab, e_dir)
        # and the matrix is symmetric in (ij) <-> (kl)
        assert np.max(np.abs(E - np.transpose(E, (2, 3, 0, 1)))) < 1e-13 * np.max(np.abs(E))


def test_shape_and_charges(tabs):
    for t in tabs.values():
        r = t.mesh.rad
        for l in range(t.l_size):
            f = np.zeros(t.mesh_size)
            f[1:] = t.shapefunc[1:, l] * r[1:t.mesh_size] ** (2 + l)
            assert abs(t.mesh.simp(f, r_for_intg=t.rshp) - 1.0) < 1e-13
        cm = t.core_mesh
        assert abs(4 * math.pi * cm.simp(t.tnvale * cm.rad ** 2) - t.zion) < 1e-12
        # sij of libpaw = the dataset's partial-wave overlap difference
        K = t.pack_index()
        for i in range(t.lmn_size):
            li, lni = t.indlmn[i][0], t.indlmn[i][2]
            f = t.phiphj[:, lni * (lni + 1) // 2 + lni] - t.tphitphj[:, lni * (lni + 1) // 2 + lni]
            assert abs(t.sij[K[i, i]] - t.mesh.simp(f)) < 1e-9


def test_atm2fft_is_band_limited_lattice_sum(tabs):
    """atm2fft of one atom = sum over periodic images of the radial density
    filtered to |G|^2 <= gsqcut, computed here by direct G-sum of the
    spline; the cell integral equals the q = 0 value (the charge)."""
    t = tabs["C"]
    nx = ny = nz = 16
    d = 0.5
    gsq = L.gsqcut_of(nx, ny, nz, d, d, d)
    assert abs(gsq - 1.0 / (2 * d) ** 2) < 1e-14
    qg = L.qgrid_vl(gsq)
    spl = L.density_spline(t, "vale", qg)
    assert abs(spl[0, 0] - t.zion) < 1e-10
    xyz = np.array([[0.3, -0.2, 0.1]])
    f = L.atm2fft(nx, ny, nz, d, d, d, xyz, [0], [spl], qg, gsq)
    assert abs(f.sum() * d ** 3 - t.zion) < 1e-9
    # direct sum at a few points over the G vectors inside the sphere
    Lc = nx * d
    g1 = np.fft.fftfreq(nx, d)
    GX, GY, GZ = np.meshgrid(g1, g1, g1, indexing="ij")
    G2 = GX ** 2 + GY ** 2 + GZ ** 2
    ok = (G2 <= gsq * 1.0000001)
    for k in range(3):
        ok &= np.abs(np.meshgrid(g1, g1, g1, indexing="ij")[k]) < 1.0 / (2 * d) - 1e-12
    nq = np.interp(np.sqrt(G2), qg, spl[:, 0])
    for (ix, iy, iz) in ((0, 0, 0), (5, 9, 3), (8, 8, 8)):
        rx, ry, rz = ix * d - Lc / 2, iy * d - Lc / 2, iz * d - Lc / 2
        ph = np.exp(2j * math.pi * (GX * (rx - xyz[0, 0]) + GY * (ry - xyz[0, 1]) + GZ * (rz - xyz[0, 2])))
        want = float(np.real(np.sum(np.where(ok, nq * ph, 0.0)))) / Lc ** 3
        got = f[ix + nx * (iy + ny * iz)]
        assert abs(got - want) < 2e-4 * np.max(np.abs(f)), (ix, iy, iz, got, want)


def test_dij_sphere_points_and_gylm(tabs):
    """pawrfgd_fft's points are exactly those within rshp (periodic
    images), and g_l S_lm integrates like the shape function: the l = 0
    column times dv sums to ~sqrt(4pi) (int g_0 r^2 dr = 1, S_00 = 1/sqrt(4pi))
    on a fine grid."""
    t = tabs["C"]
    n, d = 48, 0.12
    c = np.array([0.31, -0.17, 0.05])
    idx, gy = L.dij_sphere(n, n, n, d, d, d, c, t)
    iz, rem = np.divmod(idx, n * n)
    iy, ix = np.divmod(rem, n)
    pos = np.stack([ix, iy, iz], 1) * d - n * d / 2
    r = np.linalg.norm(pos - c, axis=1)
    assert np.all(r <= t.rshp * 1.0000001 ** 0.5 + 1e-12)
    allpts = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) * d - n * d / 2
    inside = np.sum(np.linalg.norm(allpts - c, axis=1) <= t.rshp)
    assert len(idx) == inside
    assert abs(gy[:, 0].sum() * d ** 3 - math.sqrt(4 * math.pi)) < 2e-3
