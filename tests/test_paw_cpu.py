"""CPU checks of the PAW set-up: the JTH dataset reader and radial work
(opaw_amd.paw, host code) and the oracle of the grid-sized kernels
(oracle/paw_oracle.py), each pinned by mathematics that does not restate it:
scipy's spherical harmonics and natural cubic spline, Gauss quadrature of
Gaunt integrals, orthonormality and duality of the transformed radial
functions, the adjoint identity of the spline contraction, brute-force
sphere membership.  The datasets are the reference's naphthalene inputs
(tests/golden/jth)."""
import math
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import sph_harm_y

import tests.helpers  # noqa: F401
from oracle import paw_oracle as PO
from opaw_amd import lib, paw

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jth")
PI_RATIO = math.sqrt(math.pi / paw.PI_YLM)     # the reference's Y_lm carry pi in single precision


@pytest.fixture(scope="module")
def ds():
    return {s: paw.read_dataset(os.path.join(GOLD, f"{s}.LDA_PW-JTH.xml")) for s in ("H", "C")}


def _real_ylm_scipy(l, m, th, ph):
    if m == 0:
        return sph_harm_y(l, 0, th, ph).real
    y = sph_harm_y(l, abs(m), th, ph)
    return math.sqrt(2.0) * (-1) ** m * (y.real if m > 0 else y.imag)


def test_reader_fields(ds):
    C, H = ds["C"], ds["H"]
    assert (C.symbol, C.zat, C.val, C.nr, C.ms, C.nl2) == ("C", 6.0, 4.0, 2001, 8, 9)
    assert (H.symbol, H.zat, H.val, H.nr, H.ms, H.nl2) == ("H", 1.0, 1.0, 1500, 5, 9)
    assert list(C.lstate) == [0, 0, 1, 1] and list(C.mstate) == [0, 0, -1, 0, 1, -1, 0, 1]
    assert np.allclose(C.occ, [2.0, 0.0, 2.0 / 3.0, 0.0])
    assert C.rr[0] == 1e-8 and C.rr[1] > 0
    # r = a (exp(d i) - 1) of the radial_grid element
    a, d = 9.4548737315239002e-04, 5.6729242389143399e-03
    assert np.allclose(C.rr[1:], a * (np.exp(d * np.arange(1, C.nr)) - 1), rtol=1e-12)
    # readpaw's "ncoretilde" is the pseudo valence density / sqrt(4 pi): it
    # integrates to the valence less the augmentation charge (within 3 %)
    for s in ("C", "H"):
        q = np.sum(ds[s].ncoretilde * ds[s].rr ** 2 * ds[s].dr) * 12.5663706144
        assert abs(q - ds[s].val) < 0.03 * ds[s].val
    # nrough of get_nrcut at 0.5 bohr (the reference's 14 and 10 point boxes)
    g = paw.Geometry(56, 52, 32, 0.5, 0.5, 0.5, False)
    assert list(paw.box_shape(C, g)[0]) == [14, 14, 14] and list(paw.box_shape(H, g)[0]) == [10, 10, 10]
    assert g.nfovnr == 4


def test_interpolate_is_bisection(ds):
    rr = ds["C"].rr
    rng = np.random.default_rng(0)
    t = np.concatenate([rng.uniform(0, 20, 200), rr[[0, 5, 100]], [rr[-1] * 2, -1.0]])
    i, j, wi, wj = PO.bracket(rr, t)
    for k, v in enumerate(t):
        a, b, x, y = paw.interpolate(rr, float(v))
        assert (a, b) == (i[k], j[k]) and abs(x - wi[k]) < 1e-15 and abs(y - wj[k]) < 1e-15


def test_ylm_matches_scipy_real_harmonics():
    rng = np.random.default_rng(1)
    th = np.arccos(rng.uniform(-1, 1, 50))
    ph = rng.uniform(0, 2 * np.pi, 50)
    x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    for l in range(5):
        for m in range(-l, l + 1):
            ref = _real_ylm_scipy(l, m, th, ph) * PI_RATIO
            got = np.array([paw.ylm(l, m, *v) for v in zip(x, y, z)])
            assert np.max(np.abs(got - ref)) < 1e-13, (l, m)
            if l <= 2:
                assert np.max(np.abs(PO.ylm_array(l, m, x, y, z, paw.PI_YLM) - got)) < 1e-15


def test_gaunt_matches_quadrature():
    """cg3 = integral of Y_l1m1 Y_l2m2 Y_l3m3 times sqrt(4pi / 4pi_single):
    the reference divides by 4 pi as a single-precision literal."""
    xg, wg = np.polynomial.legendre.leggauss(24)
    nph = 48
    ph = np.arange(nph) * 2 * np.pi / nph
    TH, PH = np.meshgrid(np.arccos(xg), ph, indexing="ij")
    W = np.outer(wg, np.full(nph, 2 * np.pi / nph))
    quirk = math.sqrt(4 * math.pi / paw.FOURPI_CG)
    Y = {(l, m): _real_ylm_scipy(l, m, TH, PH) for l in range(5) for m in range(-l, l + 1)}
    worst = 0.0
    for l1 in range(5):
        for l2 in range(3):
            for l3 in range(3):
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(-l2, l2 + 1):
                        for m3 in range(-l3, l3 + 1):
                            q = np.sum(W * Y[l1, m1] * Y[l2, m2] * Y[l3, m3])
                            worst = max(worst, abs(paw.cg3(l1, m1, l2, m2, l3, m3) - q * quirk))
    assert worst < 1e-13


def test_natural_spline_matches_scipy():
    rng = np.random.default_rng(2)
    x = np.arange(12) * 0.5
    y = rng.standard_normal(12)
    y2 = paw.natural_spline_y2(x, y)
    cs = CubicSpline(x, y, bc_type="natural")
    t = np.linspace(0, 5.5, 97)
    got = np.array([paw.natural_spline_eval(x, y, y2, v) for v in t])
    assert np.max(np.abs(got - cs(t))) < 1e-13


def test_b_matrix_cardinal_and_linear():
    B, dxf = paw.b_matrix(14, 56, 0.5)
    scale = dxf / 0.5
    xr = np.arange(14) * 0.5
    xf = np.arange(56) * dxf
    # the spline of a rough delta is the delta at the rough nodes
    for j, v in enumerate(xf):
        k = np.nonzero(np.isclose(xr, v, atol=1e-12))[0]
        if len(k):
            col = B[:, j] / scale
            assert abs(col[k[0]] - 1) < 1e-13 and np.sum(np.abs(col)) - 1 < 1e-12
    # natural splines reproduce straight lines
    assert np.max(np.abs((xr @ B) / scale - xf)) < 1e-12
    assert np.max(np.abs(B.sum(axis=0) / scale - 1)) < 1e-13


def test_radial_transform_orthonormal_and_dual(ds):
    for s in ("C", "H"):
        d = ds[s]
        n = d.nr_int
        w = d.rr[:n] ** 2 * d.dr[:n]
        for l in d.lstate_diff:
            idx = np.nonzero(d.lstate == l)[0]
            G = (d.ptilde1[idx, :n] * w) @ d.ptilde1[idx, :n].T
            assert np.max(np.abs(G - np.eye(len(idx)))) < 1e-12
            D0 = (d.ptilde[idx, :n] * w) @ d.phitilde[idx, :n].T
            D1 = (d.ptilde1[idx, :n] * w) @ d.phitilde1[idx, :n].T
            # p1 = O^-1/2 p, phi1 = O^1/2 phi, then one rotation: <p1|phit1> = V^T M1 <p|phit> M2 V
            assert np.max(np.abs(D1 - np.eye(len(idx)))) <= 2 * np.max(np.abs(D0 - np.eye(len(idx)))) + 1e-12
            T = d.extra["mat_t_transformed"][np.ix_(idx, idx)]
            assert np.max(np.abs(T - np.diag(np.diag(T)))) < 1e-12


def test_qijlm_monopole_is_augmentation_overlap(ds):
    """q^00_ij = delta_lm,l'm' int (phi_i phi_j - phit_i phit_j) r^2 dr / sqrt(4 pi_single)
    (int Y_00 Y Y' = delta/sqrt(4 pi)); s_ij likewise from the transformed waves."""
    for s in ("C", "H"):
        d = ds[s]
        w = d.rr ** 2 * d.dr
        w[0] = 0.0
        # difference before the sum: unbound partial waves grow to ~1e19 at
        # large r, where the all-electron and pseudo waves coincide
        full = np.array([[np.sum((d.phi[i] * d.phi[j] - d.phitilde[i] * d.phitilde[j]) * w)
                          for j in range(d.ns)] for i in range(d.ns)])
        full1 = np.array([[np.sum((d.phi1[i] * d.phi1[j] - d.phitilde1[i] * d.phitilde1[j]) * w)
                           for j in range(d.ns)] for i in range(d.ns)])
        for i in range(d.ms):
            for j in range(d.ms):
                same = d.mstate[i] == d.mstate[j] and d.lstate[d.ms_ls[i]] == d.lstate[d.ms_ls[j]]
                want = full[d.ms_ls[i], d.ms_ls[j]] / math.sqrt(paw.FOURPI_CG) if same else 0.0
                assert abs(d.qijlm[0, i, j] - want) < 1e-13
        want = np.array([full1[d.ms_ls[i], d.ms_ls[i]] for i in range(d.ms)]) * paw.SQ4PI / math.sqrt(paw.FOURPI_CG)
        assert np.allclose(np.diag(d.sij), np.maximum(want, -1 + paw.TOLLSIJ), rtol=0, atol=1e-12)
        assert np.allclose(d.qijlm, np.transpose(d.qijlm, (0, 2, 1)), atol=1e-15)


def test_shape_functions_normalised(ds):
    """sum r^(l+2) dr g_l = 1 and g_l(r) / r^l has one shape for every l."""
    for s in ("C", "H"):
        d = ds[s]
        r = d.rr[:d.nrcomp]
        for l in range(2 * d.nl - 1):
            assert abs(np.sum(r ** (l + 2) * d.dr[:d.nrcomp] * d.gl[:, l]) - 1) < 1e-13
            ratio = d.gl[1:, l] / (d.gl[1:, 0] * r[1:] ** l)
            assert np.max(np.abs(ratio / ratio[0] - 1)) < 1e-12


def _small_box(d, geom, coord):
    nrough, nfine = paw.box_shape(d, geom)
    irs, _ = paw.ir_start(coord, nrough, geom)
    Bs, dxf = [], []
    for k, dd in enumerate((geom.dx, geom.dy, geom.dz)):
        b, f = paw.b_matrix(int(nrough[k]), int(nfine[k]), dd)
        Bs.append(b)
        dxf.append(f)
    h = geom.half
    corner = [-h[k] + (irs[k] - 1) * (geom.dx, geom.dy, geom.dz)[k] for k in range(3)]
    lm = np.stack([d.lstate[d.ms_ls], d.mstate], axis=1)
    return nrough, nfine, irs, Bs, dxf, corner, lm


def test_box_contraction_adjoint_identity(ds):
    """sum_rough p3d psi dv = sum_fine p psi_interp dv_fine, psi_interp the
    tensor-product natural cubic spline (scipy) of a random rough psi."""
    d = ds["H"]
    geom = paw.Geometry(24, 24, 24, 0.5, 0.5, 0.5, False, finegrid=0.25)
    coord = np.array([0.3, -0.2, 0.1])
    nrough, nfine, irs, Bs, dxf, corner, lm = _small_box(d, geom, coord)
    F = PO.box_fine(d.rr, d.ptilde[d.ms_ls], d.ptilde1[d.ms_ls], lm, nfine, dxf, corner, coord, geom.half, paw.PI_YLM)
    p3d = PO.contract(F, *Bs)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(tuple(nrough[::-1]))            # (nr3, nr2, nr1)
    xr = [np.arange(nrough[k]) * 0.5 for k in range(3)]
    xf = [np.arange(nfine[k]) * dxf[k] for k in range(3)]
    t = CubicSpline(xr[0], psi, axis=2, bc_type="natural")(xf[0])
    t = CubicSpline(xr[1], t, axis=1, bc_type="natural")(xf[1])
    t = CubicSpline(xr[2], t, axis=0, bc_type="natural")(xf[2])
    lhs = np.einsum("szyx,zyx->s", p3d, psi) * 0.125
    rhs = np.einsum("szyx,zyx->s", F, t) * dxf[0] * dxf[1] * dxf[2]
    assert np.max(np.abs(lhs - rhs)) < 1e-12 * np.max(np.abs(rhs))


def test_sphere_oracle_membership_and_g3d(ds):
    d = ds["C"]
    nx, ny, nz, h = 20, 18, 16, 0.5
    for coord, label in (((0.3, -0.45, 0.2), "centre"), ((4.6, -4.3, 3.9), "corner: wraps through the images")):
        idx, g3d = PO.sphere(nx, ny, nz, h, h, h, coord, d.rcut, d.rcomp, d.rr, d.gl, d.nl2, paw.PI_YLM)
        # brute force over the 27 images
        ix, iy, iz = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        P = np.stack([-nx * h / 2 + ix * h, -ny * h / 2 + iy * h, -nz * h / 2 + iz * h], axis=1)
        L = np.array([nx * h, ny * h, nz * h])
        allp = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
        pos = np.stack([-nx * h / 2 + allp[:, 0] * h, -ny * h / 2 + allp[:, 1] * h, -nz * h / 2 + allp[:, 2] * h], 1)
        dmin = np.full(len(pos), np.inf)
        for sh in np.stack(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1]), -1).reshape(-1, 3):
            dmin = np.minimum(dmin, np.linalg.norm(pos - np.array(coord) + sh * L, axis=1))
        flat = allp[:, 0] + nx * (allp[:, 1] + ny * allp[:, 2])
        want = np.sort(flat[dmin < d.rcut])
        assert np.array_equal(idx, want), label
        # g3d vs numpy.interp of g_l and scipy harmonics
        diff = P - np.array(coord)
        diff -= L * np.round(diff / L)
        r = np.maximum(np.linalg.norm(diff, axis=1), 1e-8)
        th, ph = np.arccos(np.clip(diff[:, 2] / r, -1, 1)), np.arctan2(diff[:, 1], diff[:, 0])
        for Lc in range(d.nl2):
            l = int(math.floor(math.sqrt(Lc)))
            m = Lc - l * l - l
            rad = np.interp(r, d.rr[:d.nrcomp], d.gl[:, l])
            ref = np.where(r > d.rcomp, 0.0, rad * _real_ylm_scipy(l, m, th, ph) * PI_RATIO)
            assert np.max(np.abs(g3d[:, Lc] - ref)) < 1e-12 * max(1.0, np.abs(ref).max()), (label, Lc)


def test_transform_p3d_invariants(ds):
    d = ds["C"]
    rng = np.random.default_rng(4)
    P = rng.standard_normal((d.ms, 6, 5, 4)) * 0.3
    p1, s, sinv = paw.transform_p3d(P, d.sij, 0.125)
    f = p1.reshape(d.ms, -1)
    assert np.max(np.abs(f @ f.T * 0.125 - np.eye(d.ms))) < 1e-12
    P0 = P.reshape(d.ms, -1)
    lhs = (f.T * s) @ f
    rhs = P0.T @ d.sij @ P0
    assert np.max(np.abs(lhs - rhs)) < 1e-11 * np.abs(rhs).max()
    assert np.allclose(sinv, 1 / (1 + s) - 1, rtol=0, atol=1e-15)


def test_set_up_entry_points_reject_bad_arguments():
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built (run __graft_entry__.build())")
    import ctypes as C
    n = C.c_int()
    z = np.zeros(8)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    with pytest.raises(lib.OpawError, match="nl2 must be"):
        lib.call("opaw_paw_sphere", 8, 8, 8, 0.5, 0.5, 0.5, dp(z), 1.0, 0.8, 4, dp(z), 8, 3, dp(z), 0, C.byref(n),
                 None, None)
    lm = np.array([3, 5], np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    sh = np.array([4, 4, 4], np.int32)
    with pytest.raises(lib.OpawError, match="bad \\(l, m\\)"):
        lib.call("opaw_paw_box", 4, dp(z), 1, dp(z), dp(z), ip(lm), ip(sh), ip(sh), dp(z), dp(z), dp(z), dp(z),
                 dp(z), dp(z), dp(z), dp(z), dp(z))


def test_vloc_tuma_oracle_is_the_direct_sum_for_smooth_potentials():
    """The doubled-box transform route of vloc_prep reproduces the direct
    real-space sum sum_a v(|r - R_a|) for a smooth (band-limited) radial
    potential: v = -Z erf(r / rc) / r with rc well above the spacing."""
    from scipy.special import erf
    nx, ny, nz, d = 24, 20, 16, 0.4
    rr = 1e-3 * (np.exp(0.01 * np.arange(1500)) - 1) + 1e-8
    types = [(rr, -4.0 * erf(rr / 1.2) / rr), (rr, -1.0 * erf(rr / 1.0) / rr)]
    atom_type = [0, 1, 0]
    coords = np.array([[0.3, -0.2, 0.1], [-1.1, 0.9, -0.4], [1.6, 0.5, 0.7]])
    v = PO.vloc_tuma(nx, ny, nz, d, d, d, types, atom_type, coords)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.stack([(-nx / 2 + ix) * d, (-ny / 2 + iy) * d, (-nz / 2 + iz) * d], -1).reshape(-1, 3)
    ref = np.zeros(len(P))
    for t, c in zip(atom_type, coords):
        r = np.linalg.norm(P - c, axis=1)
        Z, rc = (4.0, 1.2) if t == 0 else (1.0, 1.0)
        ref += np.where(r > 1e-12, -Z * erf(r / rc) / np.maximum(r, 1e-12), -Z * 2 / (np.sqrt(np.pi) * rc))
    assert np.max(np.abs(v - ref)) < 5e-4 * np.max(np.abs(ref))   # Nyquist planes and the finite box
    # the spline restatement is scipy's natural spline
    x = rr[:50]
    y = np.sin(x * 40)
    cs = CubicSpline(x, y, bc_type="natural")
    t = np.linspace(x[0], x[-1], 77)
    assert np.max(np.abs(PO.splint(x, y, PO.natural_spline(x, y), t) - cs(t))) < 1e-12
