"""GPU: libpaw's on-site D_ij (pawdenpot + pawdij, LDA, pawxcdev = 1) on
the reference's naphthalene input (cnt.ini geometry, JTH C/H datasets,
56 x 52 x 32 at 0.5 bohr), against the CPU oracle.

Tolerance of D_ij: 1e-12 of max |D| (measured 2e-15).  pawxcm takes v'(n) and v''(n) by
differences at n (1 +- 1e-4); evaluated directly, they would turn a
last-bit difference of the cube root or the logarithm (HIP's device math
against the host libm, or invcb's iteration against cbrt) into ~1e-8 of
D^xc.  Both sides evaluate the same differences from a Taylor jet of v
instead (tests/test_libpaw_cpu.py pins the jet against the three-point
formulas in long double), so D^xc agrees to rounding.  Every other part of
D (dij0, dijhartree, dijhat) and rho_ij is held to 1e-11."""
import os

import numpy as np
import pytest

from tests.helpers import O, rel, state_from_spec
from oracle import libpaw_oracle as LO
from opaw_amd import paw, synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jth")
DIJ_TOL = 1e-12


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch


@pytest.fixture(scope="module")
def naph():
    _torch()
    geom = paw.Geometry(56, 52, 32, 0.5, 0.5, 0.5, False)
    syms, xyz = paw.read_cnt(os.path.join(GOLD, "naphthalene_cnt.ini"), geom)
    ds = {s: paw.read_dataset(os.path.join(GOLD, f"{s}.LDA_PW-JTH.xml"), half_box=geom.half) for s in ("H", "C")}
    spec = paw.system_from_datasets(geom, syms, xyz, ds, nocc=24, name="naphthalene-jth")
    return geom, spec


def _orbitals(spec, nb, seed):
    return synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=seed, envelope=3.0), spec.dv)


def test_no_standins_on_a_nonperiodic_grid(naph):
    _, spec = naph
    assert spec.extra["standins"] == ()
    assert all(t.onsite is not None for t in spec.types)
    assert all(a.dsph_idx is not None and len(a.dsph_idx) > 0 for a in spec.atoms)
    # the starting density is the neutral superposition of the atoms
    assert abs(spec.extra["dens0"].sum() * spec.dv - (4 * 10 + 8)) < 1e-6


def test_rebuild_dij_matches_oracle(naph):
    from opaw_amd.api import OpawGPU
    _, spec = naph
    nb = 6
    spec.rnel = 2.0 * nb
    phi = _orbitals(spec, nb, 5)
    st = state_from_spec(spec, nsys=2)
    phis = [phi, O.excite_psi_bar_pert(st, phi, spec.sm, spec.ipol)]
    g = OpawGPU(spec, "f64", nsys=2, nb=nb)
    g.write_orbitals(0, phis[0])
    g.write_orbitals(1, phis[1])
    print("rebuild of both systems", flush=True)
    g.ncpaw_make_hamiltonian_tddft(-1)
    dmax = worst = 0.0
    for s in (0, 1):
        O.rebuild(st, phis[s], s)
        S = st.systems[s]
        assert rel(g.read_field("vks", s), S.vks) < 1e-11
        for ia in range(1, len(spec.atoms) + 1):
            assert rel(g.read_rhoij(s, ia), S.rhoij[ia - 1]) < 1e-11, (s, ia)
            d_gpu = g.read_dij(s, ia)
            d_ref = S.dij[ia - 1]
            dmax = max(dmax, float(np.max(np.abs(d_ref))))
            err = np.max(np.abs(d_gpu - d_ref)) / max(1.0, np.max(np.abs(d_ref)))
            worst = max(worst, err)
            assert err < DIJ_TOL, (s, ia, err)
    print(f"D_ij GPU vs oracle: {worst:.2e} of max|D|", flush=True)
    assert dmax > 0.1
    g.close()


def test_onsite_terms_separately(naph):
    """D = dij0 + dijhartree + dijxc + dijhat: with the orbitals' rho_ij
    and potentials taken from the GPU, the oracle's four terms rebuild the
    GPU's D, and the XC term is not negligible."""
    from opaw_amd.api import OpawGPU
    geom, spec = naph
    nb = 6
    spec.rnel = 2.0 * nb
    phi = _orbitals(spec, nb, 11)
    g = OpawGPU(spec, "f64", nsys=1, nb=nb)
    g.write_orbitals(0, phi)
    g.ncpaw_make_hamiltonian_tddft(0)
    pot = g.read_field("vks", 0) - g.read_field("vxc", 0)
    dv = spec.dv
    worst_missing_xc = 0.0
    for ia, a in enumerate(spec.atoms, start=1):
        t = spec.types[a.itype]
        tab = dict(t.onsite)
        rho = g.read_rhoij(0, ia)
        ms = rho.shape[0]
        d0 = LO.unpack(tab["dij0_packed"], ms)
        dh = LO.unpack(LO.pawdijhartree(tab["eijkl_packed"], rho), ms)
        dxc = LO.onsite_xc_dij(tab, rho)
        dhat = LO.pawdijhat(pot, a.dsph_idx, a.dsph_g, tab["qd"], dv)
        d_gpu = g.read_dij(0, ia)
        tol = DIJ_TOL * max(1.0, np.max(np.abs(d_gpu)))
        assert np.max(np.abs(d_gpu - (d0 + dh + dxc + dhat))) < tol, ia
        worst_missing_xc = max(worst_missing_xc, float(np.max(np.abs(dxc))))
        # the packed Hartree term equals the expanded eijkl the device uses
        rsel = np.where(np.abs(rho) > 1e-10, rho, 0.0)
        assert np.max(np.abs(dh - np.einsum("ijkl,kl->ij", t.eijkl, rsel))) < 1e-12 * max(1.0, np.max(np.abs(dh)))
    assert worst_missing_xc > 1e-2
    g.close()


def test_reference_cube_root_spread(naph):
    """The oracle with the reference's invcb iteration against the GPU: the
    same bound as against the oracle's cbrt form."""
    from opaw_amd.api import OpawGPU
    _, spec = naph
    nb = 4
    spec.rnel = 2.0 * nb
    phi = _orbitals(spec, nb, 3)
    g = OpawGPU(spec, "f64", nsys=1, nb=nb)
    g.write_orbitals(0, phi)
    g.ncpaw_make_hamiltonian_tddft(0)
    first = {}
    for ia, a in enumerate(spec.atoms, start=1):
        first.setdefault(a.itype, ia)
    for ia in first.values():                # the first atom of each type
        a = spec.atoms[ia - 1]
        tab = dict(spec.types[a.itype].onsite)
        rho = g.read_rhoij(0, ia)
        d_exact = LO.onsite_xc_dij(tab, rho, exact_invcb=True)
        d_fast = LO.onsite_xc_dij(tab, rho)
        assert np.max(np.abs(d_exact - d_fast)) < DIJ_TOL * max(1.0, np.max(np.abs(d_fast)))
    g.close()
