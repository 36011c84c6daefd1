"""GPU: the host-array entry points under the Fortran drop-ins
(opaw_h_phi_host, opaw_shs_host, opaw_sn_phi_host, opaw_proj_host,
opaw_write/read_orbitals_c16, opaw_set_dt, opaw_rk4_prop_shs_host), called
through the C ABI on complex*16 host arrays and checked against the oracle
(src/hphi.f90, src/shs.f90, src/snphi.f90, src/proj.f90, src/proj1.f90,
src/rk4.f90) on the synthetic 32^3 molecule, in FP64 to the device
operators' tolerances and in the mixed context against FP64."""
import numpy as np
import pytest

from tests.helpers import O, rel, state_from_spec
from tests.test_gpu_parity import _gpu, _random_potentials, mol  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def test_h_phi_and_shs_host(mol):
    spec, st, phi = mol
    torch, g = _gpu(spec, st, phi.shape[0])
    for sys in (0, 1):
        assert rel(g.h_phi_host(phi, sys), O.h_phi(st, phi, sys)) < 1e-10, sys
        assert rel(g.shs_host(phi, sys), O.shs(st, phi, sys)) < 1e-10, sys
    # one field, as the reference calls them
    assert rel(g.h_phi_host(phi[1], 0)[0], O.h_phi(st, phi[1:2], 0)[0]) < 1e-10
    g.close()


@pytest.mark.parametrize("n", [-0.5, 0.5, 1.0])
def test_sn_phi_host(mol, n):
    spec, st, phi = mol
    torch, g = _gpu(spec, st, phi.shape[0])
    assert rel(g.sn_phi_host(phi, n), O.sn_phi(st, phi, n)) < 1e-12
    g.close()


def test_proj_host_every_atom(mol):
    spec, st, phi = mol
    torch, g = _gpu(spec, st, phi.shape[0])
    for ia, a in enumerate(spec.atoms, start=1):
        ca = g.proj_host(ia, phi[2], a.ms)
        ca1 = g.proj_host(ia, phi[2], a.ms, which_p1=True)
        assert rel(ca, O.proj(st, ia - 1, phi[2:3])[0]) < 1e-11, ia
        assert rel(ca1, O.proj1(st, ia - 1, phi[2:3])[0]) < 1e-11, ia
    from opaw_amd.lib import OpawError
    with pytest.raises(OpawError, match="projectors"):
        g.proj_host(1, phi[0], spec.atoms[0].ms + 1)
    with pytest.raises(OpawError, match="outside"):
        g.proj_host(len(spec.atoms) + 1, phi[0], spec.atoms[0].ms)
    g.close()


def test_mixed_context_host_ops(mol):
    spec, st, phi = mol
    torch, g = _gpu(spec, st, phi.shape[0], prec="mixed")
    for sys in (0, 1):
        assert rel(g.h_phi_host(phi, sys), O.h_phi(st, phi, sys)) < 1e-5, sys
        assert rel(g.shs_host(phi, sys), O.shs(st, phi, sys)) < 1e-5, sys
    assert rel(g.sn_phi_host(phi, -0.5), O.sn_phi(st, phi, -0.5)) < 1e-5
    # complex*16 orbitals round-trip through the FP32 resident copy
    g.write_orbitals_c16(0, phi)
    back = g.read_orbitals_c16(0)
    assert rel(back, phi) < 1e-6
    assert np.array_equal(back, phi.astype(np.complex64).astype(np.complex128))
    g.close()


def _kicked(spec, st, phi, prec):
    phis = [phi, O.excite_psi_bar_pert(st, phi, spec.sm, spec.ipol)]
    torch, g = _gpu(spec, st, phi.shape[0], prec=prec)
    g.write_orbitals_c16(0, phis[0])
    g.write_orbitals_c16(1, phis[1])
    return phis, g


def test_rk4_prop_shs_host_matches_oracle(mol):
    spec, _, phi = mol
    st = state_from_spec(spec, nsys=2)
    _random_potentials(st, spec, 1)
    phis, g = _kicked(spec, st, phi, "f64")
    from opaw_amd.lib import OpawError
    p = phis[0].copy()      # updated in place: never the fixture's own array
    pp = phis[1].copy()
    with pytest.raises(OpawError, match="dt not set"):
        g.rk4_prop_shs_host(p, pp)
    with pytest.raises(OpawError, match="positive"):
        g.set_dt(-1.0)
    g.set_dt(spec.dt)
    g.rk4_prop_shs_host(p, pp)
    for s, out in ((0, p), (1, pp)):
        ref = O.rk4_prop_shs(st, phis[s], s)
        assert rel(out, ref) < 1e-10, s
        assert np.array_equal(out, g.read_orbitals_c16(s)), s
    g.close()


def test_rk4_prop_shs_host_refuses_changed_orbitals(mol):
    spec, st, phi = mol
    phis, g = _kicked(spec, st, phi, "f64")
    g.set_dt(spec.dt)
    p = phis[0].copy()
    pp = phis[1].copy()
    pp[1, 7] += 1e-3
    keep_p, keep_pp = p.copy(), pp.copy()
    from opaw_amd.lib import OpawError
    with pytest.raises(OpawError, match="phi_bar_pert"):
        g.rk4_prop_shs_host(p, pp)
    assert np.array_equal(p, keep_p) and np.array_equal(pp, keep_pp)
    # nothing moved on the device either
    assert np.array_equal(g.read_orbitals_c16(0), phis[0])
    assert np.array_equal(g.read_orbitals_c16(1), phis[1])
    with pytest.raises(OpawError, match="context holds"):
        g.rk4_prop_shs_host(np.ascontiguousarray(p[:-1]), np.ascontiguousarray(pp[:-1]))
    g.close()


def test_rk4_prop_shs_host_mixed_follows_fp64(mol):
    spec, _, phi = mol
    st = state_from_spec(spec, nsys=2)
    _random_potentials(st, spec, 1)
    out = {}
    for prec in ("f64", "mixed"):
        phis, g = _kicked(spec, st, phi, prec)
        g.set_dt(spec.dt)
        p = g.read_orbitals_c16(0)
        pp = g.read_orbitals_c16(1)
        g.rk4_prop_shs_host(p, pp)
        out[prec] = (p, pp)
        g.close()
    for s in (0, 1):
        assert rel(out["mixed"][s], out["f64"][s]) < 1e-5, s
