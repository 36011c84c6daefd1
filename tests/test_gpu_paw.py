"""GPU: the PAW set-up kernels (opaw_paw_box, opaw_paw_sphere) against the
oracle on the reference's naphthalene input (cnt.ini geometry, the JTH C
and H datasets, grid 56 x 52 x 32 at 0.5 bohr), a periodic box whose
spheres and boxes wrap, and one TDDFT step of naphthalene built from the
datasets, GPU against the CPU oracle.  libpaw's part of the set-up (core
and starting densities, dij0, e_ijkl, the on-site XC of D_ij) is checked in
tests/test_gpu_onsite.py and tests/test_libpaw_cpu.py."""
import os

import numpy as np
import pytest

from tests.helpers import O, rel, state_from_spec
from oracle import paw_oracle as PO
from opaw_amd import paw, synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jth")


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch


@pytest.fixture(scope="module")
def naph():
    geom = paw.Geometry(56, 52, 32, 0.5, 0.5, 0.5, False)
    syms, xyz = paw.read_cnt(os.path.join(GOLD, "naphthalene_cnt.ini"), geom)
    ds = {s: paw.read_dataset(os.path.join(GOLD, f"{s}.LDA_PW-JTH.xml"), half_box=geom.half) for s in ("H", "C")}
    return geom, syms, xyz, ds


def _oracle_box(d, geom, c):
    nrough, nfine = paw.box_shape(d, geom)
    irs, _ = paw.ir_start(c, nrough, geom)
    Bs, dxf = [], []
    for k, dd in enumerate((geom.dx, geom.dy, geom.dz)):
        b, f = paw.b_matrix(int(nrough[k]), int(nfine[k]), dd)
        Bs.append(b)
        dxf.append(f)
    h = geom.half
    corner = [-h[k] + (irs[k] - 1) * (geom.dx, geom.dy, geom.dz)[k] for k in range(3)]
    lm = np.stack([d.lstate[d.ms_ls], d.mstate], axis=1)
    return irs, PO.box_tables(d.rr, d.ptilde[d.ms_ls], d.ptilde1[d.ms_ls], lm, nfine, dxf, corner, c, h, *Bs,
                              paw.PI_YLM)


def _check_atom(d, geom, c, label):
    irs, (p_ref, p1_ref) = _oracle_box(d, geom, c)
    p, p1 = paw.box_projectors(d, geom, c, irs)
    assert rel(p, p_ref) < 1e-12 and rel(p1, p1_ref) < 1e-12, label
    idx, g3d = paw.sphere(d, geom, c)
    idx_ref, g_ref = PO.sphere(geom.nx, geom.ny, geom.nz, geom.dx, geom.dy, geom.dz, c, d.rcut, d.rcomp, d.rr,
                               d.gl, d.nl2, paw.PI_YLM)
    assert np.array_equal(idx, idx_ref), label
    assert np.max(np.abs(g3d - g_ref)) <= 1e-13 * max(1.0, np.abs(g_ref).max()), label


def test_naphthalene_tables_match_oracle(naph):
    _torch()
    geom, syms, xyz, ds = naph
    for ia, (s, c) in enumerate(zip(syms, xyz)):
        print(f"atom {ia + 1} {s}", flush=True)
        _check_atom(ds[s], geom, c, (ia, s))


def test_periodic_wrapped_boxes_and_spheres(naph):
    _torch()
    _, _, _, ds = naph
    geom = paw.Geometry(24, 20, 28, 0.5, 0.5, 0.5, True)
    h = geom.half
    for s, c in (("C", np.array([h[0] - 0.3, -h[1] + 0.2, 0.1])), ("H", np.array([-h[0] + 0.1, 0.4, h[2] - 0.25]))):
        nrough, _ = paw.box_shape(ds[s], geom)
        irs, edge = paw.ir_start(c, nrough, geom)
        assert edge
        _check_atom(ds[s], geom, c, s)


def test_naphthalene_from_datasets_one_step(naph):
    torch = _torch()
    from opaw_amd.api import OpawGPU
    geom, syms, xyz, ds = naph
    spec = paw.system_from_datasets(geom, syms, xyz, ds, nocc=24, name="naphthalene-jth")
    # transform_p3d: every box's p1 is orthonormal on the grid
    dv = geom.dx * geom.dy * geom.dz
    for a in spec.atoms:
        f = a.p1.reshape(a.ms, -1)
        assert np.max(np.abs(f @ f.T * dv - np.eye(a.ms))) < 1e-12
    nb = 6
    spec.rnel = 2.0 * nb
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=5, envelope=3.0), spec.dv)
    st = state_from_spec(spec, nsys=2)
    phis = [phi, O.excite_psi_bar_pert(st, phi, spec.sm, spec.ipol)]
    new, d_dip, _, _ = O.tddft_step(st, phis, spec.sm)
    g = OpawGPU(spec, "f64", nsys=2, nb=nb)
    g.write_orbitals(0, phi)
    g.excite_psi_bar_pert(spec.sm, spec.ipol + 1)
    d_gpu, _ = g.opaw_tddft(1, spec.dt, spec.sm)
    for s in (0, 1):
        assert rel(g.read_orbitals(s), new[s]) < 1e-10, s
    assert np.max(np.abs(d_gpu[0] - d_dip)) <= 1e-8 * np.max(np.abs(d_dip))
    g.close()
    del torch


def test_vloc_tuma_matches_oracle(naph):
    _torch()
    geom, syms, xyz, ds = naph
    v = paw.vloc_tuma(geom, syms, xyz, ds)
    tsyms, amap = paw.type_map(syms)
    ref = PO.vloc_tuma(geom.nx, geom.ny, geom.nz, geom.dx, geom.dy, geom.dz,
                       [(ds[t].rr, ds[t].vloc) for t in tsyms], list(amap), xyz)
    assert rel(v, ref) < 1e-12
    # far from the molecule the potential is the ions' Coulomb tail -Z_val / r
    assert v.min() < -1.0
