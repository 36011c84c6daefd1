"""Adapter from a synthetic SystemSpec (opaw_amd.synth) to the oracle state.

TEST INFRASTRUCTURE (part of the oracle package): used by the tests, the smoke
entry point and the CPU-baseline leg of the benchmark only.
"""
import numpy as np

from . import opaw_oracle as O


def grid_of(spec):
    return O.Grid(spec.nx, spec.ny, spec.nz, spec.dx, spec.dy, spec.dz, spec.periodic)


def state_from_spec(spec, nsys=2, with_vk=True):
    g = grid_of(spec)
    atoms = [O.AtomBox(itype=a.itype, start=np.asarray(a.start), nr=np.asarray(a.nr), p=a.p, p1=a.p1,
                       s=np.asarray(a.s), sinv=np.asarray(a.sinv), sph_idx=np.asarray(a.sph_idx),
                       sph_g=np.asarray(a.sph_g), dsph_idx=getattr(a, "dsph_idx", None),
                       dsph_g=getattr(a, "dsph_g", None)) for a in spec.atoms]
    types = [O.AtomType(ms=t.ms, qijlm=t.qijlm, dij0=t.dij0, eijkl=t.eijkl,
                        onsite=(dict(getattr(t, "onsite")) if getattr(t, "onsite", None) is not None else None))
             for t in spec.types]
    systems = []
    for _ in range(nsys):
        vks = spec.vks0.copy() if spec.vks0 is not None else np.zeros(g.n)
        dij = [d.copy() for d in spec.dij0_atoms] if spec.dij0_atoms is not None else \
            [np.zeros((a.ms, a.ms)) for a in spec.atoms]
        systems.append(O.SysState(vks=vks, dij=dij))
    return O.OpawState(grid=g, ek=O.ek_prep(g, spec.ekcut), atoms=atoms, types=types, vloc=spec.vloc,
                       ncoret=spec.ncoret, vk=O.vk_prep(g) if with_vk else None, rnel=spec.rnel,
                       systems=systems, dt=spec.dt)
