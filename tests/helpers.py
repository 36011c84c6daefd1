"""Test helpers: adapt a synthetic SystemSpec into the oracle's state."""
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
for p in (_REPO, os.path.join(_REPO, "opaw-tddft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle.opaw_oracle as O  # noqa: E402


def grid_of(spec):
    return O.Grid(spec.nx, spec.ny, spec.nz, spec.dx, spec.dy, spec.dz, spec.periodic)


def state_from_spec(spec, nsys=2, with_vk=True):
    g = grid_of(spec)
    atoms = [O.AtomBox(itype=a.itype, start=np.asarray(a.start), nr=np.asarray(a.nr), p=a.p, p1=a.p1,
                       s=np.asarray(a.s), sinv=np.asarray(a.sinv), sph_idx=np.asarray(a.sph_idx),
                       sph_g=np.asarray(a.sph_g)) for a in spec.atoms]
    types = [O.AtomType(ms=t.ms, qijlm=t.qijlm, dij0=t.dij0, eijkl=t.eijkl) for t in spec.types]
    systems = []
    for _ in range(nsys):
        vks = spec.vks0.copy() if spec.vks0 is not None else np.zeros(g.n)
        dij = [d.copy() for d in spec.dij0_atoms] if spec.dij0_atoms is not None else \
            [np.zeros((a.ms, a.ms)) for a in spec.atoms]
        systems.append(O.SysState(vks=vks, dij=dij))
    return O.OpawState(grid=g, ek=O.ek_prep(g, spec.ekcut), atoms=atoms, types=types, vloc=spec.vloc,
                       ncoret=spec.ncoret, vk=O.vk_prep(g) if with_vk else None, rnel=spec.rnel,
                       systems=systems, dt=spec.dt)


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
