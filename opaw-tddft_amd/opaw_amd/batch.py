"""Several molecules propagated together (opaw_tddft_run_batch).

The reference runs one molecule per opaw_tddft job, with one kick direction
per run (src/tddft.f90:67-91, src/tddft_excite.f90:1-16).  Here a batch is a
list of contexts, one per molecule, each with its own grid and orbital count
and with nsys systems: system 0 unperturbed and systems 1..nsys-1 kicked
along x, y, z.  One call advances every context nt time steps; each context
runs on its own stream, so the kernels of small and large molecules overlap
on the device, and with graph=True one time step per context is captured as
a HIP graph and replayed.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import lib, synth

# grids of the polyacenes C(4n+2)H(2n+4), n = 1..6, at 0.5 bohr
POLYACENE_GRIDS: Dict[int, Tuple[int, int, int]] = {
    1: (48, 48, 48), 2: (64, 48, 48), 3: (80, 64, 48), 4: (96, 64, 48), 5: (96, 64, 48), 6: (128, 64, 48)}


def tddft_batch(gpus: Sequence, nt: int, dt: float, graph: bool = True) -> List[np.ndarray]:
    """nt time steps of every context; returns each one's dipole series
    (nt, 3 * nsys) like OpawGPU.opaw_tddft's second result."""
    n = len(gpus)
    arr = (C.c_void_p * n)(*[g.ctx.value for g in gpus])
    sizes = [3 * g.nsys * nt for g in gpus]
    dip = np.zeros(sum(sizes))
    lib.call("opaw_tddft_run_batch", C.cast(arr, C.POINTER(C.c_void_p)), n, int(nt), float(dt), int(bool(graph)),
             dip.ctypes.data_as(C.POINTER(C.c_double)))
    out, off = [], 0
    for g, m in zip(gpus, sizes):
        out.append(dip[off:off + m].reshape(nt, 3 * g.nsys))
        off += m
    return out


def kick_responses(dip: np.ndarray, sm: float) -> np.ndarray:
    """(nt, 3 * nsys) dipoles with system s kicked along axis s-1 ->
    (nt, nsys-1) responses (d_s[axis s-1] - d_0[axis s-1]) / sm, the diagonal
    of the dipole response, as plot_dip's d_dip for each kick
    (src/tddft_plot.f90:57)."""
    nsys = dip.shape[1] // 3
    return np.stack([(dip[:, 3 * s + s - 1] - dip[:, s - 1]) / sm for s in range(1, nsys)], axis=1)


def polyacene_spec(nring: int, seed: int = None):
    """Synthetic OPAW system of the polyacene with nring rings on its grid;
    9 nring + 6 doubly occupied orbitals."""
    spec = synth.molecule_system(name=f"polyacene{nring}", nring=nring, n=POLYACENE_GRIDS[nring], d=0.5,
                                 seed=100 + nring if seed is None else seed)
    spec.rnel = 2.0 * spec.nocc
    return spec


def polyacene_orbitals(spec, seed: int = 0) -> np.ndarray:
    nb = spec.nocc
    return synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=seed, envelope=4.0), spec.dv)


def kicked_context(spec, phi0: np.ndarray, precision: str, nsys: int = 4):
    """Context with the orbitals phi0 in system 0 and their kicks
    exp(-i sm r_axis) along x, y, z in systems 1..nsys-1."""
    from .api import OpawGPU
    g = OpawGPU(spec, precision, nsys=nsys, nb=phi0.shape[0])
    g.write_orbitals(0, phi0)
    for s in range(1, nsys):
        g.excite_psi_bar_pert(spec.sm, s, from_sys=0, to_sys=s)
    return g


def main(argv=None):
    """Time the batched polyacene series: python -m opaw_amd.batch."""
    import argparse
    import time
    import torch
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["f64", "mixed"], default="mixed")
    ap.add_argument("--graph", action="store_true", help="replay one captured step per molecule")
    ap.add_argument("--rings", type=int, nargs="+", default=[1, 2, 3, 4, 5, 6],
                    help="the polyacenes of the batch (rings per molecule, 1..6)")
    a = ap.parse_args(argv)
    specs = [polyacene_spec(n) for n in a.rings]
    phis = [polyacene_orbitals(s, seed=20 + k) for k, s in enumerate(specs)]
    gs = [kicked_context(s, p, a.precision, 4) for s, p in zip(specs, phis)]
    dt = specs[0].dt
    tddft_batch(gs, a.warmup, dt, graph=a.graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dips = tddft_batch(gs, a.steps, dt, graph=a.graph)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / a.steps
    for g, d in zip(gs, dips):
        assert np.all(np.isfinite(d))
        g.close()
    print(f"polyacenes n={a.rings} x 4 systems, {a.precision}, graph={a.graph}: {sec * 1e3:.3f} ms per batch step "
          f"({a.steps} steps)", flush=True)


if __name__ == "__main__":
    main()
