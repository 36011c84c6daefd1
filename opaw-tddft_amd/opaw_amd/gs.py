"""Ground-state SCF on the GPU and the learned preconditioner.

scf_chebyshev mirrors the reference's SCF loop (src/main.f90:66-95): per
iteration ncpaw_make_hamiltonian (density of the occupied orbitals, V_KS with
DIIS, D_ij), set_hminmax (power-iteration bounds for the first iterations),
chebyr_m (degree-10 Chebyshev filter on [mu, hmax]), orthog_phi, diagh, and
mu = eig(nocc+1) + dmua (src/diag.f90:108).

scf_block replaces filter + orthog + diagh by a few iterations of the block
preconditioned solver (opaw_block_solve), whose k-space preconditioner is
the Teter-Payne-Allan ratio, optionally corrected by a small MLP.  The MLP
is trained on (residual, error) pairs recorded by the solver itself
(train_preconditioner: Adam on the loss and gradient the library computes
on the GPU).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import lib


@dataclass
class ScfResult:
    eig: np.ndarray
    residual: float            # largest eigen-residual of the occupied bands at the end
    iterations: int
    seconds: float
    history: List[float] = field(default_factory=list)   # largest residual per SCF iteration
    h_applications: int = 0    # band applications of shs (power iterations, filter, residual checks, LOBPCG)


def _occupied_residual(g, eig, nocc):
    return float(np.max(g.residuals(eig)[:nocc]))


def scf_chebyshev(g, nocc: int, nscf: int = 30, tol: float = 1e-8, degree: int = 10, dmua: float = 0.1,
                  iscf_hminmax: int = 5, seed: int = 1234, dens0: Optional[np.ndarray] = None) -> ScfResult:
    """The reference's SCF (src/main.f90:66-95) on the resident orbitals of
    system 0 (h_type 1).  The first mu is the lower bound from hminmax plus
    a quarter of the interval.  With dens0 (the atomic superposition of
    get_vloc_ncoret), iteration 1 builds H from it with rho_ij = 0, as the
    reference does; otherwise from the starting orbitals."""
    g.torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    mu = None
    hlo = hhi = None
    eig = np.zeros(g.nb)
    napp = 0
    it = 0
    for it in range(1, nscf + 1):
        if it == 1 and dens0 is not None:
            g.ncpaw_make_hamiltonian_density(dens0, 0, diis=True)
        else:
            g.ncpaw_make_hamiltonian(0, diis=True)
        if it <= iscf_hminmax or hhi is None:
            hlo, hhi = g.hminmax(0, 1, 25, 1.2, seed + it)
            napp += 50
        if mu is None:
            mu = hlo + 0.25 * (hhi - hlo)
        g.chebyshev_filter(mu, hhi, degree)
        g.orthog_phi(0, 1)
        eig = g.diagh(0, 1)
        napp += degree * g.nb + g.nb
        mu = eig[nocc] + dmua if nocc < g.nb else eig[-1] + dmua
        res = _occupied_residual(g, eig, nocc)
        napp += g.nb
        hist.append(res)
        if res <= tol:
            break
    g.torch.cuda.synchronize()
    return ScfResult(eig=eig, residual=hist[-1], iterations=it, seconds=time.perf_counter() - t0, history=hist,
                     h_applications=napp)


def scf_block(g, nocc: int, nscf: int = 30, tol: float = 1e-8, inner: int = 3, mlp=None, nh: int = 16,
              record: int = 0, eta: float = 0.0, inner_max: int = 30, first_tol: float = 1e-2,
              dens0: Optional[np.ndarray] = None, record_max: int = 0) -> ScfResult:
    """SCF with the block preconditioned solver in place of the filter.
    Each iteration builds H from the current orbitals (iteration 1 from
    dens0 when given, with rho_ij = 0) and runs the block solver on it.  The
    solver's first residual is that of the orbitals it starts from, under
    the H of their own density: the SCF has converged when it is <= tol
    (iteration 2 on).
    eta = 0: `inner` solver iterations per SCF iteration; eta > 0: each
    inner solve runs (up to inner_max iterations) to max(tol, eta * the
    current SCF residual estimate), the previous iteration's, first_tol at
    iteration 1, so the preconditioner's quality sets the work.
    record_max > 0: no further training pairs once the context holds that
    many (each holds two band sets until prepared)."""
    g.torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    eig = np.zeros(g.nb)
    napp = 0
    it = 0
    g.orthog_phi(0, 1)
    prev = None
    res = np.inf
    for it in range(1, nscf + 1):
        if it == 1 and dens0 is not None:
            g.ncpaw_make_hamiltonian_density(dens0, 0, diis=True)
        else:
            g.ncpaw_make_hamiltonian(0, diis=True)
        rec = record if not (record_max and g.precond_count()[0] >= record_max) else 0
        if eta > 0.0:
            inner_tol = max(tol, first_tol) if prev is None else max(tol, eta * prev)
            eig, h, its = g.block_solve(inner_max, inner_tol, mlp=mlp, nh=nh, record=rec, nconv=nocc)
        else:
            eig, h, its = g.block_solve(inner, 0.0, mlp=mlp, nh=nh, record=rec, nconv=nocc)
        napp += g.nb * (its + 1)
        res = float(h[0])              # the start orbitals under the H of their density
        hist.append(res)
        prev = res
        if it > 1 and res <= tol:
            break
    g.torch.cuda.synchronize()
    return ScfResult(eig=eig, residual=res, iterations=it, seconds=time.perf_counter() - t0, history=hist,
                     h_applications=napp)


def mlp_init(nh: int = 16, seed: int = 0, scale: float = 0.1) -> np.ndarray:
    """Small random weights; the output bias and weights start at zero, so
    the untrained preconditioner equals Teter-Payne-Allan."""
    n = lib.load().opaw_mlp_param_count(nh)
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, n)
    p[4 * nh + nh * nh:] = 0.0       # w3, b3
    return p


def train_preconditioner(g, nh: int = 16, steps: int = 300, lr: float = 1e-2, seed: int = 0,
                         params: Optional[np.ndarray] = None, log_every: int = 0):
    """Adam on the misalignment loss of the prepared pairs of one context,
    or of several (g a list: the sum of their losses).  Returns (params,
    losses)."""
    gl = g if isinstance(g, (list, tuple)) else [g]

    class _Sum:
        @staticmethod
        def precond_loss(q, nh, grad=True):
            tot, gt = 0.0, (np.zeros_like(q) if grad else None)
            for h in gl:
                l, gr = h.precond_loss(q, nh, grad=grad)
                tot += l
                if grad:
                    gt += gr
            return tot, gt

    g = _Sum
    p = mlp_init(nh, seed) if params is None else np.array(params, dtype=np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    b1, b2, eps = 0.9, 0.999, 1e-8
    losses = []
    best = (np.inf, p.copy())
    for k in range(1, steps + 1):
        loss, grad = g.precond_loss(p, nh)
        losses.append(loss)
        if loss < best[0]:
            best = (loss, p.copy())
        m = b1 * m + (1 - b1) * grad
        v = b2 * v + (1 - b2) * grad * grad
        p = p - lr * (m / (1 - b1 ** k)) / (np.sqrt(v / (1 - b2 ** k)) + eps)
        if log_every and k % log_every == 0:
            print(f"[train] step {k} loss {loss:.6f}", flush=True)
    loss, _ = g.precond_loss(p, nh, grad=False)
    losses.append(loss)
    if loss < best[0]:
        best = (loss, p.copy())
    return best[1], losses


# polyacene training and evaluation sets of the learned preconditioner:
# rings -> grid at 0.5 bohr (the batched workload's grids for n = 1..6, 128^3
# for the held-out n = 7, 8)
ACENE_GRIDS = {1: (48, 48, 48), 2: (64, 48, 48), 3: (80, 64, 48), 4: (96, 64, 48), 5: (96, 64, 48),
               6: (128, 64, 48), 7: (128, 128, 128), 8: (128, 128, 128)}


def acene_context(nring: int, directory: str, distort=None, grid=None, seed: int = 2024):
    """GPU context of a (distorted) JTH polyacene with the reference's band
    count and seeded random start orbitals; returns (context, nocc, dens0)."""
    from . import synth, workloads as W
    from .api import OpawGPU
    spec = W.acene_system(nring, grid or ACENE_GRIDS[nring], 0.5, directory, distort=distort,
                          name=f"acene{nring}" + (f"-d{distort[1]}" if distort else ""))
    nocc = spec.nocc
    nb = W.nbands(nocc)
    spec.rnel = 2.0 * nocc
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=seed, envelope=4.0), spec.dv)
    g = OpawGPU(spec, "f64", nsys=1, nb=nb)
    g.set_nocc(nocc)
    g.set_diis(4, 0.15)
    g.write_orbitals(0, phi)
    return g, nocc, spec.extra["dens0"]


def acene_hamiltonian(nring: int, directory: str, distort=None, grid=None, seed: int = 2024):
    """A polyacene context whose H is the reference's SCF iteration 1 (the
    atomic superposition density, rho_ij = 0), with seeded random start
    orbitals: the fixed operator the preconditioned eigensolver is judged
    on.  Returns (context, nocc)."""
    g, nocc, d0 = acene_context(nring, directory, distort=distort, grid=grid, seed=seed)
    g.ncpaw_make_hamiltonian_density(d0, 0, diis=False)
    g.orthog_phi(0, 1)
    return g, nocc


def train_on_acenes(directory: str, rings=(1, 2, 3, 4, 5, 6), sigma: float = 0.1, seed: int = 7, nh: int = 8,
                    steps: int = 80, record: int = 2, tol: float = 1e-9, maxit: int = 300, log=print):
    """Training pairs from Teter-Payne-Allan solves (to tol) of the
    iteration-1 Hamiltonians of seeded sigma-bohr distortions of the
    polyacenes `rings`, every `record`-th iteration, then Adam on the sum of
    their losses.  Returns (params, losses)."""
    gs_ = []
    for n in rings:
        g, nocc = acene_hamiltonian(n, directory, distort=(sigma, seed + n))
        g.precond_keep_fields(False)
        t0 = time.perf_counter()
        _, hist, its = g.block_solve(maxit, tol, record=record, nconv=nocc)
        g.precond_prepare(0)
        log(f"[c5] train acene{n} (nb {g.nb}, nocc {nocc}): {its} iterations to {hist[-1]:.1e} in "
            f"{time.perf_counter() - t0:.2f} s, {g.precond_count()[1]} pairs")
        gs_.append(g)
    t0 = time.perf_counter()
    P, losses = train_preconditioner(gs_, nh=nh, steps=steps, lr=2e-2, seed=1)
    log(f"[c5] {steps} Adam steps in {time.perf_counter() - t0:.1f} s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    for g in gs_:
        g.close()
    return P, losses


@dataclass
class SolveResult:
    iterations: int
    seconds: float
    residual: float
    h_applications: int
    history: np.ndarray


def evaluate_on_acene(nring: int, directory: str, P, nh: int = 8, tol: float = 1e-8, maxit: int = 400, grid=None,
                      log=print):
    """The eigensolver on the undistorted polyacene's iteration-1 H, from
    the same start, to an occupied residual of tol: Teter-Payne-Allan and
    the MLP.  Returns {name: SolveResult}."""
    out = {}
    for name, p in (("tpa", None), ("mlp", P)):
        g, nocc = acene_hamiltonian(nring, directory, grid=grid)
        g.torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist, its = g.block_solve(maxit, tol, mlp=p, nh=nh, nconv=nocc)
        g.torch.cuda.synchronize()
        r = SolveResult(its, time.perf_counter() - t0, float(hist[-1]), g.nb * (its + 1), hist)
        g.close()
        log(f"[c5] acene{nring} {name}: {r.iterations} iterations, {r.seconds:.2f} s, residual {r.residual:.1e}, "
            f"{r.h_applications} band H applications")
        out[name] = r
    return out


# the held-out comparison of the solvers: grid -> (rings, bands, occupied)
SCF_SYSTEMS = {32: (1, 8, 6), 48: (2, 16, 12)}


def _scf_system(grid: int, seed: int):
    from . import synth
    from .api import OpawGPU
    nring, nb, nocc = SCF_SYSTEMS[grid]
    spec = synth.molecule_system(nring=nring, n=(grid, grid, grid), d=0.5, seed=seed)
    spec.rnel = 2.0 * nocc
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=seed + 1, envelope=3.0), spec.dv)
    g = OpawGPU(spec, "f64", nsys=1, nb=nb)
    g.set_nocc(nocc)
    g.write_orbitals(0, phi)
    return g, nocc


def main(argv=None):
    """Ground-state SCFs to 1e-8 on synthetic molecules (32^3: one ring, 8
    bands, 6 occupied; 48^3: two rings, 16 bands, 12 occupied): the
    Chebyshev-filter SCF, LOBPCG with Teter-Payne-Allan, and LOBPCG with the
    learned preconditioner, trained on the pairs of one other system's SCF.
    python -m opaw_amd.gs [--grid 48] [--methods cheb tpa mlp]
    The polyacene workload: python -m opaw_amd.gs --acenes [--eval 7 8]
    [--nh 16 --steps 300]."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--grid", type=int, choices=sorted(SCF_SYSTEMS), default=48)
    ap.add_argument("--methods", nargs="+", choices=["cheb", "tpa", "mlp"], default=["cheb", "tpa", "mlp"])
    ap.add_argument("--seeds", type=int, nargs="+", default=[33, 45], help="held-out systems")
    ap.add_argument("--train-seed", type=int, default=21)
    ap.add_argument("--nh", type=int, default=8, help="MLP hidden width")
    ap.add_argument("--steps", type=int, default=80, help="Adam steps")
    ap.add_argument("--acenes", action="store_true",
                    help="the polyacene workload: train on distorted n = 1..6, evaluate on --eval rings at 128^3")
    ap.add_argument("--eval", type=int, nargs="+", default=[7])
    ap.add_argument("--solve", type=int, default=0, metavar="ITERS",
                    help="only run ITERS Teter-Payne-Allan solver iterations on the first --eval polyacene (profiling)")
    a = ap.parse_args(argv)
    import os
    gold = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                        "tests", "golden", "jth")
    if a.solve:
        g, nocc = acene_hamiltonian(a.eval[0], gold)
        g.torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist, its = g.block_solve(a.solve, 0.0, nconv=nocc)
        g.torch.cuda.synchronize()
        print(f"[gs] acene{a.eval[0]} ({g.nb} bands): {its} iterations in {time.perf_counter() - t0:.2f} s, "
              f"residual {hist[-1]:.1e}", flush=True)
        g.close()
        return
    if a.acenes:
        P, _ = train_on_acenes(gold, nh=a.nh, steps=a.steps, log=lambda s: print(s, flush=True))
        for n in a.eval:
            out = evaluate_on_acene(n, gold, P, nh=a.nh, log=lambda s: print(s, flush=True))
            t, m = out["tpa"], out["mlp"]
            print(f"[gs] acene{n}: wall-time ratio TPA/MLP {t.seconds / m.seconds:.3f}, iterations "
                  f"{t.iterations} / {m.iterations}", flush=True)
        return
    P = None
    if "mlp" in a.methods:
        g, nocc = _scf_system(a.grid, a.train_seed)
        r = scf_block(g, nocc, nscf=80, tol=1e-8, eta=0.1, record=1)
        g.precond_prepare(0)
        t0 = time.perf_counter()
        P, losses = train_preconditioner(g, nh=a.nh, steps=a.steps, lr=2e-2, seed=1)
        print(f"[gs] {a.grid}^3 training: {g.precond_count()[1]} pairs from a {r.iterations}-iteration SCF, "
              f"{a.steps} Adam steps in {time.perf_counter() - t0:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}",
              flush=True)
        g.close()
    for seed in a.seeds:
        for name in a.methods:
            g, nocc = _scf_system(a.grid, seed)
            if name == "cheb":
                r = scf_chebyshev(g, nocc, nscf=80, tol=1e-8)
            else:
                r = scf_block(g, nocc, nscf=80, tol=1e-8, eta=0.1, mlp=(P if name == "mlp" else None), nh=a.nh)
            print(f"[gs] {a.grid}^3 seed {seed} {name}: {r.iterations} SCF iterations, {r.seconds:.3f} s, "
                  f"residual {r.residual:.1e}, {r.h_applications} band H applications", flush=True)
            g.close()


if __name__ == "__main__":
    main()
