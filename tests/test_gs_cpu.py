"""CPU pins of the ground-state oracle (oracle/gs_oracle.py): the Chebyshev
filter against numpy's Chebyshev series of the dense shs matrix, DIIS
against linear mixing (its first step) and the Lagrange conditions of its
minimisation, the preconditioner's Teter-Payne-Allan ratio at its limits,
and the loss gradient (the training kernels' oracle) against central finite
differences."""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

import tests.helpers  # noqa: F401
from oracle import gs_oracle as GO
from tests.test_oracle import dense_H, dense_S, tiny_state


def test_chebyshev_filter_is_the_chebyshev_polynomial():
    st, ekcut = tiny_state(atoms=[((-1, 3, 2), (3, 3, 2), 2)])
    n = st.grid.n
    SHS = dense_S(st, -0.5) @ dense_H(st, ekcut) @ dense_S(st, -0.5)
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    hbot, hmax, deg = -0.3, 4.0, 6
    out = GO.cheb_filter(st, psi, 0, deg, hbot, hmax)
    w, V = np.linalg.eigh(SHS)
    xs = (w - 0.5 * (hmax + hbot)) / (0.5 * (hmax - hbot))
    Tk = npcheb.chebval(xs, [0] * deg + [1])
    ref = (psi @ V.conj()) * Tk @ V.T
    assert np.max(np.abs(out - ref)) < 1e-10 * np.max(np.abs(ref))


def test_diis_first_step_is_linear_mixing_and_minimises():
    rng = np.random.default_rng(2)
    d = GO.Diis(L=4, mix=0.15, dv=0.5)
    v0 = rng.standard_normal(50)
    assert np.array_equal(d(v0), v0)
    v1 = rng.standard_normal(50)
    assert np.allclose(d(v1), v0 + 0.15 * (v1 - v0), rtol=0, atol=1e-15)
    outs = [d(rng.standard_normal(50)) for _ in range(6)]
    assert len(d.err) == 4 and len(d.bank) == 4
    # the coefficients minimise |sum C e|^2 subject to sum C = 1
    E = np.array(d.err)
    B = E @ E.T * 0.5
    C = np.linalg.solve(B, np.ones(4))
    C /= C.sum()
    g = B @ C
    assert np.allclose(g, g[0], rtol=1e-10)
    assert np.all(np.isfinite(outs[-1]))


def test_tpa_limits():
    assert GO.tpa(0.0) == 1.0
    x = 1e3
    assert abs(GO.tpa(x) * 16 * x ** 4 / (8 * x ** 3) - 1) < 1e-2
    xs = np.linspace(0, 10, 101)
    assert np.all(np.diff(GO.tpa(xs)) < 0)


@pytest.mark.parametrize("nh", [4, 16])
def test_loss_gradient_matches_finite_differences(nh):
    rng = np.random.default_rng(3)
    N, nb = 60, 3
    ek = rng.uniform(0, 2.0, N) / N
    pairs = []
    for _ in range(2):
        rk = rng.standard_normal((nb, N)) + 1j * rng.standard_normal((nb, N))
        ee = rng.standard_normal((nb, N)) + 1j * rng.standard_normal((nb, N))
        pairs.append((rk, ee, rng.uniform(0.5, 2.0, nb), rng.uniform(-1, 1, nb)))
    n = 5 * nh + nh * nh + 1
    P = rng.normal(0, 0.5, n)
    L, g = GO.precond_loss(pairs, ek, N, P, nh)
    assert 0 <= L <= 1
    h = 1e-6
    for k in rng.choice(n, 12, replace=False):
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (GO.precond_loss(pairs, ek, N, Pp, nh, False)[0] - GO.precond_loss(pairs, ek, N, Pm, nh, False)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(fd)), k


def test_scf_comparison_command_parses_without_a_gpu():
    """python -m opaw_amd.gs: the systems it builds and its arguments (the
    runs themselves need the GPU)"""
    import os
    import subprocess
    import sys
    from opaw_amd import gs
    assert gs.SCF_SYSTEMS == {32: (1, 8, 6), 48: (2, 16, 12)}
    out = subprocess.run([sys.executable, "-m", "opaw_amd.gs", "--help"], capture_output=True, text=True,
                         timeout=120, check=True,
                         cwd=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opaw-tddft_amd"))
    for flag in ("--grid", "--methods", "--seeds", "--nh", "--steps"):
        assert flag in out.stdout
