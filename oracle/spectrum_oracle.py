"""FP64 CPU restatement of the absorption-spectrum post-processing.

TEST INFRASTRUCTURE: this module is the oracle the HIP spectrum kernel is
checked against.  Only the tests and the CPU-baseline leg of the benchmark
may import it; it is never on the product path.

The reference writes the dipole series d_dip(t) = (dip_pert - dip)/sm
(src/tddft_plot.f90:57,103) and leaves the spectrum to external scripts
(README.md:7-11), so there is no reference code to follow; the definition is
the linear-response one for a delta kick:

    alpha(w) = -dt * sum_n c_n f(t_n) d(t_n) exp(i w t_n),   t_n = n dt
    c_n = 1 - [n = 0]/2 - [n = nt-1]/2 (trapezoid rule)
    f = 1, exp(-t/tau) or exp(-t^2 / (2 tau^2))
    S(w) = (2 w / pi) Im alpha(w)

The minus sign: the kick phi -> phi exp(-i sm r) (src/tddft_excite.f90:11-12)
is the potential sm r delta(t) on the electrons and d_dip is the dipole of
the electron density, so the polarizability is -d_dip(t).  A series from
the time loop starts after the first step (row it of d_dip.dat holds the
state after it steps, src/tddft.f90:67-85); the sample d(0) = 0 of the
kick instant is prepended by the caller.

Pinning (tests/test_spectrum_cpu.py):
  exponential damping  the sum is a geometric series: for
                       d = exp(-g t) cos(w0 t) the closed form of the
                       discrete sum is exact to rounding
  Gaussian damping     Re alpha of d = cos(w0 t) equals the integral
                       (sqrt(2 pi) tau / 4)(exp(-(w-w0)^2 tau^2/2) +
                       exp(-(w+w0)^2 tau^2/2)) (even integrand: the
                       trapezoid sum is spectrally accurate); Im alpha equals
                       the Dawson-function integral minus the
                       Euler-Maclaurin end term dt^2 f'(0) / 12 = dt^2 w / 12
  peaks                parabolic interpolation recovers w0 of a Lorentzian
"""
from __future__ import annotations

import numpy as np

HARTREE_EV = 27.211386245988   # CODATA 2018


def damping(t: np.ndarray, kind: int, tau: float) -> np.ndarray:
    if kind == 0:
        return np.ones_like(t)
    if kind == 1:
        return np.exp(-t / tau)
    if kind == 2:
        return np.exp(-0.5 * (t / tau) ** 2)
    raise ValueError("damp_kind must be 0, 1 or 2")


def spectrum(d: np.ndarray, dt: float, w: np.ndarray, kind: int = 1, tau: float = 50.0):
    """d: (nsig, nt) or (nt,); w: frequencies (Hartree).  Returns
    (alpha (nsig, nw) complex, strength (nsig, nw))."""
    d = np.atleast_2d(np.asarray(d, dtype=np.float64))
    nt = d.shape[1]
    t = np.arange(nt) * dt
    c = np.ones(nt)
    c[0] -= 0.5
    c[-1] -= 0.5          # a single sample has weight 0
    g = d * (c * damping(t, kind, tau))[None, :]
    alpha = np.empty((d.shape[0], len(w)), np.complex128)
    for k, wk in enumerate(w):
        ph = wk * t
        alpha[:, k] = -dt * (g @ np.cos(ph) + 1j * (g @ np.sin(ph)))
    strength = (2.0 * np.asarray(w)[None, :] / np.pi) * alpha.imag
    return alpha, strength


def geometric_alpha(g: float, w0: float, dt: float, nt: int, w: np.ndarray, tau: float) -> np.ndarray:
    """Closed form of the discrete sum for d = exp(-g t) cos(w0 t) with
    exponential damping exp(-t/tau): d f exp(i w t) = sum over z in
    (-g + i w0, -g - i w0) of exp((z - 1/tau + i w) t) / 2, and
    sum_{n<nt} c_n r^n = (1 - r^nt)/(1 - r) - (1 + r^(nt-1))/2; alpha
    carries the minus sign of the definition."""
    out = np.zeros(len(w), np.complex128)
    for z in (-g + 1j * w0, -g - 1j * w0):
        r = np.exp((z - 1.0 / tau + 1j * np.asarray(w)) * dt)
        out += 0.5 * ((1 - r ** nt) / (1 - r) - 0.5 * (1 + r ** (nt - 1)))
    return -dt * out
