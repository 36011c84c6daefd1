"""Absorption spectrum of the d_dip series (post-processing; the reference
leaves it to external scripts, README.md:7-11).  The transform runs on the
GPU through opaw_spectrum (include/opaw_hip.h); peak location is host work.

For a kick phi -> phi exp(-i sm r_pol) (src/tddft_excite.f90:11-12) the
series d_dip = (dip_pert - dip)/sm (src/tddft_plot.f90:57) is minus the
response function alpha_pol,pol(t) (d_dip is the electron density's
dipole); the dipole strength is S(w) = (2 w / pi) Im alpha(w), positive at
absorption peaks, with integral S dw = N_electrons for a local Hamiltonian
(Thomas-Reiche-Kuhn sum rule); the isotropic spectrum averages the three
kick directions.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Tuple

import numpy as np

from . import lib

HARTREE_EV = 27.211386245988   # CODATA 2018
DAMP = dict(none=0, exp=1, gauss=2)


def absorption(d: np.ndarray, dt: float, w: np.ndarray, damping: str = "exp", tau: float = 50.0
               ) -> Tuple[np.ndarray, np.ndarray]:
    """d: (nt,) or (nsig, nt) series sampled at t = n dt; w: uniformly spaced
    frequencies in Hartree.  Returns (alpha (nsig, nw) complex,
    strength (nsig, nw))."""
    d = np.ascontiguousarray(np.atleast_2d(d), dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    nw = len(w)
    dw = float(w[1] - w[0]) if nw > 1 else 0.0
    if nw > 2 and not np.allclose(np.diff(w), dw, rtol=1e-9, atol=1e-15):
        raise ValueError("frequencies must be uniformly spaced")
    nsig, nt = d.shape
    alpha = np.zeros((nsig, nw), np.complex128)
    strength = np.zeros((nsig, nw))
    dp = C.POINTER(C.c_double)
    lib.call("opaw_spectrum", d.ctypes.data_as(dp), nsig, nt, float(dt), DAMP[damping], float(tau),
             float(w[0]), dw, nw, alpha.ctypes.data_as(dp), strength.ctypes.data_as(dp))
    return alpha, strength


def kick_spectrum(d_dip: np.ndarray, dt: float, w: np.ndarray, damping: str = "exp", tau: float = 30.0
                  ) -> Tuple[np.ndarray, np.ndarray]:
    """Spectrum of d_dip series as the time loop returns them: row k holds
    the state after k+1 steps (plot_dip runs after propdt,
    src/tddft.f90:67-85), so the kick instant's sample d(0) = 0 is
    prepended.  d_dip: (nt,) or (nsig, nt)."""
    d = np.atleast_2d(np.asarray(d_dip, dtype=np.float64))
    d = np.concatenate([np.zeros((d.shape[0], 1)), d], axis=1)
    return absorption(d, dt, w, damping, tau)


def peaks(w: np.ndarray, s: np.ndarray, rel_height: float = 0.05) -> List[Tuple[float, float]]:
    """Local maxima of s above rel_height * max(s), each refined by the
    vertex of the parabola through it and its two neighbours.  Returns
    [(w_peak, s_peak)] in increasing w."""
    w = np.asarray(w, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    top = s.max()
    out = []
    for k in range(1, len(s) - 1):
        if s[k] > s[k - 1] and s[k] >= s[k + 1] and s[k] > rel_height * top:
            a, b, c = s[k - 1], s[k], s[k + 1]
            den = a - 2 * b + c
            x = 0.5 * (a - c) / den if den != 0 else 0.0
            h = w[1] - w[0]
            out.append((w[k] + x * h, b - 0.25 * (a - c) * x))
    return out
