"""The reference's on-disk formats around the propagation path, through the
C ABI (opaw_write_wf ... opaw_write_dipoles in include/opaw_hip.h).

    wf_bar.bin / wf_bar.txt   writewf / write_wf_for_sgw_debug
                              (src/write_wf.f90), read_wf_tddft
                              (src/read_wf.f90:130-228)
    eig_bar                   src/diag.f90:109-115
    d_dip.dat / dip.dat       plot_dip (src/tddft_plot.f90:60-106)

All functions are host-only (no GPU needed).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

from . import lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def write_wf(path: str, phi: np.ndarray, eig: np.ndarray, grid, spacing, ihomo: int = 0,
             binary: bool = True, short_header: bool = False):
    """phi: (nstates, nz, ny, nx) or (nstates, N) complex (stored as
    sign(Re phi)|phi|); grid = (nx, ny, nz); spacing = (dx, dy, dz)."""
    nx, ny, nz = map(int, grid)
    phi = np.ascontiguousarray(np.asarray(phi, dtype=np.complex128).reshape(-1, nx * ny * nz))
    eig = np.ascontiguousarray(eig, dtype=np.float64)
    if eig.size != phi.shape[0]:
        raise ValueError("one eigenvalue per state")
    lib.call("opaw_write_wf", str(path).encode(), int(binary), int(short_header), nx, ny, nz,
             *map(float, spacing), phi.shape[0], int(ihomo), _dp(eig), phi.ctypes.data)


def read_wf(path: str, grid, spacing, max_states: int, binary: bool = True
            ) -> Tuple[np.ndarray, np.ndarray, int]:
    """Returns (phi (nstates, N) complex, eig (nstates), ihomo)."""
    nx, ny, nz = map(int, grid)
    n = nx * ny * nz
    phi = np.zeros((max_states, n), np.complex128)
    eig = np.zeros(max_states)
    ns = C.c_int()
    ih = C.c_int()
    lib.call("opaw_read_wf", str(path).encode(), int(binary), nx, ny, nz, *map(float, spacing),
             int(max_states), C.byref(ns), C.byref(ih), _dp(eig), phi.ctypes.data)
    return phi[:ns.value].copy(), eig[:ns.value].copy(), ih.value


def write_eig(path: str, eig: np.ndarray):
    eig = np.ascontiguousarray(eig, dtype=np.float64)
    lib.call("opaw_write_eig", str(path).encode(), eig.size, _dp(eig))


def read_eig(path: str, n: int) -> np.ndarray:
    eig = np.zeros(n)
    lib.call("opaw_read_eig", str(path).encode(), int(n), _dp(eig))
    return eig


def write_dipoles(d_dip: Optional[np.ndarray], dip: Optional[np.ndarray], dt: float, it0: int = 0,
                  d_dip_path: Optional[str] = "d_dip.dat", dip_path: Optional[str] = "dip.dat",
                  append: bool = False):
    """Rows k of d_dip (nt, 3) and dip (nt, 3*nsys) are steps it = it0+k+1."""
    dd = None if d_dip is None else np.ascontiguousarray(d_dip, dtype=np.float64).reshape(-1, 3)
    dp = None if dip is None else np.ascontiguousarray(dip, dtype=np.float64)
    nt = (dd if dd is not None else dp).shape[0]
    nsys = 1 if dp is None else dp.reshape(nt, -1).shape[1] // 3
    lib.call("opaw_write_dipoles",
             None if dd is None or d_dip_path is None else str(d_dip_path).encode(),
             None if dp is None or dip_path is None else str(dip_path).encode(),
             int(append), int(it0), int(nt), float(dt),
             None if dd is None else _dp(dd), None if dp is None else _dp(dp), int(nsys))


def open_dipole_series(directory: str, it_restart: int) -> Tuple[str, str, bool]:
    """Which files plot_dip continues after a restart at it_restart
    (src/tddft_plot.f90:62-93): the *_cont.dat pair when it exists; else,
    when d_dip.dat exists, a new *_cont.dat pair seeded with its first
    it_restart lines (dip_cont.dat gets the time and the first three values
    of each dip.dat line, as the reference's `read(19,*) tmp, tmp3d` does);
    else a fresh d_dip.dat / dip.dat.  Returns (d_dip path, dip path,
    append)."""
    j = lambda f: os.path.join(directory, f)
    if it_restart <= 0:
        return j("d_dip.dat"), j("dip.dat"), False
    if os.path.exists(j("d_dip_cont.dat")):
        return j("d_dip_cont.dat"), j("dip_cont.dat"), True
    if not os.path.exists(j("d_dip.dat")):
        return j("d_dip.dat"), j("dip.dat"), False

    def first_rows(fn):
        rows = []
        with open(fn) as f:
            for line in f:
                v = line.replace("D", "E").replace("d", "e").split()
                if v:
                    rows.append([float(x) for x in v[:4]])
                if len(rows) == it_restart:
                    break
        if len(rows) < it_restart:
            raise ValueError(f"{fn}: fewer than {it_restart} lines to continue from")
        return np.array(rows)

    a = first_rows(j("d_dip.dat"))
    b = first_rows(j("dip.dat"))
    # each line keeps the time it was read with: written as step it0 = 1 of
    # a series whose "dt" is that time
    for k in range(it_restart):
        write_dipoles(a[k:k + 1, 1:], b[k:k + 1, 1:], float(a[k, 0]), 1,
                      j("d_dip_cont.dat"), None, append=k > 0)
        write_dipoles(None, b[k:k + 1, 1:], float(b[k, 0]), 1, None, j("dip_cont.dat"), append=k > 0)
    return j("d_dip_cont.dat"), j("dip_cont.dat"), True
