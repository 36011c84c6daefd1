"""FP64 CPU restatement of the grid-sized PAW set-up.

TEST INFRASTRUCTURE: the oracle the HIP set-up kernels (opaw_paw_box,
opaw_paw_sphere) are checked against.  Only the tests may import it; it is
never on the product path.

  box_tables   get_local_p3d (src/p3d.f90:2-114): radial projector times the
               real Y_lm sampled on the fine grid (linear interpolation in r,
               minimum image of the atom), then contracted onto the rough
               box with the spline matrices of get_b
  sphere       map_localgrid (src/map_localgrid.f90:4-58) and
               get_localg3d (src/g3d.f90:1-119)

Pinning (tests/test_paw_cpu.py):
  box_tables   adjoint identity: sum_rough p3d psi dv equals
               sum_fine p psi_interp dv_fine for a random rough-grid psi
               interpolated by scipy's natural cubic spline (tensor
               product); separable product functions contract axis by axis
  sphere       membership vs a brute-force distance test over all images;
               g3d vs numpy.interp of g_l and scipy's spherical harmonics
               made real (with the reference's single-precision pi)
The spline matrices come from SPLINE_dn / SPLINT_dn
(src/lib/library_NRF_spline.f, src/lib/library_NRF_splint.f), the Numerical
Recipes natural cubic spline, checked against scipy's natural CubicSpline.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np


def bracket(rr: np.ndarray, t: np.ndarray):
    """src/interpolate.f90 for many targets: rr[i] <= t < rr[i+1]."""
    k = np.searchsorted(rr, t, side="right")
    i = np.clip(k - 1, 0, len(rr) - 2)
    j = i + 1
    wi = (rr[j] - t) / (rr[j] - rr[i])
    wj = (t - rr[i]) / (rr[j] - rr[i])
    return i, j, wi, wj


def ylm_array(l: int, m: int, xx, yy, zz, pi: float):
    """src/ylm.f90 for arrays (l <= 2 written out; enough for s, p, d)."""
    p1 = 1.0 / (2.0 * math.sqrt(pi))
    if l == 0:
        return np.full_like(xx, p1)
    if l == 1:
        p2 = p1 * math.sqrt(3.0)
        return p2 * {-1: yy, 0: zz, 1: xx}[m]
    if l == 2:
        p3 = math.sqrt(15.0) / (2.0 * math.sqrt(pi))
        if m == -2:
            return p3 * xx * yy
        if m == -1:
            return p3 * yy * zz
        if m == 0:
            return p3 * (2.0 * zz ** 2 - xx ** 2 - yy ** 2) / (2.0 * math.sqrt(3.0))
        if m == 1:
            return p3 * xx * zz
        return p3 * (xx ** 2 - yy ** 2) / 2.0
    raise ValueError("oracle ylm covers l <= 2")


def min_image(x, c, h):
    """minbyabs (src/p3d.f90:230-246)."""
    v = x - c
    a = x + 2.0 * h - c
    b = x - 2.0 * h - c
    v = np.where(np.abs(a) < np.abs(v), a, v)
    return np.where(np.abs(b) < np.abs(v), b, v)


def box_fine(rr, rad, rad1, lm, nfine, dxf, corner, coord, half, pi):
    """Fine-grid projectors (S = 2 ms, nf3, nf2, nf1)."""
    ax = [corner[k] + np.arange(nfine[k]) * dxf[k] for k in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    xc, yc, zc = (min_image(v, coord[k], half[k]) for k, v in enumerate((x, y, z)))
    r = np.maximum(np.sqrt(xc * xc + yc * yc + zc * zc), 1e-8)
    xx, yy, zz = xc / r, yc / r, zc / r
    i, j, wi, wj = bracket(rr, r)
    ms = len(lm)
    out = np.empty((2 * ms,) + r.shape)
    for s in range(ms):
        y_lm = ylm_array(int(lm[s][0]), int(lm[s][1]), xx, yy, zz, pi)
        out[s] = (rad[s][i] * wi + rad[s][j] * wj) * y_lm
        out[ms + s] = (rad1[s][i] * wi + rad1[s][j] * wj) * y_lm
    return out


def contract(F, bx, by, bz):
    """local(s, k, j, i) = sum bx(i,a) by(j,b) bz(k,c) F(s, c, b, a)."""
    return np.einsum("xa,yb,zc,scba->szyx", bx, by, bz, F, optimize=True)


def box_tables(rr, rad, rad1, lm, nfine, dxf, corner, coord, half, bx, by, bz, pi):
    F = box_fine(rr, rad, rad1, lm, nfine, dxf, corner, coord, half, pi)
    out = contract(F, bx, by, bz)
    ms = len(lm)
    return out[:ms], out[ms:]


def sphere(nx, ny, nz, dx, dy, dz, coord, rcut, rcomp, rr, gl, nl2, pi):
    """(flat grid indices in grid order, g3d (ngrid, nl2))."""
    h = (nx * dx / 2.0, ny * dy / 2.0, nz * dz / 2.0)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = -h[0] + ix * dx, -h[1] + iy * dy, -h[2] + iz * dz
    d = []
    for v, c, hh in ((x, coord[0], h[0]), (y, coord[1], h[1]), (z, coord[2], h[2])):
        d.append(np.minimum(np.minimum(np.abs(v - c), np.abs(v + 2 * hh - c)), np.abs(v - 2 * hh - c)))
    inside = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2) < rcut
    idx = np.nonzero(inside.reshape(-1))[0]
    xs, ys, zs = x.reshape(-1)[idx], y.reshape(-1)[idx], z.reshape(-1)[idx]
    xc, yc, zc = min_image(xs, coord[0], h[0]), min_image(ys, coord[1], h[1]), min_image(zs, coord[2], h[2])
    r = np.maximum(np.sqrt(xc * xc + yc * yc + zc * zc), 1e-8)
    xx, yy, zz = xc / r, yc / r, zc / r
    nrc = gl.shape[0]
    i, j, wi, wj = bracket(rr, r)
    i = np.minimum(i, nrc - 1)
    j = np.minimum(j, nrc - 1)
    g3d = np.zeros((len(idx), nl2))
    for L in range(nl2):
        l = int(math.floor(math.sqrt(L)))
        m = L - l * l - l
        g3d[:, L] = (gl[i, l] * wi + gl[j, l] * wj) * ylm_array(l, m, xx, yy, zz, pi)
    g3d[r > rcomp] = 0.0
    return idx, g3d


def natural_spline(x, y):
    """SPLINE_dn (src/lib/library_NRF_spline.f): natural end conditions."""
    n = len(x)
    y2 = np.zeros(n)
    u = np.zeros(n)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[i - 1] + 2.0
        y2[i] = (sig - 1.0) / p
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]))
                / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
    for k in range(n - 2, -1, -1):
        y2[k] = y2[k] * y2[k + 1] + u[k]
    return y2


def splint(x, y, y2, t):
    """SPLINT_dn for many points."""
    khi = np.clip(np.searchsorted(x, t, side="right"), 1, len(x) - 1)
    klo = khi - 1
    h = x[khi] - x[klo]
    a = (x[khi] - t) / h
    b = (t - x[klo]) / h
    return a * y[klo] + b * y[khi] + ((a ** 3 - a) * y2[klo] + (b ** 3 - b) * y2[khi]) * h * h / 6.0


def vloc_tuma(nx, ny, nz, dx, dy, dz, types, atom_type, coords):
    """vloc_prep (src/vloc/vloc_prep.f90:15-79) for the nonperiodic box:
    types = [(rr, vr)], atom_type 0-based, coords (natom, 3)."""
    nxb, nyb, nzb = 2 * nx, 2 * ny, 2 * nz
    dv = dx * dy * dz
    iz, iy, ix = np.meshgrid(np.arange(nzb), np.arange(nyb), np.arange(nxb), indexing="ij")
    X, Y, Z = (ix - nxb // 2) * dx, (iy - nyb // 2) * dy, (iz - nzb // 2) * dz
    rs = np.maximum(1e-10, np.sqrt(X * X + Y * Y + Z * Z))

    def kfold(n, d):
        k = np.arange(n) * (2 * np.pi / (n * d))
        return np.where(k > np.pi / d, k - 2 * np.pi / d, k)

    kx, ky, kz = kfold(nxb, dx), kfold(nyb, dy), kfold(nzb, dz)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    r0 = (-nxb / 2 * dx, -nyb / 2 * dy, -nzb / 2 * dz)
    shift = np.exp(-1j * (r0[0] * KX + r0[1] * KY + r0[2] * KZ))
    nyq = (np.arange(nxb)[None, None, :] == nxb // 2) | (np.arange(nyb)[None, :, None] == nyb // 2) | \
          (np.arange(nzb)[:, None, None] == nzb // 2)
    rstart = (-0.5 * dx * nxb, -0.5 * dy * nyb, -0.5 * dz * nzb)
    cvn = np.zeros((nzb, nyb, nxb), np.complex128)
    for t, (rr, vr) in enumerate(types):
        idx = [a for a in range(len(atom_type)) if atom_type[a] == t]
        if not idx:
            continue
        dr1 = rr[1] - rr[0]
        xs = np.maximum(rr, 0.1 * dr1)
        y2 = natural_spline(xs, vr)
        charge = vr[-1] * rr[-1]
        rl = np.maximum(rs, 0.1 * dr1)
        v = splint(xs, vr, y2, rl)
        v = np.where(rs < rr[0], vr[0], np.where(rs > rr[-1], charge / rs, v))
        cvkb = np.fft.fftn(v) * dv * shift
        form = np.zeros_like(cvn)
        for a in idx:
            c = coords[a]
            form += np.exp(-1j * ((c[0] - rstart[0]) * KX + (c[1] - rstart[1]) * KY + (c[2] - rstart[2]) * KZ))
        form[nyq] = 0.0
        cvn += form * cvkb
    out = np.fft.ifftn(cvn) * (nxb * nyb * nzb) / (nxb * nyb * nzb * dv)
    mx, my, mz = nxb // 2 - nx // 2, nyb // 2 - ny // 2, nzb // 2 - nz // 2
    return out[mz:mz + nz, my:my + ny, mx:mx + nx].real.reshape(-1)
