"""FP64 CPU restatement of the ground-state solver pieces.

TEST INFRASTRUCTURE: the oracle the HIP ground-state kernels are checked
against.  Only the tests may import it; it is never on the product path.

  cheb_filter    chebyr_m (src/chebyr_m.f90:1-48) with the oracle's shs
  hminmax        get_hminmax (src/hminmax.f90:18-108), random starts from
                 the same splitmix64 stream as the library
  Diis           diis_general (src/diis.f90:12-153)
  precond_*      the k-space preconditioner M = TPA(x) exp(g) and the
                 misalignment loss with its gradient (manual backward pass)

Pinning (tests): the Chebyshev recurrence against numpy's Chebyshev
polynomial of the dense shs matrix on a tiny grid; DIIS against the closed
form of one step (linear mixing) and the Lagrange conditions of the
minimisation; the loss gradient against central finite differences; TPA
against its defining polynomial ratio values at x = 0 (1) and large x
(27/(16 x^4) decay).
"""
from __future__ import annotations

import math

import numpy as np

from oracle import opaw_oracle as O


def cheb_filter(st, phis: np.ndarray, sys: int, degree: int, hbot: float, hmax: float) -> np.ndarray:
    havg = 0.5 * (hmax + hbot)
    dh = 0.5 * (hmax - hbot)
    ca = phis.copy()
    cb = phis.copy()
    cc = None
    for k in range(1, degree + 1):
        cc = O.shs(st, cb, sys)
        if k == 1:
            cc = (cc - havg * cb) / dh
        else:
            cc = 2.0 * (cc - havg * cb) / dh - ca
        ca, cb = cb, cc
    return cc


def splitmix_uniform(seed: int, pass_index: int, n: int) -> np.ndarray:
    mask = (1 << 64) - 1
    x = (seed + 0x9E3779B97F4A7C15 * (pass_index + 1)) & mask
    out = np.empty(n)
    for r in range(n):
        x = (x + 0x9E3779B97F4A7C15) & mask
        v = x
        v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & mask
        v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & mask
        v ^= v >> 31
        out[r] = (v >> 11) * 2.0 ** -53
    return out


def hminmax(st, sys: int, niter: int, fctr: float, seed: int, n: int):
    e = [0.0, 0.0]
    for p in range(2):
        z = splitmix_uniform(seed, p, n).astype(np.complex128)[None, :]
        for _ in range(niter):
            z = z / math.sqrt(np.sum(np.abs(z) ** 2))
            hz = O.shs(st, z, sys)
            if p == 1:
                hz = hz - e[0] * z
            e[p] = float(np.real(np.vdot(z, hz)) / np.real(np.vdot(z, z)))
            z = hz
    e2 = e[1] + e[0]
    hx, hn = max(e[0], e2), min(e[0], e2)
    havg, dh = 0.5 * (hn + hx), (hx - hn) * fctr
    return havg - dh / 2.0, havg + dh / 2.0


class Diis:
    """diis_general: bank of outputs, errors, and the extrapolation."""

    def __init__(self, L: int = 4, mix: float = 0.15, dv: float = 1.0):
        self.L, self.mix, self.dv = L, mix, dv
        self.stage = 0
        self.bank = []
        self.err = []

    def __call__(self, vec: np.ndarray) -> np.ndarray:
        if self.stage == 0:
            self.stage = 1
            self.bank = [vec.copy()]
            return vec.copy()
        if len(self.err) == self.L:
            self.err.pop(0)
        self.err.append(vec - self.bank[-1])
        K = len(self.err)
        B = np.zeros((K + 1, K + 1))
        for i in range(K):
            for j in range(K):
                B[i, j] = np.dot(self.err[i], self.err[j]) * self.dv
        B[:K, K] = 1.0
        B[K, :K] = 1.0
        rhs = np.zeros(K + 1)
        rhs[K] = 1.0
        C = np.linalg.solve(B, rhs)[:K]
        C = C / C.sum()
        out = sum((self.bank[k] + self.mix * self.err[k]) * C[k] for k in range(K))
        if len(self.bank) == self.L:
            self.bank.pop(0)
        self.bank.append(out.copy())
        return out


# ---------------------------------------------------------------------------
def tpa(x):
    p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
    return p / (p + 16.0 * x ** 4)


def mlp_unpack(P, nh):
    W1 = P[:2 * nh].reshape(nh, 2)
    b1 = P[2 * nh:3 * nh]
    W2 = P[3 * nh:3 * nh + nh * nh].reshape(nh, nh)
    b2 = P[3 * nh + nh * nh:4 * nh + nh * nh]
    w3 = P[4 * nh + nh * nh:5 * nh + nh * nh]
    b3 = P[5 * nh + nh * nh]
    return W1, b1, W2, b2, w3, b3


def mlp_forward(P, nh, u0, u1):
    W1, b1, W2, b2, w3, b3 = mlp_unpack(P, nh)
    U = np.stack([u0, u1], axis=-1)
    h1 = np.tanh(U @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3, (U, h1, h2)


def multiplier(P, nh, x, y):
    m = tpa(x)
    if P is None:
        return m, None
    g, cache = mlp_forward(P, nh, np.log1p(x), np.full_like(x, y))
    return m * np.exp(g), cache


def precond_loss(pairs, ek, N, P, nh, want_grad=True):
    """pairs: list of (rk (nb, N), ek_err (nb, N), t (nb), y (nb));
    ek: the kinetic table (N,) with its 1/N.  Returns (loss, grad)."""
    sums = []
    eemax = 0.0
    for rk, ee, t, y in pairs:
        bs = []
        for i in range(rk.shape[0]):
            m, _ = multiplier(P, nh, ek * N / t[i], y[i])
            z = np.conj(ee[i]) * rk[i]
            A = np.sum(m * z)
            B = np.sum(m * m * np.abs(rk[i]) ** 2)
            Ee = np.sum(np.abs(ee[i]) ** 2)
            bs.append((A, B, Ee))
            eemax = max(eemax, Ee)
        sums.append(bs)
    count = sum(1 for bs in sums for A, B, Ee in bs if Ee > 1e-12 * eemax and B > 0)
    L = 0.0
    grad = np.zeros_like(P) if (want_grad and P is not None) else None
    for (rk, ee, t, y), bs in zip(pairs, sums):
        for i, (A, B, Ee) in enumerate(bs):
            if not (Ee > 1e-12 * eemax and B > 0):
                continue
            L += 1.0 - abs(A) ** 2 / (Ee * B)
            if grad is None:
                continue
            x = ek * N / t[i]
            m, (U, h1, h2) = multiplier(P, nh, x, y[i])
            z = np.conj(ee[i]) * rk[i]
            dm = (-2.0 * np.real(np.conj(A) * z) * B + abs(A) ** 2 * 2.0 * m * np.abs(rk[i]) ** 2) / (Ee * B * B)
            dg = dm * m / count
            W1, b1, W2, b2, w3, b3 = mlp_unpack(P, nh)
            dz2 = (dg[:, None] * w3[None, :]) * (1 - h2 ** 2)
            dh1 = dz2 @ W2
            dz1 = dh1 * (1 - h1 ** 2)
            gW1 = dz1.T @ U
            gb1 = dz1.sum(0)
            gW2 = dz2.T @ h1
            gb2 = dz2.sum(0)
            gw3 = h2.T @ dg
            gb3 = dg.sum()
            grad += np.concatenate([gW1.reshape(-1), gb1, gW2.reshape(-1), gb2, gw3, [gb3]])
    return L / count, grad
