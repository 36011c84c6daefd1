"""CPU check of the host Hermitian eigensolver of the subspace problems
(opaw-tddft_amd/csrc/opaw_linalg.h) against numpy.linalg.eigh: random,
degenerate, diagonal, already tridiagonal and tiny matrices."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "cpu", "linalg_check.cpp")
CXX = shutil.which("g++")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if CXX is None:
        pytest.skip("no g++")
    out = tmp_path_factory.mktemp("linalg") / "linalg_check"
    subprocess.run([CXX, "-O2", "-std=c++17", "-I", os.path.join(REPO, "opaw-tddft_amd", "csrc"), SRC, "-o",
                    str(out)], check=True)
    return str(out)


def _run(exe, A):
    n = A.shape[0]
    data = struct.pack("<i", n) + np.ascontiguousarray(A, dtype=np.complex128).tobytes()
    res = subprocess.run([exe], input=data, capture_output=True, check=True, timeout=60).stdout
    w = np.frombuffer(res, np.float64, n)
    V = np.frombuffer(res, np.complex128, n * n, 8 * n).reshape(n, n)
    return w, V


def _cases():
    rng = np.random.default_rng(0)
    out = []
    for n in (1, 2, 3, 5, 17, 48, 96):
        B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        out.append(("random", B + B.conj().T))
    out.append(("identity", np.eye(20, dtype=complex)))
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30)))
    lam = np.repeat([-1.0, 0.5, 0.5, 2.0, 3.0, 3.0], 5)
    out.append(("degenerate", Q @ np.diag(lam) @ Q.conj().T))
    off = rng.standard_normal(11) + 1j * rng.standard_normal(11)
    T = np.diag(rng.standard_normal(12)).astype(complex) + np.diag(off, -1) + np.diag(off.conj(), 1)
    out.append(("tridiagonal", T))
    out.append(("diagonal", np.diag(rng.standard_normal(9)).astype(complex)))
    out.append(("scaled", 1e-9 * out[3][1]))
    return out


@pytest.mark.parametrize("name,A", _cases())
def test_matches_eigh(exe, name, A):
    w, V = _run(exe, A)
    ref = np.linalg.eigvalsh(A)
    scale = max(1e-300, np.max(np.abs(ref)))
    assert np.max(np.abs(w - ref)) < 1e-12 * scale * max(1, A.shape[0] / 10), name
    n = A.shape[0]
    assert np.max(np.abs(V.conj().T @ V - np.eye(n))) < 1e-12 * max(1, n / 10), name
    assert np.max(np.abs(A @ V - V * w)) < 1e-12 * scale * max(1, n / 10), name


# ------------------------------------------------------ lowest eigenpairs
def _lowest(exe, A, nout):
    n = A.shape[0]
    data = struct.pack("<iii", -2, n, nout) + np.ascontiguousarray(A, dtype=np.complex128).tobytes()
    res = subprocess.run([exe], input=data, capture_output=True, check=True, timeout=60).stdout
    w = np.frombuffer(res, np.float64, nout)
    V = np.frombuffer(res, np.complex128, n * nout, 8 * nout).reshape(n, nout)
    return w, V


def _split_blocks():
    """block-diagonal with the same 4 x 4 block three times: the tridiagonal
    splits and every eigenvalue is threefold, across blocks"""
    rng = np.random.default_rng(5)
    B = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    B = B + B.conj().T
    return np.kron(np.eye(3), B)


@pytest.mark.parametrize("name,A", _cases() + [("split", _split_blocks())])
def test_lowest_matches_eigh(exe, name, A):
    """the nout lowest pairs by QL values and inverse iteration: eigenvalues
    as numpy's, orthonormal vectors (within degenerate clusters too) with
    residuals at rounding level"""
    n = A.shape[0]
    ref = np.linalg.eigvalsh(A)
    scale = max(1e-300, np.max(np.abs(ref)))
    for nout in sorted({1, max(1, n // 3), n}):
        w, V = _lowest(exe, A, nout)
        tol = 1e-12 * max(1, n / 10)
        assert np.max(np.abs(w - ref[:nout])) < tol * scale, (name, nout)
        assert np.max(np.abs(V.conj().T @ V - np.eye(nout))) < tol, (name, nout)
        assert np.max(np.abs(A @ V - V * w)) < tol * scale, (name, nout)


# --------------------------------------------------------------- the pencil
def _pencil(exe, S, H, nout, floor=1e-12):
    m = S.shape[0]
    data = (struct.pack("<iiid", -1, m, nout, floor) + np.ascontiguousarray(S, np.complex128).tobytes()
            + np.ascontiguousarray(H, np.complex128).tobytes())
    res = subprocess.run([exe], input=data, capture_output=True, check=True, timeout=60).stdout
    r, fast = struct.unpack("<ii", res[:8])
    if r < nout:
        return r, bool(fast), None, None
    w = np.frombuffer(res, np.float64, nout, 8)
    Cm = np.frombuffer(res, np.complex128, m * nout, 8 + 8 * nout).reshape(nout, m).T
    return r, bool(fast), w, Cm


def _basis(m, npt=300, seed=0):
    """S = Y^H Y and H = Y^H A Y for m random columns Y (npt points) with
    column norms spread over 12 decades, A random Hermitian"""
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((npt, m)) + 1j * rng.standard_normal((npt, m))
    Y *= 10.0 ** rng.uniform(-6, 6, m)
    B = rng.standard_normal((npt, npt)) + 1j * rng.standard_normal((npt, npt))
    return Y, 0.5 * (B + B.conj().T)


def _forms(Y, A):
    S = Y.conj().T @ Y
    H = Y.conj().T @ A @ Y
    return 0.5 * (S + S.conj().T), 0.5 * (H + H.conj().T)


def _check_pairs(S, H, w, Cm, tol):
    n = len(w)
    assert np.max(np.abs(Cm.conj().T @ S @ Cm - np.eye(n))) < tol
    scale = np.max(np.abs(w))
    assert np.max(np.abs(H @ Cm - S @ Cm * w)) < tol * scale * np.sqrt(np.max(np.abs(np.diag(S))))


@pytest.mark.parametrize("m,nout", [(1, 1), (5, 2), (48, 16), (96, 32)])
def test_pencil_well_conditioned_takes_cholesky(exe, m, nout):
    import scipy.linalg
    Y, A = _basis(m, seed=m)
    S, H = _forms(Y, A)
    r, fast, w, Cm = _pencil(exe, S, H, nout)
    assert r == m and fast
    d = 1.0 / np.sqrt(np.real(np.diag(S)))
    ref = scipy.linalg.eigh(d[:, None] * H * d, d[:, None] * S * d, eigvals_only=True)[:nout]
    assert np.max(np.abs(w - ref)) < 1e-10 * np.max(np.abs(ref))
    _check_pairs(d[:, None] * S * d, d[:, None] * H * d, w, Cm / d[:, None], 1e-10)


def test_pencil_dependent_column_falls_back_and_drops_it(exe):
    """an exact copy of a column: the overlap is singular, the Cholesky route
    is refused, one direction is dropped and the lowest pairs equal those of
    the basis without the copy"""
    import scipy.linalg
    m, nout = 24, 8
    Y, A = _basis(m, seed=7)
    Yd = np.concatenate([Y, Y[:, 3:4] * 2.5], axis=1)
    S, H = _forms(Yd, A)
    r, fast, w, Cm = _pencil(exe, S, H, nout)
    assert not fast and r == m
    S0, H0 = _forms(Y, A)
    ref = scipy.linalg.eigh(H0, S0, eigvals_only=True)[:nout]
    assert np.max(np.abs(w - ref)) < 1e-9 * np.max(np.abs(ref))
    d = 1.0 / np.sqrt(np.real(np.diag(S)))
    _check_pairs(d[:, None] * S * d, d[:, None] * H * d, w, Cm / d[:, None], 1e-8)


def test_pencil_near_dependent_column_falls_back(exe):
    """a column within 1e-8 (relative) of another: the unit-scaled overlap
    has an eigenvalue near 1e-16, below the floor, so the rank-revealing
    route runs and drops it"""
    m, nout = 24, 8
    Y, A = _basis(m, seed=9)
    rng = np.random.default_rng(1)
    near = Y[:, 5:6] + 1e-8 * np.linalg.norm(Y[:, 5]) * rng.standard_normal((Y.shape[0], 1)) / np.sqrt(Y.shape[0])
    S, H = _forms(np.concatenate([Y, near], axis=1), A)
    r, fast, w, Cm = _pencil(exe, S, H, nout)
    assert not fast and r == m
    S0, H0 = _forms(Y, A)
    import scipy.linalg
    ref = scipy.linalg.eigh(H0, S0, eigvals_only=True)[:nout]
    assert np.max(np.abs(w - ref)) < 1e-6 * np.max(np.abs(ref))


def test_pencil_rank_below_nout(exe):
    m = 6
    Y, A = _basis(m, seed=3)
    Yd = np.concatenate([Y[:, :2]] * 3, axis=1)
    S, H = _forms(Yd, A)
    r, fast, w, _ = _pencil(exe, S, H, 4)
    assert r == 2 and not fast and w is None
