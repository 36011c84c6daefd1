"""CPU check of the host Hermitian eigensolver of orthog_phi and diagh
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
