"""The Stockham stage index logic of csrc/fft_core.h, compiled for the host
with g++ and run on emulated threads, against numpy.fft (every radix mix and
both LDS layouts used by the kernels), including the runtime-plan stages of
lengths outside the compiled list (radices 7, 11, 13), and the two line
passes of the xy-plane kernel on rectangular planes (nx != ny, padded LDS
rows) against numpy.fft.fft2."""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_stockham_stages_match_numpy(tmp_path):
    exe = tmp_path / "fftchk"
    src = os.path.join(REPO, "tests", "cpu", "fft_core_check.cpp")
    inc = os.path.join(REPO, "opaw-tddft_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{inc}", "-include", "array", src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows = np.loadtxt(out.splitlines())
    got = {}
    for N, S, l, j, re_, im in rows:
        got.setdefault((int(N), int(S), int(l)), {})[int(j)] = re_ + 1j * im
    assert len(got) > 20
    for (N, S, l), vals in got.items():
        j = np.arange(N)
        x = np.sin(0.37 * (j + 1) + l) + 1j * np.cos(0.11 * j * j + 0.5 * l)
        ref = np.fft.fft(x) if S < 0 else np.fft.ifft(x) * N
        v = np.array([vals[i] for i in range(N)])
        assert np.max(np.abs(v - ref)) / np.max(np.abs(ref)) < 1e-13, (N, S, l)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_rectangular_plane_passes_match_fft2(tmp_path):
    exe = tmp_path / "fftchk"
    src = os.path.join(REPO, "tests", "cpu", "fft_core_check.cpp")
    inc = os.path.join(REPO, "opaw-tddft_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{inc}", "-include", "array", src, "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "plane"], capture_output=True, text=True, check=True).stdout
    rows = np.loadtxt(out.splitlines())
    planes = {}
    for NX, NY, S, y, x, re_, im in rows:
        key = (int(NX), int(NY), int(S))
        planes.setdefault(key, np.zeros((int(NY), int(NX)), complex))[int(y), int(x)] = re_ + 1j * im
    assert len(planes) == 5
    for (NX, NY, S), got in planes.items():
        x = np.arange(NX)[None, :]
        y = np.arange(NY)[:, None]
        a = np.sin(0.37 * (x + 1) + 0.7 * y) + 1j * np.cos(0.11 * x * x + 0.5 * y)
        ref = np.fft.fft2(a) if S < 0 else np.fft.ifft2(a) * a.size
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-13, (NX, NY, S)
