"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/opaw_hip.h, and the ctypes table matches the header.  No
compute call is made (no GPU here)."""
import os
import re
import subprocess

import pytest

from opaw_amd import lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "opaw_hip.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*([A-Za-z_][\w\s\*]*?)\b(opaw_\w+)\s*\(([^)]*)\)\s*;", src, flags=re.M):
        args = [a for a in m.group(3).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(2)] = len(args)
    return out


def test_header_declares_the_api():
    d = declared()
    assert len(d) >= 40
    for name in ("opaw_create", "opaw_h_phi", "opaw_shs", "opaw_sn_phi", "opaw_rk4_prop_shs",
                 "opaw_tddft_run", "opaw_get_dij", "opaw_orthog_phi", "opaw_diagh"):
        assert name in d


def test_ctypes_table_matches_header():
    d = declared()
    assert set(d) == set(lib.SIGNATURES)
    for name, nargs in d.items():
        assert len(lib.SIGNATURES[name][1]) == nargs, name


def test_library_loads_and_exports_every_symbol():
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built (run __graft_entry__.build())")
    L = lib.load()
    for name in declared():
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", str(lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (opaw_\w+)", nm))
    assert set(declared()) <= exported
    assert L.opaw_version() >= 10000
    # compiled FFT lengths cover every grid of the workloads (and their MT doubles)
    for n in (48, 64, 96, 128, 192, 256, 32, 12, 10, 8):
        assert L.opaw_fft_length_supported(n) == 1, n
    # lengths outside the compiled list run on the runtime-plan path when their
    # prime factors are <= 13 (the reference's own 56 x 52 x 32 grid and its
    # doubled 112 x 104 x 64 box)
    for n in (7, 13, 52, 56, 104, 112, 143, 224):
        assert L.opaw_fft_length_supported(n) == 1, n
    for n in (17, 34, 512, 1):
        assert L.opaw_fft_length_supported(n) == 0, n
    # xy planes with a one-visit plane kernel: the squares and the polyacene
    # rectangles; FP64 only where the plane fits the LDS
    F64, MIX = lib.OPAW_PREC_F64, lib.OPAW_PREC_MIXED
    for nx, ny in ((48, 48), (64, 64), (96, 96), (64, 48), (80, 64), (96, 64), (128, 64)):
        assert L.opaw_plane_path(nx, ny, F64) == 1 and L.opaw_plane_path(nx, ny, MIX) == 1, (nx, ny)
    assert L.opaw_plane_path(128, 128, MIX) == 1 and L.opaw_plane_path(128, 128, F64) == 0
    for nx, ny in ((48, 40), (40, 48), (56, 52), (48, 64), (192, 192)):
        assert L.opaw_plane_path(nx, ny, F64) == 0 and L.opaw_plane_path(nx, ny, MIX) == 0, (nx, ny)


def test_null_context_is_an_error_not_a_crash():
    L = lib.load()
    assert L.opaw_synchronize(None) == -1
    assert b"null" in L.opaw_last_error()
