"""CPU checks of the Fortran iso_c_binding module (opaw-tddft_amd/fortran):
it compiles, every bind(C) name is an entry point of include/opaw_hip.h with
the same argument count, and a Fortran program linked against the library
gets the library's error status back (null context, no GPU call made).
Skipped when no Fortran compiler is installed."""
import os
import re
import shutil
import subprocess

import pytest

from opaw_amd import lib
from tests.test_abi_cpu import declared

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(REPO, "opaw-tddft_amd", "fortran", "opaw_hip_mod.f90")
FC = shutil.which("amdflang") or shutil.which("flang") or shutil.which("gfortran")

pytestmark = pytest.mark.skipif(FC is None, reason="no Fortran compiler")


def bound_functions():
    src = open(MOD).read()
    out = {}
    for m in re.finditer(r"function\s+(\w+)\s*\(([^)]*)\)\s*&?\s*\n?\s*bind\(C,\s*name=\"(\w+)\"\)", src):
        args = [a for a in m.group(2).split(",") if a.strip()]
        assert m.group(1) == m.group(3)
        out[m.group(3)] = len(args)
    return out


def test_bind_names_match_header():
    b = bound_functions()
    d = declared()
    assert len(b) >= 15
    for name, nargs in b.items():
        assert name in d, name
        assert d[name] == nargs, name


def test_module_compiles_and_links(tmp_path):
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built (run __graft_entry__.build())")
    prog = tmp_path / "t.f90"
    prog.write_text(
        "program t\n"
        "  use iso_c_binding\n"
        "  use opaw_hip_mod\n"
        "  implicit none\n"
        "  integer(c_int) :: st\n"
        "  st = opaw_rk4_prop_shs(c_null_ptr, 0.1d0)\n"
        "  print '(I0)', st\n"
        "end program\n")
    libdir = str(lib.LIB_PATH.parent)
    subprocess.run([FC, "-c", MOD, "-J", str(tmp_path), "-o", str(tmp_path / "m.o")], check=True)
    subprocess.run([FC, str(prog), str(tmp_path / "m.o"), "-I", str(tmp_path), "-J", str(tmp_path),
                    "-L", libdir, "-lopaw_hip", f"-Wl,-rpath,{libdir}", "-o", str(tmp_path / "t")],
                   check=True)
    out = subprocess.run([str(tmp_path / "t")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "-1"     # OPAW_ERR_ARG: null context
