"""CPU checks of the Fortran side of the boundary (opaw-tddft_amd/fortran):
the iso_c_binding module compiles, every bind(C) name is an entry point of
include/opaw_hip.h with the same argument count, a Fortran program linked
against the library gets the library's error status back, and every
reference-named drop-in of opaw_dropin.f90 (h_phi, h_phi_pert, shs,
shs_pert, sn_phi, proj, proj1, rk4_prop_shs, with the reference's argument
lists) reaches the library and stops the run with its message on failure
(null context: no GPU call is made).  Skipped when no Fortran compiler is
installed."""
import os
import re
import shutil
import subprocess

import pytest

from opaw_amd import lib
from tests.test_abi_cpu import declared

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(REPO, "opaw-tddft_amd", "fortran", "opaw_hip_mod.f90")
DROPIN = os.path.join(REPO, "opaw-tddft_amd", "fortran", "opaw_dropin.f90")
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


# one program per drop-in: the reference's call, with its argument types
DROPIN_CALLS = {
    "h_phi": "call h_phi(a, b, 1)",
    "h_phi_pert": "call h_phi_pert(a, b, 1)",
    "shs": "call shs(a, b, 1)",
    "shs_pert": "call shs_pert(a, b, 1)",
    "sn_phi": "call sn_phi(a, b, 1, -0.5d0)",
    "proj": "call proj(1, a, ca, 4, 1)",
    "proj1": "call proj1(1, a, ca, 4, 1)",
    "rk4_prop_shs": "call rk4_prop_shs(p, pp, 8, 2)",
}


@pytest.fixture(scope="module")
def dropin_objs(tmp_path_factory):
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("fdropin")
    subprocess.run([FC, "-c", MOD, "-J", str(d), "-o", str(d / "m.o")], check=True)
    subprocess.run([FC, "-c", DROPIN, "-I", str(d), "-J", str(d), "-o", str(d / "dropin.o")], check=True)
    return d


def _run_dropin(d, name, call):
    prog = d / f"p_{name}.f90"
    prog.write_text(
        "program t\n"
        "  implicit none\n"
        "  complex*16 :: a(8), b(8), ca(4), p(8, 2), pp(8, 2)\n"
        "  a = (1d0, 0d0)\n"
        "  p = (1d0, 0d0)\n"
        "  pp = (1d0, 0d0)\n"
        f"  {call}\n"
        "  print *, 'returned'\n"
        "end program\n")
    libdir = str(lib.LIB_PATH.parent)
    exe = d / f"p_{name}"
    subprocess.run([FC, str(prog), str(d / "dropin.o"), str(d / "m.o"), "-I", str(d), "-J", str(d),
                    "-L", libdir, "-lopaw_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("name", sorted(DROPIN_CALLS))
def test_dropin_reaches_library_and_stops(dropin_objs, name):
    out = _run_dropin(dropin_objs, name, DROPIN_CALLS[name])
    text = out.stdout + out.stderr
    assert out.returncode != 0, text
    assert "returned" not in out.stdout
    assert f"opaw_hip: {name} failed" in " ".join(text.split()), text
    assert "null context" in text, text


def test_dropin_rejects_other_k_points(dropin_objs):
    out = _run_dropin(dropin_objs, "h_phi_ik2", "call h_phi(a, b, 2)")
    text = out.stdout + out.stderr
    assert out.returncode != 0 and "only ik = 1" in " ".join(text.split()), text
