"""The reference's on-disk formats (CPU): files written by this library are
byte-identical to the files the reference's own WRITE statements produce for
the same data (Fortran unformatted wf_bar.bin in both header layouts, the
(' ',4e18.8) lines of d_dip.dat, the complex*16 restart records), and each
side reads what the other wrote (wf_bar.bin / .txt, eig_bar, dip.dat,
tddft_save/info).  The Fortran side (tests/cpu/formats_check.f90) holds the
reference's statements and is built with the image's Fortran compiler;
skipped when there is none."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from opaw_amd import formats, lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "cpu", "formats_check.f90")
FC = shutil.which("amdflang") or shutil.which("flang") or shutil.which("gfortran")

pytestmark = pytest.mark.skipif(FC is None, reason="no Fortran compiler")

NX, NY, NZ, NS = 5, 4, 3, 3
N = NX * NY * NZ
DX, DY, DZ = 0.4, 0.45, 0.5
NB_R, NDS_R, IT_R = 4, 1, 20


def _data():
    rng = np.random.default_rng(7)
    tmp = rng.standard_normal((NS, N)) * 10.0 ** rng.integers(-5, 4, (NS, N))
    tmp[0, 0] = 0.0                      # Re phi = 0 is stored as -|phi| = -0.0
    tmp[1, 3] = -0.0
    eig = np.array([-0.731234567890123, -0.25, 0.125e-3])
    # d_dip rows: time, d(3) -- zeros, signs, 3-digit exponents, carries
    dd = np.array([
        [0.0, 0.0, -0.0, 1.0],
        [0.1, -1.234567891234, 9.999999999e5, -9.9999999951e-3],
        [2.5, 1e-120, -3.3e150, 123456789.0],
        [10.0, 5.5e-8, -0.999999996, 0.5],
        [1234.5, -6.02214076e23, 1.602176634e-19, 2.0 ** -30],
        [0.05, 0.123456785, -0.123456775, 0.99999999],
    ])
    return tmp, eig, dd


def _phi():
    """The complex orbitals the Fortran side folds, and the stored values
    sign(Re phi)|phi| (|phi| = hypot(Re, Im), as libm's cabs)."""
    tmp, _, _ = _data()
    phi = tmp + 0.5j * np.roll(tmp, -1, axis=1)
    h = np.hypot(phi.real, phi.imag)
    return phi, np.where(phi.real > 0, h, -h)


@pytest.fixture(scope="module")
def fortran(tmp_path_factory):
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("fortran_formats")
    exe = d / "formats_check"
    subprocess.run([FC, "-O1", SRC, "-J", str(d), "-o", str(exe)], check=True, cwd=d)
    tmp, eig, dd = _data()
    with open(d / "in.raw", "wb") as f:
        f.write(struct.pack("<8i", NX, NY, NZ, NS, dd.shape[0], NB_R, NDS_R, IT_R))
        f.write(struct.pack("<3d", DX, DY, DZ))
        f.write(eig.tobytes())
        f.write(tmp.tobytes())
        f.write(dd.tobytes())
    subprocess.run([str(exe), "write"], check=True, cwd=d, timeout=60)

    def run(*args):
        subprocess.run([str(exe), *map(str, args)], check=True, cwd=d, timeout=60)
    return d, run


def _fortran_read(d, run, path, binary):
    run("read", path, "bin" if binary else "txt", N, 16)
    raw = (d / "read.raw").read_bytes()
    nx, ny, nz, nsp, ns = struct.unpack_from("<5i", raw, 0)
    dx, dy, dz = struct.unpack_from("<3d", raw, 20)
    off = 44
    states = []
    for _ in range(ns):
        io, sp = struct.unpack_from("<2i", raw, off)
        v = np.frombuffer(raw, np.float64, N, off + 8)
        states.append((io, sp, v.copy()))
        off += 8 + 8 * N
    return (nx, ny, nz, nsp, ns, dx, dy, dz), states


@pytest.mark.parametrize("short", [False, True])
def test_wf_binary_bytes_match_reference_writer(fortran, tmp_path, short):
    d, _ = fortran
    _, eig, _ = _data()
    phi, _ = _phi()
    out = tmp_path / "wf.bin"
    formats.write_wf(out, phi, eig, (NX, NY, NZ), (DX, DY, DZ), ihomo=2, binary=True,
                     short_header=short)
    ref = (d / ("wf_short.bin" if short else "wf_long.bin")).read_bytes()
    assert out.read_bytes() == ref


def test_wf_stored_orbital_is_signed_modulus(tmp_path):
    """src/write_wf.f90:55-63: tmp = |phi| where Re phi > 0, else -|phi|."""
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built")
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    phi[0, :4] = [1j, -1j, 0.0, 3.0 - 4.0j]
    out = tmp_path / "wf.bin"
    formats.write_wf(out, phi, np.zeros(2), (NX, NY, NZ), (DX, DY, DZ), binary=True)
    got, eig, ih = formats.read_wf(out, (NX, NY, NZ), (DX, DY, DZ), 4)
    h = np.hypot(phi.real, phi.imag)
    want = np.where(phi.real > 0, h, -h)
    assert np.array_equal(got.real, want) and np.all(got.imag == 0)
    assert got[0, 3].real == 5.0 and got[0, 0].real == -1.0 and np.signbit(got[0, 2].real)


@pytest.mark.parametrize("name,binary", [("wf_long.bin", True), ("wf_short.bin", True),
                                         ("wf_long.txt", False), ("wf_short.txt", False)])
def test_wf_reader_reads_reference_files(fortran, name, binary):
    d, _ = fortran
    _, eig, _ = _data()
    _, stored = _phi()
    phi, ev, ih = formats.read_wf(d / name, (NX, NY, NZ), (DX, DY, DZ), 8, binary=binary)
    assert phi.shape == (NS, N)
    assert np.array_equal(ev, eig)
    assert ih == (2 if "long" in name else 0)
    assert np.array_equal(phi.real, stored) and np.all(phi.imag == 0)


@pytest.mark.parametrize("binary", [True, False])
def test_reference_reader_reads_our_wf(fortran, tmp_path, binary):
    """read_wf_tddft reads the 8-record layout (short_header)."""
    d, run = fortran
    _, eig, _ = _data()
    phi, stored = _phi()
    out = tmp_path / ("wf.bin" if binary else "wf.txt")
    formats.write_wf(out, phi, eig, (NX, NY, NZ), (DX, DY, DZ), ihomo=2,
                     binary=binary, short_header=True)
    head, states = _fortran_read(d, run, out, binary)
    assert head == (NX, NY, NZ, 1, NS, DX, DY, DZ)
    for i, (io, sp, v) in enumerate(states):
        assert (io, sp) == (i + 1, 1)
        assert np.array_equal(v, stored[i])


def test_wf_reader_checks_like_reference(fortran):
    """src/read_wf.f90:181-187: grid, spacing and nsp must match."""
    d, _ = fortran
    with pytest.raises(lib.OpawError, match="grid differs"):
        formats.read_wf(d / "wf_long.bin", (NX + 1, NY, NZ), (DX, DY, DZ), 8)
    with pytest.raises(lib.OpawError, match="spacing differs"):
        formats.read_wf(d / "wf_long.bin", (NX, NY, NZ), (DX, DY, DZ * 1.01), 8)
    with pytest.raises(lib.OpawError, match="max_states"):
        formats.read_wf(d / "wf_long.bin", (NX, NY, NZ), (DX, DY, DZ), NS - 1)
    with pytest.raises(lib.OpawError, match="cannot open"):
        formats.read_wf(d / "missing.bin", (NX, NY, NZ), (DX, DY, DZ), 8)


def test_eig_round_trips_with_reference(fortran, tmp_path):
    d, run = fortran
    _, eig, _ = _data()
    assert np.array_equal(formats.read_eig(d / "eig_f", NS), eig)
    out = tmp_path / "eig_bar"
    formats.write_eig(out, eig)
    run("eig", out, NS)
    assert np.array_equal(np.frombuffer((d / "eig.raw").read_bytes(), np.float64), eig)


def test_d_dip_lines_match_reference_format(fortran, tmp_path):
    """`write(15,888) (it-1)*dt, d_dip(:)` with 888 format (' ',4e18.8):
    each row is written as step it = 2 of a series whose dt is the row's
    time, so (it-1)*dt is the time itself."""
    d, _ = fortran
    _, _, dd = _data()
    out = tmp_path / "d_dip.dat"
    for k, row in enumerate(dd):
        formats.write_dipoles(row[None, 1:], None, float(row[0]), it0=1, d_dip_path=out,
                              dip_path=None, append=k > 0)
    assert out.read_text() == (d / "d_dip_f.dat").read_text()


def test_dipole_series_times_and_dip_read_by_reference(fortran, tmp_path):
    d, run = fortran
    rng = np.random.default_rng(5)
    nt, nsys, dt = 7, 2, 0.05
    d_dip = rng.standard_normal((nt, 3))
    dip = rng.standard_normal((nt, 3 * nsys)) * 1e-3
    a, b = tmp_path / "d_dip.dat", tmp_path / "dip.dat"
    formats.write_dipoles(d_dip[:4], dip[:4], dt, 0, a, b)
    formats.write_dipoles(d_dip[4:], dip[4:], dt, 4, a, b, append=True)
    run("dip", b, nt, 1 + 3 * nsys)
    rows = np.frombuffer((d / "dip.raw").read_bytes(), np.float64).reshape(nt, -1)
    assert np.array_equal(rows[:, 0], np.arange(nt) * dt)
    assert np.array_equal(rows[:, 1:], dip)
    lines = a.read_text().splitlines()
    assert len(lines) == nt and all(len(s) == 1 + 4 * 18 for s in lines)
    got = np.array([[float(s[1 + 18 * q:1 + 18 * (q + 1)]) for q in range(4)] for s in lines])
    assert np.allclose(got[:, 0], np.arange(nt) * dt, rtol=0, atol=1e-12)
    assert np.allclose(got[:, 1:], d_dip, rtol=5.1e-8, atol=0)   # 8 significant digits


def test_dipole_continuation_after_restart(fortran, tmp_path):
    """src/tddft_plot.f90:62-93: a restart at it_restart copies the first
    it_restart lines into d_dip_cont.dat / dip_cont.dat (time and three
    values per dip line) and appends there; a second restart appends."""
    rng = np.random.default_rng(6)
    dt, nt, itr = 0.1, 6, 4
    d_dip = rng.standard_normal((nt, 3))
    dip = rng.standard_normal((nt, 6))
    formats.write_dipoles(d_dip, dip, dt, 0, tmp_path / "d_dip.dat", tmp_path / "dip.dat")
    a, b, app = formats.open_dipole_series(str(tmp_path), itr)
    assert a.endswith("d_dip_cont.dat") and b.endswith("dip_cont.dat") and app
    src = (tmp_path / "d_dip.dat").read_text().splitlines()
    assert (tmp_path / "d_dip_cont.dat").read_text().splitlines() == src[:itr]
    cont = np.loadtxt(tmp_path / "dip_cont.dat")
    assert cont.shape == (itr, 4)
    assert np.array_equal(cont[:, 1:], dip[:itr, :3])
    assert np.array_equal(cont[:, 0], np.arange(itr) * dt)
    # the continuation opens the *_cont pair for appending
    a2, b2, app2 = formats.open_dipole_series(str(tmp_path), itr + 1)
    assert (a2, b2, app2) == (a, b, True)
    # no restart: fresh files
    assert formats.open_dipole_series(str(tmp_path), 0)[2] is False


def test_restart_records_and_info_match_reference(fortran, tmp_path):
    """write_to_restart writes one complex*16 record per orbital: 4-byte
    length, the nx*ny*nz values, 4-byte length.  The GPU restart test
    checks the library's files against this same framing."""
    d, run = fortran
    _, tmp = _phi()
    raw = (d / "phi_f.bin").read_bytes()
    want = b""
    for i in range(NS):
        z = (tmp[i] + 1j * (-0.5 * tmp[i] + (i + 1))).astype(np.complex128)
        want += struct.pack("<i", 16 * N) + z.tobytes() + struct.pack("<i", 16 * N)
    assert raw == want
    run("rec", "phi_f.bin", N, NS)
    assert (d / "rec.raw").read_bytes() == b"".join(
        (tmp[i] + 1j * (-0.5 * tmp[i] + (i + 1))).astype(np.complex128).tobytes() for i in range(NS))
    # tddft_save/info: what `write(305,*) nb, nds, it` wrote is read back by
    # the whitespace reader of the restart path, and the library's own info
    # line is read by the reference's `read(34,*) nb, nds_read, it_restart`
    toks = (d / "info_f").read_text().split()
    assert list(map(int, toks)) == [NB_R, NDS_R, IT_R]
    (tmp_path / "info").write_text(f" {NB_R} 1 {IT_R}\n")
    run("info", tmp_path / "info")
    assert struct.unpack("<3i", (d / "info.raw").read_bytes()) == (NB_R, 1, IT_R)


@pytest.mark.parametrize("bad_ns", [-5, 2 ** 31 - 1])
def test_wf_reader_rejects_corrupt_state_count(tmp_path, bad_ns):
    """A negative or huge nstates record is refused before anything is
    allocated: an error status, not an exception through the C ABI."""
    _, eig, _ = _data()
    phi, _ = _phi()
    out = tmp_path / "wf.bin"
    formats.write_wf(out, phi, eig, (NX, NY, NZ), (DX, DY, DZ), ihomo=2, binary=True, short_header=False)
    raw = bytearray(out.read_bytes())
    off = 0
    for _ in range(7):                       # nx ny nz, dx dy dz, nsp
        n = struct.unpack_from("<i", raw, off)[0]
        off += 8 + n
    n = struct.unpack_from("<i", raw, off)[0]   # the nstates record: tag, then the i4 value
    struct.pack_into("<i", raw, off + 4 + n - 4, bad_ns)
    out.write_bytes(bytes(raw))
    with pytest.raises(lib.OpawError, match="nstates outside"):
        formats.read_wf(out, (NX, NY, NZ), (DX, DY, DZ), 8)
