"""GPU: the restart files of write_to_restart / read_phi_bar
(src/tddft_write_to_restart.f90, src/read_phi_bar.f90) through the C ABI.
The files hold the framing the reference's Fortran writes (pinned on the CPU
against the reference's statements), the orbitals come back bit-exact, and a
run continued from the files is bit-identical to the uninterrupted run (the
time loop restarts at it_restart + 1, src/tddft.f90:26-28, and every step
rebuilds the Hamiltonian from the orbitals alone)."""
import struct

import numpy as np
import pytest

import tests.helpers  # noqa: F401  (puts the package on sys.path)
from opaw_amd import synth

pytestmark = pytest.mark.gpu


def _system():
    spec = synth.molecule_system(nring=1, n=(32, 32, 32), d=0.5, seed=11)
    nb = 4
    spec.rnel = 2.0 * nb
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=3, envelope=3.0), spec.dv)
    return spec, phi


def _gpu(spec, prec, nsys, nb):
    import torch
    from opaw_amd.api import OpawGPU
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return OpawGPU(spec, prec, nsys=nsys, nb=nb)


def _framed(z):
    n = z.nbytes
    return struct.pack("<i", n) + z.tobytes() + struct.pack("<i", n)


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_restart_continues_bit_identically(prec, tmp_path):
    spec, phi = _system()
    nb = phi.shape[0]
    print(f"restart prec={prec} grid={spec.nx}x{spec.ny}x{spec.nz} nb={nb}", flush=True)
    g = _gpu(spec, prec, 2, nb)
    g.write_orbitals(0, phi)
    g.excite_psi_bar_pert(spec.sm, spec.ipol + 1)
    g.opaw_tddft(3, spec.dt, spec.sm)
    save = tmp_path / "tddft_save"
    g.write_to_restart(save, 3)
    assert (save / "info").read_text().split() == [str(nb), "1", "3"]
    for s, fn in ((0, "phi_bar.bin"), (1, "phi_bar_pert.bin")):
        o = g.read_orbitals(s).astype(np.complex128)
        assert (save / fn).read_bytes() == b"".join(_framed(o[i]) for i in range(nb))
    d_cont, dip_cont = g.opaw_tddft(2, spec.dt, spec.sm)
    o_cont = [g.read_orbitals(s) for s in (0, 1)]
    g.close()

    h = _gpu(spec, prec, 2, nb)
    assert h.read_phi_bar(save) == 3
    d_rest, dip_rest = h.opaw_tddft(2, spec.dt, spec.sm)
    assert np.array_equal(d_rest, d_cont)
    assert np.array_equal(dip_rest, dip_cont)
    for s in (0, 1):
        assert np.array_equal(h.read_orbitals(s), o_cont[s])
    h.close()


def test_restart_more_systems_and_checks(tmp_path):
    from opaw_amd.lib import OpawError
    spec, phi = _system()
    nb = phi.shape[0]
    g = _gpu(spec, "f64", 3, nb)
    rng = np.random.default_rng(2)
    data = [phi * np.exp(1j * rng.uniform(0, 6, (nb, 1))) for _ in range(3)]
    for s in range(3):
        g.write_orbitals(s, data[s])
    g.write_to_restart(tmp_path / "save", 12)
    assert (tmp_path / "save" / "phi_bar_pert2.bin").exists()
    g.close()
    h = _gpu(spec, "f64", 3, nb)
    assert h.read_phi_bar(tmp_path / "save") == 12
    for s in range(3):
        assert np.array_equal(h.read_orbitals(s), data[s])
    h.close()
    # nb of the info record must match the context (read_phi_bar allocates nb)
    k = _gpu(spec, "f64", 3, nb - 1)
    with pytest.raises(OpawError, match="restart has 4 orbitals"):
        k.read_phi_bar(tmp_path / "save")
    k.close()
    # nds /= 1 is the reference's 'nds_read/=nds' stop
    (tmp_path / "save" / "info").write_text(f" {nb} 2 12\n")
    k = _gpu(spec, "f64", 3, nb)
    with pytest.raises(OpawError, match="nds_read/=nds"):
        k.read_phi_bar(tmp_path / "save")
    k.close()
