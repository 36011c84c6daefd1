"""GPU: the spectrum kernel (opaw_spectrum) against the oracle and the
closed form, and the accuracy of the mixed path end to end: over a
2000-step C10H8 64^3 propagation (the reference's naphthalene run length,
dt = 0.05) the mixed dipole trace stays within 1e-4 (relative L2) of FP64
and the absorption peaks of the two move by <= 0.01 eV."""
import numpy as np
import pytest

import tests.helpers  # noqa: F401
from oracle import spectrum_oracle as SO
from opaw_amd import spectrum, synth

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch


@pytest.mark.parametrize("kind,nsig,nt,nw", [("none", 1, 1, 1), ("exp", 3, 1000, 257), ("gauss", 2, 4001, 130),
                                             ("exp", 5, 63, 5)])
def test_spectrum_kernel_matches_oracle(kind, nsig, nt, nw):
    _torch()
    print(f"spectrum kind={kind} nsig={nsig} nt={nt} nw={nw}", flush=True)
    rng = np.random.default_rng(nt + nw)
    d = rng.standard_normal((nsig, nt))
    dt = 0.05
    w = np.linspace(0.0, 1.2, nw) if nw > 1 else np.array([0.3])
    alpha, s = spectrum.absorption(d, dt, w, damping=kind, tau=30.0)
    ra, rs = SO.spectrum(d, dt, w, kind=spectrum.DAMP[kind], tau=30.0)
    scale = np.abs(ra).max()
    assert np.max(np.abs(alpha - ra)) <= 1e-12 * scale
    assert np.max(np.abs(s - rs)) <= 1e-12 * max(np.abs(rs).max(), scale)


def test_spectrum_kernel_matches_closed_form():
    _torch()
    dt, nt, g, w0, tau = 0.05, 2000, 0.004, 0.45, 40.0
    t = np.arange(nt) * dt
    d = np.exp(-g * t) * np.cos(w0 * t)
    w = np.linspace(0.0, 1.0, 1001)
    alpha, s = spectrum.absorption(d, dt, w, damping="exp", tau=tau)
    ref = SO.geometric_alpha(g, w0, dt, nt, w, tau)
    assert np.max(np.abs(alpha[0] - ref)) <= 1e-11 * np.abs(ref).max()
    (wp, _), = spectrum.peaks(w, np.abs(alpha[0]) ** 2, rel_height=0.5)
    assert np.allclose(s[0], 2 * w / np.pi * alpha[0].imag, rtol=1e-13, atol=0)
    assert abs(wp - w0) < 2e-3


def test_mixed_path_dipole_and_peaks_match_fp64():
    _torch()
    from opaw_amd.api import OpawGPU
    nb, nt = 24, 2000
    spec = synth.molecule_system(nring=2, n=(64, 64, 64), d=0.5, seed=1234)
    spec.rnel = 2.0 * nb
    phi0 = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=7, envelope=4.0), spec.dv)
    dd = {}
    for prec in ("f64", "mixed"):
        print(f"propagating C10H8 64^3 {prec} for {nt} steps", flush=True)
        g = OpawGPU(spec, prec, nsys=2, nb=nb)
        g.write_orbitals(0, phi0)
        g.excite_psi_bar_pert(spec.sm, spec.ipol + 1)
        dd[prec], _ = g.opaw_tddft(nt, spec.dt, spec.sm)
        g.close()
    err = np.linalg.norm(dd["mixed"] - dd["f64"]) / np.linalg.norm(dd["f64"])
    print(f"dipole relative L2 (mixed vs f64, {nt} steps): {err:.3e}", flush=True)
    assert err <= 1e-4
    # free motion right after the kick: every electron moves with velocity
    # -sm, so d_dip(t) = -N_e t at small t (the first row is t = dt)
    ne = 2 * nb
    assert abs(dd["f64"][0, spec.ipol] / spec.dt + ne) < 0.02 * ne
    w = np.linspace(0.002, 3.0, 6000)   # Hartree
    series = np.stack([dd["f64"][:, spec.ipol], dd["mixed"][:, spec.ipol]])
    _, s = spectrum.kick_spectrum(series, spec.dt, w, damping="exp", tau=30.0)
    fsum = s.sum(axis=1) * (w[1] - w[0])
    print(f"f-sum: f64 {fsum[0]:.3f}, mixed {fsum[1]:.3f}, N_e {ne}", flush=True)
    assert np.all(np.abs(fsum - ne) < 0.05 * ne)   # nonlocal projectors move it a little
    p64 = spectrum.peaks(w, s[0], rel_height=0.1)
    p32 = spectrum.peaks(w, s[1], rel_height=0.1)
    print("peaks (eV) f64  :", [round(p * SO.HARTREE_EV, 4) for p, _ in p64], flush=True)
    print("peaks (eV) mixed:", [round(p * SO.HARTREE_EV, 4) for p, _ in p32], flush=True)
    assert len(p64) >= 3 and len(p64) == len(p32)
    shift = max(abs(a - b) for (a, _), (b, _) in zip(p64, p32)) * SO.HARTREE_EV
    print(f"largest peak shift: {shift:.2e} eV", flush=True)
    assert shift <= 0.01
