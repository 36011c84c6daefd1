"""CPU pins of the spectrum oracle (oracle/spectrum_oracle.py) by closed
forms, and of the host peak finder; the C ABI entry point is checked for
its argument errors without a GPU (every error below is raised before any
device call)."""
import numpy as np
import pytest
from scipy.special import dawsn

import tests.helpers  # noqa: F401  (puts the package and the oracle on sys.path)
from oracle import spectrum_oracle as SO
from opaw_amd import lib, spectrum


@pytest.mark.parametrize("g,w0,tau", [(0.0, 0.35, 20.0), (0.01, 0.8, 5.0), (0.05, 0.1, 1e6)])
def test_exponential_damping_matches_geometric_series(g, w0, tau):
    dt, nt = 0.05, 1501
    t = np.arange(nt) * dt
    d = np.exp(-g * t) * np.cos(w0 * t)
    w = np.linspace(0.0, 2.0, 401)
    alpha, s = SO.spectrum(d, dt, w, kind=1, tau=tau)
    ref = SO.geometric_alpha(g, w0, dt, nt, w, tau)
    assert np.max(np.abs(alpha[0] - ref)) < 1e-11 * np.max(np.abs(ref))
    assert np.allclose(s[0], 2 * w / np.pi * ref.imag, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_a_plausible_mistake_fails_the_pin():
    """Dropping the trapezoid half-weights moves alpha by dt/2 (d(0) = 1):
    far above the pin's tolerance."""
    dt, nt, tau = 0.05, 801, 10.0
    t = np.arange(nt) * dt
    d = np.cos(0.4 * t)
    w = np.linspace(0.0, 1.0, 51)
    ref = SO.geometric_alpha(0.0, 0.4, dt, nt, w, tau)
    wrong = -dt * (np.exp(-t / tau) * d) @ np.exp(1j * np.outer(t, w))
    assert np.max(np.abs(wrong - ref)) > 1e3 * 1e-11 * np.max(np.abs(ref))


def test_gaussian_damping_matches_integrals():
    """d = cos(w0 t): Re alpha is the integral of an even function (the
    trapezoid sum is spectrally accurate); Im alpha is the Dawson integral
    minus the Euler-Maclaurin end term dt^2 w / 12 (f'(0) = w)."""
    dt, tau, w0 = 0.02, 6.0, 0.7
    nt = int(12 * tau / dt) + 1       # damping exp(-72) at the end
    t = np.arange(nt) * dt
    w = np.linspace(0.05, 1.5, 59)
    alpha, _ = SO.spectrum(np.cos(w0 * t), dt, w, kind=2, tau=tau)
    re = np.sqrt(2 * np.pi) * tau / 4 * (np.exp(-0.5 * ((w - w0) * tau) ** 2) +
                                        np.exp(-0.5 * ((w + w0) * tau) ** 2))
    alpha = -alpha                    # the definition's sign
    assert np.max(np.abs(alpha[0].real - re)) < 1e-12 * re.max()
    k = np.sqrt(2) * tau
    im = 0.5 * k * (dawsn((w + w0) * tau / np.sqrt(2)) + dawsn((w - w0) * tau / np.sqrt(2)))
    im_trap = im - dt ** 2 * w / 12
    assert np.max(np.abs(alpha[0].imag - im_trap)) < 1e-7 * np.abs(im).max()
    assert np.max(np.abs(alpha[0].imag - im)) > 1e-6 * np.abs(im).max()   # the end term is seen


def test_no_damping_is_the_plain_trapezoid_dft():
    rng = np.random.default_rng(1)
    d = rng.standard_normal(64)
    dt = 0.1
    w = 2 * np.pi / (64 * dt) * np.arange(8)
    alpha, _ = SO.spectrum(d, dt, w, kind=0)
    c = np.ones(64)
    c[[0, -1]] = 0.5
    ref = -dt * np.conj(np.fft.fft(c * d))[:8]   # exp(+i w t) = conj of numpy's kernel
    assert np.max(np.abs(alpha[0] - ref)) < 1e-13 * np.abs(ref).max()


def test_peak_finder_recovers_lorentzian_centre():
    w = np.linspace(0.0, 1.0, 201)
    for w0 in (0.3, 0.3123, 0.71234):
        s = 1.0 / ((w - w0) ** 2 + 0.02 ** 2)
        p = spectrum.peaks(w, s)
        assert len(p) == 1
        assert abs(p[0][0] - w0) < 3e-4   # grid step 5e-3, half width 2e-2
    s = np.exp(-((w - 0.25) / 0.03) ** 2) + 0.5 * np.exp(-((w - 0.6) / 0.03) ** 2)
    p = spectrum.peaks(w, s)
    assert [round(x, 3) for x, _ in p] == [0.25, 0.6]


def test_spectrum_argument_errors_need_no_gpu():
    if not lib.LIB_PATH.exists():
        pytest.fail(f"{lib.LIB_PATH} not built (run __graft_entry__.build())")
    d = np.zeros((2, 10))
    with pytest.raises(lib.OpawError, match="dt must be positive"):
        spectrum.absorption(d, 0.0, np.linspace(0, 1, 5))
    with pytest.raises(lib.OpawError, match="tau must be positive"):
        spectrum.absorption(d, 0.1, np.linspace(0, 1, 5), damping="gauss", tau=0.0)
    with pytest.raises(ValueError, match="uniformly"):
        spectrum.absorption(d, 0.1, np.array([0.0, 0.1, 0.3]))
