"""The multi-molecule time loop (opaw_tddft_run_batch) through the C ABI.

Every kernel of the time step reduces in a fixed order (no atomics), so a
context advanced inside a batch must reproduce, bit for bit, the same context
advanced alone by opaw_tddft_run from the same start; that run is itself
checked against the oracle in the other GPU tests.

* Three molecules with different grids (one on the runtime-planned
  56 x 52 x 32 grid), orbital counts, system counts and precisions: graph
  replay across two batch calls with an eager step and host synchronisations
  between them (so each captured step is replayed several times, and the
  replays are interleaved with ordinary work), then the plain-enqueue batch;
  dipoles of every step and the final orbitals equal the contexts' own runs.
* One step of a batched benzene (48^3, 2 systems) against the oracle.
* The polyacene series n = 1..6 (48^3 ... 128 x 64 x 48; 15 ... 60 orbitals;
  4 systems each: unperturbed and kicks along x, y, z), mixed path, graph
  replay: equal to the molecules run alone, norms conserved.
* The time loop on the legacy null stream (twice) and on a stream of its
  own: bit-identical.
* Argument errors: a context twice, a null context, nt <= 0.
"""
import numpy as np
import pytest

from tests.helpers import O, rel, state_from_spec
from opaw_amd import batch, synth

pytestmark = pytest.mark.gpu


def _api():
    import torch
    from opaw_amd.api import OpawGPU
    assert torch.cuda.is_available(), "GPU test without a GPU"
    return torch, OpawGPU


def _norms(spec, phi):
    return np.sqrt(spec.dv * np.sum(np.abs(phi) ** 2, axis=1))


def _small_set():
    """(spec, orbitals, precision, nsys) of three uneven molecules."""
    out = []
    s1 = batch.polyacene_spec(1)
    out.append((s1, batch.polyacene_orbitals(s1, seed=3), "f64", 2))
    s2 = synth.molecule_system(nring=2, n=(56, 52, 32), d=0.5, seed=31)
    s2.rnel = 2.0 * 6
    phi2 = synth.lowdin_orthonormalise(synth.random_orbitals(s2, 6, seed=9, envelope=3.0), s2.dv)
    out.append((s2, phi2, "mixed", 3))
    s3 = batch.polyacene_spec(2)
    out.append((s3, batch.polyacene_orbitals(s3, seed=4), "f64", 4))
    return out


def _alone(spec, phi, prec, nsys, pieces):
    """dipoles per step and final orbitals of one context advanced by
    opaw_tddft in the given pieces of steps"""
    torch, _ = _api()
    g = batch.kicked_context(spec, phi, prec, nsys)
    dips = []
    for n in pieces:
        _, dip = g.opaw_tddft(n, spec.dt, spec.sm)
        torch.cuda.synchronize()
        dips.append(dip)
    orbs = [g.read_orbitals(s) for s in range(nsys)]
    g.close()
    return np.concatenate(dips), orbs


def test_batch_equals_alone_graph_and_plain():
    torch, _ = _api()
    sets = _small_set()
    dt = sets[0][0].dt
    ref = [_alone(spec, phi, prec, nsys, (3, 1, 4)) for spec, phi, prec, nsys in sets]
    print("[batch] references done", flush=True)
    gs = [batch.kicked_context(spec, phi, prec, nsys) for spec, phi, prec, nsys in sets]
    d1 = batch.tddft_batch(gs, 3, dt, graph=True)
    torch.cuda.synchronize()
    # ordinary work between the replays: one eager step per context
    d2 = []
    for g in gs:
        _, dip = g.opaw_tddft(1, dt, g.spec.sm)
        torch.cuda.synchronize()
        d2.append(dip)
    d3 = batch.tddft_batch(gs, 4, dt, graph=True)
    torch.cuda.synchronize()
    for k, (g, (rdip, rorb)) in enumerate(zip(gs, ref)):
        dip = np.concatenate([d1[k], d2[k], d3[k]])
        assert dip.shape == rdip.shape
        assert np.array_equal(dip, rdip), (k, np.max(np.abs(dip - rdip)))
        for s in range(g.nsys):
            assert np.array_equal(g.read_orbitals(s), rorb[s]), (k, s)
    for g in gs:
        g.close()
    print("[batch] graph replay equal", flush=True)
    # the plain-enqueue batch
    ref = [_alone(spec, phi, prec, nsys, (3,)) for spec, phi, prec, nsys in sets]
    gs = [batch.kicked_context(spec, phi, prec, nsys) for spec, phi, prec, nsys in sets]
    d = batch.tddft_batch(gs, 3, dt, graph=False)
    for k, (g, (rdip, rorb)) in enumerate(zip(gs, ref)):
        assert np.array_equal(d[k], rdip), k
        for s in range(g.nsys):
            assert np.array_equal(g.read_orbitals(s), rorb[s]), (k, s)
        g.close()


def test_time_loop_same_on_null_and_own_stream():
    """The time loop splits its work over the caller's stream and a side
    stream.  Run on the legacy default (null) stream twice and on a stream of
    its own, the same context must give bit-identical dipoles and orbitals
    (a missing hand-over between the streams shows up as run-to-run noise)."""
    torch, OpawGPU = _api()
    spec = batch.polyacene_spec(1)
    phi = batch.polyacene_orbitals(spec, seed=3)
    runs = []
    for mode in ("null", "null", "own"):
        s = torch.cuda.Stream() if mode == "own" else None
        g = OpawGPU(spec, "f64", nsys=2, nb=phi.shape[0], stream=(s.cuda_stream if s is not None else 0))
        g.write_orbitals(0, phi)
        g.excite_psi_bar_pert(spec.sm, 1, from_sys=0, to_sys=1)
        _, dip = g.opaw_tddft(3, spec.dt, spec.sm)
        torch.cuda.synchronize()
        runs.append((dip, g.read_orbitals(0), g.read_orbitals(1)))
        g.close()
    for k in (1, 2):
        for a, b in zip(runs[0], runs[k]):
            assert np.array_equal(a, b), k


def test_batch_step_vs_oracle():
    _api()
    spec = batch.polyacene_spec(1)
    nb = 6
    spec.rnel = 2.0 * nb
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=5, envelope=3.0), spec.dv)
    st = state_from_spec(spec, nsys=2)
    phis = [phi, O.excite_psi_bar_pert(st, phi, spec.sm, 0)]
    new, d_dip, _, _ = O.tddft_step(st, phis, spec.sm)
    other = batch.polyacene_spec(2)
    gs = [batch.kicked_context(spec, phi, "f64", 2),
          batch.kicked_context(other, batch.polyacene_orbitals(other, seed=1), "mixed", 2)]
    dip = batch.tddft_batch(gs, 1, spec.dt, graph=True)[0]
    for s in (0, 1):
        assert rel(gs[0].read_orbitals(s), new[s]) < 1e-10, s
    d_gpu = (dip[0, 3:6] - dip[0, 0:3]) / spec.sm
    assert np.max(np.abs(d_gpu - d_dip)) <= 1e-8 * np.max(np.abs(d_dip))
    for g in gs:
        g.close()


def test_polyacene_series_batched_mixed():
    torch, _ = _api()
    specs = [batch.polyacene_spec(n) for n in range(1, 7)]
    phis = [batch.polyacene_orbitals(s, seed=20 + k) for k, s in enumerate(specs)]
    dt = specs[0].dt
    nt = 3
    ref = [_alone(s, p, "mixed", 4, (nt,)) for s, p in zip(specs, phis)]
    print("[batch] polyacene references done", flush=True)
    gs = [batch.kicked_context(s, p, "mixed", 4) for s, p in zip(specs, phis)]
    dips = batch.tddft_batch(gs, nt, dt, graph=True)
    for k, (g, (rdip, rorb)) in enumerate(zip(gs, ref)):
        assert g.nb == 9 * (k + 1) + 6
        assert np.array_equal(dips[k], rdip), k
        n0 = _norms(specs[k], phis[k])
        for s in range(4):
            o = g.read_orbitals(s)
            assert np.array_equal(o, rorb[s]), (k, s)
            assert np.max(np.abs(_norms(specs[k], o.astype(np.complex128)) - n0)) < 1e-6, (k, s)
        resp = batch.kick_responses(dips[k], specs[k].sm)
        assert resp.shape == (nt, 3) and np.all(np.isfinite(resp))
        g.close()


def test_batch_argument_errors():
    _api()
    spec = batch.polyacene_spec(1)
    nb = 2
    spec.rnel = 2.0 * nb
    phi = synth.lowdin_orthonormalise(synth.random_orbitals(spec, nb, seed=1, envelope=3.0), spec.dv)
    g = batch.kicked_context(spec, phi, "f64", 2)
    with pytest.raises(Exception, match="twice"):
        batch.tddft_batch([g, g], 1, spec.dt)
    with pytest.raises(Exception, match="positive"):
        batch.tddft_batch([g], 0, spec.dt)

    class _Null:
        import ctypes
        ctx = ctypes.c_void_p(None)
        nsys = 2
    with pytest.raises(Exception, match="null context"):
        batch.tddft_batch([g, _Null()], 1, spec.dt)
    g.close()
