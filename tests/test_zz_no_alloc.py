"""GPU: opaw_spectrum_noalloc (device buffers, static scratch, no
allocation) against opaw_spectrum (host buffers, allocating) through the C
ABI: the same bits on two calls in a row, the second smaller, plus the
argument checks that keep a bad call off the device."""
import ctypes as C

import numpy as np
import pytest

import tests.helpers  # noqa: F401
from opaw_amd import lib

pytestmark = pytest.mark.gpu


def _reference(d, nsig, nt, dt, kind, tau, w0, dw, nw):
    a = np.zeros(2 * nsig * nw)
    s = np.zeros(nsig * nw)
    lib.call("opaw_spectrum", d.ctypes.data_as(C.POINTER(C.c_double)), nsig, nt, dt, kind, tau, w0, dw, nw,
             a.ctypes.data_as(C.POINTER(C.c_double)), s.ctypes.data_as(C.POINTER(C.c_double)))
    return a, s


def test_spectrum_noalloc_matches_allocating_entry_point():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    L = lib.load()
    stream = torch.cuda.Stream()
    rng = np.random.default_rng(17)
    for (nsig, nt, nw, kind) in ((3, 4001, 2500, 1), (2, 1200, 700, 2)):
        print(f"noalloc spectrum nsig={nsig} nt={nt} nw={nw} damp={kind}", flush=True)
        d = np.ascontiguousarray(rng.normal(size=nsig * nt))
        dt, tau, w0, dw = 0.05, 40.0, 0.0, 0.001
        a_ref, s_ref = _reference(d, nsig, nt, dt, kind, tau, w0, dw, nw)
        dd = torch.from_numpy(d).to("cuda")
        a = torch.full((2 * nsig * nw,), float("nan"), dtype=torch.float64, device="cuda")
        s = torch.full((nsig * nw,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            rc = L.opaw_spectrum_noalloc(C.c_void_p(dd.data_ptr()), nsig, nt, dt, kind, tau, w0, dw, nw,
                                         C.c_void_p(a.data_ptr()), C.c_void_p(s.data_ptr()),
                                         C.c_void_p(stream.cuda_stream))
        assert rc == 0, lib.last_error()
        stream.synchronize()
        assert np.array_equal(a.cpu().numpy(), a_ref)
        assert np.array_equal(s.cpu().numpy(), s_ref)


def test_spectrum_noalloc_rejects_bad_calls():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    L = lib.load()
    nsig, nt, nw = 1, 100, 10
    dd = torch.zeros(nt, dtype=torch.float64, device="cuda")
    a = torch.zeros(2 * nw, dtype=torch.float64, device="cuda")
    s = torch.zeros(nw, dtype=torch.float64, device="cuda")
    host = np.zeros(2 * nw)
    args = (0.05, 0, 0.0, 0.0, 0.001, nw)
    # a host output buffer is refused before anything is launched
    rc = L.opaw_spectrum_noalloc(C.c_void_p(dd.data_ptr()), nsig, nt, *args,
                                 C.c_void_p(host.ctypes.data), C.c_void_p(s.data_ptr()), None)
    assert rc == lib.OPAW_ERR_ARG and "device memory" in lib.last_error()
    # more samples than the scratch holds
    rc = L.opaw_spectrum_noalloc(C.c_void_p(dd.data_ptr()), 600, 4000, *args,
                                 C.c_void_p(a.data_ptr()), C.c_void_p(s.data_ptr()), None)
    assert rc == lib.OPAW_ERR_ARG and "scratch" in lib.last_error()
    torch.cuda.synchronize()
