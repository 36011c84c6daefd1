"""ctypes binding of libopaw_hip.so (the C ABI declared in include/opaw_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libopaw_hip.so"

OPAW_OK = 0
OPAW_ERR_ARG = -1
OPAW_ERR_HIP = -2
OPAW_ERR_ALLOC = -3
OPAW_ERR_STATE = -4
OPAW_ERR_UNSUPPORTED = -5
OPAW_PREC_F64 = 0
OPAW_PREC_MIXED = 1
FIELD = dict(dens=0, nhat=1, vxc=2, vh=3, vks=4)
KCLASS = dict(fft_x=0, fft_y=1, fft_z=2, fft_xh=3, hartree=4, box_proj=5, box_add=6, rk4=7,
              dens=8, pot=9, other=10)

P = C.c_void_p
I = C.c_int
D = C.c_double
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
LLP = C.POINTER(C.c_longlong)

# name: (restype, [argtypes])
SIGNATURES = {
    "opaw_last_error": (C.c_char_p, []),
    "opaw_version": (I, []),
    "opaw_fft_length_supported": (I, [I]),
    "opaw_plane_path": (I, [I, I, I]),
    "opaw_create": (I, [C.POINTER(P), I, I, I, D, D, D, I, I, I, I]),
    "opaw_destroy": (I, [P]),
    "opaw_set_stream": (I, [P, P]),
    "opaw_synchronize": (I, [P]),
    "opaw_device_bytes": (C.c_longlong, [P]),
    "opaw_ek_prep": (I, [P, D]),
    "opaw_set_ek": (I, [P, DP]),
    "opaw_set_atoms": (I, [P, I, IP, IP, I, IP, IP, I]),
    "opaw_set_projectors": (I, [P, I, DP, DP, DP, DP]),
    "opaw_set_sphere": (I, [P, I, I, IP, DP]),
    "opaw_set_type": (I, [P, I, DP, DP, DP]),
    "opaw_set_onsite": (I, [P, I, I, I, I, I, DP, DP, DP, DP, DP, DP, DP, IP, DP]),
    "opaw_set_dij_sphere": (I, [P, I, I, IP, DP]),
    "opaw_finalize_tables": (I, [P]),
    "opaw_set_potentials": (I, [P, DP, DP, D, I]),
    "opaw_vk_prep": (I, [P]),
    "opaw_write_vks": (I, [P, I, DP]),
    "opaw_read_vks": (I, [P, I, DP]),
    "opaw_write_dij": (I, [P, I, I, DP]),
    "opaw_read_dij": (I, [P, I, I, DP]),
    "opaw_read_rhoij": (I, [P, I, I, DP]),
    "opaw_read_field": (I, [P, I, I, DP]),
    "opaw_write_orbitals": (I, [P, I, P]),
    "opaw_read_orbitals": (I, [P, I, P]),
    "opaw_orbitals_device": (P, [P, I]),
    "opaw_h_phi": (I, [P, I, P, P, I]),
    "opaw_sn_phi": (I, [P, P, P, I, D]),
    "opaw_shs": (I, [P, I, P, P, I]),
    "opaw_proj": (I, [P, P, I, I, P]),
    "opaw_h_phi_host": (I, [P, I, DP, DP, I]),
    "opaw_shs_host": (I, [P, I, DP, DP, I]),
    "opaw_sn_phi_host": (I, [P, DP, DP, I, D]),
    "opaw_proj_host": (I, [P, I, DP, I, DP, I]),
    "opaw_write_orbitals_c16": (I, [P, I, DP]),
    "opaw_read_orbitals_c16": (I, [P, I, DP]),
    "opaw_set_dt": (I, [P, D]),
    "opaw_precond_keep_fields": (I, [P, I]),
    "opaw_rk4_prop_shs_host": (I, [P, DP, DP, C.c_longlong, I]),
    "opaw_fft3d": (I, [P, I, P, P, I]),
    "opaw_tddft_rhoij_2": (I, [P, I, P]),
    "opaw_tddft_dens_2": (I, [P, I, P]),
    "opaw_get_nhat": (I, [P, I]),
    "opaw_get_pot_tddft_2": (I, [P, I]),
    "opaw_get_dij": (I, [P, I]),
    "opaw_make_hamiltonian_tddft": (I, [P, I]),
    "opaw_excite_psi_bar_pert": (I, [P, I, I, D, I]),
    "opaw_rk4_prop_shs": (I, [P, D]),
    "opaw_plot_dip": (I, [P, D, DP, DP, DP]),
    "opaw_tddft_run": (I, [P, I, D, D, DP, DP]),
    "opaw_tddft_run_batch": (I, [C.POINTER(C.c_void_p), I, I, D, I, DP]),
    "opaw_orthog_phi": (I, [P, I, I, DP]),
    "opaw_diagh": (I, [P, I, I, DP]),
    "opaw_write_wf": (I, [C.c_char_p, I, I, I, I, I, D, D, D, I, I, DP, P]),
    "opaw_read_wf": (I, [C.c_char_p, I, I, I, I, D, D, D, I, IP, IP, DP, P]),
    "opaw_write_eig": (I, [C.c_char_p, I, DP]),
    "opaw_read_eig": (I, [C.c_char_p, I, DP]),
    "opaw_write_dipoles": (I, [C.c_char_p, C.c_char_p, I, I, I, D, DP, DP, I]),
    "opaw_write_restart": (I, [P, C.c_char_p, I]),
    "opaw_read_restart": (I, [P, C.c_char_p, IP]),
    "opaw_spectrum": (I, [DP, I, I, D, I, D, D, D, I, DP, DP]),
    "opaw_spectrum_noalloc": (I, [P, I, I, D, I, D, D, D, I, P, P, P]),
    "opaw_paw_box": (I, [I, DP, I, DP, DP, IP, IP, IP, DP, DP, DP, DP, DP, DP, DP, DP, DP]),
    "opaw_vloc_tuma": (I, [P, I, IP, DP, DP, I, IP, DP, DP]),
    "opaw_paw_sphere": (I, [I, I, I, D, D, D, DP, D, D, I, DP, I, I, DP, I, IP, IP, DP]),
    "opaw_set_nocc": (I, [P, I]),
    "opaw_set_diis": (I, [P, I, D]),
    "opaw_set_subspace_device": (I, [P, I]),
    "opaw_ncpaw_make_hamiltonian": (I, [P, I, I]),
    "opaw_ncpaw_make_hamiltonian_density": (I, [P, I, DP, I]),
    "opaw_hminmax": (I, [P, I, I, I, D, C.c_ulonglong, DP, DP]),
    "opaw_chebyshev_filter": (I, [P, I, I, I, D, D]),
    "opaw_residuals": (I, [P, I, I, DP, DP]),
    "opaw_block_solve": (I, [P, I, I, I, D, I, DP, I, I, DP, DP, IP]),
    "opaw_mlp_param_count": (I, [I]),
    "opaw_precond_prepare": (I, [P, I]),
    "opaw_precond_clear": (I, [P]),
    "opaw_precond_count": (I, [P, IP, IP]),
    "opaw_precond_loss": (I, [P, DP, I, DP, DP]),
    "opaw_precond_pair": (I, [P, I, P, P, DP, DP, IP]),
    "opaw_profile_enable": (I, [P, I]),
    "opaw_profile_reset": (I, [P]),
    "opaw_profile_read": (I, [P, I, DP, LLP]),
}


class OpawError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_LIB = None


def load(path: Path | None = None):
    """Load (once) and type the shared library; raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    # torch ships its own HIP/HSA runtime; it must initialise the device
    # before this library's runtime (from /opt/rocm) does, or torch then
    # finds no GPU.  Device pointers are shared between the two.
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    if not p.exists():
        raise OpawError("load", -99, f"{p} not built: run __graft_entry__.build() "
                        "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def last_error() -> str:
    msg = load().opaw_last_error()
    return msg.decode() if msg else ""


def check(fn: str, code: int):
    if code != OPAW_OK:
        raise OpawError(fn, code, last_error())
    return code


def call(name: str, *args):
    """Call an int-returning entry point and raise on a nonzero status."""
    return check(name, getattr(load(), name)(*args))
