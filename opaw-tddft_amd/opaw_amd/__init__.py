"""MI355X-native OPAW-TDDFT Kohn-Sham propagation hot path.

``opaw_amd.lib``    ctypes binding of libopaw_hip.so (C ABI in include/opaw_hip.h)
``opaw_amd.api``    OpawGPU: host mirror of the reference subroutine interface
``opaw_amd.synth``  seeded synthetic systems (naphthalene-shaped MT, periodic H.psi)
``opaw_amd.build``  in-tree hipcc build of the library
"""
from . import synth  # noqa: F401

__all__ = ["synth", "lib", "api", "build"]


def __getattr__(name):
    if name in ("lib", "api", "build"):
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
