"""CPU FP64 oracle for the OPAW-TDDFT Kohn-Sham propagation hot path.

TEST INFRASTRUCTURE ONLY. This package is the checker: the tests, the smoke
entry point and the CPU-baseline leg of the benchmark may import it; the
product path (the HIP library and its Python host mirror) never does.
"""
from .opaw_oracle import *  # noqa: F401,F403
