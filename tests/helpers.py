"""Test helpers: adapt a synthetic SystemSpec into the oracle's state."""
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
for p in (_REPO, os.path.join(_REPO, "opaw-tddft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle.opaw_oracle as O  # noqa: E402
from oracle.adapter import grid_of, state_from_spec  # noqa: E402,F401


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
