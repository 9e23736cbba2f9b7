"""Shared helpers of the tests: oracle parameters matching a gbi_params, error metrics."""
from __future__ import annotations

import numpy as np

import oracle as O


def oracle_params(p, nthreads: int = 8, dense: int = 0, box=None) -> O.Params:
    """Oracle parameters describing the same model as the ABI parameters `p` (`box`: the
    edges of the periodic box of a CutoffPeriodic system)."""
    return O.params_from_abi(p, nthreads=nthreads, dense=dense, box=box)


def rel_rms(f, ref) -> float:
    return float(np.linalg.norm(np.asarray(f) - np.asarray(ref)) / np.linalg.norm(ref))


def rel_err(a, b) -> float:
    return float(abs(a - b) / abs(b))
