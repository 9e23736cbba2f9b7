"""Shared helpers of the tests: oracle parameters matching a gbi_params, error metrics."""
from __future__ import annotations

import numpy as np

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N

_SHARED = ["n_radial", "n_lebedev", "r0", "r1", "grid_length", "buffer", "sw", "rscale", "roffset",
           "gamma0", "beta", "lambda_", "P1", "P2", "S0", "vguard", "alpha0", "alpha1", "eta", "slope",
           "shift", "prefactor"]


def oracle_params(p: N.Params, nthreads: int = 8, dense: int = 0) -> O.Params:
    """Oracle parameters describing the same model as the ABI parameters `p`."""
    op = O.gbsw_standalone_params() if p.model == N.MODEL_GBSW else O.gbmv_params(p.integral_type)
    for f in _SHARED:
        setattr(op, f, getattr(p, f))
    op.cap = p.cell_cap
    op.orders[0], op.orders[1] = p.orders[0], p.orders[1]
    op.policy = O.POLICY_BRUTE if p.lookup_policy == N.LOOKUP_BRUTE else O.POLICY_GRID
    op.nthreads = nthreads
    op.dense = dense
    return op


def rel_rms(f, ref) -> float:
    return float(np.linalg.norm(np.asarray(f) - np.asarray(ref)) / np.linalg.norm(ref))


def rel_err(a, b) -> float:
    return float(abs(a - b) / abs(b))
