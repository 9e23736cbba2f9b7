"""The adjoint formulas pinned on the FP64 oracle (CPU): closed-form derivatives of the Born
map built from the oracle's own intermediates, and the chain relations between radii and
the GBSW smoothing constants, against central finite differences of the oracle energy.

These are the expressions the device adjoint evaluates (Born map: per atom from dE/dB,
I4, I7; GBSW: dE/dR' summed with weights R + roffset and rscale), so a wrong term or sign
fails here on the CPU before any GPU comparison.
"""
import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import synthetic as S


def _fd(op, field, s, rel=1e-6):
    v = getattr(op, field)
    h = rel * abs(v) if v != 0.0 else rel
    setattr(op, field, v + h)
    ep = O.energy(op, s.pos, s.charge, s.radius)
    setattr(op, field, v - h)
    em = O.energy(op, s.pos, s.charge, s.radius)
    setattr(op, field, v)
    return (ep - em) / (2 * h)


def _born_map_derivatives(op, res):
    """dE/dtheta of the Born map constants from dE/dI4 and the integrals (FP64).
    B = slope/D + shift, D = a0 (1/eta - I4) + a1 T^(1/4), T = 1/(4 eta^4) - I7."""
    i4, i7 = (0, 1) if op.orders[0] == 4 else (1, 0)
    I4, I7 = res.integrals[i4], res.integrals[i7]
    a0, a1, eta, slope = op.alpha0, op.alpha1, op.eta, op.slope
    T = 1.0 / (4.0 * eta ** 4) - I7
    D = a0 * (1.0 / eta - I4) + a1 * T ** 0.25
    dEdB = res.dEdI[i4] / (slope * a0 / D ** 2)
    g = -slope / D ** 2 * dEdB
    dDdeta = -a0 / eta ** 2 - a1 * 0.25 * T ** -0.75 / eta ** 5
    return {"alpha0": np.sum(g * (1.0 / eta - I4)), "alpha1": np.sum(g * T ** 0.25), "eta": np.sum(g * dDdeta),
            "slope": np.sum(dEdB / D), "shift": np.sum(dEdB), "prefactor": res.energy / op.prefactor}


@pytest.mark.parametrize("model", ["gbsw", "gbmv2", "gbmv1"])
def test_born_map_adjoint_formulas(model):
    s = S.globule(40, 41)
    if model == "gbsw":
        op = O.gbsw_standalone_params()
    else:
        op = O.gbmv_params(O.INT_GBMV2 if model == "gbmv2" else O.INT_GBMV1, policy=O.POLICY_BRUTE)
    res = O.compute(op, s.pos, s.charge, s.radius)
    d = _born_map_derivatives(op, res)
    for name, val in d.items():
        fd = _fd(op, name, s)
        assert abs(val - fd) < 2e-6 * max(abs(fd), 1e-3), (model, name, val, fd)


def test_gbsw_smoothing_chain_relations():
    """dE/drscale = sum_j dE/dR'_j (R_j + roffset) and dE/droffset = sum_j dE/dR'_j rscale,
    with dE/dR'_j = (dE/dR_j) / rscale."""
    s = S.globule(30, 42)
    op = O.gbsw_standalone_params()
    h = 1e-5
    dEdR = np.zeros(s.n)
    for j in range(s.n):
        R = s.radius.copy()
        R[j] += h
        ep = O.energy(op, s.pos, s.charge, R)
        R[j] -= 2 * h
        em = O.energy(op, s.pos, s.charge, R)
        dEdR[j] = (ep - em) / (2 * h)
    dEdRp = dEdR / op.rscale
    assert abs(np.sum(dEdRp * (s.radius + op.roffset)) - _fd(op, "rscale", s)) < 1e-5 * np.sum(np.abs(dEdRp))
    assert abs(np.sum(dEdRp * op.rscale) - _fd(op, "roffset", s)) < 1e-5 * np.sum(np.abs(dEdRp))
