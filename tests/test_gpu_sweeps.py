"""Parameter sweeps of the CharmmGBMVForce integral types on the HIP path vs the FP64 oracle
(needs an MI355X): every GBIntegralType the reference implements (GBSWIntegral,
GBMVIntegralTypeI, GBMVIntegralTypeII; CharmmGBMVForce.h:99-104) across Lebedev rules 1-5
(parametersInt[1]), radial counts (parametersInt[2]), quadrature ranges, system sizes and
GB constants.  CustomIntegral (type 3) has no implementation in the reference
(CharmmReferenceKernels.cpp:129-133) and is rejected (test_abi).

Tolerances as in test_gpu_parity: force relative RMS <= 1e-5, energy relative error <= 1e-5.
"""
import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params, rel_err, rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-5

ITYPES = {"gbsw": N.INTEGRAL_GBSW, "gbmv1": N.INTEGRAL_GBMV1, "gbmv2": N.INTEGRAL_GBMV2}


def _check(tag, p, s):
    eng = N.Engine(p, s.charge, s.radius)
    e, f = eng.execute(s.pos)
    born = eng.born_radii()
    eng.close()
    ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
    ee, fe = rel_err(e, ref.energy), rel_rms(f, ref.forces)
    db = float(np.max(np.abs(born - ref.born) / ref.born))
    print(f"[sweep] {tag} N={s.n}: energy {ee:.2e} force {fe:.2e} born {db:.2e}", flush=True)
    assert np.all(np.isfinite(f))
    assert ee < ENERGY_TOL and fe < FORCE_TOL and db < 1e-5, tag


@pytest.mark.parametrize("kind", list(ITYPES))
@pytest.mark.parametrize("n_lebedev", [6, 14, 26, 38, 50])
def test_lebedev_rules(kind, n_lebedev):
    p = N.default_params(N.MODEL_GBMV, ITYPES[kind])
    p.n_lebedev = n_lebedev
    _check(f"{kind} lebedev {n_lebedev}", p, S.globule(700, 30 + n_lebedev))


@pytest.mark.parametrize("kind", list(ITYPES))
@pytest.mark.parametrize("n_radial,r0,r1", [(8, 0.05, 2.0), (16, 0.1, 1.5), (32, 0.05, 2.0), (48, 0.05, 3.0)])
def test_radial_quadratures(kind, n_radial, r0, r1):
    p = N.default_params(N.MODEL_GBMV, ITYPES[kind])
    p.n_radial, p.r0, p.r1 = n_radial, r0, r1
    _check(f"{kind} radial {n_radial} [{r0}, {r1}]", p, S.globule(600, 40 + n_radial))


@pytest.mark.parametrize("kind", list(ITYPES))
@pytest.mark.parametrize("n", [40, 333, 2000])
def test_sizes(kind, n):
    p = N.default_params(N.MODEL_GBMV, ITYPES[kind])
    _check(f"{kind} size", p, S.globule(n, 50 + n))


@pytest.mark.parametrize("consts", [dict(gamma0=0.3, beta=-60.0, lambda_=0.2),
                                    dict(gamma0=0.5, beta=-150.0, lambda_=0.05),
                                    dict(beta=-40.0, lambda_=0.3, grid_length=0.12, buffer=0.25),
                                    dict(cell_cap=16), dict(beta=-5.0, lambda_=0.5)])
def test_gbmv1_constants(consts):
    """GBMV-I (GBMVIntegralTypeI.cpp:18-57, 155-224) away from its defaults: the sigmoid
    steepness and threshold, gamma0, the lookup grid and a cell cap that truncates lists."""
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV1)
    for k, v in consts.items():
        setattr(p, k, v)
    _check(f"gbmv1 {consts}", p, S.globule(800, 61))


@pytest.mark.parametrize("consts", [dict(P1=0.4, P2=1.0, S0=0.6), dict(beta=-12.0, lambda_=0.7),
                                    dict(orders=(7, 4)), dict(beta=-2.0), dict(beta=-5.0, lambda_=1.0),
                                    dict(lambda_=6.0)])
def test_gbmv2_constants(consts):
    """GBMV-II (GBMVIntegralTypeII.cpp:19-62, 163-282) away from its defaults.  The soft
    sigmoids (beta = -2, -5) and the high threshold (lambda = 6) put points at S well above 6
    where V is neither 0 nor 1: the sigmoid is evaluated wherever exp(beta (S - lambda)) is
    representable, so V and its chain-rule term match the FP64 oracle there too."""
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    for k, v in consts.items():
        if k == "orders":
            p.orders[0], p.orders[1] = v
        else:
            setattr(p, k, v)
    _check(f"gbmv2 {consts}", p, S.globule(800, 62))
