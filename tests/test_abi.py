"""The C ABI (include/gbi.h) on the CPU: symbols, defaults, error paths (no compute).

Every entry point the header declares must be exported by libgbi.so; parameter defaults
are the reference constants (CharmmReferenceGBSW.cpp:20-52, GBSWIntegral.cpp:24-61,
GBMVIntegralTypeI.cpp:18-57, GBMVIntegralTypeII.cpp:19-62, CustomGBIntegral.cpp:12-16).
"""
import ctypes
import os
import re

import numpy as np
import pytest

from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import forces as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "gbi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:gbi_status|void|const char\*|int32_t)\s+(gbi_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = N.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(N.EXPORTS) == syms
    assert lib.gbi_abi_version() == 6


def test_defaults_are_reference_constants():
    p = N.default_params(N.MODEL_GBSW)
    assert (p.n_radial, p.n_lebedev, p.r0, p.r1) == (24, 38, 0.05, 2.0)
    assert (p.grid_length, p.buffer, p.cell_cap, p.sw) == (0.15, 0.20, 0, 0.03)
    assert (p.alpha0, p.alpha1, p.eta) == (-0.180, 1.817, 0.05)
    assert p.prefactor == pytest.approx(-138.935456 * (1 - 1 / 80), rel=1e-15)
    g2 = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    assert (g2.n_lebedev, g2.cell_cap, g2.buffer) == (26, 64, 0.21)
    assert (g2.beta, g2.lambda_, g2.P1, g2.P2, g2.S0) == (-20.0, 0.5, 0.45, 1.25, 0.7)
    g1 = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV1)
    assert (g1.gamma0, g1.beta, g1.lambda_, g1.buffer) == (0.44, -100.0, 0.1, 0.21)
    g0 = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBSW)
    assert g0.buffer == pytest.approx(0.04) and g0.cell_cap == 64


def _create(p, q=(0.1, -0.1), r=(0.15, 0.2)):
    q = np.asarray(q, float)
    r = np.asarray(r, float)
    h = ctypes.c_void_p()
    st = N.load().gbi_create(len(q), q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(p), ctypes.byref(h))
    if st == N.GBI_OK:
        N.load().gbi_destroy(h)
    return st


def test_invalid_arguments_are_rejected_before_touching_a_device():
    lib = N.load()
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    bad = N.Params.from_buffer_copy(p)
    bad.integral_type = N.INTEGRAL_CUSTOM
    assert _create(bad) == 2
    assert b"CustomIntegral" in lib.gbi_last_error()
    bad = N.Params.from_buffer_copy(p)
    bad.n_lebedev = 7
    assert _create(bad) == 1
    bad = N.Params.from_buffer_copy(p)
    bad.nonbonded_method = N.CUTOFF_NON_PERIODIC
    bad.cutoff = 0.0
    assert _create(bad) == 1
    bad = N.Params.from_buffer_copy(p)
    bad.nonbonded_method = 3
    assert _create(bad) == 1
    bad = N.Params.from_buffer_copy(p)
    bad.nonbonded_method = N.CUTOFF_PERIODIC
    bad.lookup_policy = N.LOOKUP_BRUTE
    assert _create(bad) == 2
    bad = N.Params.from_buffer_copy(p)
    bad.orders[1] = 5
    assert _create(bad) == 2
    bad = N.Params.from_buffer_copy(p)
    bad.orders[0] = 1
    assert _create(bad) == 1
    # the GBMV sigmoid needs a finite, nonzero beta (beta = 0 makes V independent of the geometry)
    for v in (0.0, np.inf):
        bad = N.Params.from_buffer_copy(p)
        bad.beta = v
        assert _create(bad) == 1
    bad = N.Params.from_buffer_copy(p)
    bad.lambda_ = np.nan
    assert _create(bad) == 1
    assert _create(p, r=(0.15, -0.2)) == 1
    assert _create(p, q=(np.nan, 0.0)) == 7
    assert N.load().gbi_params_default(N.MODEL_GBMV, N.INTEGRAL_CUSTOM, ctypes.byref(N.Params())) == 2


def test_create_reports_missing_device_or_succeeds():
    st = _create(N.default_params(N.MODEL_GBSW))
    assert st in (N.GBI_OK, 8)


def test_null_pointers():
    lib = N.load()
    assert lib.gbi_execute(None, None, 1, 1, None, None) == 1
    assert lib.gbi_get_energy(None, None) == 1
    lib.gbi_destroy(None)


def test_force_api_round_trip():
    q = np.array([0.3, -0.2, -0.1])
    r = np.array([0.17, 0.2, 0.1])
    f = F.canonical_gbmv_force(q, r, F.CharmmGBMVForce.GBMVIntegralTypeII, lebedev_rule=5)
    p, qq, rr = F.gbmv_params_from_force(f)
    assert p.n_lebedev == 50 and p.integral_type == 2 and p.model == N.MODEL_GBMV
    assert np.array_equal(qq, q) and np.array_equal(rr, r)
    assert p.prefactor == pytest.approx(-138.935456 * (1 - 1 / 80))
    assert F.CustomGBIntegralForce is F.CharmmGBMVForce
    assert p.nonbonded_method == N.NO_CUTOFF
    f.setNonbondedMethod(F.CharmmGBMVForce.CutoffPeriodic)
    f.setCutoffDistance(1.2)
    p, _, _ = F.gbmv_params_from_force(f)
    assert (p.nonbonded_method, p.cutoff) == (N.CUTOFF_PERIODIC, 1.2)


def test_force_api_rejects_what_the_hip_platform_does_not_implement():
    f = F.canonical_gbmv_force([0.1], [0.2], F.CharmmGBMVForce.GBSWIntegral)
    f._energy[1] = ("P*q1*q2/r", F.CharmmGBMVForce.ParticlePair)
    with pytest.raises(N.GbiError):
        F.gbmv_params_from_force(f)
    f = F.canonical_gbmv_force([0.1], [0.2], F.CharmmGBMVForce.CustomIntegral)
    with pytest.raises(N.GbiError):
        F.gbmv_params_from_force(f)
    g = F.CharmmGBMVForce()
    g.addPerParticleParameter("q")
    g.addGBIntegral("I4", [4])
    with pytest.raises(F.OpenMMException):
        F.gbmv_params_from_force(g)
    g = F.CharmmGBMVForce()
    g.addPerParticleParameter("radius")
    g.addGBIntegral("I4", [1])
    with pytest.raises(F.OpenMMException):
        F.gbmv_params_from_force(g)


def test_canonical_force_needs_integrals_named_by_their_order():
    """The canonical Born expression names its integrals: swapping the names of the order-4
    and order-7 integrals, or asking for energy parameter derivatives, sends the force to the
    generated kernels (GbiError status 2 from the canonical mapping)."""
    f = F.canonical_gbmv_force([0.1, -0.1], [0.2, 0.17], F.CharmmGBMVForce.GBMVIntegralTypeII)
    F.gbmv_params_from_force(f)
    f.setGBIntegralParameters(0, "I4", [7])
    f.setGBIntegralParameters(1, "I7", [4])
    with pytest.raises(N.GbiError) as err:
        F.gbmv_params_from_force(f)
    assert err.value.status == 2
    g = F.canonical_gbmv_force([0.1, -0.1], [0.2, 0.17], F.CharmmGBMVForce.GBMVIntegralTypeII)
    g.setGBIntegralParameters(0, "I7", [7])
    g.setGBIntegralParameters(1, "I4", [4])
    p, _, _ = F.gbmv_params_from_force(g)
    assert (p.orders[0], p.orders[1]) == (7, 4)
    g.addEnergyParameterDerivative("alpha0")
    with pytest.raises(N.GbiError) as err:
        F.gbmv_params_from_force(g)
    assert err.value.status == 2


def test_gbsw_force_api():
    f = F.CharmmGBSWForce()
    assert f.getSolventDielectric() == 78.3 and f.getSoluteDielectric() == 1.0
    assert f.getSurfaceAreaEnergy() == pytest.approx(2.25936)
    f.addParticle(0.5, 0.17, 1.0)
    f.addParticle(-0.5, 0.2, 1.0)
    k = F.HipCalcCharmmGBSWForceKernel()
    assert k.Name() == "CalcCharmmGBSWForce"
    with pytest.raises(F.OpenMMException):
        k.initialize(F.System(3), f)


def test_autotuner_cache_key_and_store(tmp_path):
    """The autotuner's JSON cache: the key separates models, integral types, quadratures and
    methods and buckets the atom count by powers of two; entries are added and replaced."""
    g = N.default_params(N.MODEL_GBSW)
    m = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    m.n_lebedev = 50
    k1 = N.tuning_key([m], [5000])
    assert k1 == N.tuning_key([m], [4097]) != N.tuning_key([m], [4096])  # 4097-8192 share a bucket
    assert N.tuning_key([g, m], [300, 400]) == N.tuning_key([m, g], [400, 300])  # order of systems
    assert N.tuning_key([g], [5000]) != k1
    m2 = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    assert N.tuning_key([m2], [5000]) != k1  # Lebedev 26 vs 50
    path = str(tmp_path / "tuning.json")
    assert N.load_tuning(path) == {}
    t = {"blocks_per_cu": 64, "pair_columns": 0, "tight_grid_length": 0.2, "ms_per_eval": 0.6}
    N.store_tuning(path, k1, t)
    N.store_tuning(path, "other", dict(t, blocks_per_cu=16))
    N.store_tuning(path, k1, dict(t, blocks_per_cu=128))
    data = N.load_tuning(path)
    assert data[k1]["blocks_per_cu"] == 128 and data["other"]["blocks_per_cu"] == 16
    assert data[k1]["tight_grid_length"] == 0.2
