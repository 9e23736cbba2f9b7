"""General CharmmGBMVForce expressions on the HIP path (kernels generated from the
expressions and compiled with hiprtc) vs the FP64 oracle (needs an MI355X).

Tolerances as in test_gpu_parity: force relative RMS <= 1e-5, energy relative error <= 1e-5.
The canonical forms written as expressions must also match the library's hand-written
canonical path.
"""
import numpy as np
import pytest

import oracle as O
from oracle import custom as C
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params, rel_err, rel_rms
from test_oracle_custom import BORN, G, SELF, STILL, TABLES, TAB_TERMS, TAB_VALUES, TERMS2, VALUES2

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-5


def _engine(p, s, integrals, values, terms, globals_):
    return N.Engine.custom(p, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, integrals, values, terms,
                           globals_)


@pytest.mark.parametrize("itype", [N.INTEGRAL_GBSW, N.INTEGRAL_GBMV1, N.INTEGRAL_GBMV2])
def test_canonical_forms_as_expressions(itype):
    s = S.globule(500, 70 + itype)
    p = N.default_params(N.MODEL_GBMV, itype)
    eng = _engine(p, s, ["I4", "I7"], [BORN], [SELF, (STILL, N.TERM_PAIR_NO_EXCLUSIONS)], G)
    e, f = eng.execute(s.pos)
    born = eng.custom_values()[0]
    eng.close()
    canon = N.Engine(p, s.charge, s.radius)
    ec, fc = canon.execute(s.pos)
    bc = canon.born_radii()
    canon.close()
    ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
    print(f"[custom] canonical type {itype}: energy {rel_err(e, ref.energy):.2e} force {rel_rms(f, ref.forces):.2e} "
          f"vs hand-written path: energy {rel_err(e, ec):.2e} force {rel_rms(f, fc):.2e}", flush=True)
    assert rel_err(e, ref.energy) < ENERGY_TOL and rel_rms(f, ref.forces) < FORCE_TOL
    assert np.max(np.abs(born - ref.born) / ref.born) < 1e-5
    assert rel_err(e, ec) < 1e-6 and np.max(np.abs(born - bc) / bc) < 1e-12


@pytest.mark.parametrize("method", [N.NO_CUTOFF, N.CUTOFF_NON_PERIODIC, N.CUTOFF_PERIODIC])
def test_general_expressions(method):
    s = S.globule(600, 75)
    L = np.array([4.2, 4.4, 4.3])
    pos = s.pos - s.pos.mean(axis=0) + L / 2
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV1)
    p.nonbonded_method, p.cutoff = method, 1.3
    g = dict(G, k=0.3)
    ex = np.array([(i, i + 1) for i in range(0, s.n - 1, 3)], dtype=np.int32)
    eng = _engine(p, s, ["I4", "I7"], VALUES2, TERMS2, g)
    if method == N.CUTOFF_PERIODIC:
        eng.set_periodic_box(L)
    eng.set_exclusions(ex)
    e, f = eng.execute(pos)
    vals = eng.custom_values()
    op = oracle_params(p, box=L if method == N.CUTOFF_PERIODIC else None)
    E, F, V, _ = C.compute(op, pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4", "I7"], VALUES2,
                           TERMS2, g, exclusions=ex)
    print(f"[custom] general method {method}: energy {rel_err(e, E):.2e} force {rel_rms(f, F):.2e} "
          f"values {np.max(np.abs(vals - V) / np.abs(V)):.2e}", flush=True)
    assert rel_err(e, E) < ENERGY_TOL and rel_rms(f, F) < FORCE_TOL
    assert np.max(np.abs(vals - V) / np.abs(V)) < 1e-5
    # a new global value (Context::setParameter)
    g2 = dict(g, k=-0.2, P=0.9 * g["P"])
    eng.set_custom_globals([g2[k] for k in g])
    e2, f2 = eng.execute(pos)
    E2, F2, _, _ = C.compute(op, pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4", "I7"], VALUES2,
                             TERMS2, g2, exclusions=ex)
    assert rel_err(e2, E2) < ENERGY_TOL and rel_rms(f2, F2) < FORCE_TOL
    eng.close()


def test_one_integral_and_graph_replay():
    s = S.globule(400, 76)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.orders[0] = 4
    values = [("B", "1/(1/radius-0.5*I4*radius)")]
    terms = [("0.5*P*q^2/B", N.TERM_SINGLE), (STILL, N.TERM_PAIR_NO_EXCLUSIONS)]
    eng = _engine(p, s, ["I4"], values, terms, G)
    eng.set_graph(True)
    rng = np.random.default_rng(5)
    for rep in range(3):
        x = s.pos + rng.normal(0.0, 0.01, s.pos.shape)
        e, f = eng.execute(x)
        E, F, _, _ = C.compute(oracle_params(p), x, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                               values, terms, G)
        print(f"[custom] one integral, replay {rep}: energy {rel_err(e, E):.2e} force {rel_rms(f, F):.2e}", flush=True)
        assert rel_err(e, E) < ENERGY_TOL and rel_rms(f, F) < FORCE_TOL
    eng.close()


def _lib_table(t):
    return N.Tabulated(t.name, t.kind, t.values, t.min, t.max, t.periodic, t.dims, t.yrange, t.zrange)


@pytest.mark.parametrize("method", [N.NO_CUTOFF, N.CUTOFF_PERIODIC])
def test_coordinates_and_tabulated_functions(method):
    """Computed values and single-particle terms using the atom's x, y, z (their gradients are
    forces on the atom itself), a natural and a periodic Continuous1DFunction, a periodic
    Continuous2DFunction and a natural Continuous3DFunction of the coordinates, and
    Discrete1D/2D/3DFunctions, against the sympy/scipy oracle."""
    s = S.globule(500, 77)
    L = np.array([4.0, 4.1, 4.2])
    pos = s.pos - s.pos.mean(axis=0) + (L / 2 if method == N.CUTOFF_PERIODIC else 0.0)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.orders[0] = 4
    p.nonbonded_method, p.cutoff = method, 1.4
    eng = N.Engine.custom(p, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"], TAB_VALUES, TAB_TERMS, G,
                          functions=[_lib_table(t) for t in TABLES])
    if method == N.CUTOFF_PERIODIC:
        eng.set_periodic_box(L)
    e, f = eng.execute(pos)
    vals = eng.custom_values()
    eng.close()
    E, F, V, _ = C.compute(oracle_params(p, box=L if method == N.CUTOFF_PERIODIC else None), pos, ["q", "radius"],
                           np.stack([s.charge, s.radius], 1), 1, ["I4"], TAB_VALUES, TAB_TERMS, G,
                           functions=TABLES)
    print(f"[custom] x,y,z + tables, method {method}: energy {rel_err(e, E):.2e} force {rel_rms(f, F):.2e} "
          f"values {np.max(np.abs(vals - V) / np.abs(V)):.2e}", flush=True)
    assert rel_err(e, E) < ENERGY_TOL and rel_rms(f, F) < FORCE_TOL
    assert np.max(np.abs(vals - V) / np.abs(V)) < 1e-5
