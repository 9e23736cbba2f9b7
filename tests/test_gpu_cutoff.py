"""Cutoff, periodic and exclusion paths on the HIP path vs the FP64 oracle (needs an MI355X).

Tolerances as in test_gpu_parity: force relative RMS <= 1e-5, energy relative error <= 1e-5.
Covered: CutoffNonPeriodic for both forces (the standalone GBSW chain rule truncated at
the cutoff, CharmmReferenceGBSW.cpp:285-296), cutoffs short enough that whole 64-atom
blocks are skipped, CutoffPeriodic for a globule interacting with its images and for a
liquid filling the box, moves by box vectors, batches mixing methods, graph replay across a
box change (the periodic grid regrows), excluded pairs of a ParticlePair term
(CharmmReferenceGBMV.cpp:272-284) with and without a cutoff, and the rejected cases.
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


def _params(kind: str, method: int, cutoff: float):
    if kind == "gbsw":
        p = N.default_params(N.MODEL_GBSW)
    elif kind == "gbmv2-50":
        p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
        p.n_lebedev = 50
    else:
        p = N.default_params(N.MODEL_GBMV, {"gbmv-gbsw": N.INTEGRAL_GBSW, "gbmv1": N.INTEGRAL_GBMV1,
                                            "gbmv2": N.INTEGRAL_GBMV2}[kind])
    p.nonbonded_method = method
    p.cutoff = cutoff
    return p


def _check(tag, p, s, box=None, pos=None):
    pos = s.pos if pos is None else pos
    print(f"[cutoff] {tag} N={s.n}", flush=True)
    eng = N.Engine(p, s.charge, s.radius)
    if box is not None:
        eng.set_periodic_box(box)
    e, f = eng.execute(pos)
    born = eng.born_radii()
    eng.close()
    ref = O.compute(oracle_params(p, box=box), pos, s.charge, s.radius)
    ee, fe = rel_err(e, ref.energy), rel_rms(f, ref.forces)
    db = float(np.max(np.abs(born - ref.born) / ref.born))
    print(f"[cutoff] {tag}: energy rel {ee:.2e} force rel-RMS {fe:.2e} born {db:.2e}", flush=True)
    assert np.all(np.isfinite(f))
    assert ee < ENERGY_TOL and fe < FORCE_TOL and db < 1e-5, tag
    return e, f


@pytest.mark.parametrize("kind", ["gbsw", "gbmv2-50", "gbmv1", "gbmv-gbsw"])
@pytest.mark.parametrize("cutoff", [0.8, 1.4])
def test_cutoff_non_periodic(kind, cutoff):
    s = S.globule(500, 11)
    _check(f"{kind} c={cutoff}", _params(kind, N.CUTOFF_NON_PERIODIC, cutoff), s)


def test_short_cutoff_skips_blocks():
    """3000 atoms in serpentine order: 64-atom blocks are compact, so with a 0.7 nm cutoff most
    block pairs lie beyond it and are skipped."""
    s = S.globule(3000, 12)
    for kind in ("gbsw", "gbmv2"):
        _check(f"{kind} 3000 c=0.7", _params(kind, N.CUTOFF_NON_PERIODIC, 0.7), s)


@pytest.mark.parametrize("kind", ["gbsw", "gbmv2-50", "gbmv1", "gbmv-gbsw"])
def test_periodic_globule_sees_its_images(kind):
    s = S.globule(500, 13)
    L = np.array([4.1, 3.9, 4.3])
    pos = s.pos - s.pos.mean(axis=0) + L / 2
    _check(f"{kind} globule box", _params(kind, N.CUTOFF_PERIODIC, 1.5), s, box=L, pos=pos)


@pytest.mark.parametrize("kind", ["gbsw", "gbmv2"])
def test_periodic_liquid(kind):
    L = [2.7, 2.6, 2.8]
    s = S.liquid(L, 14)
    _check(f"{kind} liquid", _params(kind, N.CUTOFF_PERIODIC, 1.2), s, box=L)


def test_box_vector_moves_gbsw():
    """Atoms moved by whole box vectors (positions far outside the box) give the same result."""
    L = np.array([2.7, 2.6, 2.8])
    s = S.liquid(L, 15)
    p = _params("gbsw", N.CUTOFF_PERIODIC, 1.2)
    eng = N.Engine(p, s.charge, s.radius)
    eng.set_periodic_box(L)
    e0, f0 = eng.execute(s.pos)
    rng = np.random.default_rng(3)
    moved = s.pos + L * rng.integers(-3, 4, size=s.pos.shape)
    e1, f1 = eng.execute(moved)
    eng.close()
    print(f"[cutoff] box-vector moves: dE/E {abs(e1 - e0) / abs(e0):.2e}", flush=True)
    assert abs(e1 - e0) < 1e-6 * abs(e0)
    assert rel_rms(f1, f0) < 1e-6


def test_batch_mixing_methods():
    L = np.array([4.1, 3.9, 4.3])
    mix = [("gbsw", N.CUTOFF_PERIODIC, 1.5, 300, 1), ("gbmv2-50", N.CUTOFF_NON_PERIODIC, 1.0, 350, 2),
           ("gbmv1", N.NO_CUTOFF, 1.0, 250, 3), ("gbmv2", N.CUTOFF_PERIODIC, 1.3, 280, 4),
           ("gbsw", N.CUTOFF_NON_PERIODIC, 0.9, 260, 5)]
    systems = [S.globule(n, seed) for *_, n, seed in mix]
    params = [_params(kind, m, c) for kind, m, c, _, _ in mix]
    pos = []
    for k, s in enumerate(systems):
        pos.append(s.pos - s.pos.mean(axis=0) + L / 2 + (0.0 if mix[k][1] == N.CUTOFF_PERIODIC else 50.0 * k))
    eng = N.Engine.batch(params, [s.charge for s in systems], [s.radius for s in systems])
    for k, (kind, m, *_rest) in enumerate(mix):
        if m == N.CUTOFF_PERIODIC:
            eng.set_periodic_box(L, system=k)
    e, f = eng.execute(np.concatenate(pos))
    for k, (s, p) in enumerate(zip(systems, params)):
        a, b = eng.offsets[k], eng.offsets[k + 1]
        box = L if mix[k][1] == N.CUTOFF_PERIODIC else None
        ref = O.compute(oracle_params(p, box=box), pos[k], s.charge, s.radius)
        ee, fe = rel_err(e[k], ref.energy), rel_rms(f[a:b], ref.forces)
        print(f"[cutoff] batch system {k} {mix[k][0]} method {mix[k][1]}: energy {ee:.2e} force {fe:.2e}", flush=True)
        assert ee < ENERGY_TOL and fe < FORCE_TOL, k
    eng.close()


def test_graph_replay_across_box_changes():
    """Graph replays while the box grows (more cells: the grid is regrown and the graph
    rebuilt) and shrinks; every evaluation is checked against the oracle."""
    s = S.globule(400, 16)
    p = _params("gbmv2", N.CUTOFF_PERIODIC, 1.3)
    eng = N.Engine(p, s.charge, s.radius)
    eng.set_graph(True)
    for k, L in enumerate(([3.6, 3.7, 3.8], [4.4, 4.6, 4.5], [3.2, 3.3, 3.1], [3.2, 3.3, 3.1])):
        L = np.array(L)
        pos = s.pos - s.pos.mean(axis=0) + L / 2
        eng.set_periodic_box(L)
        e, f = eng.execute(pos)
        ref = O.compute(oracle_params(p, box=L), pos, s.charge, s.radius)
        ee, fe = rel_err(e, ref.energy), rel_rms(f, ref.forces)
        print(f"[cutoff] replay {k} box {L}: energy {ee:.2e} force {fe:.2e}", flush=True)
        assert ee < ENERGY_TOL and fe < FORCE_TOL, k
    eng.close()


def test_rejected_cases():
    s = S.globule(50, 17)
    p = _params("gbsw", N.CUTOFF_PERIODIC, 1.2)
    eng = N.Engine(p, s.charge, s.radius)
    with pytest.raises(N.GbiError):  # no box yet
        eng.execute(s.pos)
    with pytest.raises(N.GbiError) as ex:
        eng.set_periodic_box(np.array([[3.0, 0.0, 0.0], [0.5, 3.0, 0.0], [0.0, 0.0, 3.0]]))
    assert ex.value.status == 2
    with pytest.raises(N.GbiError) as ex:
        eng.set_periodic_box([2.3, 3.0, 3.0])  # below twice the cutoff
    assert ex.value.status == 1
    eng.set_periodic_box([3.0, 3.0, 3.0])
    _, f = eng.execute(s.pos)
    # the adjoint covers periodic systems (checked against the oracle in test_gpu_adjoint.py);
    # its forces are the plain evaluation's
    assert np.array_equal(eng.execute_adjoint(s.pos)["forces"], f)
    eng.close()
    flat = N.Engine(_params("gbsw", N.NO_CUTOFF, 1.0), s.charge, s.radius)
    with pytest.raises(N.GbiError):
        flat.set_periodic_box([3.0, 3.0, 3.0])
    flat.close()


def _bonded_like_exclusions(pos, k=3):
    """Each atom with its k nearest neighbours (the shape of 1-2/1-3 exclusions)."""
    from scipy.spatial import cKDTree

    _, nb = cKDTree(pos).query(pos, k=k + 1)
    pairs = {(min(i, int(j)), max(i, int(j))) for i in range(len(pos)) for j in nb[i, 1:]}
    return np.array(sorted(pairs), dtype=np.int32)


@pytest.mark.parametrize("method", [N.NO_CUTOFF, N.CUTOFF_NON_PERIODIC, N.CUTOFF_PERIODIC])
def test_exclusions(method):
    s = S.globule(600, 18)
    L = np.array([4.4, 4.2, 4.6])
    pos = s.pos - s.pos.mean(axis=0) + L / 2
    ex = _bonded_like_exclusions(pos)
    p = _params("gbmv2-50", method, 1.3)
    eng = N.Engine(p, s.charge, s.radius)
    box = L if method == N.CUTOFF_PERIODIC else None
    if box is not None:
        eng.set_periodic_box(L)
    eng.set_exclusions(ex)
    e, f = eng.execute(pos)
    op = O.set_exclusions(oracle_params(p, box=box), ex)
    ref = O.compute(op, pos, s.charge, s.radius)
    plain = O.compute(oracle_params(p, box=box), pos, s.charge, s.radius)
    ee, fe = rel_err(e, ref.energy), rel_rms(f, ref.forces)
    print(f"[cutoff] exclusions method {method}: {len(ex)} pairs, energy {ee:.2e} force {fe:.2e} "
          f"(the exclusions move E by {abs(ref.energy - plain.energy):.1f} kJ/mol)", flush=True)
    assert ee < ENERGY_TOL and fe < FORCE_TOL
    assert abs(ref.energy - plain.energy) > 1e3 * ENERGY_TOL * abs(ref.energy)
    # cleared again: the plain energy
    eng.set_exclusions(np.zeros((0, 2), dtype=np.int32))
    e2, _ = eng.execute(pos)
    assert rel_err(e2, plain.energy) < ENERGY_TOL
    eng.close()


def test_exclusions_adjoint_and_errors():
    s = S.globule(300, 19)
    ex = _bonded_like_exclusions(s.pos)
    p = _params("gbmv1", N.NO_CUTOFF, 1.0)
    eng = N.Engine(p, s.charge, s.radius)
    eng.set_exclusions(ex)
    e, f = eng.execute(s.pos)
    out = eng.execute_adjoint(s.pos)
    print(f"[cutoff] exclusions adjoint: E {e:.6f} vs {out['energy']:.6f}", flush=True)
    assert abs(out["energy"] - e) <= 1e-12 * abs(e)
    dq = out["dEdq"]
    # E is quadratic in the charges: sum_i q_i dE/dq_i = 2E
    assert abs(float(np.dot(s.charge, dq)) - 2.0 * e) < 1e-5 * abs(e)
    with pytest.raises(N.GbiError):
        eng.set_exclusions([[0, s.n]])
    eng.close()
    g = N.Engine(_params("gbsw", N.NO_CUTOFF, 1.0), s.charge, s.radius)
    with pytest.raises(N.GbiError):
        g.set_exclusions(ex)
    g.close()
