"""Shapes that stress the lookup grid and the binning (needs an MI355X): a thin slab, a rod,
two clusters far apart (most of the grid empty), a regular lattice (every atom's points in
the same relative cells), a single isolated atom next to a cluster, atoms sharing a position
with another (r = 0 in the pair term, identical points) and a two-atom system; and a system of
one atom (no pairs, zero force).  The GBMV-form
models bin by 0.45 nm bricks of their 0.15 nm grid, the standalone GBSW force by cells of
its 0.20 nm grid; every case is compared with the FP64 oracle at 1e-5 (force relative RMS,
energy relative error, Born radii)."""
import os

import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params, rel_err, rel_rms

pytestmark = pytest.mark.gpu


def _lattice(shape, spacing, seed, jitter=0.02):
    rng = np.random.default_rng(seed)
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in shape], indexing="ij"), -1).reshape(-1, 3)
    pos = idx * spacing + rng.uniform(-jitter, jitter, idx.shape)
    n = len(pos)
    q = rng.uniform(-0.8, 0.8, n)
    q -= q.mean()
    R = rng.choice(S.RADII, n)
    return pos, q, R


def _systems():
    out = {}
    out["slab"] = _lattice((18, 18, 3), 0.22, 1)
    out["rod"] = _lattice((3, 3, 60), 0.22, 2)
    a, b = S.globule(300, 3), S.globule(300, 4)
    out["two clusters"] = (np.concatenate([a.pos, b.pos + [3.0, 0.0, 0.0]]), np.concatenate([a.charge, b.charge]),
                           np.concatenate([a.radius, b.radius]))
    out["regular lattice"] = _lattice((8, 8, 8), 0.25, 5, jitter=0.0)
    c = S.globule(400, 6)
    out["isolated atom"] = (np.concatenate([c.pos, [[4.0, 4.0, 4.0]]]), np.concatenate([c.charge, [0.5]]),
                            np.concatenate([c.radius, [0.17]]))
    d = S.globule(300, 8)
    dup = [5, 77, 150]  # three atoms duplicated at the same positions, other charges and radii
    out["coincident atoms"] = (np.concatenate([d.pos, d.pos[dup]]), np.concatenate([d.charge, [0.3, -0.4, 0.2]]),
                               np.concatenate([d.radius, [0.17, 0.2, 0.1]]))
    out["two atoms"] = (np.array([[0.0, 0.0, 0.0], [0.31, 0.0, 0.0]]), np.array([0.5, -0.5]), np.array([0.17, 0.2]))
    return out


SYSTEMS = _systems()
MODELS = {"gbmv2": (N.MODEL_GBMV, N.INTEGRAL_GBMV2, 50), "gbmv1": (N.MODEL_GBMV, N.INTEGRAL_GBMV1, 38),
          "gbsw-form": (N.MODEL_GBMV, N.INTEGRAL_GBSW, 38), "gbsw": (N.MODEL_GBSW, N.INTEGRAL_GBSW, 38)}


@pytest.mark.parametrize("model", sorted(MODELS))
@pytest.mark.parametrize("shape", sorted(SYSTEMS))
def test_shapes_match_oracle(shape, model):
    pos, q, R = SYSTEMS[shape]
    m, it, leb = MODELS[model]
    p = N.default_params(m, it)
    p.n_lebedev = leb
    print(f"[shapes] {shape} {model} N={len(q)}", flush=True)
    eng = N.Engine(p, q, R)
    e, f = eng.execute(pos)
    born = eng.born_radii()
    eng.close()
    ref = O.compute(oracle_params(p, nthreads=min(16, os.cpu_count() or 1)), pos, q, R)
    db = float(np.max(np.abs(born - ref.born) / ref.born))
    print(f"[shapes] {shape} {model}: born {db:.2e} energy {rel_err(e, ref.energy):.2e} "
          f"force {rel_rms(f, ref.forces):.2e}", flush=True)
    assert db < 1e-5
    assert rel_err(e, ref.energy) < 1e-5
    assert rel_rms(f, ref.forces) < 1e-5


@pytest.mark.parametrize("model", sorted(MODELS))
def test_single_atom(model):
    """One atom: the self energy alone, no pair, zero force (checked absolutely)."""
    pos, q, R = np.array([[0.1, 0.2, 0.3]]), np.array([0.7]), np.array([0.17])
    m, it, leb = MODELS[model]
    p = N.default_params(m, it)
    p.n_lebedev = leb
    eng = N.Engine(p, q, R)
    e, f = eng.execute(pos)
    born = eng.born_radii()
    eng.close()
    ref = O.compute(oracle_params(p), pos, q, R)
    print(f"[shapes] single atom {model}: born {born[0]:.6f} vs {ref.born[0]:.6f}, energy {e:.6f} vs "
          f"{ref.energy:.6f}, |F| {np.abs(f).max():.1e}", flush=True)
    assert abs(born[0] - ref.born[0]) < 1e-5 * ref.born[0]
    assert rel_err(e, ref.energy) < 1e-5
    assert np.abs(f).max() < 1e-6
