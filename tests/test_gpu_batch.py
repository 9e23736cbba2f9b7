"""Batched evaluation (several systems per launch), hipGraph replay and launch-shape
tuning on the HIP path, checked against the FP64 oracle system by system (needs an MI355X).

Tolerances as in test_gpu_parity: force relative RMS <= 1e-5, energy relative error
<= 1e-5 per system.  Where the same kernels run on the same inputs (graph replay vs
direct launches, a system in a batch vs alone) the results are compared bitwise or to
FP64 summation-order level.
"""
import os

import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params, rel_err, rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-5


def _params(kind: str):
    if kind == "gbsw":
        return N.default_params(N.MODEL_GBSW)
    if kind == "gbmv2-50":
        p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
        p.n_lebedev = 50
        return p
    if kind == "gbmv2-brute":
        p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
        p.lookup_policy = N.LOOKUP_BRUTE
        return p
    itype = {"gbmv-gbsw": N.INTEGRAL_GBSW, "gbmv1": N.INTEGRAL_GBMV1, "gbmv2": N.INTEGRAL_GBMV2}[kind]
    return N.default_params(N.MODEL_GBMV, itype)


MIX = [("gbsw", 300, 1), ("gbmv2-50", 420, 2), ("gbmv-gbsw", 350, 3), ("gbmv1", 260, 4), ("gbmv2-brute", 120, 5),
       ("gbsw", 510, 6), ("gbmv2", 333, 7)]


def _build(mix):
    systems = [S.globule(n, seed) for _, n, seed in mix]
    params = [_params(kind) for kind, _, _ in mix]
    # place the systems far apart: a batch does not couple them, whatever their coordinates
    pos = np.concatenate([s.pos + 50.0 * k for k, s in enumerate(systems)])
    eng = N.Engine.batch(params, [s.charge for s in systems], [s.radius for s in systems])
    return systems, params, pos, eng


def test_mixed_batch_matches_oracle_per_system():
    systems, params, pos, eng = _build(MIX)
    print(f"[batch] {len(MIX)} systems, N={eng.n}", flush=True)
    e, f = eng.execute(pos)
    born = eng.born_radii()
    assert np.all(np.isfinite(f))
    for k, ((kind, n, seed), s, p) in enumerate(zip(MIX, systems, params)):
        a, b = eng.offsets[k], eng.offsets[k + 1]
        ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
        ee, fe = rel_err(e[k], ref.energy), rel_rms(f[a:b], ref.forces)
        db = float(np.max(np.abs(born[a:b] - ref.born) / ref.born))
        print(f"[batch] system {k} {kind} N={n}: energy rel {ee:.2e} force rel-RMS {fe:.2e} born {db:.2e}",
              flush=True)
        assert ee < ENERGY_TOL and fe < FORCE_TOL and db < 1e-5, (k, kind)
    eng.close()


def test_batch_equals_single_system_runs():
    """A system inside a batch gets the same volumes and forces as alone (the pair tiles of
    a batch may split columns differently: FP64 summation-order differences only)."""
    systems, params, pos, eng = _build(MIX[:4])
    e, f = eng.execute(pos)
    integ = eng.integrals()
    eng.close()
    for k, (s, p) in enumerate(zip(systems, params)):
        one = N.Engine(p, s.charge, s.radius)
        e1, f1 = one.execute(s.pos + 50.0 * k)
        i1 = one.integrals()
        one.close()
        a, b = int(np.sum([x.n for x in systems[:k]])), int(np.sum([x.n for x in systems[:k + 1]]))
        assert np.array_equal(integ[:, a:b], i1), k
        assert abs(e[k] - e1) <= 1e-12 * abs(e1), k
        assert np.max(np.abs(f[a:b] - f1)) <= 1e-9 * np.max(np.abs(f1)), k


def test_graph_replay_matches_direct_launches():
    """Several graph replays on changing positions, with eager work and synchronisation in
    between; every replay is checked against the direct launch sequence."""
    systems, params, pos, eng = _build(MIX[:3])
    ref_eng = N.Engine.batch(params, [s.charge for s in systems], [s.radius for s in systems])
    eng.set_graph(True)
    rng = np.random.default_rng(11)
    for rep in range(4):
        x = pos + rng.normal(0.0, 0.01, pos.shape)
        e_ref, f_ref = ref_eng.execute(x)
        e_g, f_g = eng.execute(x)
        print(f"[graph] replay {rep}: max |dE| {np.max(np.abs(e_g - e_ref)):.1e}", flush=True)
        assert np.array_equal(e_g, e_ref), rep
        assert np.array_equal(f_g, f_ref), rep
        eng.born_radii()  # eager work between replays
    eng.close()
    ref_eng.close()


def test_regrow_after_expansion():
    """Positions that outgrow the sized grids (cells, list entries, point slots) are redone
    with larger buffers, in a batch and under graph replay."""
    systems, params, pos, eng = _build(MIX[:3])
    eng.set_graph(True)
    eng.execute(pos)
    big = []
    for k, s in enumerate(systems):
        c = s.pos.mean(axis=0)
        big.append((s.pos - c) * 1.6 + c)
    e, f = eng.execute(np.concatenate([b + 50.0 * k for k, b in enumerate(big)]))
    for k, (s, p) in enumerate(zip(systems, params)):
        a, b = eng.offsets[k], eng.offsets[k + 1]
        ref = O.compute(oracle_params(p), big[k], s.charge, s.radius)
        assert rel_err(e[k], ref.energy) < ENERGY_TOL and rel_rms(f[a:b], ref.forces) < FORCE_TOL, k
    eng.close()


def test_regrow_dense_cell_lists():
    """A globule compressed to ~10x the density: cells with up to 400 raw members (past
    the LDS sort size, sorted in global memory) and more list entries than the first
    sizing allows; the evaluation is redone with larger buffers and matches the oracle."""
    for kind, n, seed in (("gbmv2-50", 400, 21), ("gbsw", 400, 22)):
        s = S.globule(n, seed)
        c = s.pos.mean(axis=0)
        dense = (s.pos - c) * 0.45 + c
        p = _params(kind)
        eng = N.Engine(p, s.charge, s.radius)
        e, f = eng.execute(dense)
        ref = O.compute(oracle_params(p), dense, s.charge, s.radius)
        print(f"[dense] {kind}: energy {e:.6f} vs {ref.energy:.6f}", flush=True)
        assert rel_err(e, ref.energy) < ENERGY_TOL and rel_rms(f, ref.forces) < FORCE_TOL, kind
        eng.close()


def test_regrow_raw_cell_lists():
    """The per-cell raw member lists are sized from the positions at the first evaluation:
    a globule compressed after that (and, under graph replay, replayed), and two dense blobs
    far apart (a bounding box far emptier than the blobs) overflow them; the evaluation is
    redone with lists sized from the device's counts and matches the oracle."""
    for kind, n, seed, graph in (("gbmv2-50", 400, 23, False), ("gbsw", 400, 24, True)):
        s = S.globule(n, seed)
        c = s.pos.mean(axis=0)
        p = _params(kind)
        eng = N.Engine(p, s.charge, s.radius)
        eng.set_graph(graph)
        eng.execute(s.pos)
        dense = (s.pos - c) * 0.45 + c
        for rep in range(2):
            e, f = eng.execute(dense)
            ref = O.compute(oracle_params(p), dense, s.charge, s.radius)
            print(f"[raw lists] {kind} compressed, graph {graph}, rep {rep}: energy {e:.6f} vs {ref.energy:.6f}",
                  flush=True)
            assert rel_err(e, ref.energy) < ENERGY_TOL and rel_rms(f, ref.forces) < FORCE_TOL, (kind, rep)
        eng.close()
    a, b = S.globule(300, 25), S.globule(300, 26)
    pos = np.concatenate([(a.pos - a.pos.mean(0)) * 0.5, (b.pos - b.pos.mean(0)) * 0.5 + 8.0])
    charge, radius = np.concatenate([a.charge, b.charge]), np.concatenate([a.radius, b.radius])
    p = _params("gbmv2-50")
    eng = N.Engine(p, charge, radius)
    e, f = eng.execute(pos)
    eng.close()
    ref = O.compute(oracle_params(p), pos, charge, radius)
    print(f"[raw lists] two blobs: energy {e:.6f} vs {ref.energy:.6f}", flush=True)
    assert rel_err(e, ref.energy) < ENERGY_TOL and rel_rms(f, ref.forces) < FORCE_TOL


def test_autotune_keeps_parity(tmp_path):
    """Autotuning keeps parity; with a JSON cache the second context of the same kind takes
    the stored shapes without timing and gives the same result."""
    s = S.globule(900, 8)
    cache = str(tmp_path / "tuning.json")
    for p in (_params("gbsw"), _params("gbmv2-50")):
        eng = N.Engine(p, s.charge, s.radius)
        eng.set_positions(s.pos)
        t = eng.autotune(evals=2, cache=cache)
        print(f"[autotune] model {p.model}/{p.integral_type}: {t}", flush=True)
        assert t["ms_per_eval"] > 0.0 and not t.get("cached")
        e, f = eng.execute(s.pos)
        ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
        assert rel_err(e, ref.energy) < ENERGY_TOL and rel_rms(f, ref.forces) < FORCE_TOL
        eng.close()
        again = N.Engine(p, s.charge, s.radius)
        again.set_positions(s.pos)
        t2 = again.autotune(evals=2, cache=cache)
        assert t2.get("cached") and again.tuning()["blocks_per_cu"] == t["blocks_per_cu"]
        e2, f2 = again.execute(s.pos)
        again.close()
        assert e2 == e and np.array_equal(f2, f)


def test_batch_of_64_mixed_systems_sampled():
    """The batched workload shape (64 systems, sizes drawn from U[300, 6000] scaled down,
    alternating GBSW 24x38 / GBMV-II 24x50); every eighth system checked against the oracle."""
    sizes = S.batch_sizes(64, lo=100, hi=700, seed=3)
    systems = [S.globule(int(n), 100 + k) for k, n in enumerate(sizes)]
    params = [_params("gbsw") if k % 2 == 0 else _params("gbmv2-50") for k in range(64)]
    pos = np.concatenate([s.pos + 40.0 * k for k, s in enumerate(systems)])
    eng = N.Engine.batch(params, [s.charge for s in systems], [s.radius for s in systems])
    e, f = eng.execute(pos)
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(e))
    for k in range(0, 64, 8):
        s, p = systems[k], params[k]
        a, b = eng.offsets[k], eng.offsets[k + 1]
        ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
        assert rel_err(e[k], ref.energy) < ENERGY_TOL and rel_rms(f[a:b], ref.forces) < FORCE_TOL, k
    eng.close()


def test_batch64_full_sizes_matches_oracle():
    """The batched workload exactly as the benchmark builds it: 64 systems with
    N ~ U[300, 6000] (seeded; 205,186 atoms), alternating CharmmGBSWForce 24 x 38 and
    CharmmGBMVForce GBMVIntegralTypeII 24 x 50, one context.  Every eighth system and the
    largest are checked against the oracle; the graph replay the benchmark times gives the
    direct launches' results bitwise."""
    sizes, systems = S.batch64(0, workers=min(16, os.cpu_count() or 1))
    params = [_params("gbsw") if k % 2 == 0 else _params("gbmv2-50") for k in range(64)]
    pos = np.concatenate([s.pos + 40.0 * k for k, s in enumerate(systems)])
    print(f"[batch64] {len(systems)} systems, {int(np.sum(sizes))} atoms, largest {int(np.max(sizes))}",
          flush=True)
    eng = N.Engine.batch(params, [s.charge for s in systems], [s.radius for s in systems])
    e, f = eng.execute(pos)
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(e))
    eng.set_graph(True)
    eng.set_positions(pos)
    for _ in range(2):
        eng.compute()
        eng.synchronize()
        assert np.array_equal(eng.energy(), e) and np.array_equal(eng.forces(), f)
    eng.close()
    check = sorted(set(range(0, 64, 8)) | {int(np.argmax(sizes))})
    nthreads = min(16, os.cpu_count() or 1)
    for k in check:
        s, p = systems[k], params[k]
        a, b = int(np.sum(sizes[:k])), int(np.sum(sizes[:k + 1]))
        ref = O.compute(oracle_params(p, nthreads=nthreads), s.pos, s.charge, s.radius)
        ee, fe = rel_err(e[k], ref.energy), rel_rms(f[a:b], ref.forces)
        print(f"[batch64] system {k} N={s.n} {'gbsw' if k % 2 == 0 else 'gbmv2'}: energy rel {ee:.2e} "
              f"force rel-RMS {fe:.2e}", flush=True)
        assert ee < ENERGY_TOL and fe < FORCE_TOL, k
