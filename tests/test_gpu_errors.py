"""Error paths of an evaluation and recovery from them (needs an MI355X).  The conditions are
inputs that the library's own checks turn into errors: a non-finite position (the host check
of gbi_execute; for positions that stay on the device, the bounding-box kernel's flag, after
which every later kernel of the evaluation returns at once), a system too spread out for the
lookup grid (the reference's limit of 1000 nm per axis, CustomGBIntegral.cpp:173-187), and
charges that are all
zero (energy and forces exactly zero).  After an error the same context evaluates valid
positions again and matches the FP64 oracle."""
import os

import numpy as np
import pytest
import torch  # before the library loads the HIP runtime (device-resident positions below)

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params, rel_err, rel_rms

pytestmark = pytest.mark.gpu

MODELS = {"gbmv2": (N.MODEL_GBMV, N.INTEGRAL_GBMV2, 50), "gbsw": (N.MODEL_GBSW, N.INTEGRAL_GBSW, 38)}


def _params(model):
    m, it, leb = MODELS[model]
    p = N.default_params(m, it)
    p.n_lebedev = leb
    return p


def _check_oracle(p, eng, s, tag):
    e, f = eng.execute(s.pos)
    ref = O.compute(oracle_params(p, nthreads=min(16, os.cpu_count() or 1)), s.pos, s.charge, s.radius)
    print(f"[errors] {tag}: energy {rel_err(e, ref.energy):.2e} force {rel_rms(f, ref.forces):.2e}", flush=True)
    assert rel_err(e, ref.energy) < 1e-5 and rel_rms(f, ref.forces) < 1e-5


@pytest.mark.parametrize("model", sorted(MODELS))
def test_nonfinite_position_then_recovery(model):
    s = S.globule(300, 90)
    p = _params(model)
    eng = N.Engine(p, s.charge, s.radius)
    bad = s.pos.copy()
    bad[17, 1] = np.nan
    with pytest.raises(N.GbiError) as ex:  # first evaluation: caught while sizing the grid
        eng.execute(bad)
    assert ex.value.status == 7
    _check_oracle(p, eng, s, f"{model} after a non-finite first evaluation")
    bad[17, 1] = np.inf
    d_bad = torch.tensor(bad, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.set_positions(d_bad.data_ptr(), on_device=True)  # no host check: the device flags it
    eng.compute()
    with pytest.raises(N.GbiError) as ex:
        eng.synchronize()
    assert ex.value.status == 7
    del d_bad
    _check_oracle(p, eng, s, f"{model} after a non-finite later evaluation")
    eng.close()


@pytest.mark.parametrize("model", sorted(MODELS))
def test_grid_too_large_then_recovery(model):
    s = S.globule(200, 91)
    p = _params(model)
    eng = N.Engine(p, s.charge, s.radius)
    _check_oracle(p, eng, s, f"{model} before")
    far = s.pos.copy()
    far[0, 0] += 2000.0  # one atom 2000 nm away along x: the grid would be longer than 1000 nm
    with pytest.raises(N.GbiError) as ex:
        eng.execute(far)
    assert ex.value.status == 6
    print(f"[errors] {model} spread out: {ex.value}", flush=True)
    _check_oracle(p, eng, s, f"{model} after a spread-out evaluation")
    eng.close()


@pytest.mark.parametrize("model", sorted(MODELS))
def test_zero_charges(model):
    s = S.globule(250, 92)
    p = _params(model)
    eng = N.Engine(p, np.zeros(s.n), s.radius)
    e, f = eng.execute(s.pos)
    born = eng.born_radii()
    eng.close()
    ref = O.compute(oracle_params(p), s.pos, np.zeros(s.n), s.radius)
    print(f"[errors] {model} zero charges: energy {e}, |F| {np.abs(f).max()}", flush=True)
    assert e == 0.0 and np.all(f == 0.0)
    assert np.max(np.abs(born - ref.born) / ref.born) < 1e-5  # the radii do not depend on the charges
