"""HIP path against the golden fixtures of the 500-atom workload (needs an MI355X): the
stored FP64 oracle results, no oracle run.  Tolerances as in test_gpu_parity: force
relative RMS <= 1e-5, energy and Born radii relative error <= 1e-5."""
import os

import numpy as np
import pytest

from openmmcustomgbintegralforce_amd import _native as N

from common import rel_err, rel_rms
from test_golden import _golden, _params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["gbsw", "gbmv2", "gbmv1"])
def test_hip_matches_golden(name):
    g = _golden()
    print(f"[golden] {name} N=500", flush=True)
    eng = N.Engine(_params(name), g["charge"], g["radius"])
    e, f = eng.execute(g["pos"])
    born = eng.born_radii()
    eng.close()
    ee = rel_err(e, float(g[f"{name}_energy"][0]))
    fe = rel_rms(f, g[f"{name}_forces"])
    be = float(np.max(np.abs(born - g[f"{name}_born"]) / g[f"{name}_born"]))
    print(f"[golden] {name}: energy rel {ee:.2e} force rel-RMS {fe:.2e} born max rel {be:.2e}", flush=True)
    assert ee < 1e-5 and fe < 1e-5 and be < 1e-5, name
