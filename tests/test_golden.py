"""Golden fixtures of the 500-atom workload geometry (tests/golden/globule500_seed0.npz,
written by tests/golden/make_golden.py): the seeded generator still produces the stored
geometry, and the FP64 oracle still produces the stored energies, forces, Born radii and
integrals (CPU, no GPU needed)."""
import os

import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import synthetic as S

from common import rel_err, rel_rms

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "globule500_seed0.npz")


def _golden():
    return np.load(FIXTURE, allow_pickle=False)


def _params(name):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.params(name)


def test_generator_reproduces_fixture_geometry():
    g = _golden()
    s = S.globule(500, 0)
    assert np.allclose(s.pos, g["pos"], rtol=0.0, atol=1e-12)
    assert np.array_equal(s.charge, g["charge"])
    assert np.array_equal(s.radius, g["radius"])


@pytest.mark.parametrize("name", ["gbsw", "gbmv2", "gbmv1"])
def test_oracle_reproduces_fixture(name):
    g = _golden()
    ref = O.compute(O.params_from_abi(_params(name), nthreads=min(8, os.cpu_count() or 1)), g["pos"], g["charge"],
                    g["radius"])
    assert rel_err(ref.energy, float(g[f"{name}_energy"][0])) < 1e-11
    assert rel_rms(ref.forces, g[f"{name}_forces"]) < 1e-10
    assert np.max(np.abs(ref.born - g[f"{name}_born"]) / g[f"{name}_born"]) < 1e-11
    assert np.allclose(ref.integrals, g[f"{name}_integrals"], rtol=1e-11, atol=1e-14)
