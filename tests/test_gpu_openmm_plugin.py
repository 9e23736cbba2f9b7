"""The C++ OpenMM kernel classes end to end on the GPU: kernels created by name from the
"HIP" platform's registered factory, initialize / execute / copyParametersToContext,
compared with the FP64 oracle (force relative RMS <= 1e-5, energy relative error <= 1e-5).
"""
import subprocess

import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params, rel_err, rel_rms
from test_openmm_plugin import DRIVER, write_input

pytestmark = pytest.mark.gpu


def _run(tmp_path, model, itype, rule, s, **method):
    inp, out = tmp_path / "in.bin", tmp_path / "out.txt"
    write_input(inp, model, itype, rule, s.pos, s.charge, s.radius, **method)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    e = float(lines[0].split()[1])
    f = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + s.n]])
    rest = dict(ln.split() for ln in lines[1 + s.n:])
    return e, f, rest


@pytest.mark.parametrize("case", ["gbsw", "gbmv2"])
def test_cpp_kernel_classes_match_oracle(tmp_path, case):
    if case == "gbsw":
        s = S.globule(300, 40)
        p = N.default_params(N.MODEL_GBSW)
        e, f, rest = _run(tmp_path, 0, 0, 0, s)
    else:
        s = S.globule(400, 41)
        p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
        p.n_lebedev = 50
        e, f, rest = _run(tmp_path, 1, 2, 5, s)
    print(f"[openmm-plugin] {case} N={s.n} energy {e:.6f}", flush=True)
    ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
    assert rel_err(e, ref.energy) < 1e-5
    assert rel_rms(f, ref.forces) < 1e-5
    assert float(rest["energy_no_forces"]) == e
    # q -> 1.5 q scales E by 2.25 exactly; the pair terms are FP32 products of the new charges
    assert abs(float(rest["energy_scaled"]) - 2.25 * e) < 1e-6 * abs(e)
    assert rest["illegal_name_throws"] == "1"
    assert rest["count_change_throws"] == "1"


@pytest.mark.parametrize("case", ["gbsw", "gbmv2"])
def test_cpp_kernel_classes_cutoff_periodic(tmp_path, case):
    """CutoffPeriodic through the C++ kernel classes: the box comes from the Context at
    every execute (CharmmReferenceKernels.cpp:351, :430)."""
    L = [2.7, 2.6, 2.8]
    s = S.liquid(L, 42)
    if case == "gbsw":
        p = N.default_params(N.MODEL_GBSW)
        model, itype, rule = 0, 0, 0
    else:
        p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
        model, itype, rule = 1, 2, 0
    p.nonbonded_method, p.cutoff = N.CUTOFF_PERIODIC, 1.2
    e, f, rest = _run(tmp_path, model, itype, rule, s, method=N.CUTOFF_PERIODIC, cutoff=1.2, box=L)
    print(f"[openmm-plugin] {case} periodic N={s.n} energy {e:.6f}", flush=True)
    ref = O.compute(oracle_params(p, box=L), s.pos, s.charge, s.radius)
    assert rel_err(e, ref.energy) < 1e-5
    assert rel_rms(f, ref.forces) < 1e-5
    assert float(rest["energy_no_forces"]) == e


def test_cpp_kernel_classes_particle_pair_exclusions(tmp_path):
    """A canonical CharmmGBMVForce whose pair term is ParticlePair: the kernel class passes
    the force's exclusions (CharmmReferenceGBMV.cpp:272-284)."""
    from test_gpu_cutoff import _bonded_like_exclusions

    s = S.globule(400, 43)
    ex = _bonded_like_exclusions(s.pos)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    e, f, rest = _run(tmp_path, 1, 2, 0, s, method=N.NO_CUTOFF, cutoff=1.0, exclusions=ex)
    print(f"[openmm-plugin] gbmv2 ParticlePair with {len(ex)} exclusions: energy {e:.6f}", flush=True)
    ref = O.compute(O.set_exclusions(oracle_params(p), ex), s.pos, s.charge, s.radius)
    assert rel_err(e, ref.energy) < 1e-5
    assert rel_rms(f, ref.forces) < 1e-5


def test_cpp_kernel_classes_general_expressions(tmp_path):
    """A CharmmGBMVForce with non-canonical expressions through the C++ kernel class: kernels
    generated from its expressions (driver model 2: one integral, B = 1/(1/radius -
    0.5 I4 radius), 0.5 P q^2/B and the STILL pair term)."""
    from oracle import custom as C

    s = S.globule(350, 44)
    e, f, rest = _run(tmp_path, 2, 2, 0, s)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.orders[0] = 4
    P = -138.935456 * (1.0 - 1.0 / 80.0)
    E, F, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                           [("B", "1/(1/radius-0.5*I4*radius)")],
                           [("0.5*P*q^2/B", C.SINGLE), ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))",
                                                        C.PAIR_NO_EXCLUSIONS)], {"P": P})
    print(f"[openmm-plugin] general expressions N={s.n}: energy {e:.6f} vs {E:.6f}", flush=True)
    assert rel_err(e, E) < 1e-5
    assert rel_rms(f, F) < 1e-5
    assert abs(float(rest["energy_scaled"]) - 2.25 * e) < 1e-6 * abs(e)


def test_python_kernel_mirror_general_expressions():
    """forces.HipCalcCharmmGBMVForceKernel runs a non-canonical CharmmGBMVForce."""
    from oracle import custom as C
    from openmmcustomgbintegralforce_amd import forces as F

    s = S.globule(300, 45)
    f = F.CharmmGBMVForce()
    f.addPerParticleParameter("q")
    f.addPerParticleParameter("radius")
    f.addGlobalParameter("P", -138.935456 * (1.0 - 1.0 / 80.0))
    f.addGlobalParameter("k", 0.2)
    f.addGBIntegral("I4", [4])
    f.addGBIntegral("I7", [7])
    f.setGBIntegralType(F.CharmmGBMVForce.GBMVIntegralTypeI)
    born = "1/(-0.180*(1/0.05-I4)+1.817*(1/(4*0.05^4)-I7)^(1/4))"
    f.addComputedValue("B", born, F.CharmmGBMVForce.SingleParticle)
    f.addEnergyTerm("0.5*P*q^2/B+k*I4*radius", F.CharmmGBMVForce.SingleParticle)
    f.addEnergyTerm("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", F.CharmmGBMVForce.ParticlePair)
    for q, r in zip(s.charge, s.radius):
        f.addParticle([q, r])
    for i in range(0, s.n - 1, 4):
        f.addExclusion(i, i + 1)
    kern = F.HipCalcCharmmGBMVForceKernel()
    kern.initialize(F.System(s.n), f)
    ctx = F.ContextImpl(s.pos)
    e = kern.execute(ctx)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV1)
    E, Fo, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1,
                            ["I4", "I7"], [("B", born)],
                            [("0.5*P*q^2/B+k*I4*radius", C.SINGLE),
                             ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", C.PAIR)],
                            {"P": -138.935456 * (1.0 - 1.0 / 80.0), "k": 0.2}, exclusions=f._exclusions)
    print(f"[openmm-plugin] python mirror general expressions: energy {e:.6f} vs {E:.6f}", flush=True)
    assert rel_err(e, E) < 1e-5 and rel_rms(ctx.forces, Fo) < 1e-5
