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
    # the same evaluation with the Context's state on the device (a GPU platform's posq and
    # fixed-point force buffer, atoms in reversed slots): bitwise the host path's result
    assert rest["device_energy_equal"] == "1" and rest["device_forces_equal"] == "1"
    if case == "gbmv2":
        # Context::setParameter("alpha0", -0.2) reaches the next execute (CharmmReferenceKernels.cpp:355-357)
        pa = N.Params.from_buffer_copy(p)
        pa.alpha0 = -0.2
        ref_a = O.compute(oracle_params(pa), s.pos, s.charge, s.radius)
        print(f"[openmm-plugin] alpha0 -0.2: energy {float(rest['energy_alpha0']):.6f} vs {ref_a.energy:.6f}", flush=True)
        assert rel_err(float(rest["energy_alpha0"]), ref_a.energy) < 1e-5


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
    ctx = F.ContextImpl(s.pos, forces=[f])
    e = kern.execute(ctx)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV1)
    E, Fo, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1,
                            ["I4", "I7"], [("B", born)],
                            [("0.5*P*q^2/B+k*I4*radius", C.SINGLE),
                             ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", C.PAIR)],
                            {"P": -138.935456 * (1.0 - 1.0 / 80.0), "k": 0.2}, exclusions=f._exclusions)
    print(f"[openmm-plugin] python mirror general expressions: energy {e:.6f} vs {E:.6f}", flush=True)
    assert rel_err(e, E) < 1e-5 and rel_rms(ctx.forces, Fo) < 1e-5


DERIV_VALUES = [("B", "1/(1/radius-k*I4*radius)")]
DERIV_TERMS = [("0.5*P*q^2/B", 0), ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", 2)]
P_DEFAULT = -138.935456 * (1.0 - 1.0 / 80.0)


def _deriv_oracle(s, k, P):
    from oracle import custom as C

    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.orders[0] = 4
    E, _, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                           DERIV_VALUES, DERIV_TERMS, {"P": P, "k": k})
    return E


def _fd(s, k, P, which, h):
    if which == "k":
        return (_deriv_oracle(s, k + h, P) - _deriv_oracle(s, k - h, P)) / (2 * h)
    return (_deriv_oracle(s, k, P + h) - _deriv_oracle(s, k, P - h)) / (2 * h)


def test_cpp_kernel_classes_energy_parameter_derivatives(tmp_path):
    """CharmmGBMVForce::addEnergyParameterDerivative through the C++ kernel class (driver
    model 3: B = 1/(1/radius - k I4 radius), STILL energy, derivatives with respect to k and
    P): each execute adds dE/dk and dE/dP to the Context (CharmmReferenceKernels.cpp:360-368),
    checked against central differences of the oracle energy; Context::setParameter("k")
    reaches the next execute."""
    s = S.globule(300, 46)
    e, f, rest = _run(tmp_path, 3, 2, 0, s)
    E = _deriv_oracle(s, 0.5, P_DEFAULT)
    fd_k, fd_P = _fd(s, 0.5, P_DEFAULT, "k", 1e-4), _fd(s, 0.5, P_DEFAULT, "P", 1e-3)
    dk, dP = float(rest["dEdk"]), float(rest["dEdP"])
    print(f"[openmm-plugin] dE/dk {dk:.6f} (FD {fd_k:.6f}), dE/dP {dP:.6f} (FD {fd_P:.6f})", flush=True)
    assert rel_err(e, E) < 1e-5
    assert rel_err(dk, fd_k) < 1e-4 and rel_err(dP, fd_P) < 1e-5
    # E is linear in P: dE/dP = E/P exactly
    assert rel_err(dP, e / P_DEFAULT) < 1e-6
    e2, dk2 = float(rest["energy_k045"]), float(rest["dEdk_k045"])
    assert rel_err(e2, _deriv_oracle(s, 0.45, P_DEFAULT)) < 1e-5
    assert rel_err(dk2, _fd(s, 0.45, P_DEFAULT, "k", 1e-4)) < 1e-4


def test_python_kernel_mirror_context_parameters_and_derivatives():
    """forces.HipCalcCharmmGBMVForceKernel reads the global parameters from the Context at
    every execute (canonical force: Born-map constants through gbi_update_constants; general
    expressions: gbi_set_custom_globals) and adds energy parameter derivatives to it;
    copyParametersToContext of a general force keeps the compiled kernels."""
    from openmmcustomgbintegralforce_amd import forces as F

    s = S.globule(300, 47)
    # canonical: alpha0 and P changed in the Context
    f = F.canonical_gbmv_force(s.charge, s.radius, F.CharmmGBMVForce.GBMVIntegralTypeII, lebedev_rule=5)
    kern = F.HipCalcCharmmGBMVForceKernel()
    kern.initialize(F.System(s.n), f)
    assert not kern._custom
    ctx = F.ContextImpl(s.pos, forces=[f])
    ctx.setParameter("alpha0", -0.19)
    ctx.setParameter("P", 0.9 * P_DEFAULT)
    e = kern.execute(ctx)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.n_lebedev = 50
    p.alpha0, p.prefactor = -0.19, 0.9 * P_DEFAULT
    ref = O.compute(oracle_params(p), s.pos, s.charge, s.radius)
    print(f"[openmm-plugin] python canonical with Context alpha0/P: {e:.6f} vs {ref.energy:.6f}", flush=True)
    assert rel_err(e, ref.energy) < 1e-5 and rel_rms(ctx.forces, ref.forces) < 1e-5
    # general expressions with derivatives and a Context parameter
    g = F.CharmmGBMVForce()
    g.addPerParticleParameter("q")
    g.addPerParticleParameter("radius")
    g.addGlobalParameter("P", P_DEFAULT)
    g.addGlobalParameter("k", 0.5)
    g.addGBIntegral("I4", [4])
    g.setGBIntegralType(F.CharmmGBMVForce.GBMVIntegralTypeII)
    g.addComputedValue("B", DERIV_VALUES[0][1], F.CharmmGBMVForce.SingleParticle)
    for expr, ty in DERIV_TERMS:
        g.addEnergyTerm(expr, ty)
    g.addEnergyParameterDerivative("k")
    for q, r in zip(s.charge, s.radius):
        g.addParticle([q, r])
    kg = F.HipCalcCharmmGBMVForceKernel()
    kg.initialize(F.System(s.n), g)
    assert kg._custom
    cg = F.ContextImpl(s.pos, forces=[g])
    cg.setParameter("k", 0.45)
    e = kg.execute(cg)
    fd = _fd(s, 0.45, P_DEFAULT, "k", 1e-4)
    print(f"[openmm-plugin] python general k=0.45: dE/dk {cg.energy_param_derivs['k']:.6f} (FD {fd:.6f})", flush=True)
    assert rel_err(e, _deriv_oracle(s, 0.45, P_DEFAULT)) < 1e-5
    assert rel_err(cg.energy_param_derivs["k"], fd) < 1e-4
    # new charges through copyParametersToContext: same kernels, E scales by 1.5^2
    for i in range(s.n):
        g.setParticleParameters(i, [1.5 * s.charge[i], s.radius[i]])
    kg.copyParametersToContext(cg, g)
    e15 = kg.execute(cg, False, True)
    assert abs(e15 - 2.25 * e) < 1e-6 * abs(e)


def test_cpp_kernel_classes_tabulated_functions_and_coordinates(tmp_path):
    """Driver model 4: a CharmmGBMVForce with Continuous1D/2D/3DFunctions, Discrete1D/2D/3DFunctions
    and the coordinates x, y, z in its Born value and single-particle term, through the C++ kernel
    class (OpenMM's tabulated-function objects converted to the C ABI's tables)."""
    from oracle import custom as C

    s = S.globule(300, 45)
    e, f, rest = _run(tmp_path, 4, 2, 0, s)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.orders[0] = 4
    P = -138.935456 * (1.0 - 1.0 / 80.0)
    scale = C.Tabulated("scale", 0, 1.0 + 0.2 * np.sin(0.3 * np.arange(31)), 0.0, 3.0)
    w = C.Tabulated("w", 1, [0.9, 1.1])
    lut = C.Tabulated("lut", 2, [0.7, 0.8, 0.95, 1.05, 1.1, 1.2], dims=(2, 3, 1))
    tri3 = C.Tabulated("tri3", 3, 1.0 + 0.03 * np.arange(12), dims=(2, 2, 3))
    i2, j2 = np.meshgrid(np.arange(5), np.arange(4), indexing="ij")
    surf = C.Tabulated("surf", 4, (0.2 * np.sin(0.7 * i2) + 0.1 * j2).T.ravel(), -1.5, 1.5, dims=(5, 4, 1),
                       yrange=(-1.5, 1.5))
    i3, j3, k3 = np.meshgrid(np.arange(4), np.arange(3), np.arange(3), indexing="ij")
    vol3 = C.Tabulated("vol3", 5, (1.0 + 0.05 * i3 - 0.03 * j3 + 0.02 * k3 * i3).transpose(2, 1, 0).ravel(), -1.5, 1.5,
                       dims=(4, 3, 3), yrange=(-1.5, 1.5), zrange=(-1.5, 1.5))
    E, F, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                           [("B", "scale(10*radius)*lut(step(q), 10*radius)*(1+0.002*(x+y-z))/(1/radius-0.5*I4*radius)")],
                           [("0.5*P*q^2/B*w(step(q))*tri3(step(q), step(radius-0.18), 20*radius-2)"
                             "+0.01*q^2*(x^2+y^2+z^2)+0.2*q^2*surf(x, y)+0.3*q^2*vol3(x, z, y)", C.SINGLE),
                            ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", C.PAIR_NO_EXCLUSIONS)], {"P": P},
                           functions=[scale, w, lut, tri3, surf, vol3])
    print(f"[openmm-plugin] tables + x,y,z N={s.n}: energy {e:.6f} vs {E:.6f} force {rel_rms(f, F):.2e}", flush=True)
    assert rel_err(e, E) < 1e-5
    assert rel_rms(f, F) < 1e-5


def test_python_kernel_mirror_tabulated_functions():
    """forces.CharmmGBMVForce.addTabulatedFunction through the Python kernel mirror (a
    Continuous1DFunction, a Discrete2DFunction, a Continuous2DFunction); a table class the
    platform does not know is refused with OpenMMException."""
    from oracle import custom as C
    from openmmcustomgbintegralforce_amd import forces as FM

    s = S.globule(250, 46)
    P = -138.935456 * (1.0 - 1.0 / 80.0)
    vals = 1.0 + 0.2 * np.sin(0.3 * np.arange(31))

    def force(fn, arg="10*radius"):
        f = FM.CharmmGBMVForce()
        f.addPerParticleParameter("q")
        f.addPerParticleParameter("radius")
        f.addGlobalParameter("P", P)
        f.addGBIntegral("I4", [4])
        f.setGBIntegralType(FM.CharmmGBMVForce.GBMVIntegralTypeII)
        f.addTabulatedFunction("scale", fn)
        f.addComputedValue("B", f"scale({arg})/(1/radius-0.5*I4*radius)", FM.CharmmGBMVForce.SingleParticle)
        f.addEnergyTerm("0.5*P*q^2/B+0.01*q^2*z^2", FM.CharmmGBMVForce.SingleParticle)
        f.addEnergyTerm("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", FM.CharmmGBMVForce.ParticlePairNoExclusions)
        for q, r in zip(s.charge, s.radius):
            f.addParticle([q, r])
        return f

    k = FM.HipCalcCharmmGBMVForceKernel()
    f1 = force(FM.Continuous1DFunction(vals, 0.0, 3.0))
    k.initialize(FM.System(s.n), f1)
    ctx = FM.ContextImpl(s.pos, forces=(f1,))
    e = k.execute(ctx, True, True)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.orders[0] = 4
    E, F, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                           [("B", "scale(10*radius)/(1/radius-0.5*I4*radius)")],
                           [("0.5*P*q^2/B+0.01*q^2*z^2", C.SINGLE),
                            ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", C.PAIR_NO_EXCLUSIONS)], {"P": P},
                           functions=[C.Tabulated("scale", 0, vals, 0.0, 3.0)])
    print(f"[openmm-plugin] python mirror tables: energy {e:.6f} vs {E:.6f}", flush=True)
    assert rel_err(e, E) < 1e-5 and rel_rms(ctx.forces, F) < 1e-5
    d2 = [0.8, 0.9, 1.0, 1.1, 1.15, 1.25]
    f2 = force(FM.Discrete2DFunction(2, 3, d2), "step(q), 10*radius")
    k2 = FM.HipCalcCharmmGBMVForceKernel()
    k2.initialize(FM.System(s.n), f2)
    ctx2 = FM.ContextImpl(s.pos, forces=(f2,))
    e2 = k2.execute(ctx2, True, True)
    E2, F2, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                             [("B", "scale(step(q), 10*radius)/(1/radius-0.5*I4*radius)")],
                             [("0.5*P*q^2/B+0.01*q^2*z^2", C.SINGLE),
                              ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", C.PAIR_NO_EXCLUSIONS)], {"P": P},
                             functions=[C.Tabulated("scale", 2, d2, dims=(2, 3, 1))])
    print(f"[openmm-plugin] python mirror Discrete2D: energy {e2:.6f} vs {E2:.6f}", flush=True)
    assert rel_err(e2, E2) < 1e-5 and rel_rms(ctx2.forces, F2) < 1e-5
    # Continuous2DFunction of (10 radius, 10 B-free argument) on [0.5, 2.5] x [0, 1]
    gx, gy = np.meshgrid(np.linspace(0.5, 2.5, 6), np.linspace(0.0, 1.0, 4), indexing="ij")
    c2 = (1.0 + 0.1 * np.sin(gx) * np.cos(2 * gy)).T.ravel()
    f3 = force(FM.Continuous2DFunction(6, 4, c2, 0.5, 2.5, 0.0, 1.0), "10*radius, 0.4")
    k3 = FM.HipCalcCharmmGBMVForceKernel()
    k3.initialize(FM.System(s.n), f3)
    ctx3 = FM.ContextImpl(s.pos, forces=(f3,))
    e3 = k3.execute(ctx3, True, True)
    E3, F3, _, _ = C.compute(oracle_params(p), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1, ["I4"],
                             [("B", "scale(10*radius, 0.4)/(1/radius-0.5*I4*radius)")],
                             [("0.5*P*q^2/B+0.01*q^2*z^2", C.SINGLE),
                              ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))", C.PAIR_NO_EXCLUSIONS)], {"P": P},
                             functions=[C.Tabulated("scale", 4, c2, 0.5, 2.5, dims=(6, 4, 1), yrange=(0.0, 1.0))])
    print(f"[openmm-plugin] python mirror Continuous2D: energy {e3:.6f} vs {E3:.6f}", flush=True)
    assert rel_err(e3, E3) < 1e-5 and rel_rms(ctx3.forces, F3) < 1e-5

    class Unknown(FM.TabulatedFunction):
        pass

    with pytest.raises(FM.OpenMMException):
        FM.HipCalcCharmmGBMVForceKernel().initialize(FM.System(s.n), force(Unknown()))
