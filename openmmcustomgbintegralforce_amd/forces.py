"""Python mirror of the reference's force API and kernel interface for the GB-integral path.

Classes and method names follow ``openmmapi/include/openmm/CharmmGBSWForce.h``,
``CharmmGBMVForce.h`` and ``charmmKernels.h``; the HIP kernels behind them are reached
through the C ABI (``include/gbi.h``).  ``CustomGBIntegralForce`` is an alias of
``CharmmGBMVForce`` (the "custom GB integral force").

The canonical GB forms (integrals of order 4 and 7, the Born radius from them, the STILL
self + pair energy; see ``canonical_gbmv_force``) run on the hand-written kernels; any other
``CharmmGBMVForce`` expressions run on kernels generated from them (``Engine.custom``,
``gbmv_custom_from_force``).
"""
from __future__ import annotations

import numpy as np

from . import _native as N

ONE_4PI_EPS0 = 138.935456  # OpenMM 7.x


class OpenMMException(RuntimeError):
    pass


# --------------------------------------------------------------------------------------
# CharmmGBSWForce (CharmmGBSWForce.h:58-222)
# --------------------------------------------------------------------------------------
class CharmmGBSWForce:
    NoCutoff, CutoffNonPeriodic, CutoffPeriodic = 0, 1, 2

    def __init__(self):
        self._particles = []  # (charge, radius, scalingFactor)
        self._solvent = 78.3
        self._solute = 1.0
        self._sa = 2.25936
        self._method = self.NoCutoff
        self._cutoff = 1.0

    def getNumParticles(self) -> int:
        return len(self._particles)

    def addParticle(self, charge: float, radius: float, scalingFactor: float = 1.0) -> int:
        self._particles.append((float(charge), float(radius), float(scalingFactor)))
        return len(self._particles) - 1

    def getParticleParameters(self, index: int):
        return self._particles[index]

    def setParticleParameters(self, index: int, charge: float, radius: float, scalingFactor: float) -> None:
        self._particles[index] = (float(charge), float(radius), float(scalingFactor))

    def getSolventDielectric(self) -> float:
        return self._solvent

    def setSolventDielectric(self, v: float) -> None:
        self._solvent = float(v)

    def getSoluteDielectric(self) -> float:
        return self._solute

    def setSoluteDielectric(self, v: float) -> None:
        self._solute = float(v)

    def getSurfaceAreaEnergy(self) -> float:
        return self._sa

    def setSurfaceAreaEnergy(self, v: float) -> None:
        self._sa = float(v)

    def getNonbondedMethod(self) -> int:
        return self._method

    def setNonbondedMethod(self, m: int) -> None:
        self._method = int(m)

    def getCutoffDistance(self) -> float:
        return self._cutoff

    def setCutoffDistance(self, d: float) -> None:
        self._cutoff = float(d)

    def arrays(self):
        a = np.array(self._particles, dtype=np.float64).reshape(-1, 3)
        return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


# --------------------------------------------------------------------------------------
# CharmmGBMVForce (CharmmGBMVForce.h:63-626)
# --------------------------------------------------------------------------------------
class TabulatedFunction:
    """OpenMM's tabulated-function base (openmm/TabulatedFunction.h)."""

    _periodic = False

    def getPeriodic(self) -> bool:
        return self._periodic


class Continuous1DFunction(TabulatedFunction):
    """A cubic spline through values equally spaced on [min, max] (natural, zero outside the
    range; or periodic with period max - min)."""

    def __init__(self, values, min: float, max: float, periodic: bool = False):
        self._values, self._min, self._max, self._periodic = list(values), float(min), float(max), bool(periodic)

    def getFunctionParameters(self):
        return list(self._values), self._min, self._max

    def setFunctionParameters(self, values, min: float, max: float) -> None:
        self._values, self._min, self._max = list(values), float(min), float(max)


class Discrete1DFunction(TabulatedFunction):
    """values[round(x)]."""

    def __init__(self, values):
        self._values = list(values)

    def getFunctionParameters(self):
        return list(self._values)

    def setFunctionParameters(self, values) -> None:
        self._values = list(values)


class Discrete2DFunction(TabulatedFunction):
    """values[x + xsize y] at the rounded arguments."""

    def __init__(self, xsize: int, ysize: int, values):
        self._args = (int(xsize), int(ysize), list(values))

    def getFunctionParameters(self):
        return self._args


class Discrete3DFunction(TabulatedFunction):
    """values[x + xsize (y + ysize z)] at the rounded arguments."""

    def __init__(self, xsize: int, ysize: int, zsize: int, values):
        self._args = (int(xsize), int(ysize), int(zsize), list(values))

    def getFunctionParameters(self):
        return self._args


class Continuous2DFunction(TabulatedFunction):
    """values[x + xsize y] equally spaced on [xmin, xmax] x [ymin, ymax]: the bicubic patch
    with spline knot derivatives (natural and zero outside, or periodic)."""

    def __init__(self, xsize, ysize, values, xmin, xmax, ymin, ymax, periodic=False):
        self._args = (int(xsize), int(ysize), list(values), float(xmin), float(xmax), float(ymin), float(ymax))
        self._periodic = bool(periodic)

    def getFunctionParameters(self):
        return self._args


class Continuous3DFunction(TabulatedFunction):
    """values[x + xsize (y + ysize z)] equally spaced on [xmin, xmax] x [ymin, ymax] x
    [zmin, zmax]: the tricubic patch with spline knot derivatives."""

    def __init__(self, xsize, ysize, zsize, values, xmin, xmax, ymin, ymax, zmin, zmax, periodic=False):
        self._args = (int(xsize), int(ysize), int(zsize), list(values), float(xmin), float(xmax), float(ymin),
                      float(ymax), float(zmin), float(zmax))
        self._periodic = bool(periodic)

    def getFunctionParameters(self):
        return self._args


class CharmmGBMVForce:
    NoCutoff, CutoffNonPeriodic, CutoffPeriodic = 0, 1, 2
    SingleParticle, ParticlePair, ParticlePairNoExclusions = 0, 1, 2
    GBSWIntegral, GBMVIntegralTypeI, GBMVIntegralTypeII, CustomIntegral = 0, 1, 2, 3

    def __init__(self):
        self._method = self.NoCutoff
        self._cutoff = 1.0
        self._integral_type = self.GBSWIntegral  # CharmmGBMVForce.cpp:51
        self._param_names = []
        self._globals = []  # (name, default)
        self._particles = []
        self._exclusions = []
        self._values = []  # (name, expression, type)
        self._integrals = []  # (name, ints, reals)
        self._energy = []  # (expression, type)
        self._param_derivs = []
        self._functions = []  # (name, TabulatedFunction)

    # -- bookkeeping (CharmmGBMVForce.h:119-410)
    def getNumParticles(self):
        return len(self._particles)

    def getNumExclusions(self):
        return len(self._exclusions)

    def getNumPerParticleParameters(self):
        return len(self._param_names)

    def getNumGlobalParameters(self):
        return len(self._globals)

    def getNumComputedValues(self):
        return len(self._values)

    def getNumGBIntegrals(self):
        return len(self._integrals)

    def getNumEnergyTerms(self):
        return len(self._energy)

    def getNonbondedMethod(self):
        return self._method

    def setNonbondedMethod(self, m):
        self._method = int(m)

    def getCutoffDistance(self):
        return self._cutoff

    def setCutoffDistance(self, d):
        self._cutoff = float(d)

    def addPerParticleParameter(self, name: str) -> int:
        self._param_names.append(name)
        return len(self._param_names) - 1

    def getPerParticleParameterName(self, i):
        return self._param_names[i]

    def addGlobalParameter(self, name: str, defaultValue: float) -> int:
        self._globals.append((name, float(defaultValue)))
        return len(self._globals) - 1

    def getGlobalParameterName(self, i):
        return self._globals[i][0]

    def getGlobalParameterDefaultValue(self, i):
        return self._globals[i][1]

    def addEnergyParameterDerivative(self, name: str) -> None:
        self._param_derivs.append(name)

    def getNumEnergyParameterDerivatives(self) -> int:
        return len(self._param_derivs)

    def getEnergyParameterDerivativeName(self, i) -> str:
        return self._param_derivs[i]

    def addParticle(self, parameters=()) -> int:
        self._particles.append([float(v) for v in parameters])
        return len(self._particles) - 1

    def getParticleParameters(self, i):
        return list(self._particles[i])

    def setParticleParameters(self, i, parameters):
        self._particles[i] = [float(v) for v in parameters]

    def addComputedValue(self, name: str, expression: str, type: int) -> int:
        self._values.append((name, expression, int(type)))
        return len(self._values) - 1

    def getComputedValueParameters(self, i):
        return self._values[i]

    def addGBIntegral(self, name: str, parametersInt=(), parametersReal=()) -> int:
        self._integrals.append((name, [int(v) for v in parametersInt], [float(v) for v in parametersReal]))
        return len(self._integrals) - 1

    def getGBIntegralParameters(self, i):
        return self._integrals[i]

    def setGBIntegralParameters(self, i, name: str, parametersInt=(), parametersReal=()) -> None:
        self._integrals[i] = (name, [int(v) for v in parametersInt], [float(v) for v in parametersReal])

    def setGBIntegralType(self, t: int) -> None:
        self._integral_type = int(t)

    def getGBIntegralType(self) -> int:
        return self._integral_type

    def addEnergyTerm(self, expression: str, type: int) -> int:
        self._energy.append((expression, int(type)))
        return len(self._energy) - 1

    def getEnergyTermParameters(self, i):
        return self._energy[i]

    def addExclusion(self, p1: int, p2: int) -> int:
        self._exclusions.append((int(p1), int(p2)))
        return len(self._exclusions) - 1

    # -- tabulated functions (CharmmGBMVForce.h:437-483)
    def addTabulatedFunction(self, name: str, function: TabulatedFunction) -> int:
        self._functions.append((name, function))
        return len(self._functions) - 1

    def getNumTabulatedFunctions(self) -> int:
        return len(self._functions)

    getNumFunctions = getNumTabulatedFunctions

    def getTabulatedFunction(self, i) -> TabulatedFunction:
        return self._functions[i][1]

    def getTabulatedFunctionName(self, i) -> str:
        return self._functions[i][0]

    def addFunction(self, name: str, values, min: float, max: float) -> int:
        return self.addTabulatedFunction(name, Continuous1DFunction(values, min, max))


CustomGBIntegralForce = CharmmGBMVForce


def _norm(expr: str) -> str:
    return "".join(expr.split())


# Canonical expressions of the CharmmGBMVForce family recognised by the HIP platform.
CANONICAL_BORN = ("1/(alpha0*(1/eta-I4)+alpha1*(1/(4*eta^4)-I7)^(1/4))", )
CANONICAL_SELF = "0.5*P*q^2/B"
CANONICAL_PAIR = "P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))"


def canonical_gbmv_force(charges, radii, integral_type: int, *, prefactor: float | None = None,
                         lebedev_rule: int | None = None, n_radial: int = 24,
                         alpha0=-0.180, alpha1=1.817, eta=0.05, slope=1.0, shift=0.0,
                         solute_dielectric=1.0, solvent_dielectric=80.0,
                         pair_type: int = 2) -> CharmmGBMVForce:
    """A CharmmGBMVForce holding the canonical GB energy the HIP platform implements:

    integrals I4 (order 4) and I7 (order 7); computed value
    B = slope/(alpha0 (1/eta - I4) + alpha1 (1/(4 eta^4) - I7)^(1/4)) + shift;
    energy 0.5 P q^2/B (single particle) + P q1 q2/f (pairs: ``pair_type``
    ParticlePairNoExclusions by default, or ParticlePair, which honours ``addExclusion``),
    P = -ONE_4PI_EPS0 (1/solute - 1/solvent).
    ``lebedev_rule`` (parametersInt[1]) and ``n_radial`` (parametersInt[2]) extend the
    reference, which fixes rule 3 and 24 radial points (GBSWIntegral.cpp:37-40).
    """
    f = CharmmGBMVForce()
    f.addPerParticleParameter("q")
    f.addPerParticleParameter("radius")
    P = prefactor if prefactor is not None else -ONE_4PI_EPS0 * (1.0 / solute_dielectric - 1.0 / solvent_dielectric)
    for name, val in (("alpha0", alpha0), ("alpha1", alpha1), ("eta", eta), ("slope", slope),
                      ("shift", shift), ("P", P)):
        f.addGlobalParameter(name, val)
    ints4 = [4] + ([lebedev_rule, n_radial] if lebedev_rule is not None else [])
    ints7 = [7] + ([lebedev_rule, n_radial] if lebedev_rule is not None else [])
    f.addGBIntegral("I4", ints4)
    f.addGBIntegral("I7", ints7)
    f.setGBIntegralType(integral_type)
    f.addComputedValue("B", "slope*" + CANONICAL_BORN[0] + "+shift", CharmmGBMVForce.SingleParticle)
    f.addEnergyTerm(CANONICAL_SELF, CharmmGBMVForce.SingleParticle)
    f.addEnergyTerm(CANONICAL_PAIR, pair_type)
    for q, r in zip(charges, radii):
        f.addParticle([q, r])
    return f


_LEBEDEV_RULE_NPTS = {1: 6, 2: 14, 3: 26, 4: 38, 5: 50}


def gbmv_params_from_force(force: CharmmGBMVForce) -> tuple:
    """(gbi Params, charges, radii) of a canonical CharmmGBMVForce; raises otherwise."""
    t = force.getGBIntegralType()
    if t == CharmmGBMVForce.CustomIntegral:
        raise N.GbiError(2, "CustomIntegral (type 3) is not implemented by the reference either")
    names = [force.getPerParticleParameterName(i) for i in range(force.getNumPerParticleParameters())]
    if "radius" not in names:  # GBSWIntegral.cpp:79-80
        raise OpenMMException("GBSWIntegral: the per perticle parameter 'radius' must be defined")
    ints = [force.getGBIntegralParameters(i)[1] for i in range(force.getNumGBIntegrals())]
    for pi in ints:
        if len(pi) == 0:
            raise OpenMMException("GBSWIntegral: the order of integral must be given")
        if pi[0] < 2:
            raise OpenMMException("GBSWIntegral: the order of integral must be greater or equal to 2")
    orders = [pi[0] for pi in ints]
    # the canonical Born expression names its integrals: I4 must be of order 4, I7 of order 7
    named = {force.getGBIntegralParameters(i)[0]: o for i, o in enumerate(orders)}
    if len(orders) != 2 or named != {"I4": 4, "I7": 7}:
        raise N.GbiError(2, "the canonical Born map takes an integral I4 of order 4 and an integral I7 of order 7")
    if force.getNumEnergyParameterDerivatives() > 0:
        raise N.GbiError(2, "energy parameter derivatives run on the generated kernels")
    values = [(n, _norm(e), ty) for n, e, ty in force._values]
    energies = [(_norm(e), ty) for e, ty in force._energy]
    canon_values = [("B", _norm("slope*" + CANONICAL_BORN[0] + "+shift"), CharmmGBMVForce.SingleParticle)]
    canon_energy = [[(_norm(CANONICAL_SELF), CharmmGBMVForce.SingleParticle), (_norm(CANONICAL_PAIR), t)]
                    for t in (CharmmGBMVForce.ParticlePairNoExclusions, CharmmGBMVForce.ParticlePair)]
    if values != canon_values or energies not in canon_energy:
        raise N.GbiError(2, "only the canonical GB expressions run on the HIP platform")
    g = {force.getGlobalParameterName(i): force.getGlobalParameterDefaultValue(i)
         for i in range(force.getNumGlobalParameters())}
    p = N.default_params(N.MODEL_GBMV, t)
    first = ints[0]
    if len(first) > 1:
        p.n_lebedev = _LEBEDEV_RULE_NPTS[first[1]]
    if len(first) > 2:
        p.n_radial = first[2]
    p.orders[0], p.orders[1] = orders[0], orders[1]
    for k in ("alpha0", "alpha1", "eta", "slope", "shift"):
        setattr(p, k, g[k])
    p.prefactor = g["P"]
    p.nonbonded_method = force.getNonbondedMethod()
    p.cutoff = force.getCutoffDistance()
    iq, ir = names.index("q"), names.index("radius")
    parts = np.array([force.getParticleParameters(i) for i in range(force.getNumParticles())], dtype=np.float64)
    return p, parts[:, iq].copy(), parts[:, ir].copy()


# --------------------------------------------------------------------------------------
# Stand-ins for the OpenMM objects the kernel interface takes
# --------------------------------------------------------------------------------------
class System:
    def __init__(self, n: int):
        self._n = int(n)

    def getNumParticles(self) -> int:
        return self._n


class ContextImpl:
    """Holds positions (N x 3 nm), the periodic box vectors and the force accumulator, as
    ReferencePlatform::PlatformData; the global parameters of ``forces`` starting at their
    defaults (OpenMM's ContextImpl constructor); and the energy-parameter derivatives the
    kernels add into (``energy_param_derivs``)."""

    def __init__(self, positions, box_vectors=None, forces=()):
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.forces = np.zeros_like(self.positions)
        self.box_vectors = None if box_vectors is None else np.asarray(box_vectors, dtype=np.float64).reshape(3, 3)
        self.parameters = {}
        for f in forces:
            for i in range(f.getNumGlobalParameters() if hasattr(f, "getNumGlobalParameters") else 0):
                self.parameters[f.getGlobalParameterName(i)] = f.getGlobalParameterDefaultValue(i)
        self.energy_param_derivs = {}

    def getPeriodicBoxVectors(self):
        return self.box_vectors

    def getParameter(self, name: str) -> float:
        if name not in self.parameters:
            raise OpenMMException("Called getParameter() with invalid parameter name")
        return self.parameters[name]

    def setParameter(self, name: str, value: float) -> None:
        self.parameters[name] = float(value)


def _set_box(engine, method: int, context: ContextImpl) -> None:
    """The box is read from the context at every evaluation (CharmmReferenceKernels.cpp:351, :430)."""
    if method == N.CUTOFF_PERIODIC:
        if context.box_vectors is None:
            raise OpenMMException("CutoffPeriodic needs the context's periodic box vectors")
        engine.set_periodic_box(context.box_vectors)


# --------------------------------------------------------------------------------------
# Kernel interface (charmmKernels.h:49-129) implemented on HIP
# --------------------------------------------------------------------------------------
class HipCalcCharmmGBSWForceKernel:
    """CalcCharmmGBSWForceKernel on HIP: initialize / execute / copyParametersToContext."""

    @staticmethod
    def Name() -> str:
        return "CalcCharmmGBSWForce"

    def __init__(self, honor_dielectrics: bool = False, **overrides):
        self._engine = None
        self._honor = honor_dielectrics
        self._overrides = overrides

    def initialize(self, system: System, force: CharmmGBSWForce) -> None:
        if system.getNumParticles() != force.getNumParticles():  # CharmmGBSWForceImpl.cpp:47-48
            raise OpenMMException("CharmmGBSWForce must have exactly as many particles as the System it belongs to.")
        q, r, _ = force.arrays()
        p = N.default_params(N.MODEL_GBSW)
        p.nonbonded_method = force.getNonbondedMethod()
        p.cutoff = force.getCutoffDistance()
        self._method = p.nonbonded_method
        if self._honor:
            p.prefactor = -ONE_4PI_EPS0 * (1.0 / force.getSoluteDielectric() - 1.0 / force.getSolventDielectric())
        for k, v in self._overrides.items():
            setattr(p, "lambda_" if k == "lambda" else k, v)
        self._engine = N.Engine(p, q, r)

    def execute(self, context: ContextImpl, includeForces: bool = True, includeEnergy: bool = True) -> float:
        _set_box(self._engine, self._method, context)
        e, _ = self._engine.execute(context.positions, includeForces, includeEnergy, forces=context.forces)
        return e if includeEnergy else 0.0

    def copyParametersToContext(self, context: ContextImpl, force: CharmmGBSWForce) -> None:
        q, r, _ = force.arrays()
        self._engine.update_parameters(q, r)

    @property
    def engine(self) -> N.Engine:
        return self._engine


def gbmv_custom_from_force(force: CharmmGBMVForce) -> dict:
    """Arguments of ``Engine.custom`` for a CharmmGBMVForce with general expressions: its
    per-particle and global parameters, integrals (orders, Lebedev rule and radial count from
    parametersInt, GBSWIntegral.cpp:37-40), computed values (evaluated as SingleParticle, as
    the reference does, CharmmReferenceGBMV.cpp:173-176), energy terms, cutoff and method."""
    t = force.getGBIntegralType()
    if t == CharmmGBMVForce.CustomIntegral:
        raise N.GbiError(2, "CustomIntegral (type 3) is not implemented by the reference either")
    names = [force.getPerParticleParameterName(i) for i in range(force.getNumPerParticleParameters())]
    if "radius" not in names:  # GBSWIntegral.cpp:79-80
        raise OpenMMException("GBSWIntegral: the per perticle parameter 'radius' must be defined")
    integ = [tuple(force.getGBIntegralParameters(i)[:2]) for i in range(force.getNumGBIntegrals())]
    for _, pi in integ:
        if len(pi) == 0:
            raise OpenMMException("GBSWIntegral: the order of integral must be given")
        if pi[0] < 2:
            raise OpenMMException("GBSWIntegral: the order of integral must be greater or equal to 2")
    if not 1 <= len(integ) <= 2:
        raise N.GbiError(2, "general expressions take one or two integrals on the HIP platform")
    p = N.default_params(N.MODEL_GBMV, t)
    first = integ[0][1]
    if len(first) > 1:
        p.n_lebedev = _LEBEDEV_RULE_NPTS[first[1]]
    if len(first) > 2:
        p.n_radial = first[2]
    for k, (_, pi) in enumerate(integ):
        p.orders[k] = pi[0]
    p.nonbonded_method = force.getNonbondedMethod()
    p.cutoff = force.getCutoffDistance()
    values = [(n, e) for n, e, _ in force._values]
    terms = [(e, ty) for e, ty in force._energy]
    g = {force.getGlobalParameterName(i): force.getGlobalParameterDefaultValue(i)
         for i in range(force.getNumGlobalParameters())}
    pv = np.array([force.getParticleParameters(i) for i in range(force.getNumParticles())], dtype=np.float64)
    functions = []
    for name, fn in force._functions:
        if isinstance(fn, Continuous1DFunction):
            v, lo, hi = fn.getFunctionParameters()
            functions.append(N.Tabulated(name, N.TAB_CONTINUOUS_1D, v, lo, hi, fn.getPeriodic()))
        elif isinstance(fn, Discrete1DFunction):
            functions.append(N.Tabulated(name, N.TAB_DISCRETE_1D, fn.getFunctionParameters()))
        elif isinstance(fn, Discrete2DFunction):
            nx, ny, v = fn.getFunctionParameters()
            functions.append(N.Tabulated(name, N.TAB_DISCRETE_2D, v, dims=(nx, ny, 1)))
        elif isinstance(fn, Discrete3DFunction):
            nx, ny, nz, v = fn.getFunctionParameters()
            functions.append(N.Tabulated(name, N.TAB_DISCRETE_3D, v, dims=(nx, ny, nz)))
        elif isinstance(fn, Continuous2DFunction):
            nx, ny, v, x0, x1, y0, y1 = fn.getFunctionParameters()
            functions.append(N.Tabulated(name, N.TAB_CONTINUOUS_2D, v, x0, x1, fn.getPeriodic(), dims=(nx, ny, 1),
                                         yrange=(y0, y1)))
        elif isinstance(fn, Continuous3DFunction):
            nx, ny, nz, v, x0, x1, y0, y1, z0, z1 = fn.getFunctionParameters()
            functions.append(N.Tabulated(name, N.TAB_CONTINUOUS_3D, v, x0, x1, fn.getPeriodic(), dims=(nx, ny, nz),
                                         yrange=(y0, y1), zrange=(z0, z1)))
        else:
            raise OpenMMException(f"CharmmGBMVForce: tabulated function '{name}': unknown kind")
    return dict(params=p, param_names=names, param_values=pv.reshape(-1, len(names)), radius_param=names.index("radius"),
                integrals=[n for n, _ in integ], values=values, terms=terms, globals_=g,
                param_derivs=list(force._param_derivs), functions=functions)


def gbmv_exclusions(force: CharmmGBMVForce) -> np.ndarray:
    """Excluded pairs (k x 2) the canonical pair term honours: those of a ParticlePair term,
    none for ParticlePairNoExclusions (CharmmReferenceGBMV.cpp:196-199, :272-284)."""
    pair_type = force._energy[1][1] if len(force._energy) > 1 else CharmmGBMVForce.ParticlePairNoExclusions
    if pair_type != CharmmGBMVForce.ParticlePair or not force._exclusions:
        return np.zeros((0, 2), dtype=np.int32)
    return np.asarray(force._exclusions, dtype=np.int32).reshape(-1, 2)


class HipCalcCharmmGBMVForceKernel:
    """CalcCharmmGBMVForceKernel on HIP for the canonical GB energy."""

    @staticmethod
    def Name() -> str:
        return "CalcCharmmGBMVForce"

    def __init__(self, **overrides):
        self._engine = None
        self._overrides = overrides

    def initialize(self, system: System, force: CharmmGBMVForce) -> None:
        if system.getNumParticles() != force.getNumParticles():  # CharmmGBMVForceImpl.cpp:58-59
            raise OpenMMException("CharmmGBMVForce must have exactly as many particles as the System it belongs to.")
        self._globals = {force.getGlobalParameterName(i): force.getGlobalParameterDefaultValue(i)
                         for i in range(force.getNumGlobalParameters())}
        self._derivs = list(force._param_derivs)
        try:
            p, q, r = gbmv_params_from_force(force)
            self._custom = False
        except N.GbiError as err:
            if err.status != 2 or force.getGBIntegralType() == CharmmGBMVForce.CustomIntegral:
                raise
            self._custom = True  # general expressions: kernels generated from the force's expressions
        if self._custom:
            spec = gbmv_custom_from_force(force)
            p = spec.pop("params")
            for k, v in self._overrides.items():
                setattr(p, "lambda_" if k == "lambda" else k, v)
            self._method = p.nonbonded_method
            self._engine = N.Engine.custom(p, **spec)
            ex = np.asarray(force._exclusions, dtype=np.int32).reshape(-1, 2)
        else:
            for k, v in self._overrides.items():
                setattr(p, "lambda_" if k == "lambda" else k, v)
            self._method = p.nonbonded_method
            self._params = p
            self._engine = N.Engine(p, q, r)
            ex = gbmv_exclusions(force)
        if len(ex):
            self._engine.set_exclusions(ex)

    def _pull_globals(self, context: ContextImpl) -> None:
        """The global parameters are read from the Context at every execute
        (CharmmReferenceKernels.cpp:355-357); the engine gets those that changed."""
        now = {k: context.getParameter(k) for k in self._globals}
        if now == self._globals:
            return
        self._globals = now
        if self._custom:
            self._engine.set_custom_globals([now[k] for k in self._engine.global_names])
            return
        for k, field in (("alpha0", "alpha0"), ("alpha1", "alpha1"), ("eta", "eta"), ("slope", "slope"),
                         ("shift", "shift"), ("P", "prefactor")):
            setattr(self._params, field, now[k])
        self._engine.update_constants(self._params)

    def execute(self, context: ContextImpl, includeForces: bool = True, includeEnergy: bool = True) -> float:
        self._pull_globals(context)
        _set_box(self._engine, self._method, context)
        e, _ = self._engine.execute(context.positions, includeForces, includeEnergy, forces=context.forces)
        if self._custom and self._derivs:  # CharmmReferenceKernels.cpp:360-368
            for k, v in self._engine.custom_param_derivs().items():
                context.energy_param_derivs[k] = context.energy_param_derivs.get(k, 0.0) + v
        return e if includeEnergy else 0.0

    def copyParametersToContext(self, context: ContextImpl, force: CharmmGBMVForce) -> None:
        if self._custom:
            if force.getNumParticles() != self._engine.n:  # CharmmReferenceKernels.cpp:373-374
                raise OpenMMException("updateParametersInContext: The number of particles has changed")
            # per-particle parameters only (CharmmReferenceKernels.cpp:376-385); the kernels stay
            self._engine.update_custom_parameters(
                [force.getParticleParameters(i) for i in range(force.getNumParticles())])
            return
        _, q, r = gbmv_params_from_force(force)
        self._engine.update_parameters(q, r)

    @property
    def engine(self) -> N.Engine:
        return self._engine


HipCalcCustomGBIntegralForceKernel = HipCalcCharmmGBMVForceKernel
