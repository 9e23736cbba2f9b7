"""The general-expression oracle (oracle/custom.py) pinned on the CPU.

* the canonical GB forms written as expressions (Born value, 0.5 P q^2/B, the STILL pair
  term) reproduce the C oracle's canonical energy, forces and Born radii to 1e-12, for the
  three integral types, with a cutoff, in a periodic box and with exclusions of a
  ParticlePair term;
* forces = -grad E by central differences for non-canonical expressions (two computed
  values, a value depending on an earlier one, single-particle terms using integrals and
  values, a pair term using parameters and values of both atoms, globals, one integral);
* the tabulated functions reproduce what a spline fixes: a natural spline through samples of
  a linear function is that function (and zero outside the range), a periodic spline through
  samples of sin is sin to the sampling error, a discrete table rounds half away from zero;
* forces = -grad E by central differences with the coordinates x, y, z in computed values and
  single-particle terms and with tabulated functions in every kind of expression;
* the library rejects malformed expressions before any device work (status 1).
"""
import ctypes

import numpy as np
from scipy.interpolate import CubicSpline
import pytest

import oracle as O
from oracle import custom as C
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

G = dict(alpha0=-0.180, alpha1=1.817, eta=0.05, slope=1.0, shift=0.0, P=-138.935456 * (1 - 1 / 80))
BORN = ("B", "slope*1/(alpha0*(1/eta-I4)+alpha1*(1/(4*eta^4)-I7)^(1/4))+shift")
SELF = ("0.5*P*q^2/B", C.SINGLE)
STILL = "P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))"

# a non-canonical form: a second value built on B, terms mixing integrals, values and globals
VALUES2 = [BORN, ("C", "B*(1+0.05*tanh(q))+0.001*I7^0.25")]
TERMS2 = [("0.5*P*q^2/C+k*radius^2*I4", C.SINGLE),
          ("P*q1*q2/sqrt(r^2+C1*C2*exp(-r^2/(2*C1*C2)))*(1+0.1*step(r-0.5)*square(radius1-radius2))", C.PAIR)]


def _canonical(s, itype, pair_type=C.PAIR_NO_EXCLUSIONS, **kw):
    return C.compute(O.gbmv_params(itype, **kw), s.pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1,
                     ["I4", "I7"], [BORN], [SELF, (STILL, pair_type)], G)


@pytest.mark.parametrize("itype", [0, 1, 2])
def test_canonical_forms_as_expressions(itype):
    s = S.globule(150, 3)
    ref = O.compute(O.gbmv_params(itype), s.pos, s.charge, s.radius)
    E, F, V, _ = _canonical(s, itype)
    assert abs(E - ref.energy) < 1e-12 * abs(ref.energy)
    assert np.abs(F - ref.forces).max() < 1e-11 * np.abs(ref.forces).max()
    assert np.abs(V[0] - ref.born).max() < 1e-12


def test_canonical_forms_with_cutoff_box_and_exclusions():
    s = S.globule(160, 4)
    L = [3.6, 3.8, 3.7]
    pos = s.pos - s.pos.mean(axis=0) + np.array(L) / 2
    ex = [(0, 5), (3, 9), (10, 11), (40, 41), (7, 100)]
    kw = dict(method=O.CUTOFF_PERIODIC, cutoff=1.2, box=L)
    ref = O.compute(O.set_exclusions(O.gbmv_params(2, **kw), ex), pos, s.charge, s.radius)
    E, F, _, _ = C.compute(O.gbmv_params(2, **kw), pos, ["q", "radius"], np.stack([s.charge, s.radius], 1), 1,
                           ["I4", "I7"], [BORN], [SELF, (STILL, C.PAIR)], G, exclusions=ex)
    assert abs(E - ref.energy) < 1e-12 * abs(ref.energy)
    assert np.abs(F - ref.forces).max() < 1e-11 * np.abs(ref.forces).max()


def _general(s, pos, one_integral=False, **kw):
    names = ["q", "radius"]
    pv = np.stack([s.charge, s.radius], 1)
    if one_integral:
        values = [("B", "1/(1/radius-0.5*I4*radius)")]
        terms = [("0.5*P*q^2/B", C.SINGLE), (STILL, C.PAIR_NO_EXCLUSIONS)]
        return C.compute(O.gbmv_params(2, **kw), pos, names, pv, 1, ["I4"], values, terms, dict(G))
    return C.compute(O.gbmv_params(1, **kw), pos, names, pv, 1, ["I4", "I7"], VALUES2, TERMS2, dict(G, k=0.3))


@pytest.mark.parametrize("one_integral", [False, True])
def test_general_expressions_forces_are_the_gradient(one_integral):
    s = S.globule(90, 8)
    E, F, _, _ = _general(s, s.pos, one_integral)
    rng = np.random.default_rng(4)
    h = 1e-6
    for _ in range(5):
        i, d = rng.integers(s.n), rng.integers(3)
        pp, pm = s.pos.copy(), s.pos.copy()
        pp[i, d] += h
        pm[i, d] -= h
        fd = -(_general(s, pp, one_integral)[0] - _general(s, pm, one_integral)[0]) / (2 * h)
        assert abs(fd - F[i, d]) <= 2e-5 * max(1.0, abs(F[i, d])), (i, d, fd, F[i, d])


def _create(spec_kw, params=None):
    p = params or N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    pv = np.ascontiguousarray([[0.1, 0.2], [-0.1, 0.15]], dtype=np.float64)
    base = dict(param_names=["q", "radius"], param_values=pv, radius_param=1, integrals=["I4", "I7"],
                values=[BORN], terms=[SELF, (STILL, C.PAIR_NO_EXCLUSIONS)], globals_=G)
    base.update(spec_kw)
    try:
        N.Engine.custom(p, **base).close()
        return 0
    except N.GbiError as e:
        return e.status


def test_library_rejects_malformed_expressions():
    assert _create(dict(values=[("B", "1/(alpha0*(1/eta-I4)")])) == 1      # unbalanced
    assert _create(dict(values=[("B", "foo(I4)")])) == 1                   # unknown function
    assert _create(dict(terms=[("0.5*P*q^2/Bx", C.SINGLE)])) == 1          # unknown variable
    assert _create(dict(terms=[("P*q1*q2/r*I41", C.PAIR)])) == 1           # integrals are not pair variables
    assert _create(dict(values=[("B", "max(I4)")])) == 1                   # arity
    assert _create(dict(values=[("r", "I4")])) == 1                        # reserved name
    # well formed: only the device is missing here (or the context is created on a GPU)
    assert _create({}) in (0, 8)


# expressions using the atom's coordinates and tabulated functions
SCALE = C.Tabulated("scale", 0, 1.0 + 0.2 * np.sin(0.3 * np.arange(31)), 0.0, 3.0)
WAVE = C.Tabulated("wave", 0, np.cos(np.linspace(0.0, 2 * np.pi, 41)), 0.0, 2 * np.pi, periodic=True)
W = C.Tabulated("w", 1, [0.9, 1.1])
# values[x + xsize (y + ysize z)] at the rounded arguments; the synthetic radii are 0.1, 0.17,
# 0.185 and 0.2 nm, so 10*radius rounds to 1 or 2 and 20*radius-2 to 0, 1 or 2
LUT = C.Tabulated("lut", 2, [0.7, 0.8, 0.95, 1.05, 1.1, 1.2], dims=(2, 3, 1))
TRI3 = C.Tabulated("tri3", 3, 1.0 + 0.03 * np.arange(12), dims=(2, 2, 3))


def _grid(n, lo, hi):
    return np.linspace(lo, hi, n)


# Continuous2D, periodic with period 2 on both axes (first and last values of every line equal),
# and Continuous3D, natural on [-1.5, 1.5]^3 (values[x + xsize (y + ysize z)])
_SX, _SY = np.meshgrid(_grid(9, 0.0, 2.0), _grid(7, 0.0, 2.0), indexing="ij")
_SURF = np.sin(np.pi * _SX) * np.cos(np.pi * _SY) + 0.5 * np.cos(np.pi * _SX)
_SURF[-1, :], _SURF[:, -1] = _SURF[0, :], _SURF[:, 0]
SURF = C.Tabulated("surf", 4, _SURF.T.ravel(), 0.0, 2.0, periodic=True, dims=(9, 7, 1), yrange=(0.0, 2.0))
_VX, _VY, _VZ = np.meshgrid(_grid(6, -1.5, 1.5), _grid(5, -1.5, 1.5), _grid(7, -1.5, 1.5), indexing="ij")
_VOL = np.exp(-0.5 * (_VX ** 2 + 0.7 * _VY ** 2)) * np.cos(_VZ) + 0.1 * _VX * _VZ
VOL3 = C.Tabulated("vol3", 5, _VOL.transpose(2, 1, 0).ravel(), -1.5, 1.5, dims=(6, 5, 7), yrange=(-1.5, 1.5),
                   zrange=(-1.5, 1.5))
TABLES = [SCALE, WAVE, W, LUT, TRI3, SURF, VOL3]
TAB_VALUES = [("B", "scale(10*radius)*lut(step(q), 10*radius)*(1+0.002*(x+y-z))/(1/radius-0.5*I4*radius)")]
TAB_TERMS = [("0.5*P*q^2/B*w(step(q))*tri3(step(q), step(radius-0.18), 20*radius-2)"
              "+0.01*q^2*(x^2+y^2+z^2)+0.05*q^2*wave(3*x)+0.2*q^2*surf(x, y)+0.3*q^2*vol3(x, z, y)", C.SINGLE),
             ("P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))*(1+0.05*wave(r))*(1+0.02*surf(r, 10*B1))",
              C.PAIR_NO_EXCLUSIONS)]


def test_tabulated_functions_pinned_by_closed_forms():
    lin = C.Tabulated("lin", 0, 0.5 + 2.0 * np.linspace(-1.0, 2.0, 7), -1.0, 2.0)
    f, df = lin.numeric()
    t = np.array([-1.0, -0.37, 0.0, 1.234, 2.0])
    assert np.allclose(f(t), 0.5 + 2.0 * t, atol=1e-13) and np.allclose(df(t), 2.0, atol=1e-12)
    assert f(np.array([-1.01, 2.01])).tolist() == [0.0, 0.0]  # zero outside [min, max]
    per = C.Tabulated("per", 0, np.sin(np.linspace(0.0, 2 * np.pi, 201)), 0.0, 2 * np.pi, periodic=True)
    f, df = per.numeric()
    t = np.array([-7.0, -0.3, 0.1, 3.0, 9.9])  # wrapped by the period
    assert np.allclose(f(t), np.sin(t), atol=2e-6) and np.allclose(df(t), np.cos(t), atol=2e-4)
    f, df = C.Tabulated("d", 1, [10.0, 20.0, 30.0]).numeric()
    assert f(np.array([0.49, 0.5, 1.5, 2.4])).tolist() == [10.0, 20.0, 30.0, 30.0]
    with pytest.raises(ValueError):
        f(np.array([2.5]))
    # 2D and 3D tables: x fastest, then y, then z (a row-major array read as [z][y][x])
    grid = np.arange(24.0).reshape(4, 3, 2)  # [z][y][x] with xsize 2, ysize 3, zsize 4
    f3, _ = C.Tabulated("c", 3, grid.ravel(), dims=(2, 3, 4)).numeric()
    x, y, z = np.array([0.4, 1.0, 0.6]), np.array([2.2, 0.5, 1.49]), np.array([3.0, 0.0, 2.5])
    assert f3(x, y, z).tolist() == [grid[3, 2, 0], grid[0, 1, 1], grid[3, 1, 1]]
    f2, _ = C.Tabulated("s", 2, grid[0].ravel(), dims=(2, 3, 1)).numeric()
    assert f2(x, y).tolist() == [grid[0, 2, 0], grid[0, 1, 1], grid[0, 1, 1]]
    with pytest.raises(ValueError):
        f2(np.array([0.0]), np.array([2.6]))


def test_continuous_2d_3d_tables_pinned_by_closed_forms():
    """Continuous2D/3DFunction: bilinear and trilinear data are reproduced exactly with their
    gradients (natural splines reproduce linear lines, and a Hermite patch with exact knot
    derivatives reproduces polynomials of degree <= 3 per axis); along a grid line the patch is
    the 1-D spline of that line; zero outside the ranges; a periodic table of sin x cos y
    converges to it."""
    xs, ys, zs = np.linspace(-1.0, 2.0, 7), np.linspace(0.5, 1.5, 5), np.linspace(-0.5, 0.5, 4)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    f, (fx, fy) = C.Tabulated("b", 4, (0.3 + 1.2 * X - 0.7 * Y + 0.4 * X * Y).T.ravel(), -1.0, 2.0, dims=(7, 5, 1),
                              yrange=(0.5, 1.5))._patch()
    px, py, pz = np.array([-1, -0.37, 0.3, 1.99, 2.0]), np.array([0.5, 0.77, 1.2, 1.5, 0.9]), np.array(
        [-0.5, 0.1, 0.33, 0.5, -0.2])
    assert np.allclose(f(px, py), 0.3 + 1.2 * px - 0.7 * py + 0.4 * px * py, atol=1e-13)
    assert np.allclose(fx(px, py), 1.2 + 0.4 * py, atol=1e-12) and np.allclose(fy(px, py), -0.7 + 0.4 * px, atol=1e-12)
    W = np.sin(1.3 * X) * np.cos(2 * Y) + 0.1 * X ** 3
    f2, (f2x, _) = C.Tabulated("w", 4, W.T.ravel(), -1.0, 2.0, dims=(7, 5, 1), yrange=(0.5, 1.5))._patch()
    cs = CubicSpline(xs, W[:, 2], bc_type="natural")
    q = np.linspace(-1.0, 2.0, 13)
    assert np.allclose(f2(q, np.full_like(q, ys[2])), cs(q), atol=1e-13)
    assert np.allclose(f2x(q, np.full_like(q, ys[2])), cs(q, 1), atol=1e-12)
    assert f2(np.array([2.01]), np.array([1.0]))[0] == 0.0 and f2(np.array([0.0]), np.array([0.49]))[0] == 0.0
    X3, Y3, Z3 = np.meshgrid(xs, ys, zs, indexing="ij")
    V3 = 1 + X3 - 2 * Y3 + 0.5 * Z3 + 0.3 * X3 * Y3 * Z3 - 0.2 * X3 * Z3
    f3, (g1, g2, g3) = C.Tabulated("c", 5, V3.transpose(2, 1, 0).ravel(), -1.0, 2.0, dims=(7, 5, 4),
                                   yrange=(0.5, 1.5), zrange=(-0.5, 0.5))._patch()
    assert np.allclose(f3(px, py, pz), 1 + px - 2 * py + 0.5 * pz + 0.3 * px * py * pz - 0.2 * px * pz, atol=1e-13)
    assert np.allclose(g1(px, py, pz), 1 + 0.3 * py * pz - 0.2 * pz, atol=1e-12)
    assert np.allclose(g2(px, py, pz), -2 + 0.3 * px * pz, atol=1e-12)
    assert np.allclose(g3(px, py, pz), 0.5 + 0.3 * px * py - 0.2 * px, atol=1e-12)
    n = 41
    g = np.linspace(0.0, 2 * np.pi, n)
    PX, PY = np.meshgrid(g, g, indexing="ij")
    P = np.sin(PX) * np.cos(PY)
    P[-1, :], P[:, -1] = P[0, :], P[:, 0]
    fp, (fpx, fpy) = C.Tabulated("p", 4, P.T.ravel(), 0.0, 2 * np.pi, periodic=True, dims=(n, n, 1),
                                 yrange=(0.0, 2 * np.pi))._patch()
    a, b = np.array([-7.0, 0.3, 3.0, 9.9]), np.array([1.0, -2.0, 7.5, 0.2])  # wrapped by the period
    assert np.allclose(fp(a, b), np.sin(a) * np.cos(b), atol=2e-6)
    assert np.allclose(fpx(a, b), np.cos(a) * np.cos(b), atol=2e-5)
    assert np.allclose(fpy(a, b), -np.sin(a) * np.sin(b), atol=2e-5)


def _tabulated(s, pos):
    pv = np.stack([s.charge, s.radius], 1)
    p = O.gbmv_params(2)
    p.orders[0] = 4
    return C.compute(p, pos, ["q", "radius"], pv, 1, ["I4"], TAB_VALUES, TAB_TERMS, dict(G), functions=TABLES)


def test_coordinates_and_tabulated_functions_forces_are_the_gradient():
    s = S.globule(80, 9)
    E, F, V, _ = _tabulated(s, s.pos)
    rng = np.random.default_rng(6)
    h = 1e-6
    for _ in range(6):
        i, d = rng.integers(s.n), rng.integers(3)
        pp, pm = s.pos.copy(), s.pos.copy()
        pp[i, d] += h
        pm[i, d] -= h
        fd = -(_tabulated(s, pp)[0] - _tabulated(s, pm)[0]) / (2 * h)
        assert abs(fd - F[i, d]) <= 2e-5 * max(1.0, abs(F[i, d])), (i, d, fd, F[i, d])


def test_library_rejects_bad_tables_and_coordinates():
    tab = N.Tabulated("scale", N.TAB_CONTINUOUS_1D, [1.0, 1.2, 1.1], 0.0, 3.0)
    val = [("B", "scale(10*radius)/(1/radius-0.5*I4*radius)")]
    assert _create(dict(values=val, functions=[N.Tabulated("scale", 6, [1.0, 2.0])])) == 2  # unknown kind
    c2 = dict(values=[("B", "surf(1, 0.5)/(1/radius-0.5*I4*radius)")])
    assert _create(dict(c2, functions=[N.Tabulated("surf", 4, [1.0] * 6, 0, 1, dims=(3, 3, 1))])) == 1  # sizes
    assert _create(dict(c2, functions=[N.Tabulated("surf", 4, [1.0] * 3, 0, 1, dims=(1, 3, 1))])) == 1  # >= 2 per axis
    assert _create(dict(c2, functions=[N.Tabulated("surf", 4, [1.0] * 4, 0, 1, dims=(2, 2, 1),
                                                   yrange=(1.0, 1.0))])) == 1                        # ymin < ymax
    assert _create(dict(c2, functions=[N.Tabulated("surf", 4, [1.0, 2.0, 1.5, 2.5], 0, 1, True, dims=(2, 2, 1))])) == 1
    assert _create(dict(c2, functions=[N.Tabulated("surf", 4, [1.0, 2.0, 1.5, 2.5], 0, 1, dims=(2, 2, 1))])) in (0, 8)
    assert _create(dict(values=[("B", "vol(1, 0.5, 0.5)/(1/radius-0.5*I4*radius)")],
                        functions=[N.Tabulated("vol", 5, np.arange(8.0), 0, 1, dims=(2, 2, 2))])) in (0, 8)
    d2 = N.Tabulated("lut", N.TAB_DISCRETE_2D, [1.0, 2.0, 3.0, 4.0], dims=(2, 2, 1))
    d3 = N.Tabulated("lut", N.TAB_DISCRETE_3D, [1.0] * 8, dims=(2, 2, 2))
    val2 = [("B", "lut(1, 0)/(1/radius-0.5*I4*radius)")]
    assert _create(dict(values=val2, functions=[N.Tabulated("lut", 2, [1.0, 2.0, 3.0], dims=(2, 2, 1))])) == 1
    assert _create(dict(values=val2, functions=[N.Tabulated("lut", 2, [1.0, 2.0], dims=(0, 2, 1))])) == 1
    assert _create(dict(values=[("B", "lut(1)/(1/radius-0.5*I4*radius)")], functions=[d2])) == 1   # two arguments
    assert _create(dict(values=val2, functions=[d3])) == 1                                        # three arguments
    assert _create(dict(values=val, functions=[N.Tabulated("scale", 0, [1.0, 2.0, 1.5], 0, 3, True)])) == 1
    assert _create(dict(values=val, functions=[N.Tabulated("scale", 0, [1.0, 2.0], 3.0, 3.0)])) == 1  # min < max
    assert _create(dict(values=[("B", "scale(I4, I7)")], functions=[tab])) == 1   # one argument
    assert _create(dict(values=val, functions=[tab, tab])) == 1                    # defined twice
    assert _create(dict(values=[("B", "cube(1, 0, 0)/(1/radius-0.5*I4*radius)")],
                        functions=[N.Tabulated("cube", 3, [1.0] * 8, dims=(2, 2, 2))])) == 1  # a built-in's name
    assert _create(dict(terms=[SELF, ("P*q1*q2/r*x", C.PAIR)])) == 1               # x is not a pair variable
    assert _create(dict(param_names=["q", "x"])) == 1                              # reserved coordinate
    # well formed: only the device is missing here (or the context is created on a GPU)
    assert _create(dict(values=[("B", "scale(10*radius)*(1+0.01*x)/(1/radius-0.5*I4*radius)")],
                        terms=[("0.5*P*q^2/B+0.01*z^2", C.SINGLE), (STILL, C.PAIR_NO_EXCLUSIONS)],
                        functions=[tab])) in (0, 8)
    assert _create(dict(values=val2, functions=[d2])) in (0, 8)
    assert _create(dict(values=[("B", "lut(1, 0, step(q))/(1/radius-0.5*I4*radius)")], functions=[d3])) in (0, 8)
