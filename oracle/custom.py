"""FP64 restatement of CharmmGBMVForce with general expressions (test infrastructure).

TEST INFRASTRUCTURE ONLY: the checker of the HIP library's general-expression path.  It
follows the Reference platform's CharmmReferenceGBMV::calculateIxn
(CharmmReferenceGBMV.cpp:150-204): integrals and their gradients (the C oracle's
orc_integral_chain, CustomGBIntegral / GBSWIntegral / GBMVIntegralType{I,II}), computed
values as SingleParticle (:173-176, :206-227), single-particle energy terms (:229-266), pair
terms over i < j with the cutoff rule r >= c skipped and exclusions of ParticlePair terms
(:268-348), and the chain rule dE/dV -> dE/dI -> forces (:350-446), including the atom's own
coordinates x, y, z in computed values and single-particle terms (the explicit gradients and
dE/dV_j dV_j/dx, :255-259, :413-435).  The expressions are parsed by sympy and differentiated
symbolically, independently of the library's dual-number code generation.  Tabulated
functions (CharmmReferenceKernels.cpp:187-191 hands them to OpenMM's tabulated-function
classes, which are not in the reference tree) follow OpenMM's documented behaviour:
Continuous1DFunction a natural (or periodic) cubic spline through values equally spaced on
[min, max], zero outside it when not periodic, evaluated with scipy's CubicSpline;
Discrete1D/2D/3DFunction the value at the rounded arguments, round half away from zero.
"""
from __future__ import annotations

import ctypes
import itertools
import re

import numpy as np
import sympy as sp
from scipy.interpolate import CubicSpline

from . import CUTOFF_PERIODIC, NO_CUTOFF, Params, _dp, lib

SINGLE, PAIR, PAIR_NO_EXCLUSIONS = 0, 1, 2


def _functions():
    x, y, z = sp.symbols("x y z")
    return {
        "sqrt": sp.sqrt, "exp": sp.exp, "log": sp.log, "sin": sp.sin, "cos": sp.cos, "tan": sp.tan,
        "asin": sp.asin, "acos": sp.acos, "atan": sp.atan, "sinh": sp.sinh, "cosh": sp.cosh, "tanh": sp.tanh,
        "erf": sp.erf, "erfc": sp.erfc, "abs": sp.Abs,
        "step": lambda a: sp.Piecewise((1, a >= 0), (0, True)),
        "delta": lambda a: sp.Piecewise((1, sp.Eq(a, 0)), (0, True)),
        "square": lambda a: a ** 2, "cube": lambda a: a ** 3, "recip": lambda a: 1 / a,
        "min": sp.Min, "max": sp.Max,
        "select": lambda a, b, c: sp.Piecewise((c, sp.Eq(a, 0)), (b, True)),
    }


class Tabulated:
    """A tabulated function: kind 0 Continuous1DFunction (values equally spaced on [min, max],
    periodic or natural spline), kinds 1, 2, 3 Discrete1D/2D/3DFunction (dims: xsize, ysize,
    zsize; values[x + xsize (y + ysize z)] at the rounded arguments), kinds 4, 5
    Continuous2D/3DFunction (values[x + xsize (y + ysize z)] equally spaced on [min, max] x
    yrange (x zrange)).

    Continuous2D/3D follow the published method of OpenMM's classes (third party, absent here):
    at every knot the derivatives d/dx, d/dy (d/dz) come from cubic splines through the table's
    lines along each axis (natural, or periodic), the mixed ones from splines of those along the
    other axes, and between knots the function is the bicubic (tricubic) Hermite patch that
    matches them, the unique one of its degree; zero outside the ranges unless periodic."""

    def __init__(self, name, kind, values, min=0.0, max=1.0, periodic=False, dims=(0, 1, 1), yrange=(0.0, 1.0),
                 zrange=(0.0, 1.0)):
        self.name, self.kind = name, int(kind)
        self.values = np.asarray(values, dtype=np.float64)
        self.min, self.max, self.periodic = float(min), float(max), bool(periodic)
        self.dims = tuple(int(d) for d in dims)
        self.yrange = (float(yrange[0]), float(yrange[1]))
        self.zrange = (float(zrange[0]), float(zrange[1]))

    def _patch(self):
        """(value, [d/dx, d/dy(, d/dz)]) callables of a Continuous2D/3DFunction."""
        nd = 2 if self.kind == 4 else 3
        n = self.dims[:nd]
        lo = [self.min, self.yrange[0], self.zrange[0]][:nd]
        hi = [self.max, self.yrange[1], self.zrange[1]][:nd]
        h = [(hi[a] - lo[a]) / (n[a] - 1) for a in range(nd)]
        V = self.values.reshape(n[::-1]).transpose()  # V[x, y(, z)]
        bc = "periodic" if self.periodic else "natural"

        def along(A, a):
            xs = lo[a] + np.arange(n[a]) * h[a]
            return CubicSpline(xs, A, axis=a, bc_type=bc)(xs, 1)

        tables = {(0,) * nd: V}
        for p in sorted(itertools.product((0, 1), repeat=nd), key=sum):
            if sum(p) == 0:
                continue
            a = max(i for i in range(nd) if p[i])  # differentiate the table without axis a along a
            q = tuple(p[i] if i != a else 0 for i in range(nd))
            tables[p] = along(tables[q], a)

        def hermite(t):
            t2, t3 = t * t, t * t * t
            return ([2 * t3 - 3 * t2 + 1, 3 * t2 - 2 * t3, t3 - 2 * t2 + t, t3 - t2],
                    [6 * t2 - 6 * t, 6 * t - 6 * t2, 3 * t2 - 4 * t + 1, 3 * t2 - 2 * t])

        def evaluate(which, *args):
            args = np.broadcast_arrays(*[np.asarray(x, dtype=np.float64) for x in args])
            inside = np.ones(args[0].shape, dtype=bool)
            ks, ws = [], []
            for a in range(nd):
                v = args[a]
                if self.periodic:
                    v = v - (hi[a] - lo[a]) * np.floor((v - lo[a]) / (hi[a] - lo[a]))
                else:
                    inside &= (v >= lo[a]) & (v <= hi[a])
                    v = np.clip(v, lo[a], hi[a])
                k = np.clip(np.floor((v - lo[a]) / h[a]).astype(np.int64), 0, n[a] - 2)
                ks.append(k)
                ws.append(hermite((v - (lo[a] + k * h[a])) / h[a]))
            out = np.zeros(args[0].shape)
            for corner in itertools.product((0, 1), repeat=nd):
                idx = tuple(ks[a] + corner[a] for a in range(nd))
                for p, T in tables.items():
                    term = T[idx] * np.prod([h[a] for a in range(nd) if p[a]])
                    for a in range(nd):
                        term = term * ws[a][1 if a == which else 0][2 * p[a] + corner[a]]
                    out = out + term
            if which is not None:
                out = out / h[which]
            return np.where(inside, out, 0.0)

        return (lambda *a: evaluate(None, *a)), [(lambda *a, w=w: evaluate(w, *a)) for w in range(nd)]

    def numeric(self):
        """(f, f') as numpy callables (f' of the one-argument kinds; the discrete tables are
        piecewise constant)."""
        if self.kind in (2, 3):
            vals, (nx, ny, nz) = self.values, self.dims

            def idx(t, n):
                k = (np.sign(t) * np.floor(np.abs(t) + 0.5)).astype(np.int64)
                if np.any((k < 0) | (k >= n)):
                    raise ValueError(f"{self.name}: argument out of range")
                return k

            def f(*a):
                a = [np.asarray(x, dtype=np.float64) for x in a]
                k = idx(a[0], nx) + nx * idx(a[1], ny)
                if self.kind == 3:
                    k = k + nx * ny * idx(a[2], nz)
                return vals[k]

            return f, None
        if self.kind == 1:
            vals = self.values

            def f(t):
                t = np.asarray(t, dtype=np.float64)
                k = (np.sign(t) * np.floor(np.abs(t) + 0.5)).astype(np.int64)
                if np.any((k < 0) | (k >= len(vals))):
                    raise ValueError(f"{self.name}: argument out of range")
                return vals[k]

            return f, lambda t: np.zeros_like(np.asarray(t, dtype=np.float64))
        xs = np.linspace(self.min, self.max, len(self.values))
        cs = CubicSpline(xs, self.values, bc_type="periodic" if self.periodic else "natural")
        lo, hi = self.min, self.max

        def make(nu):
            def g(t):
                t = np.asarray(t, dtype=np.float64)
                if self.periodic:
                    return cs(t - (hi - lo) * np.floor((t - lo) / (hi - lo)), nu)
                inside = (t >= lo) & (t <= hi)
                return np.where(inside, cs(np.clip(t, lo, hi), nu), 0.0)
            return g

        return make(0), make(1)

    def sympy_function(self):
        if self.kind in (4, 5):
            f, parts = self._patch()
            pcls = []
            for w, pf in enumerate(parts):
                c = sp.Function(f"{self.name}__d{w}")
                c._imp_ = staticmethod(pf)
                pcls.append(c)

            def fdiffn(obj, argindex=1):
                return pcls[argindex - 1](*obj.args)

            return type(self.name, (sp.Function,), {"fdiff": fdiffn, "_imp_": staticmethod(f)})
        f, df = self.numeric()
        if df is None:  # piecewise constant in every argument
            def fdiff0(obj, argindex=1):
                return sp.Integer(0)

            return type(self.name, (sp.Function,), {"fdiff": fdiff0, "_imp_": staticmethod(f)})
        dcls = sp.Function(self.name + "__d")
        dcls._imp_ = staticmethod(df)

        def fdiff(obj, argindex=1):
            return dcls(obj.args[0])

        return type(self.name, (sp.Function,), {"fdiff": fdiff, "_imp_": staticmethod(f)})


def parse(expr: str, names, functions=()) -> sp.Expr:
    """Lepton syntax -> sympy ('^' is the power operator)."""
    local = dict(_functions())
    for t in functions:
        local[t.name] = t.sympy_function()
    for n in names:
        local[n] = sp.Symbol(n)
    text = re.sub(r"\^", "**", expr)
    return sp.sympify(text, locals=local, rational=False)


def _lam(args, expr):
    return sp.lambdify(args, expr, modules=["numpy", {"Heaviside": lambda a, h=0.5: np.heaviside(a, h)}])


def integral_chain(params: Params, pos, radius, dEdI=None):
    """(integrals [2 x N], forces [N x 3] or None) from the C oracle."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    n = pos.shape[0]
    integ = np.zeros((2, n))
    forces = np.zeros((n, 3)) if dEdI is not None else None
    d = None if dEdI is None else np.ascontiguousarray(dEdI, dtype=np.float64)
    L = lib()
    L.orc_integral_chain.argtypes = [ctypes.POINTER(Params), ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 5
    rc = L.orc_integral_chain(ctypes.byref(params), n, _dp(pos), _dp(radius), _dp(d) if d is not None else None,
                              _dp(integ), _dp(forces) if forces is not None else None)
    if rc != 0:
        raise RuntimeError(f"oracle orc_integral_chain failed with {rc}")
    return integ, forces


def compute(params: Params, pos, param_names, param_values, radius_param, integrals, values, terms, globals_=None,
            exclusions=None, functions=()):
    """Energy, forces and computed values of a general CharmmGBMVForce.

    params: oracle Params (integral type, quadrature, grid, orders, method, cutoff, box);
    integrals: names (1 or 2); values: [(name, expr)]; terms: [(expr, type)]; functions:
    Tabulated functions."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    P = np.asarray(param_values, dtype=np.float64).reshape(n, len(param_names))
    G = dict(globals_ or {})
    radius = P[:, radius_param].copy()
    if len(integrals) == 1:
        params.orders[1] = params.orders[0]
    I, _ = integral_chain(params, pos, radius)
    nI, nV = len(integrals), len(values)
    gsub = {sp.Symbol(k): v for k, v in G.items()}
    psym = [sp.Symbol(k) for k in param_names]
    isym = [sp.Symbol(k) for k in integrals]
    vsym = [sp.Symbol(k) for k, _ in values]
    xyz = list(sp.symbols("x y z"))
    fn = list(functions)

    # computed values, in order, and their total derivatives dV_k/dI_m and dV_k/dx_d
    V = np.zeros((nV, n))
    dVdI = np.zeros((nV, 2, n))
    dVdX = np.zeros((nV, 3, n))
    args = xyz + psym + isym + vsym
    cols = [pos[:, 0], pos[:, 1], pos[:, 2]] + [P[:, k] for k in range(len(psym))] + [I[m] for m in range(nI)]
    for k, (name, expr) in enumerate(values):
        e = parse(expr, ["x", "y", "z"] + list(param_names) + list(G) + list(integrals) + [v for v, _ in values[:k]],
                  fn).subs(gsub)
        a = args[: 3 + len(psym) + nI + k]
        vals_in = cols + [V[j] for j in range(k)]
        V[k] = np.broadcast_to(_lam(a, e)(*vals_in), (n,))
        for m in range(nI):
            tot = np.broadcast_to(_lam(a, sp.diff(e, isym[m]))(*vals_in), (n,)).astype(float)
            for j in range(k):
                tot = tot + np.broadcast_to(_lam(a, sp.diff(e, vsym[j]))(*vals_in), (n,)) * dVdI[j, m]
            dVdI[k, m] = tot
        for d in range(3):
            tot = np.broadcast_to(_lam(a, sp.diff(e, xyz[d]))(*vals_in), (n,)).astype(float)
            for j in range(k):
                tot = tot + np.broadcast_to(_lam(a, sp.diff(e, vsym[j]))(*vals_in), (n,)) * dVdX[j, d]
            dVdX[k, d] = tot

    # single-particle terms: energy, total dE/dI, explicit coordinate gradients (the force's
    # first part, CharmmReferenceGBMV.cpp:257-259) and dE/dV_j (chained through dV_j/dx below)
    E = 0.0
    dEdI = np.zeros((2, n))
    forces = np.zeros((n, 3))
    dEdV = np.zeros((nV, n))
    sp_args = xyz + psym + isym + vsym
    sp_in = cols + [V[j] for j in range(nV)]
    for expr, ttype in terms:
        if ttype != SINGLE:
            continue
        e = parse(expr, ["x", "y", "z"] + list(param_names) + list(G) + list(integrals) + [v for v, _ in values],
                  fn).subs(gsub)
        E += float(np.sum(np.broadcast_to(_lam(sp_args, e)(*sp_in), (n,))))
        for m in range(nI):
            tot = np.broadcast_to(_lam(sp_args, sp.diff(e, isym[m]))(*sp_in), (n,)).astype(float)
            for j in range(nV):
                tot = tot + np.broadcast_to(_lam(sp_args, sp.diff(e, vsym[j]))(*sp_in), (n,)) * dVdI[j, m]
            dEdI[m] += tot
        for d in range(3):
            forces[:, d] -= np.broadcast_to(_lam(sp_args, sp.diff(e, xyz[d]))(*sp_in), (n,))
        for j in range(nV):
            dEdV[j] += np.broadcast_to(_lam(sp_args, sp.diff(e, vsym[j]))(*sp_in), (n,))

    # pair terms over i < j (atom1 = i)
    iu, ju = np.triu_indices(n, 1)
    d = pos[ju] - pos[iu]  # x_2 - x_1
    if params.method == CUTOFF_PERIODIC:
        box = np.array([params.box[0], params.box[1], params.box[2]])
        d = d - box * np.floor(d / box + 0.5)
    r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    keep = np.ones(len(iu), dtype=bool) if params.method == NO_CUTOFF else r2 < params.cutoff ** 2
    excluded = np.zeros(len(iu), dtype=bool)
    if exclusions is not None and len(exclusions):
        ex = np.sort(np.asarray(exclusions, dtype=np.int64).reshape(-1, 2), axis=1)
        key = set((ex[:, 0] * n + ex[:, 1]).tolist())
        excluded = np.fromiter(((a * n + b) in key for a, b in zip(iu, ju)), dtype=bool, count=len(iu))
    r = np.sqrt(r2)
    dEdV_pair = np.zeros((nV, n))
    rs = sp.Symbol("r")
    p1 = [sp.Symbol(k + "1") for k in param_names]
    p2 = [sp.Symbol(k + "2") for k in param_names]
    v1 = [sp.Symbol(k + "1") for k, _ in values]
    v2 = [sp.Symbol(k + "2") for k, _ in values]
    pair_args = [rs] + p1 + p2 + v1 + v2
    pair_names = ["r"] + [k + s for s in "12" for k in param_names] + [k + s for s in "12" for k, _ in values] + list(G)
    for expr, ttype in terms:
        if ttype == SINGLE:
            continue
        e = parse(expr, pair_names, fn).subs(gsub)
        m = keep & (~excluded if ttype == PAIR else True)
        a, b = iu[m], ju[m]
        ins = [r[m]] + [P[a, k] for k in range(len(p1))] + [P[b, k] for k in range(len(p2))] + \
              [V[k, a] for k in range(nV)] + [V[k, b] for k in range(nV)]
        cnt = len(a)
        E += float(np.sum(np.broadcast_to(_lam(pair_args, e)(*ins), (cnt,))))
        dEdr = np.broadcast_to(_lam(pair_args, sp.diff(e, rs))(*ins), (cnt,))
        g = (dEdr / r[m])[:, None] * d[m]  # dE/dr (x2 - x1)/r: force on atom1, minus on atom2
        np.add.at(forces, a, g)
        np.add.at(forces, b, -g)
        for k in range(nV):
            np.add.at(dEdV_pair[k], a, np.broadcast_to(_lam(pair_args, sp.diff(e, v1[k]))(*ins), (cnt,)))
            np.add.at(dEdV_pair[k], b, np.broadcast_to(_lam(pair_args, sp.diff(e, v2[k]))(*ins), (cnt,)))

    # dE/dV_k: the pair terms' part feeds dE/dI (the single-particle part is already in the
    # total dE/dI above); both feed the coordinate chain rule dE/dV_k dV_k/dx
    for m in range(nI):
        for k in range(nV):
            dEdI[m] += dEdV_pair[k] * dVdI[k, m]
    for k in range(nV):
        forces -= (dEdV[k] + dEdV_pair[k])[:, None] * dVdX[k].T
    _, fchain = integral_chain(params, pos, radius, dEdI)
    forces += fchain
    return float(E), forces, V, I
