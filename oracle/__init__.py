"""FP64 CPU oracle for the integral Generalized-Born path (TEST INFRASTRUCTURE ONLY).

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product library never does.
It wraps ``libgb_oracle.so`` (built from ``gb_oracle.c`` by ``make -C oracle``),
a C restatement of the reference's Reference platform:

* ``platforms/reference/src/CharmmReferenceGBSW.cpp`` (standalone GBSW),
* ``platforms/reference/src/{CustomGBIntegral,GBSWIntegral,GBMVIntegralTypeI,
  GBMVIntegralTypeII,CharmmReferenceGBMV}.cpp`` (CharmmGBMVForce integrals),
* ``openmmapi/include/openmm/internal/CharmmQuadrature.h`` (quadrature).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgb_oracle.so")

MODEL_GBSW_STANDALONE = 0
MODEL_GBMV_FORM = 1
INT_GBSW, INT_GBMV1, INT_GBMV2 = 0, 1, 2
POLICY_GRID, POLICY_BRUTE, POLICY_CORNERS = 0, 1, 2


class Params(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int),
        ("integral_type", ctypes.c_int),
        ("policy", ctypes.c_int),
        ("dense", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("n_radial", ctypes.c_int),
        ("n_lebedev", ctypes.c_int),
        ("r0", ctypes.c_double),
        ("r1", ctypes.c_double),
        ("grid_length", ctypes.c_double),
        ("buffer", ctypes.c_double),
        ("cap", ctypes.c_int),
        ("n_int", ctypes.c_int),
        ("orders", ctypes.c_int * 4),
        ("sw", ctypes.c_double),
        ("rscale", ctypes.c_double),
        ("roffset", ctypes.c_double),
        ("gamma0", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("P1", ctypes.c_double),
        ("P2", ctypes.c_double),
        ("S0", ctypes.c_double),
        ("vguard", ctypes.c_double),
        ("alpha0", ctypes.c_double),
        ("alpha1", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("slope", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("prefactor", ctypes.c_double),
        ("method", ctypes.c_int),
        ("cutoff", ctypes.c_double),
        ("box", ctypes.c_double * 3),
        ("n_excl", ctypes.c_int),
        ("excl", ctypes.POINTER(ctypes.c_int)),
    ]


NO_CUTOFF, CUTOFF_NON_PERIODIC, CUTOFF_PERIODIC = 0, 1, 2


class Stats(ctypes.Structure):
    _fields_ = [
        ("energy", ctypes.c_double),
        ("n_points", ctypes.c_double),
        ("n_vol_pairs", ctypes.c_double),
        ("n_vol_exact", ctypes.c_double),
        ("n_frc_pairs", ctypes.c_double),
        ("n_cells", ctypes.c_double),
        ("n_entries", ctypes.c_double),
    ]


_lib = None


def build() -> str:
    """Compile libgb_oracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_gauss_legendre.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_lebedev.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.orc_quad_table.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp]
        L.orc_defaults_gbsw_standalone.argtypes = [ctypes.POINTER(Params)]
        L.orc_defaults_gbmv.argtypes = [ctypes.POINTER(Params), ctypes.c_int]
        L.orc_compute.argtypes = [ctypes.POINTER(Params), ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, dp,
                                  ctypes.POINTER(Stats)]
        L.orc_grid_inspect.argtypes = [ctypes.POINTER(Params), ctypes.c_int, dp, dp, dp,
                                       ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_longlong, ctypes.c_longlong]
        L.orc_grid_query.argtypes = [dp, dp]
        L.orc_grid_query.restype = ctypes.c_longlong
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gauss_legendre(x0: float, x1: float, n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().orc_gauss_legendre(x0, x1, n, _dp(x), _dp(w)) == 0
    return x, w


def lebedev(npts: int):
    a = [np.zeros(npts) for _ in range(4)]
    assert lib().orc_lebedev(npts, *[_dp(v) for v in a]) == 0
    return np.stack(a[:3], axis=1), a[3]


def quad_table(n_radial=24, r0=0.05, r1=2.0, n_lebedev=38):
    out = np.zeros((n_radial * n_lebedev, 5))
    assert lib().orc_quad_table(n_radial, r0, r1, n_lebedev, _dp(out)) == 0
    return out


def gbsw_standalone_params(**kw) -> Params:
    p = Params()
    lib().orc_defaults_gbsw_standalone(ctypes.byref(p))
    return _apply(p, kw)


def gbmv_params(integral_type: int, **kw) -> Params:
    p = Params()
    lib().orc_defaults_gbmv(ctypes.byref(p), integral_type)
    return _apply(p, kw)


def _apply(p: Params, kw) -> Params:
    for k, v in kw.items():
        key = "lambda_" if k == "lambda" else k
        if key in ("orders", "box"):
            arr = getattr(p, key)
            for i, o in enumerate(v):
                arr[i] = o
        else:
            setattr(p, key, v)
    return p


_SHARED_FIELDS = ["n_radial", "n_lebedev", "r0", "r1", "grid_length", "buffer", "sw", "rscale", "roffset",
                  "gamma0", "beta", "lambda_", "P1", "P2", "S0", "vguard", "alpha0", "alpha1", "eta", "slope",
                  "shift", "prefactor"]


def params_from_abi(p, nthreads: int = 8, dense: int = 0, box=None) -> Params:
    """Oracle parameters for the model described by a gbi_params structure `p`
    (read by attribute name: model, integral_type, lookup_policy, cell_cap, orders,
    nonbonded_method, cutoff, ...); `box`: the three edges of a periodic box."""
    op = gbsw_standalone_params() if p.model == 0 else gbmv_params(p.integral_type)
    for f in _SHARED_FIELDS:
        setattr(op, f, getattr(p, f))
    op.cap = p.cell_cap
    op.orders[0], op.orders[1] = p.orders[0], p.orders[1]
    op.policy = POLICY_BRUTE if p.lookup_policy == 1 else POLICY_GRID
    op.nthreads = nthreads
    op.dense = dense
    op.method = p.nonbonded_method
    op.cutoff = p.cutoff
    if box is not None:
        for d in range(3):
            op.box[d] = float(box[d])
    return op


@dataclass
class Result:
    energy: float
    forces: np.ndarray
    born: np.ndarray
    integrals: np.ndarray
    dEdI: np.ndarray
    stats: dict


def set_exclusions(params: Params, pairs) -> Params:
    """Excluded pairs (k x 2 atom indices) of a ParticlePair term; the array is kept on the
    Params object so it outlives the call."""
    arr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    params._excl_keep = arr
    params.n_excl = arr.shape[0]
    params.excl = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if arr.shape[0] else None
    return params


def compute(params: Params, pos, charge, radius, want_stats: bool = False) -> Result:
    """Energy, forces and intermediates of one evaluation (FP64)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    charge = np.ascontiguousarray(charge, dtype=np.float64)
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    n = pos.shape[0]
    energy = np.zeros(1)
    forces = np.zeros((n, 3))
    born = np.zeros(n)
    integ = np.zeros((2, n))
    dedi = np.zeros((2, n))
    st = Stats()
    rc = lib().orc_compute(ctypes.byref(params), n, _dp(pos), _dp(charge), _dp(radius), _dp(energy),
                           _dp(forces), _dp(born), _dp(integ), _dp(dedi),
                           ctypes.byref(st) if want_stats else None)
    if rc != 0:
        raise RuntimeError(f"oracle orc_compute failed with {rc}")
    stats = {f: getattr(st, f) for f, _ in Stats._fields_} if want_stats else {}
    return Result(float(energy[0]), forces, born, integ, dedi, stats)


def energy(params: Params, pos, charge, radius) -> float:
    return compute(params, pos, charge, radius).energy


def grid_inspect(params: Params, pos, radius):
    """(geom, counts, lists) of the model's lookup grid; geom = min[3], max[3], g, n[3]."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    geom = np.zeros(10)
    ncell = ctypes.c_longlong()
    nent = ctypes.c_longlong()
    L = lib()
    rc = L.orc_grid_inspect(ctypes.byref(params), pos.shape[0], _dp(pos), _dp(radius), _dp(geom),
                            ctypes.byref(ncell), ctypes.byref(nent), None, None, 0, 0)
    if rc != 0:
        raise RuntimeError(f"oracle grid_inspect failed with {rc}")
    counts = np.zeros(ncell.value, dtype=np.int32)
    lists = np.zeros(max(nent.value, 1), dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = L.orc_grid_inspect(ctypes.byref(params), pos.shape[0], _dp(pos), _dp(radius), _dp(geom),
                            ctypes.byref(ncell), ctypes.byref(nent), counts.ctypes.data_as(ip),
                            lists.ctypes.data_as(ip), ncell.value, nent.value)
    if rc != 0:
        raise RuntimeError(f"oracle grid_inspect failed with {rc}")
    return geom, counts, lists[: nent.value]


def grid_query(geom: np.ndarray, point) -> int:
    pt = np.ascontiguousarray(point, dtype=np.float64)
    return int(lib().orc_grid_query(_dp(np.ascontiguousarray(geom)), _dp(pt)))
