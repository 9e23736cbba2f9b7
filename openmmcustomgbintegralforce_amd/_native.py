"""ctypes binding of the C ABI in ``include/gbi.h`` (``lib/libgbi.so``).

There is no fallback: if the library is missing or cannot be loaded, every entry
point raises.  Build it with ``python -m openmmcustomgbintegralforce_amd.build``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libgbi.so")

GBI_OK = 0
STATUS_NAMES = {
    0: "GBI_OK", 1: "GBI_ERR_INVALID_ARGUMENT", 2: "GBI_ERR_UNSUPPORTED", 3: "GBI_ERR_PARTICLE_COUNT",
    4: "GBI_ERR_HIP", 5: "GBI_ERR_OUT_OF_MEMORY", 6: "GBI_ERR_GRID_TOO_LARGE", 7: "GBI_ERR_NONFINITE",
    8: "GBI_ERR_NO_DEVICE", 9: "GBI_ERR_CAPACITY",
}
MODEL_GBSW, MODEL_GBMV = 0, 1
INTEGRAL_GBSW, INTEGRAL_GBMV1, INTEGRAL_GBMV2, INTEGRAL_CUSTOM = 0, 1, 2, 3
LOOKUP_REFERENCE, LOOKUP_BRUTE = 0, 1
NO_CUTOFF, CUTOFF_NON_PERIODIC, CUTOFF_PERIODIC = 0, 1, 2
FORCES, ENERGY, ADJOINT = 1, 2, 4

# dE/dtheta order of gbi_get_adjoint (gbi_theta in gbi.h)
THETA = ["prefactor", "alpha0", "alpha1", "eta", "slope", "shift", "sw", "rscale", "roffset", "gamma0",
         "beta", "lambda", "P1", "P2", "S0"]

# Every symbol the header declares (checked by the CPU tests).
EXPORTS = [
    "gbi_params_default", "gbi_create", "gbi_batch_create", "gbi_get_layout", "gbi_execute",
    "gbi_update_parameters", "gbi_set_positions", "gbi_compute", "gbi_synchronize", "gbi_get_energy",
    "gbi_get_forces", "gbi_get_born_radii", "gbi_get_integrals", "gbi_get_stats", "gbi_get_device_view",
    "gbi_set_stream", "gbi_get_adjoint", "gbi_execute_adjoint", "gbi_set_graph", "gbi_autotune",
    "gbi_set_tuning", "gbi_get_tuning", "gbi_destroy", "gbi_last_error", "gbi_abi_version", "gbi_quadrature",
    "gbi_profile_enable", "gbi_profile_read", "gbi_profile_pass_name", "gbi_set_periodic_box",
    "gbi_set_exclusions", "gbi_create_custom", "gbi_set_custom_globals", "gbi_get_custom_values",
    "gbi_get_custom_param_derivs", "gbi_update_custom_parameters", "gbi_update_constants", "gbi_execute_noalloc",
    "gbi_set_positions_posq", "gbi_add_forces_fixed",
]
PROFILE_PASSES = ["grid", "bin", "volume", "integrals", "pairs", "chain", "force", "owner"]


class GbiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32),
        ("integral_type", ctypes.c_int32),
        ("lookup_policy", ctypes.c_int32),
        ("nonbonded_method", ctypes.c_int32),
        ("cutoff", ctypes.c_double),
        ("n_radial", ctypes.c_int32),
        ("n_lebedev", ctypes.c_int32),
        ("r0", ctypes.c_double),
        ("r1", ctypes.c_double),
        ("grid_length", ctypes.c_double),
        ("buffer", ctypes.c_double),
        ("cell_cap", ctypes.c_int32),
        ("orders", ctypes.c_int32 * 2),
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
        ("tight_grid_length", ctypes.c_double),
        ("block_volume", ctypes.c_int32),
        ("block_force", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("n_points", "n_vol_pairs", "n_cells", "n_entries", "n_atoms", "n_quad", "n_active", "n_frc_pairs",
                 "slot_bytes")]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("blocks_per_cu", ctypes.c_int32),
        ("pair_columns", ctypes.c_int32),
        ("tight_grid_length", ctypes.c_double),
        ("ms_per_eval", ctypes.c_double),
    ]


class DeviceView(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p),
        ("forces", ctypes.c_void_p),
        ("energy", ctypes.c_void_p),
        ("born_radii", ctypes.c_void_p),
        ("integrals", ctypes.c_void_p),
        ("stream", ctypes.c_void_p),
    ]


class CustomSpec(ctypes.Structure):
    """gbi_custom_spec (include/gbi.h)."""
    _fields_ = [
        ("n_params", ctypes.c_int32),
        ("param_names", ctypes.POINTER(ctypes.c_char_p)),
        ("param_values", ctypes.POINTER(ctypes.c_double)),
        ("radius_param", ctypes.c_int32),
        ("n_globals", ctypes.c_int32),
        ("global_names", ctypes.POINTER(ctypes.c_char_p)),
        ("global_values", ctypes.POINTER(ctypes.c_double)),
        ("n_integrals", ctypes.c_int32),
        ("integral_names", ctypes.POINTER(ctypes.c_char_p)),
        ("n_values", ctypes.c_int32),
        ("value_names", ctypes.POINTER(ctypes.c_char_p)),
        ("value_exprs", ctypes.POINTER(ctypes.c_char_p)),
        ("n_terms", ctypes.c_int32),
        ("term_exprs", ctypes.POINTER(ctypes.c_char_p)),
        ("term_types", ctypes.POINTER(ctypes.c_int32)),
        ("n_param_derivs", ctypes.c_int32),
        ("param_deriv_names", ctypes.POINTER(ctypes.c_char_p)),
        ("n_functions", ctypes.c_int32),
        ("function_names", ctypes.POINTER(ctypes.c_char_p)),
        ("function_kinds", ctypes.POINTER(ctypes.c_int32)),
        ("function_sizes", ctypes.POINTER(ctypes.c_int32)),
        ("function_periodic", ctypes.POINTER(ctypes.c_int32)),
        ("function_ranges", ctypes.POINTER(ctypes.c_double)),
        ("function_values", ctypes.POINTER(ctypes.c_double)),
        ("function_dims", ctypes.POINTER(ctypes.c_int32)),
        ("function_ranges_yz", ctypes.POINTER(ctypes.c_double)),
    ]


TAB_CONTINUOUS_1D, TAB_DISCRETE_1D, TAB_DISCRETE_2D, TAB_DISCRETE_3D, TAB_CONTINUOUS_2D, TAB_CONTINUOUS_3D = 0, 1, 2, 3, 4, 5


class Tabulated:
    """A tabulated function for general expressions (gbi_custom_spec): kind
    TAB_CONTINUOUS_1D (cubic spline through values equally spaced on [min, max], natural and
    zero outside, or periodic), TAB_DISCRETE_1D/2D/3D (values[x + xsize (y + ysize z)] at the
    rounded arguments; dims = (xsize, ysize, zsize)) or TAB_CONTINUOUS_2D/3D (values[x + xsize
    (y + ysize z)] equally spaced on [min, max] x yrange (x zrange): the bicubic / tricubic
    Hermite patch with spline knot derivatives)."""

    def __init__(self, name, kind, values, min=0.0, max=1.0, periodic=False, dims=(0, 1, 1), yrange=(0.0, 1.0),
                 zrange=(0.0, 1.0)):
        self.name, self.kind = str(name), int(kind)
        self.values = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self.min, self.max, self.periodic = float(min), float(max), bool(periodic)
        self.dims = tuple(int(d) for d in dims)
        self.yrange = (float(yrange[0]), float(yrange[1]))
        self.zrange = (float(zrange[0]), float(zrange[1]))


TERM_SINGLE, TERM_PAIR, TERM_PAIR_NO_EXCLUSIONS = 0, 1, 2


def _strings(items):
    arr = (ctypes.c_char_p * max(1, len(items)))(*[x.encode() for x in items])
    return arr


_lib = None


def load() -> ctypes.CDLL:
    """Load libgbi.so (raises when it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with "
                           "`python -m openmmcustomgbintegralforce_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    L.gbi_params_default.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Params)]
    ip = ctypes.POINTER(ctypes.c_int32)
    L.gbi_create.argtypes = [ctypes.c_int32, dp, dp, ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.gbi_batch_create.argtypes = [ctypes.c_int32, ip, dp, dp, ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.gbi_get_layout.argtypes = [vp, ip, ip]
    L.gbi_get_adjoint.argtypes = [vp, dp, dp, dp]
    L.gbi_execute_adjoint.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.gbi_set_graph.argtypes = [vp, ctypes.c_int32]
    L.gbi_autotune.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Tuning)]
    L.gbi_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.gbi_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.gbi_execute.argtypes = [vp, dp, ctypes.c_int32, ctypes.c_int32, dp, dp]
    L.gbi_update_parameters.argtypes = [vp, ctypes.c_int32, dp, dp]
    L.gbi_set_periodic_box.argtypes = [vp, ctypes.c_int32, dp]
    L.gbi_set_exclusions.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.gbi_create_custom.argtypes = [ctypes.c_int32, ctypes.POINTER(Params), ctypes.POINTER(CustomSpec),
                                    ctypes.POINTER(vp)]
    L.gbi_set_custom_globals.argtypes = [vp, ctypes.c_int32, dp]
    L.gbi_get_custom_values.argtypes = [vp, dp]
    L.gbi_get_custom_param_derivs.argtypes = [vp, ctypes.c_int32, dp]
    L.gbi_update_custom_parameters.argtypes = [vp, ctypes.c_int32, dp]
    L.gbi_update_constants.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Params)]
    L.gbi_execute_noalloc.argtypes = [ctypes.c_int32, dp, dp, ctypes.POINTER(Params), dp, ctypes.c_int32,
                                      ctypes.c_int32, dp, dp, vp]
    L.gbi_set_positions.argtypes = [vp, vp, ctypes.c_int32]
    L.gbi_compute.argtypes = [vp, ctypes.c_uint32]
    L.gbi_synchronize.argtypes = [vp]
    L.gbi_get_energy.argtypes = [vp, dp]
    L.gbi_get_forces.argtypes = [vp, dp, ctypes.c_int32]
    L.gbi_get_born_radii.argtypes = [vp, dp]
    L.gbi_get_integrals.argtypes = [vp, dp]
    L.gbi_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.gbi_get_device_view.argtypes = [vp, ctypes.POINTER(DeviceView)]
    L.gbi_set_stream.argtypes = [vp, vp]
    L.gbi_set_positions_posq.argtypes = [vp, vp, ctypes.c_int32, vp]
    L.gbi_add_forces_fixed.argtypes = [vp, vp, ctypes.c_int32, vp]
    L.gbi_destroy.argtypes = [vp]
    L.gbi_destroy.restype = None
    L.gbi_last_error.restype = ctypes.c_char_p
    L.gbi_abi_version.restype = ctypes.c_int32
    L.gbi_quadrature.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int32, dp]
    L.gbi_profile_enable.argtypes = [vp, ctypes.c_int32]
    L.gbi_profile_read.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int32)]
    L.gbi_profile_pass_name.argtypes = [ctypes.c_int32]
    L.gbi_profile_pass_name.restype = ctypes.c_char_p
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is ctypes.c_int:  # default -> gbi_status
            fn.restype = ctypes.c_int32
    _lib = L
    return L


def _check(status: int) -> None:
    if status != GBI_OK:
        msg = load().gbi_last_error()
        raise GbiError(status, msg.decode() if msg else "")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(model: int, integral_type: int = INTEGRAL_GBSW) -> Params:
    p = Params()
    _check(load().gbi_params_default(model, integral_type, ctypes.byref(p)))
    return p


def quadrature(n_radial: int = 24, r0: float = 0.05, r1: float = 2.0, n_lebedev: int = 38) -> np.ndarray:
    out = np.zeros((n_radial * n_lebedev, 5))
    _check(load().gbi_quadrature(n_radial, r0, r1, n_lebedev, _dp(out)))
    return out


def tuning_key(params_list, n_atoms) -> str:
    """Autotuner cache key: each system's model, integral type, quadrature (radial x Lebedev)
    and nonbonded method, sorted, and the power-of-two bucket of the total atom count."""
    kinds = sorted({f"m{p.model}t{p.integral_type}q{p.n_radial}x{p.n_lebedev}c{p.nonbonded_method}"
                    for p in params_list})
    total = int(np.sum(n_atoms))
    bucket = 1 << max(0, int(total - 1).bit_length())
    return f"{'+'.join(kinds)}|N<={bucket}"


def load_tuning(path: str) -> dict:
    """The autotuner cache (JSON: key -> tuning), empty when the file does not exist."""
    import json

    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        data = json.load(fh)
    return data if isinstance(data, dict) else {}


def store_tuning(path: str, key: str, tuning: dict) -> None:
    """Adds or replaces one entry of the autotuner cache (written to a temporary file and
    renamed, so a reader never sees a partial file)."""
    import json

    data = load_tuning(path)
    data[key] = {k: (float(v) if isinstance(v, float) else int(v)) for k, v in tuning.items()}
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


class Engine:
    """One gbi_ctx: the HIP counterpart of a CalcCharmmGB{SW,MV}ForceKernel instance.

    ``params`` may be a list (one Params per system) with ``n_atoms`` giving each system's
    atom count: a batch whose per-atom arrays are the systems' arrays concatenated."""

    def __init__(self, params, charge, radius, n_atoms=None):
        self._lib = load()
        self.charge = np.ascontiguousarray(charge, dtype=np.float64)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.n = int(self.charge.shape[0])
        handle = ctypes.c_void_p()
        if isinstance(params, Params):
            self.params = [params]
            self.n_atoms = np.array([self.n], dtype=np.int32)
            _check(self._lib.gbi_create(self.n, _dp(self.charge), _dp(self.radius), ctypes.byref(params),
                                        ctypes.byref(handle)))
        else:
            self.params = list(params)
            self.n_atoms = np.ascontiguousarray(n_atoms, dtype=np.int32)
            arr = (Params * len(self.params))(*self.params)
            _check(self._lib.gbi_batch_create(len(self.params),
                                              self.n_atoms.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                              _dp(self.charge), _dp(self.radius), arr, ctypes.byref(handle)))
        self.n_systems = len(self.params)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_atoms)]).astype(np.int64)
        self._h = handle

    @classmethod
    def custom(cls, params, param_names, param_values, radius_param, integrals, values, terms, globals_=None,
               param_derivs=(), functions=()):
        """A context for general CharmmGBMVForce expressions (gbi_create_custom).

        param_values: N x len(param_names); integrals: names (1 or 2; orders from
        params.orders); values: [(name, expression)]; terms: [(expression, TERM_*)];
        globals_: {name: value}; param_derivs: names of globals whose energy derivative each
        evaluation returns (custom_param_derivs); functions: Tabulated functions."""
        eng = cls.__new__(cls)
        eng._lib = load()
        pv = np.ascontiguousarray(np.asarray(param_values, dtype=np.float64).reshape(-1, len(param_names)))
        n = pv.shape[0]
        globals_ = dict(globals_ or {})
        gnames = list(globals_)
        gvals = np.ascontiguousarray([float(globals_[k]) for k in gnames], dtype=np.float64)
        ttypes = np.ascontiguousarray([int(t) for _, t in terms], dtype=np.int32)
        keep = [_strings(list(param_names)), _strings(gnames), _strings(list(integrals)),
                _strings([v for v, _ in values]), _strings([e for _, e in values]), _strings([e for e, _ in terms]),
                _strings(list(param_derivs))]
        functions = list(functions)
        i32 = ctypes.POINTER(ctypes.c_int32)
        fk = np.ascontiguousarray([f.kind for f in functions] or [0], dtype=np.int32)
        fs = np.ascontiguousarray([len(f.values) for f in functions] or [0], dtype=np.int32)
        fp = np.ascontiguousarray([int(f.periodic) for f in functions] or [0], dtype=np.int32)
        fr = np.ascontiguousarray([v for f in functions for v in (f.min, f.max)] or [0.0], dtype=np.float64)
        fv = np.ascontiguousarray(np.concatenate([f.values for f in functions]) if functions else [0.0],
                                  dtype=np.float64)
        fdm = np.ascontiguousarray([d for f in functions for d in f.dims] or [0, 0, 0], dtype=np.int32)
        fyz = np.ascontiguousarray([v for f in functions for v in (*f.yrange, *f.zrange)] or [0.0] * 4,
                                   dtype=np.float64)
        keep += [_strings([f.name for f in functions]), fk, fs, fp, fr, fv, fdm, fyz]
        spec = CustomSpec(len(param_names), keep[0], _dp(pv), int(radius_param), len(gnames), keep[1],
                          _dp(gvals) if len(gnames) else None, len(integrals), keep[2], len(values), keep[3], keep[4],
                          len(terms), keep[5], ttypes.ctypes.data_as(i32),
                          len(param_derivs), keep[6], len(functions), keep[7], fk.ctypes.data_as(i32),
                          fs.ctypes.data_as(i32), fp.ctypes.data_as(i32), _dp(fr), _dp(fv), fdm.ctypes.data_as(i32),
                          _dp(fyz))
        handle = ctypes.c_void_p()
        _check(eng._lib.gbi_create_custom(n, ctypes.byref(params), ctypes.byref(spec), ctypes.byref(handle)))
        eng._h = handle
        eng.n = n
        eng.charge = np.zeros(n)
        eng.radius = pv[:, radius_param].copy()
        eng.params = [params]
        eng.n_atoms = np.array([n], dtype=np.int32)
        eng.n_systems = 1
        eng.offsets = np.array([0, n], dtype=np.int64)
        eng.n_values = len(values)
        eng.n_params = len(param_names)
        eng.global_names = gnames
        eng.param_derivs = list(param_derivs)
        return eng

    def custom_param_derivs(self) -> dict:
        """dE/d(global) of the last evaluation for each name given as ``param_derivs``."""
        out = np.zeros(max(1, len(self.param_derivs)))
        _check(self._lib.gbi_get_custom_param_derivs(self._h, len(self.param_derivs), _dp(out)))
        return {k: float(out[m]) for m, k in enumerate(self.param_derivs)}

    def update_custom_parameters(self, param_values) -> None:
        """New per-particle parameters (N x n_params) of a general-expression context."""
        pv = np.ascontiguousarray(np.asarray(param_values, dtype=np.float64).reshape(self.n, self.n_params))
        _check(self._lib.gbi_update_custom_parameters(self._h, self.n, _dp(pv)))

    def update_constants(self, params, system: int = 0) -> None:
        """New constants (Born map, prefactor, volume constants) of one system."""
        _check(self._lib.gbi_update_constants(self._h, int(system), ctypes.byref(params)))

    def custom_values(self) -> np.ndarray:
        """Computed values of the last evaluation (n_values x N)."""
        out = np.zeros((self.n_values, self.n))
        _check(self._lib.gbi_get_custom_values(self._h, _dp(out)))
        return out

    def set_custom_globals(self, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(self._lib.gbi_set_custom_globals(self._h, int(v.shape[0]), _dp(v)))

    @classmethod
    def batch(cls, params_list, charges, radii):
        """Batch of systems from per-system parameter, charge and radius lists."""
        n = [len(c) for c in charges]
        return cls(params_list, np.concatenate(charges), np.concatenate(radii), n_atoms=n)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.gbi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, pos, include_forces: bool = True, include_energy: bool = True, forces=None):
        """(energy, forces) of one evaluation; forces are ADDED to `forces` when given.
        The energy is a float for one system and an array of S energies for a batch."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        out = np.zeros((self.n, 3)) if forces is None else forces
        e = np.zeros(self.n_systems)
        _check(self._lib.gbi_execute(self._h, _dp(pos), int(include_forces), int(include_energy), _dp(e),
                                     _dp(out)))
        return (float(e[0]) if self.n_systems == 1 else e), out

    def execute_adjoint(self, pos):
        """dict(energy, forces, dEdR, dEdq, dEdtheta {name: value or per-system array})."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        e = np.zeros(self.n_systems)
        f = np.zeros((self.n, 3))
        dr = np.zeros(self.n)
        dq = np.zeros(self.n)
        th = np.zeros((self.n_systems, len(THETA)))
        _check(self._lib.gbi_execute_adjoint(self._h, _dp(pos), _dp(e), _dp(f), _dp(dr), _dp(dq), _dp(th)))
        one = self.n_systems == 1
        theta = {k: (float(th[0, i]) if one else th[:, i].copy()) for i, k in enumerate(THETA)}
        return {"energy": float(e[0]) if one else e, "forces": f, "dEdR": dr, "dEdq": dq, "dEdtheta": theta}

    def adjoint(self):
        """(dE/dR, dE/dq, dE/dtheta [S x 15]) of the last GBI_ADJOINT evaluation."""
        dr = np.zeros(self.n)
        dq = np.zeros(self.n)
        th = np.zeros((self.n_systems, len(THETA)))
        _check(self._lib.gbi_get_adjoint(self._h, _dp(dr), _dp(dq), _dp(th)))
        return dr, dq, th

    def update_parameters(self, charge, radius) -> None:
        charge = np.ascontiguousarray(charge, dtype=np.float64)
        radius = np.ascontiguousarray(radius, dtype=np.float64)
        _check(self._lib.gbi_update_parameters(self._h, int(charge.shape[0]), _dp(charge), _dp(radius)))
        self.charge, self.radius = charge, radius

    def set_periodic_box(self, box, system: int = 0) -> None:
        """Box of a CutoffPeriodic system: three edges (rectangular) or 3 x 3 box vectors."""
        b = np.asarray(box, dtype=np.float64)
        v = np.diag(b) if b.shape == (3,) else np.ascontiguousarray(b.reshape(3, 3))
        _check(self._lib.gbi_set_periodic_box(self._h, int(system), _dp(np.ascontiguousarray(v))))

    def set_exclusions(self, pairs, system: int = 0) -> None:
        """Excluded pairs (k x 2 local atom indices) of a ParticlePair pair term."""
        a = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if a.shape[0] else None
        _check(self._lib.gbi_set_exclusions(self._h, int(system), int(a.shape[0]), ptr))

    def set_positions(self, pos, on_device: bool = False) -> None:
        if on_device:
            _check(self._lib.gbi_set_positions(self._h, ctypes.c_void_p(int(pos)), 1))
        else:
            pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
            _check(self._lib.gbi_set_positions(self._h, pos.ctypes.data_as(ctypes.c_void_p), 0))

    def compute(self, flags: int = FORCES | ENERGY) -> None:
        _check(self._lib.gbi_compute(self._h, flags))

    def synchronize(self) -> None:
        _check(self._lib.gbi_synchronize(self._h))

    def energy(self):
        e = np.zeros(self.n_systems)
        _check(self._lib.gbi_get_energy(self._h, _dp(e)))
        return float(e[0]) if self.n_systems == 1 else e

    def forces(self) -> np.ndarray:
        f = np.zeros((self.n, 3))
        _check(self._lib.gbi_get_forces(self._h, _dp(f), 0))
        return f

    def born_radii(self) -> np.ndarray:
        b = np.zeros(self.n)
        _check(self._lib.gbi_get_born_radii(self._h, _dp(b)))
        return b

    def integrals(self) -> np.ndarray:
        out = np.zeros((2, self.n))
        _check(self._lib.gbi_get_integrals(self._h, _dp(out)))
        return out

    def stats(self) -> dict:
        s = Stats()
        _check(self._lib.gbi_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def device_view(self) -> DeviceView:
        v = DeviceView()
        _check(self._lib.gbi_get_device_view(self._h, ctypes.byref(v)))
        return v

    def set_graph(self, enable: bool = True) -> None:
        _check(self._lib.gbi_set_graph(self._h, int(enable)))

    def autotune(self, evals: int = 3, cache: str | None = None) -> dict:
        """Time candidate launch shapes on the current positions and keep the fastest
        (gbi_autotune).  With `cache` (a JSON file), a result stored for the same key
        (tuning_key: the systems' models, integral types and quadratures and the total atom
        count's power-of-two bucket) is applied without timing, and a new result is stored."""
        key = tuning_key(self.params, self.n_atoms)
        if cache is not None:
            hit = load_tuning(cache).get(key)
            if hit is not None:
                self.set_tuning(int(hit["blocks_per_cu"]), int(hit["pair_columns"]), float(hit["tight_grid_length"]))
                return dict(hit, cached=True)
        t = Tuning()
        _check(self._lib.gbi_autotune(self._h, int(evals), ctypes.byref(t)))
        out = {f: getattr(t, f) for f, _ in Tuning._fields_}
        if cache is not None:
            store_tuning(cache, key, out)
        return out

    def set_tuning(self, blocks_per_cu: int = 128, pair_columns: int = 0, tight_grid_length: float = 0.0) -> None:
        t = Tuning(blocks_per_cu, pair_columns, tight_grid_length, 0.0)
        _check(self._lib.gbi_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        _check(self._lib.gbi_get_tuning(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def profile_enable(self, enable: bool = True) -> None:
        _check(self._lib.gbi_profile_enable(self._h, int(enable)))

    def profile_read(self) -> tuple:
        """({pass name: average ms per evaluation}, number of profiled evaluations)."""
        ms = np.zeros(len(PROFILE_PASSES))
        n = ctypes.c_int32()
        _check(self._lib.gbi_profile_read(self._h, _dp(ms), ctypes.byref(n)))
        return dict(zip(PROFILE_PASSES, ms.tolist())), n.value

    def set_stream(self, stream_ptr: int | None) -> None:
        _check(self._lib.gbi_set_stream(self._h, ctypes.c_void_p(stream_ptr or 0)))
