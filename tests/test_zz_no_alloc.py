"""gbi_execute_noalloc: one evaluation with every device buffer carved from the library's
static scratch array (no device allocation), through the C ABI, against the allocating entry
point gbi_execute on the same inputs (needs an MI355X).

Both run the same kernels on the same inputs, so energies and forces agree to FP64
summation-order level (checked at 1e-9 relative, far inside the 1e-5 force tolerance of the
oracle comparisons).  Each case makes two calls in a row, the second with a smaller system,
so the second reuses the array the first filled.
"""
import ctypes

import numpy as np
import pytest

from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _noalloc(p, s):
    lib = N.load()
    fn = lib.gbi_execute_noalloc
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_int32, dp, dp, ctypes.POINTER(N.Params), dp, ctypes.c_int32, ctypes.c_int32, dp, dp,
                   ctypes.c_void_p]
    fn.restype = ctypes.c_int32
    e = np.zeros(1)
    f = np.zeros((s.n, 3))
    pos = np.ascontiguousarray(s.pos)
    N._check(fn(s.n, N._dp(s.charge), N._dp(s.radius), ctypes.byref(p), N._dp(pos), 1, 1, N._dp(e), N._dp(f), None))
    return float(e[0]), f


def _allocating(p, s):
    eng = N.Engine(p, s.charge, s.radius)
    e, f = eng.execute(s.pos)
    eng.close()
    return e, f


def _params(kind):
    if kind == "gbsw":
        return N.default_params(N.MODEL_GBSW)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.n_lebedev = 50
    return p


@pytest.mark.parametrize("kind,sizes", [("gbsw", (1500, 600)), ("gbmv2", (1000, 400))])
def test_noalloc_matches_execute(kind, sizes):
    p = _params(kind)
    for n, seed in zip(sizes, (71, 72)):
        s = S.globule(n, seed)
        print(f"[noalloc] {kind} N={n}: gbi_execute_noalloc vs gbi_execute", flush=True)
        e1, f1 = _noalloc(p, s)
        e0, f0 = _allocating(p, s)
        de = abs(e1 - e0) / abs(e0)
        df = float(np.linalg.norm(f1 - f0) / np.linalg.norm(f0))
        print(f"[noalloc] {kind} N={n}: energy rel {de:.2e} force rel {df:.2e} "
              f"bitwise {e1 == e0 and np.array_equal(f1, f0)}", flush=True)
        assert np.all(np.isfinite(f1))
        assert de < TOL and df < TOL, (kind, n)


def test_noalloc_rejects_what_does_not_fit():
    """Sizes beyond the scratch array and other nonbonded methods are errors, not allocations."""
    lib = N.load()
    fn = lib.gbi_execute_noalloc
    s = S.globule(20000, 73)
    p = _params("gbmv2")
    e = np.zeros(1)
    f = np.zeros((s.n, 3))
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_int32, dp, dp, ctypes.POINTER(N.Params), dp, ctypes.c_int32, ctypes.c_int32, dp, dp,
                   ctypes.c_void_p]
    st = fn(s.n, N._dp(s.charge), N._dp(s.radius), ctypes.byref(p), N._dp(np.ascontiguousarray(s.pos)), 1, 1,
            N._dp(e), N._dp(f), None)
    assert st == 1 and b"too large for gbi_execute_noalloc" in lib.gbi_last_error()
    q = _params("gbsw")
    q.nonbonded_method, q.cutoff = N.CUTOFF_NON_PERIODIC, 1.0
    small = S.globule(50, 74)
    st = fn(small.n, N._dp(small.charge), N._dp(small.radius), ctypes.byref(q), N._dp(np.ascontiguousarray(small.pos)),
            1, 1, N._dp(e), N._dp(f), None)
    assert st == 2
