"""Adjoint of the HIP path -- dE/dR_j, dE/dq_j and dE/d(constants) -- against central finite
differences of the FP64 oracle energy (needs an MI355X).

The adjoint has no reference counterpart; the oracle energy is the function being
differentiated.  Steps: h = 1e-4 nm for radii, 1e-3 e for charges (E is quadratic in q,
so the central difference is exact up to rounding), 1e-5 relative for constants.
Tolerances, per tested set: |GPU - FD| <= 2e-4 * max|FD| over the set (the GPU sums FP32
terms in FP64; the FD truncation error at these steps is far below that), and element by
element |GPU_i - FD_i| <= 1e-3 * max(|FD_i|, 1e-2 max|FD|), so a derivative much smaller than
the largest of its set is still checked to 0.1 % of itself.  Lookup grids whose
membership depends on R (the GBMV-form reference grid) are differentiated with the brute
policy; the exact GBSW grid is smooth in R.

dE/deta gets 2e-3: it weighs every atom by T^(-3/4)/eta^5 with T = 1/(4 eta^4) - I7, a
difference of two numbers near 4e4 for buried atoms, so the FP32 volume error in I7 is
amplified by the conditioning of T (the closed form itself is pinned in FP64 by
test_oracle_adjoint.py).
"""
import os

import numpy as np
import pytest

import oracle as O
from openmmcustomgbintegralforce_amd import _native as N
from openmmcustomgbintegralforce_amd import synthetic as S

from common import oracle_params

pytestmark = pytest.mark.gpu

TOL = 2e-4
TOL_ETA = 2e-3

# oracle field of each constant (gbi_theta order)
ORACLE_FIELD = {"prefactor": "prefactor", "alpha0": "alpha0", "alpha1": "alpha1", "eta": "eta", "slope": "slope",
                "shift": "shift", "sw": "sw", "rscale": "rscale", "roffset": "roffset", "gamma0": "gamma0",
                "beta": "beta", "lambda": "lambda_", "P1": "P1", "P2": "P2", "S0": "S0"}
MODEL_THETA = {
    "gbsw": ["prefactor", "alpha0", "alpha1", "eta", "slope", "shift", "sw", "rscale", "roffset"],
    "gbmv1": ["prefactor", "alpha0", "alpha1", "eta", "slope", "shift", "gamma0", "beta", "lambda"],
    "gbmv2": ["prefactor", "alpha0", "alpha1", "eta", "slope", "shift", "beta", "lambda", "P1", "P2", "S0"],
}


def _energy(op, pos, q, R):
    return O.energy(op, pos, q, R)


def _fd_radius(op, s, atoms, h=1e-4):
    out = []
    for j in atoms:
        R = s.radius.copy()
        R[j] += h
        ep = _energy(op, s.pos, s.charge, R)
        R[j] -= 2 * h
        em = _energy(op, s.pos, s.charge, R)
        out.append((ep - em) / (2 * h))
    return np.array(out)


def _fd_charge(op, s, atoms, h=1e-3):
    out = []
    for j in atoms:
        q = s.charge.copy()
        q[j] += h
        ep = _energy(op, s.pos, q, s.radius)
        q[j] -= 2 * h
        em = _energy(op, s.pos, q, s.radius)
        out.append((ep - em) / (2 * h))
    return np.array(out)


def _fd_theta(op, s, name):
    field = ORACLE_FIELD[name]
    v = getattr(op, field)
    h = 1e-5 * abs(v) if v != 0.0 else 1e-5
    setattr(op, field, v + h)
    ep = _energy(op, s.pos, s.charge, s.radius)
    setattr(op, field, v - h)
    em = _energy(op, s.pos, s.charge, s.radius)
    setattr(op, field, v)
    return (ep - em) / (2 * h)


TOL_ELEM = 1e-3


def _check(tag, gpu, fd, tol=TOL, tol_elem=TOL_ELEM):
    gpu, fd = np.atleast_1d(gpu), np.atleast_1d(fd)
    scale = max(np.max(np.abs(fd)), 1e-30)
    err = np.max(np.abs(gpu - fd)) / scale
    elem = np.max(np.abs(gpu - fd) / np.maximum(np.abs(fd), 1e-2 * scale))
    print(f"[adjoint] {tag}: max err / max|FD| = {err:.2e}, max element rel err {elem:.2e} "
          f"(max|FD| {scale:.3e})", flush=True)
    assert err < tol, tag
    assert elem < max(tol, tol_elem), tag


def _run(p, s, model, r_atoms, q_atoms, theta_names, tag, box=None):
    print(f"[adjoint] {tag} N={s.n}", flush=True)
    eng = N.Engine(p, s.charge, s.radius)
    if box is not None:
        eng.set_periodic_box(box)
    out = eng.execute_adjoint(s.pos)
    # the adjoint evaluation's forces are the plain evaluation's, bit for bit
    _, f = eng.execute(s.pos)
    eng.close()
    assert np.array_equal(out["forces"], f), tag
    op = oracle_params(p, box=box)
    ref_e = _energy(op, s.pos, s.charge, s.radius)
    assert abs(out["energy"] - ref_e) < 1e-5 * abs(ref_e)
    # size-independent identities: E is homogeneous of degree 2 in q and linear in P
    assert abs(np.dot(s.charge, out["dEdq"]) - 2 * out["energy"]) < 1e-5 * abs(out["energy"])
    assert abs(out["dEdtheta"]["prefactor"] * p.prefactor - out["energy"]) < 1e-9 * abs(out["energy"])
    if q_atoms is not None:
        _check(f"{tag} dE/dq", out["dEdq"][q_atoms], _fd_charge(op, s, q_atoms))
    if r_atoms is not None:
        _check(f"{tag} dE/dR", out["dEdR"][r_atoms], _fd_radius(op, s, r_atoms))
    for name in theta_names:
        _check(f"{tag} dE/d{name}", out["dEdtheta"][name], _fd_theta(op, s, name),
               TOL_ETA if name == "eta" else TOL)
    for name in N.THETA:
        if name not in MODEL_THETA[model]:
            assert out["dEdtheta"][name] == 0.0, name
    return out


def test_adjoint_gbsw_standalone():
    s = S.globule(120, 21)
    p = N.default_params(N.MODEL_GBSW)
    _run(p, s, "gbsw", np.arange(0, 120, 4), np.arange(0, 120, 6), MODEL_THETA["gbsw"], "gbsw")


def test_adjoint_gbmv2_brute():
    s = S.globule(90, 22)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.n_lebedev = 50
    p.lookup_policy = N.LOOKUP_BRUTE
    _run(p, s, "gbmv2", np.arange(0, 90, 3), np.arange(0, 90, 9), MODEL_THETA["gbmv2"], "gbmv2 brute")


def test_adjoint_gbmv1_brute():
    s = S.globule(90, 23)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV1)
    p.lookup_policy = N.LOOKUP_BRUTE
    _run(p, s, "gbmv1", np.arange(0, 90, 3), None, MODEL_THETA["gbmv1"], "gbmv1 brute")


def test_adjoint_gbsw_form_brute():
    s = S.globule(90, 24)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBSW)
    p.lookup_policy = N.LOOKUP_BRUTE
    _run(p, s, "gbsw", np.arange(0, 90, 3), None, MODEL_THETA["gbsw"], "gbsw-form brute")


def test_adjoint_gbmv2_reference_grid_constants():
    """Constants that do not move the reference grid's membership, on that grid."""
    s = S.globule(400, 25)
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.n_lebedev = 50
    _run(p, s, "gbmv2", None, np.arange(0, 400, 40), MODEL_THETA["gbmv2"], "gbmv2 reference grid")


def test_adjoint_conformations_2000():
    """GBSW 24x38, 2,000 atoms, the 8 conformations of one base geometry (sigma 0.02 nm):
    the q-homogeneity identity over every atom of each, and central FD of the oracle energy
    on each: 12 radii, 6 charges and 5 constants on the first two, 4 radii, 2 charges and
    2 constants (alpha0, rscale) on the other six."""
    base = S.globule(2000, 26)
    confs = S.conformations(base, 8, seed=27)
    p = N.default_params(N.MODEL_GBSW)
    eng = N.Engine(p, base.charge, base.radius)
    rng = np.random.default_rng(28)
    for k, c in enumerate(confs):
        out = eng.execute_adjoint(c.pos)
        assert abs(np.dot(c.charge, out["dEdq"]) - 2 * out["energy"]) < 1e-5 * abs(out["energy"]), k
        op = oracle_params(p, nthreads=min(16, os.cpu_count() or 1))
        full = k < 2
        atoms = np.sort(rng.choice(2000, 12 if full else 4, replace=False))
        _check(f"conf {k} dE/dR", out["dEdR"][atoms], _fd_radius(op, c, atoms))
        _check(f"conf {k} dE/dq", out["dEdq"][atoms[:6 if full else 2]], _fd_charge(op, c, atoms[:6 if full else 2]))
        for name in (("alpha0", "eta", "sw", "rscale", "roffset") if full else ("alpha0", "rscale")):
            _check(f"conf {k} dE/d{name}", out["dEdtheta"][name], _fd_theta(op, c, name),
                   TOL_ETA if name == "eta" else TOL)
    eng.close()


def test_adjoint_batch_matches_single():
    """Adjoint of a mixed batch equals each system's own adjoint (same kernels)."""
    mix = [(N.default_params(N.MODEL_GBSW), S.globule(150, 30))]
    p2 = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p2.n_lebedev = 50
    mix.append((p2, S.globule(200, 31)))
    eng = N.Engine.batch([m[0] for m in mix], [m[1].charge for m in mix], [m[1].radius for m in mix])
    out = eng.execute_adjoint(np.concatenate([m[1].pos + 30.0 * k for k, m in enumerate(mix)]))
    eng.close()
    for k, (p, s) in enumerate(mix):
        one = N.Engine(p, s.charge, s.radius)
        o1 = one.execute_adjoint(s.pos + 30.0 * k)
        one.close()
        a, b = eng.offsets[k], eng.offsets[k + 1]
        assert np.allclose(out["dEdR"][a:b], o1["dEdR"], rtol=1e-9, atol=1e-9 * np.max(np.abs(o1["dEdR"]))), k
        assert np.allclose(out["dEdq"][a:b], o1["dEdq"], rtol=1e-9, atol=1e-9 * np.max(np.abs(o1["dEdq"]))), k
        for name in N.THETA:
            assert abs(out["dEdtheta"][name][k] - o1["dEdtheta"][name]) <= 1e-9 * max(1.0, abs(o1["dEdtheta"][name]))


def test_adjoint_gbsw_standalone_cutoff():
    """CharmmGBSWForce with CutoffNonPeriodic: pairs within the cutoff, the chain rule truncated
    at it (forces only); the energy's derivatives take every list atom."""
    s = S.globule(200, 32)
    p = N.default_params(N.MODEL_GBSW)
    p.nonbonded_method, p.cutoff = N.CUTOFF_NON_PERIODIC, 1.0
    _run(p, s, "gbsw", np.arange(0, 200, 7), np.arange(0, 200, 10), MODEL_THETA["gbsw"], "gbsw cutoff")


def test_adjoint_gbsw_standalone_periodic():
    s = S.globule(220, 33)
    L = np.array([3.4, 3.5, 3.6])
    s.pos = s.pos - s.pos.mean(axis=0) + L / 2
    p = N.default_params(N.MODEL_GBSW)
    p.nonbonded_method, p.cutoff = N.CUTOFF_PERIODIC, 1.1
    _run(p, s, "gbsw", np.arange(0, 220, 8), np.arange(0, 220, 11), ["prefactor", "alpha0", "eta", "sw", "rscale"],
         "gbsw periodic", box=L)


def test_adjoint_gbmv2_periodic_constants():
    """CharmmGBMVForce GBMV-II in a periodic box on its reference lookup grid: charges and the
    constants that do not move the grid's membership."""
    s = S.globule(300, 34)
    L = np.array([3.8, 3.9, 4.0])
    s.pos = s.pos - s.pos.mean(axis=0) + L / 2
    p = N.default_params(N.MODEL_GBMV, N.INTEGRAL_GBMV2)
    p.n_lebedev = 50
    p.nonbonded_method, p.cutoff = N.CUTOFF_PERIODIC, 1.2
    _run(p, s, "gbmv2", None, np.arange(0, 300, 30), MODEL_THETA["gbmv2"], "gbmv2 periodic", box=L)
