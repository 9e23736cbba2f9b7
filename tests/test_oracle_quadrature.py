"""Quadrature of the oracle and of the library, pinned by numpy / scipy.

Reference: CharmmQuadrature.h:14-75 (Gauss-Legendre by Newton iteration), :183-951
(Lebedev orbits), order_table :6801-6862; table assembly GBSWIntegral.cpp:41-56.
"""
import numpy as np
import pytest
from scipy.integrate import lebedev_rule

import oracle as O

# Lebedev rule index -> (points, polynomial degree) (order_table, precision_table)
RULES = {1: (6, 3), 2: (14, 5), 3: (26, 7), 4: (38, 9), 5: (50, 11)}


def test_gauss_legendre_matches_numpy():
    x, w = O.gauss_legendre(0.05, 2.0, 24)
    t, v = np.polynomial.legendre.leggauss(24)
    assert np.allclose(x, 0.975 * t + 1.025, rtol=0, atol=2e-15)
    assert np.allclose(w, 0.975 * v, rtol=0, atol=2e-15)
    assert abs(w.sum() - 1.95) < 1e-14
    assert np.all(np.diff(x) > 0)


@pytest.mark.parametrize("n", [1, 2, 5, 24, 40])
def test_gauss_legendre_exact_polynomials(n):
    x, w = O.gauss_legendre(-1.0, 3.0, n)
    for deg in range(2 * n):
        exact = (3.0 ** (deg + 1) - (-1.0) ** (deg + 1)) / (deg + 1)
        assert abs((w * x ** deg).sum() - exact) <= 1e-11 * max(1.0, abs(exact))


@pytest.mark.parametrize("rule", sorted(RULES))
def test_lebedev_matches_scipy(rule):
    npts, degree = RULES[rule]
    assert O.lib().orc_lebedev_rule_npts(rule) == npts
    pts, w = O.lebedev(npts)
    sp, sw = lebedev_rule(degree)
    sp = sp.T
    sw = sw / (4 * np.pi)
    key = lambda p: np.lexsort(np.round(p, 12).T[::-1])  # noqa: E731
    a, b = pts[key(pts)], sp[key(sp)]
    assert np.allclose(a, b, atol=1e-14)
    assert np.allclose(w[key(pts)], sw[key(sp)], atol=1e-15)
    assert abs(w.sum() - 1.0) < 1e-14
    assert np.allclose(np.linalg.norm(pts, axis=1), 1.0, atol=1e-15)


def _sphere_monomial_mean(a, b, c):
    """(1/4pi) * integral over the unit sphere of x^a y^b z^c (closed form)."""
    from math import gamma

    if a % 2 or b % 2 or c % 2:
        return 0.0
    num = gamma((a + 1) / 2) * gamma((b + 1) / 2) * gamma((c + 1) / 2)
    return 2.0 * num / gamma((a + b + c + 3) / 2) / (4 * np.pi)


@pytest.mark.parametrize("rule", sorted(RULES))
def test_lebedev_integrates_harmonics_exactly(rule):
    npts, degree = RULES[rule]
    pts, w = O.lebedev(npts)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                got = (w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum()
                assert abs(got - _sphere_monomial_mean(a, b, c)) < 1e-14, (a, b, c)


def test_quad_table_layout():
    t = O.quad_table(24, 0.05, 2.0, 38)
    assert t.shape == (912, 5)
    x, wr = O.gauss_legendre(0.05, 2.0, 24)
    u, ws = O.lebedev(38)
    for k in (0, 7, 23):
        for m in (0, 11, 37):
            row = t[k * 38 + m]
            assert np.allclose(row[:3], x[k] * u[m], atol=1e-16)
            assert row[3] == x[k]
            assert abs(row[4] - wr[k] * ws[m]) < 1e-18
    assert abs(t[:, 4].sum() - 1.95) < 1e-13


@pytest.mark.parametrize("npts", [26, 38, 50])
def test_library_quadrature_matches_oracle(npts):
    """The product's own generator (quadrature.cpp) equals the oracle's table."""
    from openmmcustomgbintegralforce_amd import _native as N

    a = N.quadrature(24, 0.05, 2.0, npts)
    b = O.quad_table(24, 0.05, 2.0, npts)
    assert np.allclose(a, b, rtol=0, atol=1e-15)
