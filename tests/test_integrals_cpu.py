"""Known-answer checks of the oracle's panel integrals against mathematics that
does not reuse the reference formula: far-field limits, an independent polar
integration of the in-plane potential, brute-force quadrature, divided
differences of the potential, and the 2D closed forms' limits."""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_lib as orc

TRI = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.2, 0.8, 0.0], dtype=np.float32)


def area(v):
    a, b, c = v.reshape(3, 3).astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a))


def brute_potential(r, v, n=400):
    """Mean of 1/|p-r| over the triangle by a fine midpoint rule on barycentric cells."""
    a, b, c = v.reshape(3, 3).astype(np.float64)
    tot = 0.0
    h = 1.0 / n
    for i in range(n):
        s = (np.arange(n - i) + 1.0 / 3.0) * h
        t = np.full_like(s, (i + 1.0 / 3.0) * h)
        p = a + np.outer(s, b - a) + np.outer(t, c - a)
        tot += np.sum(1.0 / np.linalg.norm(p - r, axis=1))
        s2 = (np.arange(n - i - 1) + 2.0 / 3.0) * h
        t2 = np.full_like(s2, (i + 2.0 / 3.0) * h)
        p2 = a + np.outer(s2, b - a) + np.outer(t2, c - a)
        tot += np.sum(1.0 / np.linalg.norm(p2 - r, axis=1))
    return tot / (n * n)


def polar_inplane(r, v):
    """In-plane point inside the triangle: integral of 1/rho dA = integral of R(theta)
    dtheta with R the distance to the boundary along theta (no singularity)."""
    P = v.reshape(3, 3).astype(np.float64)[:, :2]
    r2 = np.asarray(r[:2], dtype=np.float64)

    def R(th):
        d = np.array([math.cos(th), math.sin(th)])
        best = np.inf
        for i in range(3):
            p0, p1 = P[i], P[(i + 1) % 3]
            e = p1 - p0
            m = np.array([[d[0], -e[0]], [d[1], -e[1]]])
            if abs(np.linalg.det(m)) < 1e-15:
                continue
            tt, ss = np.linalg.solve(m, p0 - r2)
            if tt > 0 and -1e-12 <= ss <= 1 + 1e-12:
                best = min(best, tt)
        return best

    val, _ = integrate.quad(R, 0, 2 * math.pi, limit=400, epsabs=1e-13, epsrel=1e-12)
    return val / area(v)


@pytest.mark.parametrize("r", [[0.3, 0.3, 0.0], [0.5, 0.2, 0.0], [0.25, 0.5, 0.0]])
def test_potential_opt_inplane_matches_polar_integral(r):
    got = orc.potential_opt_3d(r, TRI)
    ref = polar_inplane(np.array(r), TRI)
    assert got == pytest.approx(ref, rel=1e-9)


@pytest.mark.parametrize("r", [[0.3, 0.2, 0.5], [1.5, -0.7, 0.3], [0.4, 0.4, -2.0]])
def test_potential_opt_offplane_matches_brute_force(r):
    got = orc.potential_opt_3d(r, TRI)
    ref = brute_potential(np.array(r), TRI, n=600)
    assert got == pytest.approx(ref, rel=2e-5)


def test_potential_far_field_limit():
    c = TRI.reshape(3, 3).astype(np.float64).mean(axis=0)
    for d in (50.0, 500.0):
        r = c + np.array([0.3, -0.2, 1.0]) / np.linalg.norm([0.3, -0.2, 1.0]) * d
        assert orc.potential_opt_3d(r, TRI) * d == pytest.approx(1.0, rel=1e-3 / d)


@pytest.mark.parametrize("rule,tol", [(2, 3e-3), (3, 3e-3), (5, 2e-3)])
def test_quadrature_rules_converge_to_analytic(rule, tol):
    r = np.array([0.4, 0.3, 2.5])
    assert orc.potential_num_3d(r, TRI, rule) == pytest.approx(orc.potential_opt_3d(r, TRI), rel=tol)


@pytest.mark.parametrize("rule", [2, 3, 5])
def test_quadrature_exact_on_constant_and_linear(rule):
    # the rules integrate degree <= 2 polynomials exactly: check on a far point where the
    # kernel is nearly linear is not exact, so check sum of weights via a huge distance
    r = np.array([0.0, 0.0, 1e6])
    assert orc.potential_num_3d(r, TRI, rule) * 1e6 == pytest.approx(1.0, rel=1e-9)


@pytest.mark.parametrize("r", [[0.3, 0.3, 0.7], [2.0, 1.0, -0.5]])
def test_enfield_is_minus_normal_derivative_of_potential(r):
    r = np.array(r)
    n = np.array([0.2, -0.3, 0.9])
    n /= np.linalg.norm(n)
    h = 1e-4
    dphi = (orc.potential_opt_3d(r + h * n, TRI) - orc.potential_opt_3d(r - h * n, TRI)) / (2 * h)
    e5 = orc.enfield_num_3d(r, TRI, n, 5)
    # quadrature error dominates; at these distances rule 5 is within a few 1e-3
    assert e5 == pytest.approx(-dphi, rel=2e-2)
    far = r + np.array([0, 0, 30.0])
    dphi_far = (orc.potential_opt_3d(far + h * n, TRI) - orc.potential_opt_3d(far - h * n, TRI)) / (2 * h)
    assert orc.enfield_num_3d(far, TRI, n, 2) == pytest.approx(-dphi_far, rel=1e-4)


SEG = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0], dtype=np.float32)


def brute_2d(r):
    f = lambda x: -0.5 * math.log((r[0] - x) ** 2 + r[1] ** 2)
    val, _ = integrate.quad(f, 0.0, 1.0, points=[r[0]] if 0 < r[0] < 1 else None, limit=200)
    return val


@pytest.mark.parametrize("r", [[0.3, 0.4], [1.7, -0.2], [-0.5, 2.0], [0.5, 0.0]])
def test_potential_2d_matches_log_integral(r):
    assert orc.potential_opt_2d(np.array(r), SEG) == pytest.approx(brute_2d(r), rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("r", [[0.3, 0.4], [1.7, -0.2], [-0.5, 2.0]])
def test_enfield_2d_is_minus_normal_derivative(r):
    r = np.array(r, dtype=np.float64)
    n = np.array([0.6, 0.8])
    h = 1e-5
    dphi = (orc.potential_opt_2d(r + h * n, SEG) - orc.potential_opt_2d(r - h * n, SEG)) / (2 * h)
    assert orc.enfield_2d(r, SEG, n) == pytest.approx(-dphi, rel=1e-6)
