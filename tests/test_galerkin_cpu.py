"""The Galerkin scheme of the oracle (-g; Solver/Potential.cpp:994-1017, 2485-2964,
3138-4377, Solver/Autorefine.cpp:2399-2401, 3570-3674), pinned without the
reference's own arithmetic: every double integral is compared with a brute-force
quadrature (midpoints of a uniform subdivision of each triangle, or the mean of
the analytic potential, which the collocation tests pin), plus the symmetry of the
double integrals and closed-form capacitances solved with the Galerkin fill."""
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads

E0 = 8.854187817e-12
K = 4 * math.pi * E0

TRIS = {
    "equilateral": [0, 0, 0, 1, 0, 0, 0.5, 0.8660254, 0],
    "right": [0, 0, 0, 1.2, 0, 0, 0, 0.7, 0],
    "obtuse": [0, 0, 0, 1.0, 0, 0, 1.6, 0.35, 0.1],
}


def _subdiv_points(v, level):
    """Centroids and areas of the 4**level congruent subtriangles of triangle v."""
    v = np.asarray(v, dtype=np.float32).astype(np.float64).reshape(3, 3)  # the stored float vertices
    n = 2 ** level
    pts, w = [], []
    a, b, c = v
    for i in range(n):
        for j in range(n - i):
            p0 = a + (b - a) * i / n + (c - a) * j / n
            p1 = a + (b - a) * (i + 1) / n + (c - a) * j / n
            p2 = a + (b - a) * i / n + (c - a) * (j + 1) / n
            pts.append((p0 + p1 + p2) / 3)
            w.append(1.0)
            if i + j < n - 1:
                p3 = a + (b - a) * (i + 1) / n + (c - a) * (j + 1) / n
                pts.append((p1 + p2 + p3) / 3)
                w.append(1.0)
    w = np.array(w)
    return np.array(pts), w / w.sum()


def _mean_potential(v1, v2, level=5):
    """Mean over triangle 1 of the analytic potential of triangle 2 (unit charge)."""
    pts, w = _subdiv_points(v1, level)
    t2 = np.asarray(v2, dtype=np.float32)
    return sum(wi * orc.potential_opt_3d(p, t2) for p, wi in zip(pts, w)) / K


def _double_mean(v1, v2, kernel, level=4):
    p1, w1 = _subdiv_points(v1, level)
    p2, w2 = _subdiv_points(v2, level)
    d = p1[:, None, :] - p2[None, :, :]
    return float((w1[:, None] * w2[None, :] * kernel(d)).sum()) / K


def _shift(v, dz=0.0, dx=0.0):
    v = np.array(v, dtype=np.float64).reshape(3, 3)
    v[:, 2] += dz
    v[:, 0] += dx
    return v.ravel()


@pytest.mark.parametrize("name", list(TRIS))
def test_auto_is_the_mean_self_potential(name):
    """Auto = the triangle's own analytic potential averaged over it."""
    v = TRIS[name]
    ref = _mean_potential(v, v, level=6)
    assert abs(orc.auto_3d(v) / ref - 1.0) < 5e-4


def test_auto_equilateral_closed_form():
    # the mean of 1/r over an equilateral triangle of side a is 4 ln 3 / a
    for a in (1.0, 0.25):
        v = np.array(TRIS["equilateral"]) * a
        assert abs(orc.auto_3d(v) * K * a / (4 * math.log(3.0)) - 1.0) < 1e-6


@pytest.mark.parametrize("rule", [17, 10])
@pytest.mark.parametrize("dz,tol", [(1.5, 5e-5), (0.4, 3e-4)])
def test_mutual_half_matches_brute_force(rule, dz, tol):
    v1 = TRIS["obtuse"]
    v2 = _shift(TRIS["right"], dz=dz, dx=0.2)
    assert abs(orc.mutual_half_3d(v1, v2, rule) / _mean_potential(v1, v2) - 1.0) < tol


def test_mutual_half_self_limit():
    v = TRIS["right"]
    assert abs(orc.mutual_half_3d(v, v, 17) / orc.auto_3d(v) - 1.0) < 5e-3


def _inv_r(d):
    return 1.0 / np.sqrt((d * d).sum(-1))


@pytest.mark.parametrize("order,dz,tol", [(4, 2.5, 2e-4), (2, 6.0, 2e-4)])
def test_mutual_full_matches_brute_force(order, dz, tol):
    v1 = TRIS["equilateral"]
    v2 = _shift(TRIS["obtuse"], dz=dz, dx=0.3)
    got = orc.mutual_full_3d(v1, v2, order)
    assert abs(got / _double_mean(v1, v2, _inv_r) - 1.0) < tol
    # a double mean of a symmetric kernel is symmetric in the two triangles
    assert abs(orc.mutual_full_3d(v2, v1, order) / got - 1.0) < 1e-12


def _dot_kernel(n):
    n = np.asarray(n, dtype=np.float64)

    def k(d):
        r = np.sqrt((d * d).sum(-1))
        return (d @ n) / (r * r * r)

    return k


@pytest.mark.parametrize("rule,dz,tol", [(10, 1.6, 5e-4), (5, 3.0, 5e-4), (0, 7.0, 5e-4)])
def test_mutual_dfull_matches_brute_force(rule, dz, tol):
    v1 = _shift(TRIS["equilateral"], dz=dz, dx=0.1)
    v2 = TRIS["right"]
    n = np.array([0.0, 0.6, 0.8])
    got = orc.mutual_dfull_3d(v1, v2, n, rule)
    ref = _double_mean(v1, v2, _dot_kernel(n))
    assert abs(got / ref - 1.0) < tol
    # the field of triangle 2 pointing away from it: positive along +z above it
    assert got > 0


def test_mutual_dhalf_matches_brute_force():
    v1 = _shift(TRIS["equilateral"], dz=0.8, dx=0.1)
    v2 = TRIS["right"]
    n = np.array([0.0, 0.0, 1.0])
    c1 = np.asarray(v1).reshape(3, 3).mean(0)
    c2 = np.asarray(v2).reshape(3, 3).mean(0)
    h = float(np.linalg.norm(c1 - c2)) / 20
    got = orc.mutual_dhalf_3d(v1, v2, n, h, 17)
    ref = _double_mean(v1, v2, _dot_kernel(n), level=5)
    assert abs(got / ref - 1.0) < 5e-3


def _sphere_cap(geo_dir, scheme, name):
    d = os.path.join(geo_dir, name)
    workloads.sphere_in_shell(d, subdiv=3)
    path = os.path.join(d, "ball_only.lst")
    with open(path, "w") as f:
        f.write("* conductor sphere\nC ball.txt 1.0 0 0 0\n")
    m = fc.Mesh.build(path, mesh_eps=0.02)
    m.scheme = scheme
    coef, selfc = orc.fill(m)
    cap, _ = orc.solve(m, coef, selfc, tol=1e-10)
    return m, coef, selfc, cap[0, 0]


def test_galerkin_sphere_closed_form(geo_dir):
    """A unit conductor sphere: C = 4 pi e0 within 1 % with the Galerkin fill; its
    coefficients and self terms differ from collocation's (the scheme is read)."""
    _, cg, sg, cap_g = _sphere_cap(geo_dir, fc.FCG_SCHEME_GALERKIN, "gball")
    _, cc, sc, cap_c = _sphere_cap(geo_dir, fc.FCG_SCHEME_COLLOCATION, "cball")
    assert abs(cap_g / K - 1.0) < 0.01
    assert np.max(np.abs(sg / sc - 1.0)) > 1e-3
    assert np.max(np.abs(cg / cc - 1.0)) > 1e-4


@pytest.mark.parametrize("b,eps", [(2.0, 3.0), (3.0, 3.0), (3.0, 1.0 / 3.0)])
def test_galerkin_sphere_in_dielectric_shell(geo_dir, b, eps):
    """Dielectric interface rows with the Galerkin field integrals, every leaf pair
    evaluated directly: C = 4 pi e0 / ((1/a - 1/b)/eps + 1/b) within 1.5 % on 640
    flat panels (collocation: within 3 %)."""
    from test_physics_cpu import _dense_caps

    m = fc.Mesh.build(workloads.sphere_in_shell(os.path.join(geo_dir, f"gshell{b}{eps:.3f}"), b=b, eps_shell=eps,
                                                subdiv=2), mesh_eps=0.1)
    m.scheme = fc.FCG_SCHEME_GALERKIN
    C = _dense_caps(m, orc.leaf_dense(m))
    exact = 4 * math.pi * E0 / ((1 - 1 / b) / eps + 1 / b)
    assert abs(C[0, 0] / exact - 1.0) < 0.015


def test_scheme_ignored_in_2d(geo_dir):
    """2D jobs are solved by collocation whatever the scheme (SolveCapacitance.cpp:159-161)."""
    m = fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "gcoax")), mesh_eps=0.1)
    c0, s0 = orc.fill(m)
    m.scheme = fc.FCG_SCHEME_GALERKIN
    c1, s1 = orc.fill(m)
    assert np.array_equal(c0, c1) and np.array_equal(s0, s1)
