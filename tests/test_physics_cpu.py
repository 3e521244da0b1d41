"""Closed-form and dense-algebra checks of the oracle (CPU only).

These pin the oracle's formulation where the reference's own tests hold no
vectors: dielectric interface rows and self terms against closed-form
capacitances (3D sphere in a dielectric shell, the 2D coated coax of the
reference's 2D input-file help page), the GMRES solve against a dense LU solve
and the tree-pass product against the dense compressed matrix.
"""
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads
from test_golden_cpu import two_cubes

E0 = 8.854187817e-12


def _groups(mesh):
    gs = mesh.group_leaf_start
    return [(g, gs[g], gs[g + 1], bool(mesh.group_is_diel[g])) for g in range(mesh.G)]


def _dense_caps(mesh, P):
    """C matrix from a dense solve of P x = b with the reference's right-hand
    sides (1 V on one conductor, 0 elsewhere) and 2D balance row: rows from the
    first conductor panel on subtract the last row, the last row becomes the total
    charge (Solver/MultiplyHierarchical.cpp:305-330, SolveCapacitance.cpp:1233-1240)."""
    N = mesh.N
    conds = [(g, lo, hi) for g, lo, hi, diel in _groups(mesh) if not diel]
    A = P.copy()
    if mesh.dim == 2:
        first = conds[0][1]
        A[first:N - 1, :] -= P[N - 1, :][None, :]
        A[N - 1, :] = 1.0
    rows = conds[:-1] if mesh.dim == 2 else conds
    C = np.zeros((len(rows), len(conds)))
    for r, (_, lo, hi) in enumerate(rows):
        b = np.zeros(N)
        b[lo:hi] = 1.0
        x = np.linalg.solve(A, b)
        for c, (_, clo, chi) in enumerate(conds):
            C[r, c] = float((mesh.leaf_outperm[clo:chi] * x[clo:chi]).sum())
    return C


@pytest.mark.parametrize("b,eps", [(3.0, 3.0), (3.0, 1.0 / 3.0), (2.0, 3.0)])
def test_sphere_in_dielectric_shell_closed_form(geo_dir, b, eps):
    """C = 4 pi e0 / ((1/a - 1/b)/eps + 1/b): the leaf-level assembly of the
    interface rows (field kernels, self term, normal orientation) is within the
    discretisation error of 640 flat panels.  Swapping the two permittivities
    would move the result by 25-40 %."""
    m = fc.Mesh.build(workloads.sphere_in_shell(os.path.join(geo_dir, f"shell{b}{eps:.3f}"), b=b, eps_shell=eps,
                                                subdiv=2), mesh_eps=0.1)
    C = _dense_caps(m, orc.leaf_dense(m))
    exact = 4 * math.pi * E0 / ((1 - 1 / b) / eps + 1 / b)
    assert abs(C[0, 0] / exact - 1.0) < 0.03


def test_conductor_sphere_closed_form(geo_dir):
    d = os.path.join(geo_dir, "ball")
    workloads.sphere_in_shell(d, subdiv=3)
    path = os.path.join(d, "ball_only.lst")
    with open(path, "w") as f:
        f.write("* conductor sphere\nC ball.txt 1.0 0 0 0\n")
    m = fc.Mesh.build(path, mesh_eps=0.02)
    coef, selfc = orc.fill(m)
    cap, _ = orc.solve(m, coef, selfc, tol=1e-10)
    assert abs(cap[0, 0] / (4 * math.pi * E0) - 1.0) < 0.01


def test_coated_coax_2d_closed_form(geo_dir):
    """The reference's 2D example: C = C1 C2 / (C1 + C2) = 113.4 pF/m
    (hlp/InputFile/2D/InputFile_2D_ch02.htm), 40 segments per circle."""
    m = fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "ccoax")), mesh_eps=0.1)
    c1 = 2 * math.pi * E0 * 2.0 / math.log(1.5)
    c2 = 2 * math.pi * E0 * 1.0 / math.log(2.0 / 1.5)
    exact = c1 * c2 / (c1 + c2)
    assert abs(exact - 113.4e-12) < 0.05e-12
    dense = _dense_caps(m, orc.leaf_dense(m)) * fc.TWO_PI_TIMES_E0
    assert abs(dense[0, 0] / exact - 1.0) < 0.01
    coef, selfc = orc.fill(m)
    cap, _ = orc.solve(m, coef, selfc, tol=1e-10)
    assert abs(cap[0, 0] * fc.TWO_PI_TIMES_E0 / exact - 1.0) < 0.025


def _small_meshes(geo_dir):
    os.makedirs(os.path.join(geo_dir, "pc"), exist_ok=True)
    return [
        fc.Mesh.build(two_cubes(os.path.join(geo_dir, "pc")), mesh_eps=0.05),
        fc.Mesh.build(workloads.two_plates(os.path.join(geo_dir, "pp")), mesh_eps=0.1),
        fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "pcc")), mesh_eps=0.1),
        fc.Mesh.build(workloads.coplanar_lines_2d(os.path.join(geo_dir, "pl"), seed=3), mesh_eps=0.1),
        fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "pb"), n=4, seed=2), mesh_eps=0.2),
    ]


def test_gmres_matches_dense_lu(geo_dir):
    for m in _small_meshes(geo_dir):
        coef, selfc = orc.fill(m)
        P = orc.compressed_dense(m, coef, selfc)
        ref = _dense_caps(m, P)
        cap, _ = orc.solve(m, coef, selfc, tol=1e-13)
        assert cap.shape == ref.shape
        assert np.max(np.abs(cap - ref) / np.abs(np.diag(ref))[:, None]) < 1e-9, m.N


def test_tree_product_matches_dense_matrix(geo_dir):
    for m in _small_meshes(geo_dir):
        coef, selfc = orc.fill(m)
        P = orc.compressed_dense(m, coef, selfc)
        q = np.random.default_rng(m.N).standard_normal(m.N)
        v = orc.mvp(m, coef, selfc, q)
        ref = P @ q
        if m.dim == 2:
            first = [lo for _, lo, _, diel in _groups(m) if not diel][0]
            ref[first:m.N - 1] -= ref[m.N - 1]
            ref[m.N - 1] = q.sum()
        assert np.max(np.abs(v - ref)) / np.max(np.abs(ref)) < 1e-12, m.N
