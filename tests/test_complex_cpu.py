"""Complex permittivities in the oracle (CPU only), pinned by closed forms and
dense complex algebra.

The reference solves a job with a complex permittivity (a lossy dielectric,
eps' - j eps'') as the real 2N system [R -D; D R] (Solver/MultiplyHierarchical.cpp:
332-455), with complex self terms on the dielectric panels (Solver/Autorefine.cpp:
2358-2393, 2435-2471) and complex charge sums (Solver/SolveCapacitance.cpp:
1305-1366).  These tests check that restatement against:
  * the closed-form capacitance of a conductor sphere in a lossy dielectric shell
    and of the lossy coated coax of the reference's 2D help page;
  * a dense complex matrix (product) and a dense complex LU solve (GMRES on 2N);
  * the real solve when every imaginary part is zero.
"""
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads

E0 = 8.854187817e-12


def _conds(mesh):
    gs = mesh.group_leaf_start
    return [(gs[g], gs[g + 1]) for g in range(mesh.G) if not mesh.group_is_diel[g]]


def _dense_caps_cplx(mesh, P):
    """Complex C from a dense complex solve with the reference's right-hand sides
    and 2D balance row, charges weighted by the complex outer permittivity."""
    N = mesh.N
    conds = _conds(mesh)
    A = P.astype(np.complex128).copy()
    if mesh.dim == 2:
        first = conds[0][0]
        A[first:N - 1, :] -= A[N - 1, :][None, :]
        A[N - 1, :] = 1.0
    rows = conds[:-1] if mesh.dim == 2 else conds
    w = mesh.leaf_outperm + 1j * mesh.leaf_outperm_im
    C = np.zeros((len(rows), len(conds)), dtype=np.complex128)
    for r, (lo, hi) in enumerate(rows):
        b = np.zeros(N, dtype=np.complex128)
        b[lo:hi] = 1.0
        x = np.linalg.solve(A, b)
        for c, (clo, chi) in enumerate(conds):
            C[r, c] = (w[clo:chi] * x[clo:chi]).sum()
    return C


def _leaf_dense_cplx(mesh):
    """Leaf-level dense complex matrix: every leaf pair evaluated directly, the
    complex self terms on the diagonal."""
    P = orc.leaf_dense(mesh).astype(np.complex128)
    sre, sim = orc.self_cplx(mesh)
    P[np.arange(mesh.N), np.arange(mesh.N)] = sre + 1j * sim
    return P


def _shell(geo_dir, eps, tag):
    return fc.Mesh.build(workloads.sphere_in_shell(os.path.join(geo_dir, f"cshell{tag}"), b=3.0, eps_shell=eps,
                                                   subdiv=2), mesh_eps=0.1)


def _shell_cap_error(geo_dir, eps, subdiv):
    m = fc.Mesh.build(workloads.sphere_in_shell(os.path.join(geo_dir, f"cshell{eps}{subdiv}"), b=3.0, eps_shell=eps,
                                                subdiv=subdiv), mesh_eps=0.1)
    assert m.complex and np.any(m.group_outperm_im != 0)
    C = _dense_caps_cplx(m, _leaf_dense_cplx(m))[0, 0]
    exact = 4 * math.pi * E0 / ((1 - 1 / 3.0) / eps + 1 / 3.0)
    return C, exact


@pytest.mark.parametrize("eps", [3.0 - 1.5j, 0.5 - 0.8j])
def test_sphere_in_lossy_shell_closed_form(geo_dir, eps):
    """C = 4 pi e0 / ((1/a - 1/b)/eps + 1/b) with complex eps, from the
    leaf-level assembly (complex self terms of the interface panels): within the
    discretisation error of 640 flat panels per sphere (a few %, as for real eps)
    in modulus and in the imaginary part.  A wrong sign of the imaginary self
    term or of the coupling would move Im C by about twice its value (over 30%
    of |C| here)."""
    C, exact = _shell_cap_error(geo_dir, eps, 2)
    assert abs(exact.imag) > 0.1 * abs(exact)  # the loss is visible
    assert abs(C - exact) < 0.05 * abs(exact)
    assert abs(C.imag - exact.imag) < 0.05 * abs(exact)


def test_lossy_shell_error_shrinks_with_refinement(geo_dir):
    """The remaining difference is discretisation: four times the panels cut it
    (5% -> 3.7% -> 2.2% for eps = 3 - 1.5j at 160 / 640 / 2560 panels)."""
    e2 = abs(np.subtract(*_shell_cap_error(geo_dir, 3.0 - 1.5j, 2)))
    e3 = abs(np.subtract(*_shell_cap_error(geo_dir, 3.0 - 1.5j, 3)))
    assert e3 < 0.75 * e2


def test_lossy_coated_coax_2d_closed_form(geo_dir):
    """The reference's 2D coated coax with a lossy insulator eps_in = 2 - 1j:
    C = C1 C2 / (C1 + C2), C1 = 2 pi e0 eps_in / ln(b/a); leaf level within 1%,
    the hierarchical complex GMRES within 2.5%."""
    eps_in = 2.0 - 1.0j
    m = fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "lossy_coax"), eps_in=eps_in), mesh_eps=0.1)
    assert m.complex
    c1 = 2 * math.pi * E0 * eps_in / math.log(1.5)
    c2 = 2 * math.pi * E0 * 1.0 / math.log(2.0 / 1.5)
    exact = c1 * c2 / (c1 + c2)
    dense = _dense_caps_cplx(m, _leaf_dense_cplx(m))[0, 0] * fc.TWO_PI_TIMES_E0
    assert abs(dense - exact) < 0.01 * abs(exact)
    coef, sre, sim = orc.fill_cplx(m)
    cap, _ = orc.solve_cplx(m, coef, sre, sim, tol=1e-10)
    assert abs(cap[0, 0] * fc.TWO_PI_TIMES_E0 - exact) < 0.025 * abs(exact)


def _cplx_meshes(geo_dir):
    return [
        _shell(geo_dir, 3.0 - 1.5j, "m"),
        fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "cc2"), eps_in=2.0 - 1.0j, eps_out=1.5 - 0.2j),
                      mesh_eps=0.1),
    ]


def test_complex_product_matches_dense_matrix(geo_dir):
    for m in _cplx_meshes(geo_dir):
        coef, sre, sim = orc.fill_cplx(m)
        P = orc.compressed_dense(m, coef, sre).astype(np.complex128)
        P[np.arange(m.N), np.arange(m.N)] += 1j * sim
        q = np.random.default_rng(m.N).standard_normal(2 * m.N)
        qc = q[: m.N] + 1j * q[m.N:]
        ref = P @ qc
        if m.dim == 2:
            first = _conds(m)[0][0]
            ref[first:m.N - 1] -= ref[m.N - 1]
            ref[m.N - 1] = qc.sum()
        v = orc.mvp_cplx(m, coef, sre, sim, q)
        got = v[: m.N] + 1j * v[m.N:]
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12, m.N


def test_complex_gmres_matches_dense_lu(geo_dir):
    for m in _cplx_meshes(geo_dir):
        coef, sre, sim = orc.fill_cplx(m)
        P = orc.compressed_dense(m, coef, sre).astype(np.complex128)
        P[np.arange(m.N), np.arange(m.N)] += 1j * sim
        ref = _dense_caps_cplx(m, P)
        cap, _ = orc.solve_cplx(m, coef, sre, sim, tol=1e-13)
        assert cap.shape == ref.shape
        assert np.max(np.abs(cap - ref) / np.abs(np.diag(ref))[:, None]) < 1e-9, m.N


def test_zero_loss_equals_real_solve(geo_dir):
    """'3.0-j0.0' marks the job complex; with every imaginary part zero the
    complex solve gives the real solve's C and a zero imaginary part."""
    real = _shell(geo_dir, 3.0, "r")
    cplx = _shell(geo_dir, 3.0 - 0.0j, "z")
    assert cplx.complex and not real.complex
    coef, selfc = orc.fill(real)
    c_real, it_real = orc.solve(real, coef, selfc, tol=1e-10)
    coef2, sre, sim = orc.fill_cplx(cplx)
    np.testing.assert_array_equal(sim, 0.0)
    np.testing.assert_allclose(sre, selfc, rtol=1e-15)
    c_cplx, it_cplx = orc.solve_cplx(cplx, coef2, sre, sim, tol=1e-10)
    assert np.max(np.abs(c_cplx.real - c_real) / np.abs(c_real)) < 1e-12
    assert np.max(np.abs(c_cplx.imag)) == 0.0
    assert list(it_cplx) == list(it_real)
