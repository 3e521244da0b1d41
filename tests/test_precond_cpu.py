"""The oracle's block preconditioner (-pb), pinned on the CPU.

The partition must cover the leaves with maximal subtrees of at most b leaves
(RecurseComputePrecond, Solver/SolveCapacitance.cpp:1876-1993); every block's
stored inverse times its dense collocation matrix must be the identity
(ComputeBlockPrecond + InvertMatrix, :2068-2258); and a preconditioner never
changes the solution, so the block-preconditioned C equals the Jacobi C.
"""
import ctypes as C
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads


@pytest.fixture(scope="module")
def meshes(geo_dir):
    os.makedirs(geo_dir, exist_ok=True)
    return {
        "plates": fc.Mesh.build(workloads.two_plates(os.path.join(geo_dir, "pc_plates")), mesh_eps=0.05),
        "boxes": fc.Mesh.build(workloads.random_boxes(os.path.join(geo_dir, "pc_boxes"), n=4, seed=2), mesh_eps=0.3),
        "coax2d": fc.Mesh.build(workloads.coated_coax_2d(os.path.join(geo_dir, "pc_coax")), mesh_eps=0.1),
    }


@pytest.mark.parametrize("bsize", [2, 7, 32, 128])
@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_partition_is_maximal_subtrees(meshes, name, bsize):
    m = meshes[name]
    starts = orc.block_partition(m, bsize)
    assert starts[0] == 0 and starts[-1] == m.N
    assert np.all(np.diff(starts) > 0)
    # every block is the leaf range of one node with at most bsize leaves whose
    # parent (if any) has more than bsize leaves (or the block is a lone leaf)
    lo, hi = m.leaf_lo, m.leaf_hi
    by_range = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(lo, hi))}
    for s, e in zip(starts[:-1], starts[1:]):
        node = by_range.get((int(s), int(e)))
        assert node is not None, f"block [{s},{e}) is not a subtree"
        n = e - s
        assert n <= bsize or m.left[node] < 0
        p = m.parent[node]
        if p >= 0:
            assert hi[p] - lo[p] > bsize


@pytest.mark.parametrize("bsize", [3, 32])
@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_blocks_invert_their_collocation_matrix(meshes, name, bsize):
    m = meshes[name]
    bp = orc.BlockPrecond(m, bsize)
    desc = C.byref(m.desc)
    worst = 0.0
    for b in range(0, bp.nblocks, max(1, bp.nblocks // 25)):
        s0 = int(bp.starts[b])
        n = int(bp.starts[b + 1]) - s0
        A = np.empty((n, n))
        for i in range(n):
            for j in range(n):
                A[i, j] = (orc.lib().orc_self(desc, s0 + i) if i == j else
                           orc.pot_estimate(m, int(m.leaf_node[s0 + i]), int(m.leaf_node[s0 + j])))
        # the inverse is checked against numpy's, which shares no code with the oracle
        np.testing.assert_allclose(bp.block(b), np.linalg.inv(A), rtol=1e-9, atol=1e-9 * np.abs(bp.block(b)).max())
        worst = max(worst, np.abs(A @ bp.block(b) - np.eye(n)).max())
    assert worst < 1e-11


@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_block_preconditioned_c_equals_jacobi_c(meshes, name):
    m = meshes[name]
    coef, selfc = orc.fill(m)
    cap_j, it_j = orc.solve(m, coef, selfc, tol=1e-9)
    cap_b, it_b = orc.solve_block(m, coef, selfc, 32, tol=1e-9)
    err = np.abs(cap_b - cap_j).max() / np.abs(np.diag(cap_j)).max()
    print(f"{name}: iterations Jacobi {list(it_j)} block {list(it_b)}; C diff {err:.2e}")
    assert err < 1e-7


# ---------------------------------------------------------------------------
# Two-level preconditioner (-ps): MakePreconditioner and the super branch of
# RecurseComputePrecond (Solver/SolveCapacitance.cpp:1067-1104, 1876-1993),
# ComputeSuperPrecond (:2007-2066), ComputePrecondVectFast (:3545-3592).


def _depths(m):
    d = np.zeros(m.M, dtype=np.int64)
    for i in range(m.M - 1, -1, -1):  # post order: parents follow their children
        p = m.parent[i]
        if p >= 0:
            d[i] = d[p] + 1
    return d


def _roots(m):
    return int((m.parent < 0).sum())


@pytest.mark.parametrize("dim", [16, 64, 128, 1024])
@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_super_elements_are_the_nodes_at_the_level(meshes, name, dim):
    m = meshes[name]
    sp = orc.SuperPrecond(m, dim)
    R = _roots(m)
    if R >= dim:
        assert not sp.available
        return
    assert sp.available
    level = (dim // R).bit_length() - 1  # floor(log2(dim / R)), integer division first
    depth = _depths(m)
    assert sp.starts[0] == 0 and sp.starts[-1] == m.N and np.all(np.diff(sp.starts) > 0)
    assert sp.m <= R << level
    for s, node in enumerate(sp.nodes):
        assert m.leaf_lo[node] == sp.starts[s] and m.leaf_hi[node] == sp.starts[s + 1]
        # a node at the level, or a leaf above it
        assert depth[node] == level or (m.left[node] < 0 and depth[node] < level)


def test_super_unavailable_with_as_many_roots_as_dim(meshes):
    m = meshes["boxes"]
    R = _roots(m)
    assert not orc.SuperPrecond(m, R).available
    assert orc.SuperPrecond(m, R + 1).available


@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_super_inverse_is_numpy_inverse_of_element_matrix(meshes, name):
    m = meshes[name]
    sp = orc.SuperPrecond(m, 64)
    n = sp.m
    A = np.empty((n, n))
    for i in range(n):
        for j in range(n):
            A[i, j] = orc.node_self(m, int(sp.nodes[i])) if i == j else orc.pot_estimate(m, int(sp.nodes[i]),
                                                                                         int(sp.nodes[j]))
    np.testing.assert_allclose(sp.Sinv, np.linalg.inv(A), rtol=1e-8, atol=1e-9 * np.abs(sp.Sinv).max())


def test_node_self_scales_with_the_panel_size(meshes):
    # the potential of a uniformly charged panel at its centroid, per unit
    # charge, scales as 1 / size: nodes of one shape at different refinement
    # levels (the plates' right isosceles triangles, each half of its parent)
    # must give the same self term times longest side
    m = meshes["plates"]
    g = m.node_geom()
    shape = g["area"] / g["maxside"] ** 2
    leaf = int(m.leaf_node[0])
    same = [i for i in range(m.M) if not (g["flags"][i] & 130) and abs(shape[i] - shape[leaf]) < 1e-9 * shape[leaf]]
    sizes = {round(float(g["maxside"][i]), 9) for i in same}
    assert len(sizes) >= 3, sizes
    r = np.array([orc.node_self(m, i) * g["maxside"][i] for i in same])
    assert np.abs(r / r[0] - 1).max() < 1e-6


@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_super_apply_of_element_indicator_gives_inverse_column(meshes, name):
    # q = 1 on the leaves of element t: the collect step gives e_t when the
    # element's leaf areas add up to its own, so summing the output over element
    # s gives Sinv[s, t] (the diagonal term supplies s = t)
    m = meshes[name]
    sp = orc.SuperPrecond(m, 128)
    g = m.node_geom()
    for t in range(0, sp.m, max(1, sp.m // 7)):
        lo, hi = sp.starts[t], sp.starts[t + 1]
        la = g["area"][m.leaf_node[lo:hi]].sum()
        assert abs(la / g["area"][sp.nodes[t]] - 1) < 1e-9
        q = np.zeros(m.N)
        q[lo:hi] = 1.0
        out = sp.apply(q)
        agg = np.add.reduceat(out, sp.starts[:-1])
        np.testing.assert_allclose(agg, sp.Sinv[:, t], rtol=1e-9, atol=1e-12 * np.abs(sp.Sinv[:, t]).max())


def test_super_with_leaf_elements_is_the_exact_inverse(geo_dir):
    # when every element is a leaf the two-level preconditioner is the inverse of
    # the whole collocation matrix
    m = fc.Mesh.build(workloads.two_plates(os.path.join(geo_dir, "pc_tiny")), mesh_eps=0.3)
    sp = orc.SuperPrecond(m, 4096)
    assert sp.available and sp.m == m.N, (sp.m, m.N)
    desc = C.byref(m.desc)
    P = np.array([[orc.lib().orc_self(desc, i) if i == j else orc.pot_estimate(m, int(m.leaf_node[i]),
                                                                               int(m.leaf_node[j]))
                   for j in range(m.N)] for i in range(m.N)])
    q = np.random.default_rng(3).standard_normal(m.N)
    np.testing.assert_allclose(P @ sp.apply(q), q, rtol=1e-9, atol=1e-9 * np.abs(q).max())


@pytest.mark.parametrize("name", ["plates", "boxes", "coax2d"])
def test_super_preconditioned_c_equals_jacobi_c(meshes, name):
    m = meshes[name]
    coef, selfc = orc.fill(m)
    cap_j, it_j = orc.solve(m, coef, selfc, tol=1e-9)
    cap_s, it_s = orc.solve_super(m, coef, selfc, 128, tol=1e-9)
    err = np.abs(cap_s - cap_j).max() / np.abs(np.diag(cap_j)).max()
    print(f"{name}: iterations Jacobi {list(it_j)} two-level {list(it_s)}; C diff {err:.2e}")
    assert err < 1e-7
