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
