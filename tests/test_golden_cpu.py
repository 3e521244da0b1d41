"""CPU checks of the host builder and the oracle against the reference's own
published output and against mathematics (no GPU needed)."""
import json
import math
import os

import numpy as np
import pytest

import fastercap_amd as fc
import oracle_lib as orc
from fastercap_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

CUBE = """* 1mX1mX1m unit cube
Q mycube 1.0 1.0 0.0  1.0 0.0 0.0  1.0 0.0 1.0  1.0 1.0 1.0
Q mycube 0.0 1.0 0.0  1.0 1.0 0.0  1.0 1.0 1.0  0.0 1.0 1.0
Q mycube 1.0 0.0 0.0  0.0 0.0 0.0  0.0 0.0 1.0  1.0 0.0 1.0
Q mycube 0.0 0.0 0.0  0.0 1.0 0.0  0.0 1.0 1.0  0.0 0.0 1.0
Q mycube 0.0 0.0 0.0  1.0 0.0 0.0  1.0 1.0 0.0  0.0 1.0 0.0
Q mycube 0.0 0.0 1.0  1.0 0.0 1.0  1.0 1.0 1.0  0.0 1.0 1.0
"""
CUBES = """* Array of two cubes, in air
C cube.txt 1.000000 0.0 0.0 0.0
C cube.txt 1.000000 2.0 0.0 0.0
"""


def two_cubes(d):
    with open(os.path.join(d, "cube.txt"), "w") as f:
        f.write(CUBE)
    p = os.path.join(d, "cubes.txt")
    with open(p, "w") as f:
        f.write(CUBES)
    return p


def test_two_cubes_auto_chain_matches_reference_output(geo_dir):
    """The -a0.01 run documented in the reference help: every iteration's refined
    panel count, link counts, recursion depth, max mesh eps, GMRES iteration count
    and printed capacitance digits."""
    gold = json.load(open(os.path.join(GOLDEN, "two_cubes_doc.json")))
    path = two_cubes(geo_dir)
    m = 1e32
    for it in gold["iterations"]:
        mesh = fc.Mesh.build(path, mesh_eps=m)
        assert mesh.N == it["panels"]
        assert mesh.L == it["links_computed"]
        assert mesh.L + mesh.N == it["links_total"]
        assert mesh.stats.max_level == it["max_level"]
        assert float("%.6g" % mesh.stats.max_mesh_eps) == pytest.approx(it["max_mesh_eps"], rel=1e-9)
        coef, selfc = orc.fill(mesh)
        cap, iters = orc.solve(mesh, coef, selfc, tol=gold["gmres_tol"])
        assert list(iters) == [it["gmres_iters"]] * 2
        # the help prints 6 significant digits
        for a, b in zip(cap.ravel(), np.array(it["cap"]).ravel()):
            assert float("%.6g" % a) == pytest.approx(b, rel=1e-12), (a, b)
        assert mesh.group_names == ["g1_mycube", "g2_mycube"]
        m = mesh.stats.max_mesh_eps / math.sqrt(2.0)


@pytest.mark.parametrize("eps", [1e32, 0.02, 0.005])
def test_links_partition_the_matrix(geo_dir, eps):
    """Every ordered leaf pair i != j is covered by exactly one link, no pair i == j."""
    mesh = fc.Mesh.build(two_cubes(geo_dir), mesh_eps=eps)
    assert mesh.N <= 3000
    assert orc.link_coverage_errors(mesh) == 0
    # links are recorded in symmetric pairs
    a = np.repeat(np.arange(mesh.M), np.diff(mesh.row_ptr))
    fwd = set(zip(a.tolist(), mesh.col.tolist()))
    assert all((b, c) in fwd for (c, b) in fwd)


def test_plates_structure_and_post_order(geo_dir):
    mesh = fc.Mesh.build(workloads.two_plates(geo_dir), mesh_eps=0.02)
    assert orc.link_coverage_errors(mesh) == 0
    # post-order: children precede parents, leaf ranges nest
    for i in range(mesh.M):
        l, r = mesh.left[i], mesh.right[i]
        if l >= 0:
            assert l < i and r < i and mesh.parent[l] == i and mesh.parent[r] == i
            assert mesh.leaf_lo[i] == mesh.leaf_lo[l] and mesh.leaf_hi[l] == mesh.leaf_lo[r]
            assert mesh.leaf_hi[i] == mesh.leaf_hi[r]
    assert np.all(mesh.leaf_lo[mesh.leaf_node] == np.arange(mesh.N))


def test_compressed_matrix_is_symmetric_in_structure(geo_dir):
    mesh = fc.Mesh.build(workloads.two_plates(geo_dir), mesh_eps=0.05)
    coef, selfc = orc.fill(mesh)
    P = orc.compressed_dense(mesh, coef, selfc)
    rng = np.random.default_rng(1)
    for _ in range(3):
        q = rng.standard_normal(mesh.N)
        v = orc.mvp(mesh, coef, selfc, q)
        np.testing.assert_allclose(v, P @ q, rtol=1e-12, atol=1e-12 * np.abs(P @ q).max())
    # vacuum collocation matrix is close to symmetric (same pair, swapped roles)
    assert np.max(np.abs(P - P.T)) / np.max(np.abs(P)) < 0.05
